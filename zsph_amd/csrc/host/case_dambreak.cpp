// Dam-break case generator.  The reference ships no cases (its driver reads
// ../input/input.json and input.vtk, deltasph/main.cu:13, io/read_vkt.cu:5);
// this builds the lattice cases of the configurations directly: fluid at
// (i+1/2)dx, `layers` dummy layers outside the tank on the floor and side
// walls, hydrostatic initial pressure and the matching EOS density.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "zsph.h"
#include "host_internal.hpp"

namespace {

inline int lattice_count(double len, double dx) { return (int)std::floor(len / dx + 0.5); }

struct Emit {
    zsph_host_state *st;
    int n = 0;
    void put(double x, double y, double z, int type, double rho, double p) {
        if (st) {
            if (st->pos) { st->pos[3 * n] = (float)x; st->pos[3 * n + 1] = (float)y; st->pos[3 * n + 2] = (float)z; }
            if (st->vel) { st->vel[3 * n] = st->vel[3 * n + 1] = st->vel[3 * n + 2] = 0.0f; }
            if (st->rhop) { st->rhop[2 * n] = (float)rho; st->rhop[2 * n + 1] = (float)p; }
            if (st->acc_drhodt) for (int c = 0; c < 4; ++c) st->acc_drhodt[4 * n + c] = 0.0f;
            if (st->type) st->type[n] = type;
            if (st->table) st->table[n] = 0;
            if (st->hash) st->hash[n] = 0;
        }
        ++n;
    }
};

}  // namespace

extern "C" int zsph_case_dambreak(const zsph_dambreak *c, zsph_json *j, zsph_host_state *st,
                                  int *n_out) {
    if (!c || !j || !n_out) return ZSPH_E_ARG;
    if ((c->dim != 2 && c->dim != 3) || !(c->dx > 0.0f) || c->layers < 0) return ZSPH_E_ARG;
    const double dx = c->dx;
    const int L = c->layers;
    const int nx = lattice_count(c->lx, dx);
    const int ny = (c->dim == 3) ? lattice_count(c->wy, dx) : 1;
    const int nz = lattice_count(c->lz, dx);
    const int ncx = std::min(nx, lattice_count(c->cw, dx));
    const int ncz = std::min(nz, lattice_count(c->ch, dx));
    const int nzw = (c->wall_height > 0.0f) ? lattice_count(c->wall_height, dx) : nz;
    if (nx < 1 || ny < 1 || nz < 1 || ncx < 1 || ncz < 1) return ZSPH_E_ARG;

    const double rho0 = j->rho0;
    const double g = std::fabs((double)j->g);
    const double Hc = ncz * dx;
    if (!(j->c0 > 0.0f)) j->c0 = (float)(10.0 * std::sqrt(2.0 * g * Hc));
    const double c2 = (double)j->c0 * (double)j->c0;
    const double gamma = (j->gamma > 0.0f) ? j->gamma : 1.0;

    Emit e{st};
    const bool three = (c->dim == 3);
    auto ycoord = [&](int k) { return three ? (k + 0.5) * dx : 0.0; };

    // fluid column, x fastest
    for (int kz = 0; kz < ncz; ++kz)
        for (int ky = 0; ky < ny; ++ky)
            for (int kx = 0; kx < ncx; ++kx) {
                const double z = (kz + 0.5) * dx;
                const double p = rho0 * g * (Hc - z);
                e.put((kx + 0.5) * dx, ycoord(ky), z, ZSPH_FLUID,
                      zsph_host_rho_of_p(p, rho0, c2, gamma), p);
            }
    const int nfluid = e.n;

    const int ylo = three ? -L : 0, yhi = three ? ny + L : 1;
    // floor
    for (int l = 0; l < L; ++l)
        for (int ky = ylo; ky < yhi; ++ky)
            for (int kx = -L; kx < nx + L; ++kx)
                e.put((kx + 0.5) * dx, ycoord(ky), -(l + 0.5) * dx, ZSPH_WALL, rho0, 0.0);
    // x = 0 and x = lx walls
    for (int l = 0; l < L; ++l)
        for (int kz = 0; kz < nzw; ++kz)
            for (int ky = ylo; ky < yhi; ++ky) {
                e.put(-(l + 0.5) * dx, ycoord(ky), (kz + 0.5) * dx, ZSPH_WALL, rho0, 0.0);
                e.put((nx + l + 0.5) * dx, ycoord(ky), (kz + 0.5) * dx, ZSPH_WALL, rho0, 0.0);
            }
    // y = 0 and y = wy walls (3-D only)
    if (three)
        for (int l = 0; l < L; ++l)
            for (int kz = 0; kz < nzw; ++kz)
                for (int kx = 0; kx < nx; ++kx) {
                    e.put((kx + 0.5) * dx, -(l + 0.5) * dx, (kz + 0.5) * dx, ZSPH_WALL, rho0, 0.0);
                    e.put((kx + 0.5) * dx, (ny + l + 0.5) * dx, (kz + 0.5) * dx, ZSPH_WALL, rho0, 0.0);
                }

    *n_out = e.n;
    j->dim = c->dim;
    j->dx = c->dx;
    j->xmin = (float)(-L * dx);
    j->xmax = (float)((nx + L) * dx);
    j->ymin = three ? (float)(-L * dx) : 0.0f;
    j->ymax = three ? (float)((ny + L) * dx) : 0.0f;
    j->zmin = (float)(-L * dx);
    j->zmax = (float)(std::max(nz, nzw) * dx);
    j->ptc_num = e.n;
    j->water_ptc_num = nfluid;
    j->wall_ptc_num = e.n - nfluid;
    j->air_ptc_num = 0;
    j->rigid_ptc_num = 0;
    return ZSPH_OK;
}
