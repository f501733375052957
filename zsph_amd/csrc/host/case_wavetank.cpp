// Wave-tank case with a piston wavemaker and a floating box (the floating-
// body configuration).  One lattice at (i+1/2)dx is split between fluid and
// the box (snapped to lattice faces); the piston is `layers` dummy layers
// behind x = 0 that slide between the side walls, over a floor that extends
// `piston_travel` behind it.
#include <algorithm>
#include <cmath>

#include "zsph.h"
#include "host_internal.hpp"

namespace {

inline int lattice_count(double len, double dx) { return (int)std::floor(len / dx + 0.5); }

struct Emit {
    zsph_host_state *st;
    int n = 0;
    void put(double x, double y, double z, int type, double rho, double p) {
        if (st) {
            st->pos[3 * n] = (float)x; st->pos[3 * n + 1] = (float)y; st->pos[3 * n + 2] = (float)z;
            if (st->vel) st->vel[3 * n] = st->vel[3 * n + 1] = st->vel[3 * n + 2] = 0.0f;
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

extern "C" int zsph_case_wavetank(zsph_wavetank *c, zsph_json *j, zsph_host_state *st, int *n_out) {
    if (!c || !j || !n_out || !(c->dx > 0.0f) || c->layers < 1) return ZSPH_E_ARG;
    if (st && !st->pos) return ZSPH_E_ARG;
    const double dx = c->dx;
    const int L = c->layers;
    const int nx = lattice_count(c->lx, dx), ny = lattice_count(c->wy, dx), nz = lattice_count(c->lz, dx);
    const int nd = std::min(nz, lattice_count(c->depth, dx));
    const int tc = (int)std::ceil(std::max(0.0f, c->piston_travel) / dx);
    if (nx < 1 || ny < 1 || nz < 1 || nd < 1) return ZSPH_E_ARG;
    // box snapped to the nearest lattice faces: cells b0..b1 along each axis,
    // so the particle set is symmetric about the box centre; the snapped
    // extents are written back for the body's mass properties
    int b0[3], b1[3];
    bool has_box = true;
    for (int a = 0; a < 3; ++a) {
        b0[a] = (int)std::floor(c->box_lo[a] / dx + 0.5);
        b1[a] = (int)std::floor(c->box_hi[a] / dx + 0.5) - 1;
        if (b1[a] < b0[a]) has_box = false;
    }
    if (has_box)
        for (int a = 0; a < 3; ++a) {
            c->box_lo[a] = (float)(b0[a] * dx);
            c->box_hi[a] = (float)((b1[a] + 1) * dx);
        }
    auto in_box = [&](int i, int k, int m) {
        return has_box && i >= b0[0] && i <= b1[0] && k >= b0[1] && k <= b1[1] && m >= b0[2] && m <= b1[2];
    };
    const double rho0 = j->rho0, g = std::fabs((double)j->g), H = nd * dx;
    if (!(j->c0 > 0.0f)) j->c0 = (float)(10.0 * std::sqrt(2.0 * g * H));
    const double c2 = (double)j->c0 * j->c0;
    const double gamma = (j->gamma > 0.0f) ? j->gamma : 1.0;
    Emit e{st};
    for (int m = 0; m < nd; ++m)
        for (int k = 0; k < ny; ++k)
            for (int i = 0; i < nx; ++i) {
                if (in_box(i, k, m)) continue;
                const double z = (m + 0.5) * dx, p = rho0 * g * (H - z);
                e.put((i + 0.5) * dx, (k + 0.5) * dx, z, ZSPH_FLUID, zsph_host_rho_of_p(p, rho0, c2, gamma), p);
            }
    const int nfluid = e.n;
    if (has_box)
        for (int m = b0[2]; m <= b1[2]; ++m)
            for (int k = b0[1]; k <= b1[1]; ++k)
                for (int i = b0[0]; i <= b1[0]; ++i) {
                    const int d = std::min({i - b0[0], b1[0] - i, k - b0[1], b1[1] - k, m - b0[2], b1[2] - m});
                    if (d < L) e.put((i + 0.5) * dx, (k + 0.5) * dx, (m + 0.5) * dx, ZSPH_BODY, rho0, 0.0);
                }
    const int nbody = e.n - nfluid;
    // piston: behind x = 0, across the tank width, over the tank height
    for (int l = 0; l < L; ++l)
        for (int m = 0; m < nz; ++m)
            for (int k = 0; k < ny; ++k) e.put(-(l + 0.5) * dx, (k + 0.5) * dx, (m + 0.5) * dx, ZSPH_PISTON, rho0, 0.0);
    // floor (under the piston's travel too), far wall, side walls
    for (int l = 0; l < L; ++l)
        for (int k = -L; k < ny + L; ++k)
            for (int i = -L - tc; i < nx + L; ++i)
                e.put((i + 0.5) * dx, (k + 0.5) * dx, -(l + 0.5) * dx, ZSPH_WALL, rho0, 0.0);
    for (int l = 0; l < L; ++l)
        for (int m = 0; m < nz; ++m)
            for (int k = -L; k < ny + L; ++k)
                e.put((nx + l + 0.5) * dx, (k + 0.5) * dx, (m + 0.5) * dx, ZSPH_WALL, rho0, 0.0);
    for (int l = 0; l < L; ++l)
        for (int m = 0; m < nz; ++m)
            for (int i = -L - tc; i < nx; ++i) {
                e.put((i + 0.5) * dx, -(l + 0.5) * dx, (m + 0.5) * dx, ZSPH_WALL, rho0, 0.0);
                e.put((i + 0.5) * dx, (ny + l + 0.5) * dx, (m + 0.5) * dx, ZSPH_WALL, rho0, 0.0);
            }
    *n_out = e.n;
    j->dim = 3;
    j->dx = c->dx;
    j->xmin = (float)(-(L + tc) * dx);
    j->xmax = (float)((nx + L) * dx);
    j->ymin = (float)(-L * dx);
    j->ymax = (float)((ny + L) * dx);
    j->zmin = (float)(-L * dx);
    j->zmax = (float)(nz * dx);
    j->ptc_num = e.n;
    j->water_ptc_num = nfluid;
    j->rigid_ptc_num = nbody;
    j->wall_ptc_num = e.n - nfluid - nbody;
    j->air_ptc_num = 0;
    return ZSPH_OK;
}
