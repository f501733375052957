// CPU oracle of the delta-SPH step: TEST INFRASTRUCTURE ONLY.
//
// A serial/OpenMP restatement of the reference's equations, templated on
// the arithmetic type (double = the parity oracle, float = the CPU
// baseline).  It is imported only by tests/, the smoke check and the
// bench's cpu_baseline leg; the product library never links it.
//
// Followed reference lines (relative to the reference root):
//   hash            single_gpu_sph_src/lib/hash.cu:3-26
//   sort            single_gpu_sph_src/lib/sort.cu:3-6 (defined stable)
//   cells, reorder  single_gpu_sph_src/lib/sort.cu:8-84
//   boundary        single_gpu_sph_src/equation/delta_ns.cu:5-177,
//                   single_gpu_sph_src/equation/simple_ns.cu:3-97
//   governing       single_gpu_sph_src/equation/delta_ns.cu:179-267,
//                   single_gpu_sph_src/equation/simple_ns.cu:99-180
//   integrator      single_gpu_sph_src/lib/time_int.cu:3-91
//   step order      single_gpu_sph_src/deltasph/main.cu:58-90
// Intended-equation decisions (not the reference's defects): full +-k cell
// rows in every pass; the Antuono psi term uses (G_i + G_j).x_ij; wall
// (rho, p) are final before any pass reads them; the last sorted particle
// belongs to its cell; per-axis grid bounds.
//
// The integer stages (cell hash, neighbour predicate) are evaluated in FP32
// on FP32 positions in every instantiation, so they are bit-identical to a
// correct FP32 implementation.  Physics uses the template type.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "mlp_oracle.hpp"
#include "zsph.h"

namespace orc {

// FP32 neighbour predicate inputs: r^2 = fma(dz,dz,fma(dy,dy,dx*dx)).
inline float r2_f32(float dx, float dy, float dz) { return std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx)); }

// floor((x - o) / size) in IEEE FP32; false outside [0, dim) or NaN.
inline bool cell_f32(float x, float o, float size, int dim, int *c) {
    volatile float num = x - o;  // keep the FP32 rounding of the difference
    const float t = std::floor(num / size);
    if (!(t >= 0.0f && t < (float)dim)) return false;
    *c = (int)t;
    return true;
}

template <class R>
struct Buffer {
    std::vector<R> posr, velp, tposr, tvelp;  // 4 per particle: (x,y,z,rho) (v,p)
    std::vector<int> type, table;
    void resize(size_t n) {
        posr.assign(4 * n, R(0)); velp.assign(4 * n, R(0));
        tposr.assign(4 * n, R(0)); tvelp.assign(4 * n, R(0));
        type.assign(n, 0); table.assign(n, 0);
    }
};

template <class R>
struct Oracle {
    zsph_params p{};
    int N = 0, G = 0, X = 1, Y = 1, Z = 1, k = 1;
    uint32_t flag = 2;  // sort-key bit of non-fluid particles: smallest power of two >= G
    Buffer<R> buf[2];
    std::vector<uint32_t> key;  // unsorted after hash, sorted after sort
    std::vector<int> idx;       // permutation after sort
    std::vector<int> cs;        // G+1
    std::vector<R> acc, gd;     // 4 per particle
    std::vector<R> acc_scale, gd_scale;  // sum |term| per particle: (acc, drho), (G, D)
    std::vector<double> terms;           // 8 per particle: cont, diff, pressure + g (3), viscous (3)
    std::vector<double> feat;            // ZSPH_NF per particle (closure features)
    std::vector<double> scale;           // 2 per particle: closure scale factors (empty: plain)
    bool closure_on = false;
    bool closure_bf16 = true;
    zsph_closure closure{};
    long long killed = 0;
    bool hashed = false, sorted = false;

    explicit Oracle(const zsph_params &prm) : p(prm) {
        N = p.ptc_num;
        X = p.grid_xdim; Y = p.grid_ydim; Z = p.grid_zdim;
        G = X * Y * Z;
        k = p.grid_factor;
        buf[0].resize(N); buf[1].resize(N);
        while (flag < (uint32_t)G) flag <<= 1;
        key.assign(N, 0); idx.assign(N, 0); cs.assign(2 * ((size_t)G + 1), 0);
        acc.assign(4 * (size_t)N, R(0)); gd.assign(4 * (size_t)N, R(0));
        acc_scale.assign(2 * (size_t)N, R(0)); gd_scale.assign(2 * (size_t)N, R(0));
        terms.assign(8 * (size_t)N, 0.0);
        feat.assign((size_t)ZSPH_NF * N, 0.0);
    }

    // ---- constants in the working precision --------------------------------
    R rho0() const { return R(p.rho0); }
    R c2() const { return R(p.cs2); }
    R gam() const { return R(p.gamma); }
    R Bcoef() const { return R(p.rho0) * R(p.cs2) / R(p.gamma); }

    // EOS p(rho): linear c^2 (rho - rho0) for gamma == 1 (time_int.cu:39-40),
    // Tait B ((rho/rho0)^gamma - 1) otherwise.
    R eos_p(R rho) const {
        if (p.gamma == 1.0f) return c2() * (rho - rho0());
        return Bcoef() * (std::pow(rho / rho0(), gam()) - R(1));
    }
    // Inverse EOS (delta_ns.cu:129 for gamma == 1).
    R eos_rho(R pr) const {
        if (p.gamma == 1.0f) return pr / c2() + rho0();
        return rho0() * std::pow(R(1) + pr / Bcoef(), R(1) / gam());
    }
    // Wendland C2: W = aD (1-q/2)^4 (2q+1), F = (1/r) dW/dr = -5 aD (1-q/2)^3 / h^2.
    R kernel_W(R r) const {
        const R q = r / R(p.h);
        const R t = R(1) - q / R(2);
        return R(p.adh) * t * t * t * t * (R(2) * q + R(1));
    }
    R kernel_F(R r) const {
        const R q = r / R(p.h);
        const R t = R(1) - q / R(2);
        return R(-5) * R(p.adh) * t * t * t / (R(p.h) * R(p.h));
    }

    // ---- neighbour walk --------------------------------------------------------
    // Candidates of particle i (sorted order) in buffer b: the (2k+1)^3 cells
    // around its cell, clipped to the grid, ascending within each x-run.
    template <class F>
    void walk(const Buffer<R> &b, int i, F &&f) const {
        const uint32_t h = key[i] & (flag - 1u);
        const int cz = (int)(h / (uint32_t)(X * Y));
        const int cy = (int)((h / (uint32_t)X) % (uint32_t)Y);
        const int cx = (int)(h % (uint32_t)X);
        const int x0 = std::max(cx - k, 0), x1 = std::min(cx + k, X - 1);
        for (int z = std::max(cz - k, 0); z <= std::min(cz + k, Z - 1); ++z)
            for (int y = std::max(cy - k, 0); y <= std::min(cy + k, Y - 1); ++y) {
                const int row = (z * Y + y) * X;
                for (int j = cs[row + x0]; j < cs[row + x1 + 1]; ++j) f(j);                    // fluid
                for (int j = cs[G + 1 + row + x0]; j < cs[G + 1 + row + x1 + 1]; ++j) f(j);  // others
            }
        (void)b;
    }
    bool is_nbr(const Buffer<R> &b, int i, int j) const {
        const float dx = (float)b.posr[4 * i] - (float)b.posr[4 * j];
        const float dy = (float)b.posr[4 * i + 1] - (float)b.posr[4 * j + 1];
        const float dz = (float)b.posr[4 * i + 2] - (float)b.posr[4 * j + 2];
        return r2_f32(dx, dy, dz) <= p.rc2;
    }

    // ---- stages ----------------------------------------------------------------
    // hash.cu:3-26
    void hash(int bi) {
        Buffer<R> &b = buf[bi];
        long long kills = 0;
#pragma omp parallel for schedule(static) reduction(+ : kills)
        for (int i = 0; i < N; ++i) {
            int cx, cy, cz;
            const bool in = cell_f32((float)b.posr[4 * i], p.grid_xmin, p.grid_size, X, &cx) &&
                            cell_f32((float)b.posr[4 * i + 1], p.grid_ymin, p.grid_size, Y, &cy) &&
                            cell_f32((float)b.posr[4 * i + 2], p.grid_zmin, p.grid_size, Z, &cz);
            if (in) {
                key[i] = (uint32_t)(cz * X * Y + cy * X + cx) | (b.type[i] == ZSPH_FLUID ? 0u : flag);
            } else {
                b.type[i] = ZSPH_KILLED;
                b.posr[4 * i] = R(p.grid_xmin); b.posr[4 * i + 1] = R(p.grid_ymin);
                b.posr[4 * i + 2] = R(p.grid_zmin); b.posr[4 * i + 3] = R(p.rho0);
                for (int c = 0; c < 4; ++c) b.velp[4 * i + c] = b.tvelp[4 * i + c] = R(0);
                key[i] = flag;
                ++kills;
            }
            idx[i] = i;
        }
        killed += kills;
        hashed = true;
        sorted = false;
    }
    // sort.cu:3-6, defined as a stable ascending sort of (hash, index).
    void sort_index() {
        std::vector<int> perm(N);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int c) { return key[a] < key[c]; });
        std::vector<uint32_t> k2(N);
        std::vector<int> i2(N);
        for (int t = 0; t < N; ++t) { k2[t] = key[perm[t]]; i2[t] = idx[perm[t]]; }
        key.swap(k2);
        idx.swap(i2);
        hashed = false;
        sorted = true;
    }
    // sort.cu:8-84: cell c holds sorted positions [cs[c], cs[c+1]).
    void sort_data(int from, int to) {
        // fluid starts of cells 0..G, then the starts of the other particles
        for (int e = 0; e <= 2 * G + 1; ++e) {
            const uint32_t t = (e <= G) ? (uint32_t)e : (flag + (uint32_t)(e - G - 1));  // c runs to G; G may equal flag
            cs[e] = (int)(std::lower_bound(key.begin(), key.end(), t) - key.begin());
        }
        const Buffer<R> &a = buf[from];
        Buffer<R> &b = buf[to];
#pragma omp parallel for schedule(static)
        for (int t = 0; t < N; ++t) {
            const int s = idx[t];
            for (int c = 0; c < 4; ++c) {
                b.posr[4 * t + c] = a.posr[4 * s + c];
                b.velp[4 * t + c] = a.velp[4 * s + c];
                b.tposr[4 * t + c] = a.tposr[4 * s + c];
                b.tvelp[4 * t + c] = a.tvelp[4 * s + c];
            }
            b.type[t] = a.type[s];
            b.table[t] = a.table[s] + s - t;
        }
    }

    // Dummy pressure (delta_ns.cu:81-85, :116-131): Shepard sum over fluid
    // neighbours of p_j - rho_j (a_w - g).x_wj, g = (0,0,g).  a_w is the
    // dummy's prescribed acceleration kept in tvelp.xyz (zero in the
    // reference, which has no moving dummies; zero here for fixed walls).
    void wall(int bi) {
        Buffer<R> &b = buf[bi];
        std::vector<R> newr(N), newp(N);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            if (b.type[i] == ZSPH_FLUID) continue;
            R P = 0, S = 0;
            walk(b, i, [&](int j) {
                if (b.type[j] != ZSPH_FLUID || !is_nbr(b, i, j)) return;
                const R dx = b.posr[4 * i] - b.posr[4 * j];
                const R dy = b.posr[4 * i + 1] - b.posr[4 * j + 1];
                const R dz = b.posr[4 * i + 2] - b.posr[4 * j + 2];
                const R W = kernel_W(std::sqrt(dx * dx + dy * dy + dz * dz));
                const R gx = 0, gy = 0, gz = R(p.g);
                const R *aw = &b.tvelp[4 * i];
                const R aw_minus_g_dot_x = (aw[0] - gx) * dx + (aw[1] - gy) * dy + (aw[2] - gz) * dz;
                P += (b.velp[4 * j + 3] - b.posr[4 * j + 3] * aw_minus_g_dot_x) * W;
                S += W;
            });
            R pw = (std::fabs(S) > R(1e-8)) ? P / S : R(0);
            if (pw < R(0)) pw = R(0);
            newp[i] = pw;
            newr[i] = eos_rho(pw);
        }
        for (int i = 0; i < N; ++i)
            if (b.type[i] != ZSPH_FLUID) { b.posr[4 * i + 3] = newr[i]; b.velp[4 * i + 3] = newp[i]; }
    }

    // Solve L x = g by Gaussian elimination with partial pivoting; returns det.
    static R solve3(R A[3][3], R g[3], R out[3]) {
        R M[3][4];
        for (int r = 0; r < 3; ++r) { for (int c = 0; c < 3; ++c) M[r][c] = A[r][c]; M[r][3] = g[r]; }
        R det = 1;
        for (int c = 0; c < 3; ++c) {
            int piv = c;
            for (int r = c + 1; r < 3; ++r) if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
            if (piv != c) { for (int t = 0; t < 4; ++t) std::swap(M[c][t], M[piv][t]); det = -det; }
            det *= M[c][c];
            if (M[c][c] == R(0)) return R(0);
            for (int r = c + 1; r < 3; ++r) {
                const R f = M[r][c] / M[c][c];
                for (int t = c; t < 4; ++t) M[r][t] -= f * M[c][t];
            }
        }
        for (int r = 2; r >= 0; --r) {
            R s = M[r][3];
            for (int t = r + 1; t < 3; ++t) s -= M[r][t] * out[t];
            out[r] = s / M[r][r];
        }
        return det;
    }

    // delta_ns.cu:5-177 (after the wall pass): L_i, D_i = div v, G_i = L_i^-1 grad rho.
    // MC form (simple_ns.cu:3-97): D only.
    void ldg(int bi) {
        const Buffer<R> &b = buf[bi];
        const bool delta = p.diffusion == ZSPH_DIFF_ANTUONO;
        const R m = R(p.m);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            for (int c = 0; c < 4; ++c) gd[4 * i + c] = 0;
            gd_scale[2 * i] = gd_scale[2 * i + 1] = 0;
            if (b.type[i] != ZSPH_FLUID) continue;
            R L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            R g[3] = {0, 0, 0};
            R D = 0, gs = 0, ds = 0, ls = 0;
            const R rhoi = b.posr[4 * i + 3];
            walk(b, i, [&](int j) {
                if (j == i || !is_nbr(b, i, j)) return;
                const R x[3] = {b.posr[4 * i] - b.posr[4 * j], b.posr[4 * i + 1] - b.posr[4 * j + 1],
                                b.posr[4 * i + 2] - b.posr[4 * j + 2]};
                const R r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                const R F = kernel_F(r);
                const R rhoj = b.posr[4 * j + 3];
                const R Vj = m / rhoj;
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) L[a][c] -= F * Vj * x[a] * x[c];
                for (int a = 0; a < 3; ++a) g[a] += F * (rhoj - rhoi) * Vj * x[a];
                gs += std::fabs(F * (rhoj - rhoi) * Vj) * r;
                ls += std::fabs(F * Vj) * r * r;
                if (b.type[j] == ZSPH_FLUID) {
                    R dvdx = 0;
                    for (int a = 0; a < 3; ++a) dvdx += (b.velp[4 * i + a] - b.velp[4 * j + a]) * x[a];
                    D -= F * dvdx * Vj;
                    ds += std::fabs(F * dvdx * Vj);
                }
            });
            R G[3] = {0, 0, 0};
            // Error scale of G = L^-1 g: to first order dG = L^-1 (dg - dL G), with
            // |dg| <~ u gs and |dL| <~ u ls (u: unit round-off of the summed terms),
            // so |dG| <~ u |L^-1| (gs + ls |G|) in the max norm; with L singular G = g.
            R gscale = gs;
            if (delta) {
                if (p.dim == 2) L[1][1] += R(1);
                R Lc[3][3];
                std::memcpy(Lc, L, sizeof(L));
                R sol[3] = {0, 0, 0};
                const R det = solve3(Lc, g, sol);
                if (std::fabs(det) > R(1.8e-8)) {
                    G[0] = sol[0]; G[1] = sol[1]; G[2] = sol[2];
                    R inv_norm = 0;  // max row sum of |L^-1|, from L^-1 e_c column by column
                    R col[3][3];
                    for (int c = 0; c < 3; ++c) {
                        R e[3] = {R(c == 0), R(c == 1), R(c == 2)};
                        std::memcpy(Lc, L, sizeof(L));
                        solve3(Lc, e, col[c]);
                    }
                    for (int r = 0; r < 3; ++r)
                        inv_norm = std::max(inv_norm, std::fabs(col[0][r]) + std::fabs(col[1][r]) + std::fabs(col[2][r]));
                    const R gmax = std::max(std::fabs(G[0]), std::max(std::fabs(G[1]), std::fabs(G[2])));
                    gscale = inv_norm * (gs + ls * gmax);
                } else {
                    G[0] = g[0]; G[1] = g[1]; G[2] = g[2];
                }
            }
            gd[4 * i] = G[0]; gd[4 * i + 1] = G[1]; gd[4 * i + 2] = G[2]; gd[4 * i + 3] = D;
            gd_scale[2 * i] = delta ? gscale : R(0);
            gd_scale[2 * i + 1] = ds;
        }
    }

    // Closure features of every fluid particle of buffer bi (after ldg): the
    // same definitions as the device's gradient pass, in the working precision.
    void features(int bi) {
        const Buffer<R> &b = buf[bi];
        const R m = R(p.m), h = R(p.h), rho0 = R(p.rho0), ic0 = R(1) / std::sqrt(R(p.cs2));
        const int dim = p.dim;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            double *f = &feat[(size_t)ZSPH_NF * i];
            for (int k = 0; k < ZSPH_NF; ++k) f[k] = 0.0;
            if (b.type[i] != ZSPH_FLUID) continue;
            R L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, K[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            R ws = 0, wv2 = 0, wdr = 0;
            int nnb = 0, ndum = 0;
            const R rhoi = b.posr[4 * i + 3];
            walk(b, i, [&](int j) {
                if (j == i || !is_nbr(b, i, j)) return;
                const R x[3] = {b.posr[4 * i] - b.posr[4 * j], b.posr[4 * i + 1] - b.posr[4 * j + 1],
                                b.posr[4 * i + 2] - b.posr[4 * j + 2]};
                const R r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                const R rhoj = b.posr[4 * j + 3];
                const R FV = kernel_F(r) * m / rhoj, WV = kernel_W(r) * m / rhoj;
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) L[a][c] -= FV * x[a] * x[c];
                ws += WV;
                wdr += WV * (rhoj - rhoi);
                ++nnb;
                if (b.type[j] == ZSPH_FLUID) {
                    R u[3], u2 = 0;
                    for (int a = 0; a < 3; ++a) { u[a] = b.velp[4 * j + a] - b.velp[4 * i + a]; u2 += u[a] * u[a]; }
                    for (int a = 0; a < 3; ++a)
                        for (int c = 0; c < 3; ++c) K[a][c] += FV * u[a] * x[c];
                    wv2 += WV * u2;
                } else {
                    ++ndum;
                }
            });
            const R trL = L[0][0] + L[1][1] + L[2][2];
            if (dim == 2) L[1][1] += R(1);
            const R det = L[0][0] * (L[1][1] * L[2][2] - L[1][2] * L[2][1]) -
                          L[0][1] * (L[1][0] * L[2][2] - L[1][2] * L[2][0]) +
                          L[0][2] * (L[1][0] * L[2][1] - L[1][1] * L[2][0]);
            const R *G = &gd[4 * i];
            const R D = gd[4 * i + 3];
            const R hc = h * ic0, hr = h / rho0;
            const R w[3] = {K[2][1] - K[1][2], K[0][2] - K[2][0], K[1][0] - K[0][1]};
            const R third = (K[0][0] + K[1][1] + K[2][2]) / R(dim);
            const R S00 = K[0][0] - third, S11 = dim == 3 ? K[1][1] - third : R(0), S22 = K[2][2] - third;
            const R S01 = R(0.5) * (K[0][1] + K[1][0]), S02 = R(0.5) * (K[0][2] + K[2][0]),
                    S12 = R(0.5) * (K[1][2] + K[2][1]);
            const R *v = &b.velp[4 * i];
            const R v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            f[0] = double(R(1) - (ws + R(p.adh) * m / rhoi));
            f[1] = double(trL / R(dim));
            f[2] = double(det);
            f[3] = double(std::sqrt(L[0][1] * L[0][1] + L[0][2] * L[0][2] + L[1][2] * L[1][2]));
            for (int a = 0; a < 3; ++a) f[4 + a] = double(G[a] * hr);
            f[7] = double(std::sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]) * hr);
            f[8] = double(D * hc);
            for (int a = 0; a < 3; ++a) f[9 + a] = double(w[a] * hc);
            f[12] = double(std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) * hc);
            f[13] = double(std::sqrt(S00 * S00 + S11 * S11 + S22 * S22 + R(2) * (S01 * S01 + S02 * S02 + S12 * S12)) * hc);
            for (int a = 0; a < 3; ++a) f[14 + a] = double(v[a] * ic0);
            f[17] = double(std::sqrt(v2) * ic0);
            f[18] = double((rhoi - rho0) / rho0);
            f[19] = double(v[3] / (rho0 * R(p.cs2)));
            f[20] = nnb;
            f[21] = ws > 0 ? double(std::sqrt(wv2 / ws) * ic0) : 0.0;
            f[22] = ws > 0 ? double(wdr / (ws * rho0)) : 0.0;
            f[23] = nnb ? double(ndum) / nnb : 0.0;
        }
    }

    // Closure scale factors of buffer bi's fluid particles (features, then the network).
    void closure_scales(int bi) {
        features(bi);
        const MlpOracle net(closure, closure_bf16);
        scale.assign(2 * (size_t)N, 1.0);
        const Buffer<R> &b = buf[bi];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            if (b.type[i] != ZSPH_FLUID) continue;
            float x[ZSPH_NF];
            for (int k = 0; k < ZSPH_NF; ++k) x[k] = (float)feat[(size_t)ZSPH_NF * i + k];
            net.infer(x, &scale[2 * (size_t)i]);
        }
    }

    // delta_ns.cu:179-267 / simple_ns.cu:99-180.  The continuity, density-
    // diffusion, pressure (+ gravity) and viscous parts are kept apart in
    // `terms`; with the closure on, the diffusion and viscous parts are scaled
    // by the network's factors.
    void governing(int bi) {
        if (closure_on) closure_scales(bi);
        else scale.clear();
        const Buffer<R> &b = buf[bi];
        const bool delta = p.diffusion == ZSPH_DIFF_ANTUONO;
        const R m = R(p.m), eta2 = R(p.eta2), dhc = R(p.delta_h_cs), hcrm = R(p.h_cs_rho_m), al = R(p.alpha);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            const int ti = b.type[i];
            for (int c = 0; c < 4; ++c) acc[4 * i + c] = 0;
            for (int c = 0; c < 8; ++c) terms[8 * (size_t)i + c] = 0.0;
            acc_scale[2 * i] = acc_scale[2 * i + 1] = 0;
            if (ti != ZSPH_FLUID && ti != ZSPH_BODY && ti != ZSPH_PISTON) continue;
            const bool fi = ti == ZSPH_FLUID;
            const R rhoi = b.posr[4 * i + 3], pi = b.velp[4 * i + 3];
            R ap[3] = {0, 0, 0}, av[3] = {0, 0, 0}, cont = 0, diff = 0, as = 0, drs = 0;
            walk(b, i, [&](int j) {
                if (j == i) return;
                const bool fj = b.type[j] == ZSPH_FLUID;
                if (!fi && !fj) return;
                if (!is_nbr(b, i, j)) return;
                const R x[3] = {b.posr[4 * i] - b.posr[4 * j], b.posr[4 * i + 1] - b.posr[4 * j + 1],
                                b.posr[4 * i + 2] - b.posr[4 * j + 2]};
                const R r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
                const R r = std::sqrt(r2);
                const R F = kernel_F(r);
                const R rhoj = b.posr[4 * j + 3], pj = b.velp[4 * j + 3];
                R dvdx = 0;
                for (int c = 0; c < 3; ++c) dvdx += (b.velp[4 * i + c] - b.velp[4 * j + c]) * x[c];
                const R cterm = rhoi * dvdx * F * m / rhoj;
                cont += cterm;
                drs += std::fabs(cterm);
                const R pt = -(pi + pj) * m * F / (rhoi * rhoj);
                R pv = 0;
                R ptabs = std::fabs(pt);
                if (fi && fj) {
                    if (delta) {
                        R gx = 0;
                        for (int c = 0; c < 3; ++c) gx += (gd[4 * i + c] + gd[4 * j + c]) * x[c];
                        const R d1 = dhc * F * R(2) * (rhoi - rhoj) * r2 / (r2 + eta2) * m / rhoj;
                        const R d2 = -dhc * F * gx * m / rhoj;
                        diff += d1 + d2;
                        drs += std::fabs(d1) + std::fabs(d2);
                    }
                    pv = hcrm * F * (gd[4 * i + 3] + gd[4 * j + 3] + al * dvdx / (r2 + eta2)) / (rhoi * rhoj);
                    ptabs += std::fabs(pv);
                }
                if (!delta) {
                    const R d = (rhoi - rhoj) * r2 * F * dhc / (r2 + eta2) * m / rhoj;
                    diff += d;
                    drs += std::fabs(d);
                }
                for (int c = 0; c < 3; ++c) { ap[c] += pt * x[c]; av[c] += pv * x[c]; }
                as += ptabs * r;
            });
            if (fi) ap[2] += R(p.g);
            const R sd = scale.empty() ? R(1) : R(scale[2 * (size_t)i]);
            const R sv = scale.empty() ? R(1) : R(scale[2 * (size_t)i + 1]);
            for (int c = 0; c < 3; ++c) acc[4 * i + c] = ap[c] + sv * av[c];
            acc[4 * i + 3] = cont + sd * diff;
            double *t = &terms[8 * (size_t)i];
            t[0] = double(cont); t[1] = double(diff);
            for (int c = 0; c < 3; ++c) { t[2 + c] = double(ap[c]); t[5 + c] = double(av[c]); }
            acc_scale[2 * i] = as + (fi ? std::fabs(R(p.g)) : R(0));
            acc_scale[2 * i + 1] = drs;
        }
    }

    // ---- moving dummies: floating body and piston ---------------------------------
    // Rigid equations of Test_6DOF/equation/acc_alpha.cpp:3-35 and
    // time_int.cpp:3-22 in FP64, integrated with the fluid's midpoint
    // scheme; restated here independently of the product (body-frame
    // Gaussian elimination instead of an explicit inverse).
    zsph_rigid bs{}, bs0{};  // body state now / at the start of the step
    double t = 0.0, t0 = 0.0;
    bool has_body = false, has_piston = false;
    double p_amp = 0.0, p_w = 0.0, p_phase = 0.0;
    int p_axis = 0;

    static void rotm(const double q[4], double M[3][3]) {
        const double a = q[0], b = q[1], c = q[2], d = q[3];
        M[0][0] = a * a + b * b - c * c - d * d; M[0][1] = 2 * (b * c - a * d); M[0][2] = 2 * (b * d + a * c);
        M[1][0] = 2 * (b * c + a * d); M[1][1] = a * a - b * b + c * c - d * d; M[1][2] = 2 * (c * d - a * b);
        M[2][0] = 2 * (b * d - a * c); M[2][1] = 2 * (c * d + a * b); M[2][2] = a * a - b * b - c * c + d * d;
    }
    static void mv(const double M[3][3], const double v[3], double o[3], bool transpose = false) {
        for (int r = 0; r < 3; ++r) {
            o[r] = 0.0;
            for (int c = 0; c < 3; ++c) o[r] += (transpose ? M[c][r] : M[r][c]) * v[c];
        }
    }
    static void crs(const double a[3], const double b[3], double o[3]) {
        const double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
        o[0] = x; o[1] = y; o[2] = z;
    }
    // Solve A x = y (3x3, partial pivoting).
    static void gauss3(const double Ain[9], const double y[3], double x[3]) {
        double A[3][4];
        for (int r = 0; r < 3; ++r) { for (int c = 0; c < 3; ++c) A[r][c] = Ain[3 * r + c]; A[r][3] = y[r]; }
        for (int k = 0; k < 3; ++k) {
            int piv = k;
            for (int r = k + 1; r < 3; ++r) if (std::fabs(A[r][k]) > std::fabs(A[piv][k])) piv = r;
            for (int c = 0; c < 4; ++c) std::swap(A[k][c], A[piv][c]);
            for (int r = k + 1; r < 3; ++r) {
                const double f = A[r][k] / A[k][k];
                for (int c = k; c < 4; ++c) A[r][c] -= f * A[k][c];
            }
        }
        for (int k = 2; k >= 0; --k) {
            double v = A[k][3];
            for (int c = k + 1; c < 3; ++c) v -= A[k][c] * x[c];
            x[k] = v / A[k][k];
        }
    }
    // a = (F + M g) / M;  alpha: I (R^T alpha) = R^T (T - w x (R I R^T w))   (acc_alpha.cpp:5-31)
    static void body_accelerations(zsph_rigid &s, double g) {
        for (int k = 0; k < 3; ++k) s.acc[k] = s.F[k] / s.mass;
        s.acc[2] += g;
        double Rm[3][3], wb[3], Iwb[3], Lw[3], wxL[3], rhs[3], tb[3], ab[3];
        rotm(s.q, Rm);
        mv(Rm, s.omega, wb, true);  // w in the body frame
        for (int r = 0; r < 3; ++r) Iwb[r] = s.I[3 * r] * wb[0] + s.I[3 * r + 1] * wb[1] + s.I[3 * r + 2] * wb[2];
        mv(Rm, Iwb, Lw);            // angular momentum, world frame
        crs(s.omega, Lw, wxL);
        for (int k = 0; k < 3; ++k) rhs[k] = s.T[k] - wxL[k];
        mv(Rm, rhs, tb, true);
        gauss3(s.I, tb, ab);
        mv(Rm, ab, s.alpha);
    }
    // (c, v, w, q) = s0 + dt * rates(s): c' = v, v' = a, w' = alpha, q' = 1/2 (0,w) q.
    static void body_advance(const zsph_rigid &s0, const zsph_rigid &r, double dt, zsph_rigid &out) {
        const double w[3] = {r.omega[0], r.omega[1], r.omega[2]};
        const double *q = r.q;
        const double dq[4] = {0.5 * (-w[0] * q[1] - w[1] * q[2] - w[2] * q[3]),
                              0.5 * (w[0] * q[0] + w[1] * q[3] - w[2] * q[2]),
                              0.5 * (w[1] * q[0] + w[2] * q[1] - w[0] * q[3]),
                              0.5 * (w[2] * q[0] + w[0] * q[2] - w[1] * q[1])};
        zsph_rigid o = r;
        for (int k = 0; k < 3; ++k) {
            o.c[k] = s0.c[k] + r.vel[k] * dt;
            o.vel[k] = s0.vel[k] + r.acc[k] * dt;
            o.omega[k] = s0.omega[k] + r.alpha[k] * dt;
        }
        double nq = 0.0;
        for (int k = 0; k < 4; ++k) { o.q[k] = s0.q[k] + dq[k] * dt; nq += o.q[k] * o.q[k]; }
        nq = std::sqrt(nq);
        for (int k = 0; k < 4; ++k) o.q[k] /= nq;
        out = o;
    }
    // Fluid force and torque on the body particles of buffer bi (last governing).
    void body_force(int bi, double F[3], double T[3]) const {
        const Buffer<R> &b = buf[bi];
        for (int k = 0; k < 3; ++k) F[k] = T[k] = 0.0;
        for (int i = 0; i < N; ++i) {
            if (b.type[i] != ZSPH_BODY) continue;
            double f[3], d[3], tq[3];
            for (int k = 0; k < 3; ++k) {
                f[k] = (double)p.m * (double)acc[4 * i + k];
                d[k] = (double)b.posr[4 * i + k] - bs.c[k];
            }
            crs(d, f, tq);
            for (int k = 0; k < 3; ++k) { F[k] += f[k]; T[k] += tq[k]; }
        }
    }
    void piston_motion(double tt, double &d, double &dv, double &da) const {
        const double ph = p_w * tt + p_phase;
        d = p_amp * (std::sin(ph) - std::sin(p_phase));
        dv = p_amp * p_w * std::cos(ph);
        da = -p_amp * p_w * p_w * std::sin(ph);
    }
    // Body and piston dummies of buffer bi at the current time.
    void kinematics(int bi) {
        Buffer<R> &b = buf[bi];
        double Rm[3][3];
        rotm(bs.q, Rm);
        double d = 0, dv = 0, da = 0;
        if (has_piston) piston_motion(t, d, dv, da);
        for (int i = 0; i < N; ++i) {
            double x[3], v[3], a[3];
            if (b.type[i] == ZSPH_BODY && has_body) {
                const double r0[3] = {(double)b.tposr[4 * i], (double)b.tposr[4 * i + 1], (double)b.tposr[4 * i + 2]};
                double r[3], wr[3], wwr[3], ar[3];
                mv(Rm, r0, r);
                crs(bs.omega, r, wr);
                crs(bs.omega, wr, wwr);
                crs(bs.alpha, r, ar);
                for (int k = 0; k < 3; ++k) {
                    x[k] = bs.c[k] + r[k];
                    v[k] = bs.vel[k] + wr[k];
                    a[k] = bs.acc[k] + ar[k] + wwr[k];
                }
            } else if (b.type[i] == ZSPH_PISTON && has_piston) {
                for (int k = 0; k < 3; ++k) { x[k] = (double)b.tposr[4 * i + k]; v[k] = a[k] = 0.0; }
                x[p_axis] += d; v[p_axis] = dv; a[p_axis] = da;
            } else {
                continue;
            }
            for (int k = 0; k < 3; ++k) { b.posr[4 * i + k] = R(x[k]); b.velp[4 * i + k] = R(v[k]); b.tvelp[4 * i + k] = R(a[k]); }
        }
    }
    // Body force, accelerations, midpoint update and time of one half step.
    void body_half(int bi, bool predictor) {
        if (!has_body && !has_piston) return;
        if (predictor) t0 = t;
        const double dt = predictor ? (double)p.half_dt : (double)p.half_dt * 2.0;
        t = t0 + dt;
        if (has_body) {
            body_force(bi, bs.F, bs.T);
            body_accelerations(bs, (double)p.g);
            if (predictor) bs0 = bs;
            body_advance(bs0, bs, dt, bs);
        }
    }
    // Body-frame offsets r0 = R^T (x - c) (stored in FP32, as the product
    // stores them), v = V + w x d, a_w = A + alpha x d + w x (w x d) on buffer 0.
    int set_body(const zsph_rigid &r) {
        bs = r;
        double qn = 0.0;
        for (int k = 0; k < 4; ++k) qn += r.q[k] * r.q[k];
        qn = std::sqrt(qn);
        for (int k = 0; k < 4; ++k) bs.q[k] = r.q[k] / qn;
        double Rm[3][3];
        rotm(bs.q, Rm);
        Buffer<R> &b = buf[0];
        int n = 0;
        for (int i = 0; i < N; ++i) {
            if (b.type[i] != ZSPH_BODY) continue;
            ++n;
            double d[3], r0[3], wd[3], wwd[3], ad[3];
            for (int k = 0; k < 3; ++k) d[k] = (double)b.posr[4 * i + k] - bs.c[k];
            mv(Rm, d, r0, true);
            crs(bs.omega, d, wd);
            crs(bs.omega, wd, wwd);
            crs(bs.alpha, d, ad);
            for (int k = 0; k < 3; ++k) {
                b.tposr[4 * i + k] = R((float)r0[k]);
                b.velp[4 * i + k] = R(bs.vel[k] + wd[k]);
                b.tvelp[4 * i + k] = R(bs.acc[k] + ad[k] + wwd[k]);
            }
        }
        if (!n) return ZSPH_E_ARG;
        bs0 = bs;
        has_body = true;
        return ZSPH_OK;
    }
    int set_piston(double amp, double freq, double phase, int axis) {
        p_amp = amp; p_w = 2.0 * M_PI * freq; p_phase = phase; p_axis = axis;
        double d, dv, da;
        piston_motion(t, d, dv, da);
        Buffer<R> &b = buf[0];
        int n = 0;
        for (int i = 0; i < N; ++i) {
            if (b.type[i] != ZSPH_PISTON) continue;
            ++n;
            for (int k = 0; k < 3; ++k) {
                b.tposr[4 * i + k] = b.posr[4 * i + k];
                b.velp[4 * i + k] = b.tvelp[4 * i + k] = R(0);
            }
            b.tposr[4 * i + axis] = R((float)((double)(float)b.posr[4 * i + axis] - d));
            b.velp[4 * i + axis] = R(dv);
            b.tvelp[4 * i + axis] = R(da);
        }
        if (!n) return ZSPH_E_ARG;
        has_piston = true;
        return ZSPH_OK;
    }

    // time_int.cu:3-48
    void prediction(int bi) {
        body_half(bi, true);
        Buffer<R> &b = buf[bi];
        const R hdt = R(p.half_dt);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            if (b.type[i] != ZSPH_FLUID) continue;
            for (int c = 0; c < 4; ++c) { b.tposr[4 * i + c] = b.posr[4 * i + c]; b.tvelp[4 * i + c] = b.velp[4 * i + c]; }
            for (int c = 0; c < 3; ++c) {
                b.posr[4 * i + c] = b.tposr[4 * i + c] + b.tvelp[4 * i + c] * hdt;
                b.velp[4 * i + c] = b.tvelp[4 * i + c] + acc[4 * i + c] * hdt;
            }
            R rho = b.tposr[4 * i + 3] + acc[4 * i + 3] * hdt;
            if (rho < rho0()) rho = rho0();
            b.posr[4 * i + 3] = rho;
            b.velp[4 * i + 3] = eos_p(rho);
        }
        if (has_body || has_piston) kinematics(bi);
    }
    // time_int.cu:50-91 (dt = 2 half_dt)
    void correction(int bi) {
        body_half(bi, false);
        Buffer<R> &b = buf[bi];
        const R dt = R(p.half_dt) * R(2);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            if (b.type[i] != ZSPH_FLUID) continue;
            R v_star[3] = {b.velp[4 * i], b.velp[4 * i + 1], b.velp[4 * i + 2]};
            for (int c = 0; c < 3; ++c) {
                b.posr[4 * i + c] = b.tposr[4 * i + c] + v_star[c] * dt;
                b.velp[4 * i + c] = b.tvelp[4 * i + c] + acc[4 * i + c] * dt;
            }
            R rho = b.tposr[4 * i + 3] + acc[4 * i + 3] * dt;
            if (rho < rho0()) rho = rho0();
            b.posr[4 * i + 3] = rho;
            b.velp[4 * i + 3] = eos_p(rho);
        }
        if (has_body || has_piston) kinematics(bi);
    }

    void boundary(int bi) { wall(bi); ldg(bi); }

    // main.cu:58-90: half-step A old->new + prediction, B new->old + correction.
    void step(int nsteps) {
        for (int s = 0; s < nsteps; ++s) {
            hash(0); sort_index(); sort_data(0, 1); boundary(1); governing(1); prediction(1);
            hash(1); sort_index(); sort_data(1, 0); boundary(0); governing(0); correction(0);
        }
    }

    // Run diagnostics (count, kinetic energy, momentum, front, top) of buffer bi.
    void diagnostics(int bi, float front_zmax, zsph_diag *d) const {
        const Buffer<R> &b = buf[bi];
        std::memset(d, 0, sizeof(*d));
        d->front_x = d->zmax_fluid = -INFINITY;
        const double m = p.m;
        for (int i = 0; i < N; ++i) {
            if (b.type[i] != ZSPH_FLUID) continue;
            ++d->n_fluid;
            const double u[3] = {(double)b.velp[4 * i], (double)b.velp[4 * i + 1], (double)b.velp[4 * i + 2]};
            d->kinetic += 0.5 * m * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            for (int c = 0; c < 3; ++c) d->momentum[c] += m * u[c];
            const double x = (double)b.posr[4 * i], z = (double)b.posr[4 * i + 2];
            if ((float)z < front_zmax) d->front_x = std::max(d->front_x, x);
            d->zmax_fluid = std::max(d->zmax_fluid, z);
        }
        d->mass_fluid = d->n_fluid * m;
        d->n_killed = (int)killed;
    }
    // Shepard fluid pressure at points (brute force over all particles).
    void probes(int bi, const float *pts, int np, double *out) const {
        const Buffer<R> &b = buf[bi];
        for (int k = 0; k < np; ++k) {
            R pw = 0, ws = 0;
            for (int j = 0; j < N; ++j) {
                if (b.type[j] != ZSPH_FLUID) continue;
                const float dx = pts[3 * k] - (float)b.posr[4 * j];
                const float dy = pts[3 * k + 1] - (float)b.posr[4 * j + 1];
                const float dz = pts[3 * k + 2] - (float)b.posr[4 * j + 2];
                if (!(r2_f32(dx, dy, dz) <= p.rc2)) continue;
                const R ex = R(pts[3 * k]) - b.posr[4 * j], ey = R(pts[3 * k + 1]) - b.posr[4 * j + 1],
                        ez = R(pts[3 * k + 2]) - b.posr[4 * j + 2];
                const R W = kernel_W(std::sqrt(ex * ex + ey * ey + ez * ez));
                pw += b.velp[4 * j + 3] * W;
                ws += W;
            }
            out[k] = (ws > R(1e-8)) ? (double)(pw / ws) : 0.0;
        }
    }

    static uint64_t mix64(uint64_t z) {
        z += 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // Neighbour count and digest sum mix64(id_j) over the neighbours' original
    // ids id_j = table[j] + j (independent of the particle order): by the cell
    // walk, or by brute force over all j (O(N^2), small cases only).
    void nbr_stats(int bi, bool brute, int *count, uint64_t *digest) const {
        const Buffer<R> &b = buf[bi];
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            int c = 0;
            uint64_t h = 0;
            auto visit = [&](int j) {
                if (j == i || !is_nbr(b, i, j)) return;
                ++c;
                h += mix64((uint64_t)(b.table[j] + j));
            };
            if (brute)
                for (int j = 0; j < N; ++j) visit(j);
            else
                walk(b, i, visit);
            count[i] = c;
            digest[i] = h;
        }
    }
};

}  // namespace orc
