// Neighbour-loop physics for gfx950: dummy-wall extrapolation, renormalised
// density gradient + velocity divergence, continuity + momentum, and the
// midpoint predictor-corrector.
//   wall + ldg  <- single_gpu_sph_src/equation/delta_ns.cu:5-177 (cuda_boundary_ns),
//                  equation/simple_ns.cu:3-97 (MC variant)
//   governing   <- equation/delta_ns.cu:179-267, equation/simple_ns.cu:99-180
//   predict /
//   correct     <- lib/time_int.cu:3-48, :50-91
// Particles are in cell order with every fluid particle first: fluid
// [0, Nf), dummies [Nf, N), Nf = cs[G].  Each kernel launches over the
// segment it updates, and a neighbour's kind follows from its index, so no
// type array is read inside a neighbour loop.
// The reference's boundary kernel writes wall (rho, p) while other threads
// of the same launch read them; here the wall pass is its own launch, and the
// gradient pass runs only after every wall value is final.
#include "common.hpp"
#include "launch.hpp"

namespace zs {

static inline int nblocks(int n) { return (n + ZS_BLOCK - 1) / ZS_BLOCK; }

__device__ __forceinline__ float *wptr(float4 *a, int i) { return reinterpret_cast<float *>(a + i) + 3; }

// Cell of a sort key, relative to the particle's scene (cell_base = 0 for one scene).
__device__ __forceinline__ uint32_t cell_of(const Prm &q, uint32_t key) {
    return (key & (q.flag - 1u)) - (uint32_t)q.cell_base;
}

// Dummy particles: Shepard extrapolation of fluid pressure with the
// hydrostatic and wall-acceleration correction,
//   p_w = max(sum (p_j - rho_j (a_w - g) . x_wj) W / sum W, 0),
// rho_w from the inverse EOS (delta_ns.cu:81-85, :116-131; a_w = 0 there).
// a_w is the dummy's prescribed acceleration, kept in tvelp.xyz (zero for
// fixed walls).
template <int WALK, bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_wall(Prm q0, const Prm *__restrict__ sp, Buf b,
        const uint32_t *__restrict__ skey,
                                                   const int *__restrict__ cs) {
    __shared__ int qbuf[WalkLds<WALK>::ints];
    const int nf = cs[q0.G];
    const int i = nf + xcd_block(gridDim.x) * ZS_BLOCK + threadIdx.x;
    if (i >= q0.N) return;
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    const float4 xi = b.posr[i];
    const float4 aw = b.tvelp[i];
    float P = 0.0f, S = 0.0f;
    for_each_neighbor<WALK, true, false>(q, i, xi, cell_of(q, skey[i]), cs, b.posr, qbuf,
                                   [&](int j, float dx, float dy, float dz, float r2, float4 xj) {
                                       const float qq = sqrtf(r2) * q.inv_h;
                                       const float t = 1.0f - 0.5f * qq;
                                       const float W = q.adh * t * t * t * t * (2.0f * qq + 1.0f);
                                       const float ax = aw.x * dx + aw.y * dy + (aw.z - q.g) * dz;
                                       P += (b.velp[j].w - xj.w * ax) * W;
                                       S += W;
                                   });
    float p = (fabsf(S) > 1e-8f) ? P / S : 0.0f;
    p = fmaxf(p, 0.0f);
    *wptr(b.posr, i) = eos_rho(q, p);
    *wptr(b.velp, i) = p;
}

// Fluid particles: L_i = -sum F V_j x_ij (x) x_ij, D_i = -sum_fluid F V_j v_ij.x_ij,
// G_i = L_i^-1 sum F (rho_j - rho_i) V_j x_ij (delta_ns.cu:75-79, :86-90,
// :99-114, :133-174).  L^-1 is applied once to the accumulated sum (it does
// not depend on j).  MC form: D only.  Output gd = (G, D); dummies get 0.
template <int WALK, bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_ldg(Prm q0, const Prm *__restrict__ sp, Buf b,
        const uint32_t *__restrict__ skey,
                                                  const int *__restrict__ cs, float4 *__restrict__ gd,
                                                  int *__restrict__ nlist, int *__restrict__ ncount) {
    __shared__ int qbuf[WalkLds<WALK>::ints];
    const int nf = cs[q0.G];
    const int i = xcd_block(gridDim.x) * ZS_BLOCK + threadIdx.x;
    if (i >= q0.N) return;
    if (i >= nf) {
        gd[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    const bool delta = q.diffusion == ZSPH_DIFF_ANTUONO;
    const float4 xi = b.posr[i];
    const float4 vi = b.velp[i];
    float lxx = 0, lxy = 0, lxz = 0, lyy = 0, lyz = 0, lzz = 0;
    float gx = 0, gy = 0, gz = 0, D = 0;
    auto body = [&](int j, float dx, float dy, float dz, float r2, float4 xj) {
        const float qq = sqrtf(r2) * q.inv_h;
        const float t = 1.0f - 0.5f * qq;
        const float FV = q.fk * t * t * t * q.m / xj.w;
        if (delta) {
            lxx -= FV * dx * dx; lxy -= FV * dx * dy; lxz -= FV * dx * dz;
            lyy -= FV * dy * dy; lyz -= FV * dy * dz; lzz -= FV * dz * dz;
            const float c = FV * (xj.w - xi.w);
            gx += c * dx; gy += c * dy; gz += c * dz;
        }
        if (j < nf) {
            const float4 vj = b.velp[j];
            D -= FV * ((vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz);
        }
    };
    const uint32_t cell = cell_of(q, skey[i]);
    if (WALK == WALK_LIST) {
        const int cnt = build_list<true, true>(q, i, xi, cell, cs, b.posr, nlist);
        ncount[i] = cnt;
        visit_list<true, true>(q, i, xi, cell, cs, b.posr, nlist, cnt, body);
    } else {
        for_each_neighbor<WALK, true, true>(q, i, xi, cell, cs, b.posr, qbuf, body);
    }
    if (delta) {
        if (q.dim == 2) lyy += 1.0f;  // x-z plane: unit yy entry keeps L invertible
        const float c00 = lyy * lzz - lyz * lyz;
        const float c01 = lxz * lyz - lxy * lzz;
        const float c02 = lxy * lyz - lxz * lyy;
        const float det = lxx * c00 + lxy * c01 + lxz * c02;
        if (fabsf(det) > 1.8e-8f) {
            const float c11 = lxx * lzz - lxz * lxz;
            const float c12 = lxy * lxz - lxx * lyz;
            const float c22 = lxx * lyy - lxy * lxy;
            const float id = 1.0f / det;
            const float ox = (c00 * gx + c01 * gy + c02 * gz) * id;
            const float oy = (c01 * gx + c11 * gy + c12 * gz) * id;
            const float oz = (c02 * gx + c12 * gy + c22 * gz) * id;
            gx = ox; gy = oy; gz = oz;
        }
    } else {
        gx = gy = gz = 0.0f;
    }
    gd[i] = make_float4(gx, gy, gz, D);
}

// Continuity + momentum of fluid particles (delta_ns.cu:204-265; MC form
// simple_ns.cu:123-177).  Pressure and continuity over all neighbours,
// delta term and viscosity over fluid pairs, gravity added.
template <int WALK, bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_governing(Prm q0, const Prm *__restrict__ sp, Buf b,
        const uint32_t *__restrict__ skey,
                                                        const int *__restrict__ cs, const float4 *__restrict__ gd,
                                                        float4 *__restrict__ acc, const int *__restrict__ nlist,
                                                        const int *__restrict__ ncount) {
    __shared__ int qbuf[WalkLds<WALK>::ints];
    const int nf = cs[q0.G];
    const int i = xcd_block(gridDim.x) * ZS_BLOCK + threadIdx.x;
    if (i >= nf) return;
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    const bool delta = q.diffusion == ZSPH_DIFF_ANTUONO;
    const float4 xi = b.posr[i];
    const float4 vi = b.velp[i];
    const float4 gi = gd[i];
    const float rhoi = xi.w, pi = vi.w;
    float ax = 0, ay = 0, az = 0, dr = 0;
    auto body = [&](int j, float dx, float dy, float dz, float r2, float4 xj) {
        const float qq = sqrtf(r2) * q.inv_h;
        const float t = 1.0f - 0.5f * qq;
        const float F = q.fk * t * t * t;
        const float4 vj = b.velp[j];
        const float rhoj = xj.w, pj = vj.w;
        const float Vj = q.m / rhoj;
        const float dvdx = (vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz;
        dr += rhoi * dvdx * F * Vj;
        const float irr = 1.0f / (rhoi * rhoj);
        float pt = -(pi + pj) * q.m * F * irr;
        if (j < nf) {
            const float ire = 1.0f / (r2 + q.eta2);
            const float4 gj = gd[j];
            if (delta) {
                const float psi = 2.0f * (rhoi - rhoj) * r2 * ire -
                                  ((gi.x + gj.x) * dx + (gi.y + gj.y) * dy + (gi.z + gj.z) * dz);
                dr += q.delta_h_cs * F * psi * Vj;
            }
            pt += q.h_cs_rho_m * F * (gi.w + gj.w + q.alpha * dvdx * ire) * irr;
        }
        if (!delta) dr += (rhoi - rhoj) * r2 * F * q.delta_h_cs / (r2 + q.eta2) * Vj;
        ax += pt * dx; ay += pt * dy; az += pt * dz;
    };
    const uint32_t cell = cell_of(q, skey[i]);
    if (WALK == WALK_LIST)
        visit_list<true, true>(q, i, xi, cell, cs, b.posr, nlist, ncount[i], body);
    else
        for_each_neighbor<WALK, true, true>(q, i, xi, cell, cs, b.posr, qbuf, body);
    acc[i] = make_float4(ax, ay, az + q.g, dr);
}

// Dummies: moving ones (body, piston) get the pressure force of their fluid
// neighbours, without gravity (their sum is the fluid force on the body);
// fixed walls and removed particles get 0.
template <int WALK, bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_governing_dummy(Prm q0, const Prm *__restrict__ sp, Buf b,
        const uint32_t *__restrict__ skey,
                                                              const int *__restrict__ cs, float4 *__restrict__ acc) {
    __shared__ int qbuf[WalkLds<WALK>::ints];
    const int nf = cs[q0.G];
    const int i = nf + blockIdx.x * ZS_BLOCK + threadIdx.x;
    if (i >= q0.N) return;
    const int ti = b.type[i];
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    float ax = 0, ay = 0, az = 0, dr = 0;
    if (ti == ZSPH_BODY || ti == ZSPH_PISTON) {
        const float4 xi = b.posr[i];
        const float4 vi = b.velp[i];
        const float rhoi = xi.w, pi = vi.w;
        for_each_neighbor<WALK, true, false>(q, i, xi, cell_of(q, skey[i]), cs, b.posr, qbuf,
                                       [&](int j, float dx, float dy, float dz, float r2, float4 xj) {
            const float qq = sqrtf(r2) * q.inv_h;
            const float t = 1.0f - 0.5f * qq;
            const float F = q.fk * t * t * t;
            const float4 vj = b.velp[j];
            const float rhoj = xj.w;
            const float dvdx = (vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz;
            dr += rhoi * dvdx * F * q.m / rhoj;
            const float pt = -(pi + vj.w) * q.m * F / (rhoi * rhoj);
            ax += pt * dx; ay += pt * dy; az += pt * dz;
        });
    }
    acc[i] = make_float4(ax, ay, az, dr);
}

// Midpoint predictor (time_int.cu:3-48), fluid only: save state n, then
// x += dt/2 v, v += dt/2 a, rho = max(rho + dt/2 drho, rho0), p = EOS(rho).
template <bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_predict(Prm q0, const Prm *__restrict__ sp, Buf b,
        const int *__restrict__ cs,
                                                      const float4 *__restrict__ acc) {
    const int i = blockIdx.x * ZS_BLOCK + threadIdx.x;
    if (i >= cs[q0.G]) return;
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    float4 x = b.posr[i], v = b.velp[i];
    const float4 a = acc[i];
    b.tposr[i] = x;
    b.tvelp[i] = v;
    x.x += v.x * q.half_dt; x.y += v.y * q.half_dt; x.z += v.z * q.half_dt;
    v.x += a.x * q.half_dt; v.y += a.y * q.half_dt; v.z += a.z * q.half_dt;
    x.w = fmaxf(x.w + a.w * q.half_dt, q.rho0);
    v.w = eos_p(q, x.w);
    b.posr[i] = x;
    b.velp[i] = v;
}

// Midpoint corrector (time_int.cu:50-91): x = x^n + dt v*, v = v^n + dt a*,
// rho = max(rho^n + dt drho*, rho0), p = EOS(rho).
template <bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_correct(Prm q0, const Prm *__restrict__ sp, Buf b,
        const int *__restrict__ cs,
                                                      const float4 *__restrict__ acc) {
    const int i = blockIdx.x * ZS_BLOCK + threadIdx.x;
    if (i >= cs[q0.G]) return;
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    float4 x = b.tposr[i], v = b.tvelp[i];
    const float4 vs = b.velp[i];
    const float4 a = acc[i];
    x.x += vs.x * q.dt; x.y += vs.y * q.dt; x.z += vs.z * q.dt;
    v.x += a.x * q.dt; v.y += a.y * q.dt; v.z += a.z * q.dt;
    x.w = fmaxf(x.w + a.w * q.dt, q.rho0);
    v.w = eos_p(q, x.w);
    b.posr[i] = x;
    b.velp[i] = v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Neighbour count and set digest sum_j mix64(j) (mod 2^64) per particle,
// through the same walk as the physics kernels.
template <int WALK, bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_nbr_stats(Prm q0, const Prm *__restrict__ sp, Buf b,
        const uint32_t *__restrict__ skey,
                                                        const int *__restrict__ cs, int *__restrict__ count,
                                                        uint64_t *__restrict__ digest) {
    __shared__ int qbuf[WalkLds<WALK>::ints];
    const int i = blockIdx.x * ZS_BLOCK + threadIdx.x;
    if (i >= q0.N) return;
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    const float4 xi = b.posr[i];
    int c = 0;
    uint64_t h = 0;
    for_each_neighbor<WALK, true, true>(q, i, xi, cell_of(q, skey[i]), cs, b.posr, qbuf,
                                  [&](int j, float, float, float, float, float4) {
                                      ++c;
                                      h += mix64((uint64_t)j);
                                  });
    count[i] = c;
    digest[i] = h;
}

// Walks other than the queue use the mask walk where no list applies; a
// batch (sp != nullptr) runs the per-scene-parameter instantiation.
#define ZS_WALK_LAUNCH(kern, grid, ...)                                                         \
    do {                                                                                        \
        if (sp) {                                                                               \
            if (walk == WALK_QUEUE)                                                             \
                kern<WALK_QUEUE, true><<<(grid), ZS_BLOCK, 0, s>>>(__VA_ARGS__);                \
            else                                                                                \
                kern<WALK_MASK, true><<<(grid), ZS_BLOCK, 0, s>>>(__VA_ARGS__);                 \
        } else {                                                                                \
            if (walk == WALK_QUEUE)                                                             \
                kern<WALK_QUEUE, false><<<(grid), ZS_BLOCK, 0, s>>>(__VA_ARGS__);               \
            else                                                                                \
                kern<WALK_MASK, false><<<(grid), ZS_BLOCK, 0, s>>>(__VA_ARGS__);                \
        }                                                                                       \
    } while (0)

void launch_wall(hipStream_t s, int walk, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                 const int *cs) {
    if (q.N) ZS_WALK_LAUNCH(k_wall, nblocks(q.N), q, sp, b, skey, cs);
}
void launch_ldg(hipStream_t s, int walk, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                const int *cs, float4 *gd, int *nlist, int *ncount) {
    if (!q.N) return;
    if (walk == WALK_LIST) {
        if (sp)
            k_ldg<WALK_LIST, true><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, gd, nlist, ncount);
        else
            k_ldg<WALK_LIST, false><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, gd, nlist, ncount);
    } else {
        ZS_WALK_LAUNCH(k_ldg, nblocks(q.N), q, sp, b, skey, cs, gd, nlist, ncount);
    }
}
void launch_governing(hipStream_t s, int walk, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                      const int *cs, const float4 *gd, float4 *acc, const int *nlist, const int *ncount) {
    if (!q.N) return;
    if (walk == WALK_LIST) {
        if (sp)
            k_governing<WALK_LIST, true><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, gd, acc, nlist,
                                                                           ncount);
        else
            k_governing<WALK_LIST, false><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, gd, acc, nlist,
                                                                            ncount);
    } else {
        ZS_WALK_LAUNCH(k_governing, nblocks(q.N), q, sp, b, skey, cs, gd, acc, nlist, ncount);
    }
    ZS_WALK_LAUNCH(k_governing_dummy, nblocks(q.N), q, sp, b, skey, cs, acc);
}
void launch_predict(hipStream_t s, const Prm &q, const Prm *sp, const Buf &b, const int *cs, const float4 *acc) {
    if (!q.N) return;
    if (sp)
        k_predict<true><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, cs, acc);
    else
        k_predict<false><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, cs, acc);
}
void launch_correct(hipStream_t s, const Prm &q, const Prm *sp, const Buf &b, const int *cs, const float4 *acc) {
    if (!q.N) return;
    if (sp)
        k_correct<true><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, cs, acc);
    else
        k_correct<false><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, cs, acc);
}
void launch_nbr_stats(hipStream_t s, int walk, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                      const int *cs, int *count, uint64_t *digest) {
    if (q.N) ZS_WALK_LAUNCH(k_nbr_stats, nblocks(q.N), q, sp, b, skey, cs, count, digest);
}

}  // namespace zs
