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

// Waves per SIMD the single-scene plain gradient and continuity+momentum
// passes are compiled for (a register budget of 512 / waves VGPRs; the batch
// forms keep the compiler's own budget, about 90 VGPRs and 5 waves: capped
// like the single-scene forms they spill and measured slower), and the
// records per wave the continuity+momentum pass stages in LDS (positions and
// velocities of a row pair; larger pairs are gathered from L2): 6 waves
// without spilling registers, with 176 records (192 measured 8 % faster than
// 7 waves with 160 records, which spills, and than 5 waves with 240; 176 runs
// as fast as 192 on the dam break and 0.8 % faster on the wave tank, 160
// slower).
#ifndef ZS_LDG_WAVES
#define ZS_LDG_WAVES 7
#endif
#ifndef ZS_GOV_WAVES
#define ZS_GOV_WAVES 6
#endif
#ifndef ZS_GOV_SLOTS
#define ZS_GOV_SLOTS 176
#endif

__device__ __forceinline__ float *wptr(float4 *a, int i) { return reinterpret_cast<float *>(a + i) + 3; }

// Cell of a sort key, relative to the particle's scene (cell_base = 0 for one scene).
__device__ __forceinline__ uint32_t cell_of(const Prm &q, uint32_t key) {
    return (key & (q.flag - 1u)) - (uint32_t)q.cell_base;
}

// Diffusion form of a launch (template argument DF): ZSPH_DIFF_ANTUONO or
// ZSPH_DIFF_MC fixed at compile time for one scene, so the pair bodies carry
// no per-pair test of it; DIFF_PER_SCENE reads it from each particle's scene
// parameters (a batch's scenes may differ).
constexpr int DIFF_PER_SCENE = -1;
template <int DF>
__device__ __forceinline__ bool is_antuono(const Prm &q) {
    if constexpr (DF == DIFF_PER_SCENE) return q.diffusion == ZSPH_DIFF_ANTUONO;
    else return DF == ZSPH_DIFF_ANTUONO;
}

// Dummy particles: Shepard extrapolation of fluid pressure with the
// hydrostatic and wall-acceleration correction,
//   p_w = max(sum (p_j - rho_j (a_w - g) . x_wj) W / sum W, 0),
// rho_w from the inverse EOS (delta_ns.cu:81-85, :116-131; a_w = 0 there).
// a_w is the dummy's prescribed acceleration, kept in tvelp.xyz (zero for
// fixed walls).
template <int WALK, bool BATCH, int BS>
__global__ __launch_bounds__(BS) void k_wall(Prm q0, const Prm *__restrict__ sp, Buf b,
                                             const uint32_t *__restrict__ skey, const int *__restrict__ cs) {
    stage_stamp(q0, -1, ZSPH_ST_WALL);
    const int nf = cs[q0.G];
    // dummy blocks round-robin over the XCDs: dummies far from the fluid
    // cost little, so contiguous ranges of them would load the XCDs unevenly
    const int i = nf + blockIdx.x * BS + threadIdx.x;
    if (i >= q0.N) return;
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    const float4 xi = b.posr[i];
    const float4 aw = b.tvelp[i];
    float P = 0.0f, S = 0.0f;
    const float nhh = -0.5f * q.inv_h, ih2 = 2.0f * q.inv_h;  // t = 1 - r/(2h), 1 + 2r/h: one multiply-add each
    for_each_neighbor<WALK, true, false>(q, i, xi, cell_of(q, skey[i]), cs, b.posr,
                                   [&](int j, float dx, float dy, float dz, float r2, float4 xj) {
                                       const float r = hw_sqrt(r2);
                                       const float t = fmaf(r, nhh, 1.0f);
                                       const float W = q.adh * t * t * t * t * fmaf(r, ih2, 1.0f);
                                       const float ax = aw.x * dx + aw.y * dy + (aw.z - q.g) * dz;
                                       P += (ld(b.velp, j).w - xj.w * ax) * W;
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
// FEAT: also the closure network's ZSPH_NF features of each fluid particle
// (feat[i * ZSPH_NF ...], see closure_features) from the same neighbour walk.
// Compiled for at least 7 waves per SIMD (at most 72 VGPRs; the compiler's
// own choice, 77, allows 6) in the single-scene mask and hit-mask walks:
// measured 5 % faster.  8 waves (64 VGPRs) spill and measured slower.
template <int WALK, bool BATCH, bool FEAT, int DF, int BS>
__global__ __launch_bounds__(BS, ((WALK != WALK_LIST && !BATCH && !FEAT) ? ZS_LDG_WAVES : 1)) void k_ldg(Prm q0, const Prm *__restrict__ sp, Buf b,
                                                  const uint32_t *__restrict__ skey, const int *__restrict__ cs,
                                                  float4 *__restrict__ gd, int *__restrict__ nlist,
                                                  int *__restrict__ ncount, float *__restrict__ feat, SkinRec sr) {
    // the stage clocks: the recording form of WALK_SKIN is launched first, the replaying form right after it
    if constexpr (WALK != WALK_SKIN_REPLAY) stage_stamp(q0, ZSPH_ST_WALL, ZSPH_ST_LDG);
    // WALK_SKIN: the recording form runs on rebuild half steps, the replaying form on the others
    if constexpr (WALK == WALK_SKIN_REC) {
        if (*sr.rebuilt == 0) return;
    } else if constexpr (WALK == WALK_SKIN_REPLAY) {
        if (*sr.rebuilt != 0) return;
    }
    const int nf = cs[q0.G];
    const int i = pass_particle<BS>(q0, nf);
    if (i >= q0.N) return;
    if (i >= nf) {
        gd[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    // the wave's LDS slice for staging its candidate rows (a batch's wave
    // stages only when its lanes share one scene's grid)
    __shared__ float4 stage_lds[(BS / 64) * (STAGE_SLOTS + STAGE_PAD)];
    with_scene<BATCH>(q0, sp, b, i, [&](const Prm &q) {
        const bool delta = is_antuono<DF>(q);
        const bool ell = delta || FEAT;  // L is needed
        const float4 xi = b.posr[i];
        const float4 vi = b.velp[i];
        float lxx = 0, lxy = 0, lxz = 0, lyy = 0, lyz = 0, lzz = 0;
        float gx = 0, gy = 0, gz = 0, D = 0;
        // FEAT only: velocity-gradient tensor K_ab = sum F V (v_j - v_i)_a x_b (fluid
        // neighbours), kernel sum, neighbour counts, velocity spread, density contrast
        float kxx = 0, kxy = 0, kxz = 0, kyx = 0, kyy = 0, kyz = 0, kzx = 0, kzy = 0, kzz = 0;
        float ws = 0, wv2 = 0, wdr = 0;
        int nnb = 0, ndum = 0;
        // t = 1 - r/(2h) in one multiply-add; F V_j = (fk m) t^3 / rho_j
        const float nhh = -0.5f * q.inv_h;
        const float fkm = q.fk * q.m;
        // vj: velp[j] (read only for fluid j)
        auto body = [&](int j, float dx, float dy, float dz, float r2, float4 xj, float4 vj, auto kind) {
            const float r = hw_sqrt(r2);
            const float t = fmaf(r, nhh, 1.0f);
            const float FV = fkm * t * t * t * hw_rcp(xj.w);
            // F V_j x_ij, shared by L, the density gradient and D
            const float fx = FV * dx, fy = FV * dy, fz = FV * dz;
            if (ell) {
                lxx -= fx * dx; lxy -= fx * dy; lxz -= fx * dz;
                lyy -= fy * dy; lyz -= fy * dz; lzz -= fz * dz;
            }
            if (delta) {
                const float drho = xj.w - xi.w;
                gx += fx * drho; gy += fy * drho; gz += fz * drho;
            }
            float WV = 0.0f;
            if (FEAT) {
                const float Vj = q.m * hw_rcp(xj.w);
                WV = q.adh * t * t * t * t * (2.0f * r * q.inv_h + 1.0f) * Vj;
                ws += WV;
                wdr += WV * (xj.w - xi.w);
                ++nnb;
            }
            if (fluid_j(kind, j, nf)) {
                // D -= F V_j (v_i - v_j) . x_ij, one multiply-add per component
                const float ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
                D += fx * ux;
                D += fy * uy;
                D += fz * uz;
                if (FEAT) {
                    kxx += FV * ux * dx; kxy += FV * ux * dy; kxz += FV * ux * dz;
                    kyx += FV * uy * dx; kyy += FV * uy * dy; kyz += FV * uy * dz;
                    kzx += FV * uz * dx; kzy += FV * uz * dy; kzz += FV * uz * dz;
                    wv2 += WV * (ux * ux + uy * uy + uz * uz);
                }
            } else if (FEAT) {
                ++ndum;
            }
        };
        // gathered per neighbour besides its position: velp[j] for fluid j
        auto gath = [&](int j, auto kind) {
            if constexpr (decltype(kind)::value == 1) {
                return Empty{};
            } else if constexpr (decltype(kind)::value == 0) {
                return ld(b.velp, j);
            } else {
                float4 v;  // loaded, and read by the body, for fluid neighbours only
                if (j < nf) v = ld(b.velp, j);
                return v;
            }
        };
        auto walk_body = [&](int j, float dx, float dy, float dz, float r2, float4 xj, auto kind, const auto &e) {
            float4 vj;
            if constexpr (!std::is_same_v<std::decay_t<decltype(e)>, Empty>) vj = e;
            body(j, dx, dy, dz, r2, xj, vj, kind);
        };
        const uint32_t cell = cell_of(q, skey[i]);
        LdsStage stg{stage_lds + (threadIdx.x >> 6) * (STAGE_SLOTS + STAGE_PAD)};
        if constexpr (BATCH) stg.tag = b.scene[i];
        if (WALK == WALK_LIST) {  // the mask walk, recording what it visits
            ListWriter lw(nlist, i);
            for_each_neighbor<WALK_MASK, true, true>(q, i, xi, cell, cs, b.posr,
                                                     [&](int j, float dx, float dy, float dz, float r2, float4 xj,
                                                         auto kind, const auto &e) {
                                                         lw.put(j);
                                                         walk_body(j, dx, dy, dz, r2, xj, kind, e);
                                                     },
                                                     NoRec{}, gath);
            ncount[i] = lw.finish();
        } else if constexpr (WALK == WALK_SKIN_REC || WALK == WALK_SKIN_REPLAY) {
            uint32_t *hits = reinterpret_cast<uint32_t *>(nlist);  // this half step's exact masks
            const size_t n = (size_t)q0.N;
            Prm qr = q;  // rows around the rebuild position, trimmed to rc + s
            qr.trim2 = q.trim2_skin;
            const float4 xc = sr.xref[i];
            auto exact = [&](int j, float dx, float dy, float dz, float r2, float4 xj, auto kind,
                             const auto &e) -> bool {
                hold(xj.w);  // every load of the hit before the test (one L2 round trip)
                hold_xyz(e);
                if (!(r2 <= q.rc2)) return false;
                walk_body(j, dx, dy, dz, r2, xj, kind, e);
                return true;
            };
            auto rec_exact = [&](int r, int c, uint32_t m) {
                if (c < 2) hits[(size_t)(2 * r + c) * n + (size_t)i] = m;
            };
            if constexpr (WALK == WALK_SKIN_REC) {
                bool over = false;  // a row of more than 64 candidates: this particle is not recorded
                skin_walk<false>(qr, i, xi, xc, cell, cs, b.posr, sr.skin, n, exact,
                                 [&](int r, int c, uint32_t m) {
                                     if (c < 2) sr.skin[(size_t)(2 * r + c) * n + (size_t)i] = m;
                                     else over = true;
                                 },
                                 rec_exact, gath);
                ncount[i] = over ? 1 : 0;
            } else if (ncount[i]) {  // not recorded at the rebuild: the mask walk (and so its successor)
                for_each_neighbor<WALK_MASK, true, true>(q, i, xi, cell, cs, b.posr, walk_body, NoRec{}, gath);
            } else {
                skin_walk<true>(qr, i, xi, xc, cell, cs, b.posr, sr.skin, n, exact, [](int, int, uint32_t) {},
                                rec_exact, gath);
            }
        } else if (WALK == WALK_HITS) {  // the mask walk, recording each chunk's hit mask
            uint32_t *hits = reinterpret_cast<uint32_t *>(nlist);
            bool over = false;  // a row with more than 64 candidates: not recorded
            for_each_neighbor<WALK_MASK, true, true>(q, i, xi, cell, cs, b.posr, walk_body,
                                                     [&](int r, int c, uint32_t m) {
                                                         if (c < 2) hits[(size_t)(2 * r + c) * (size_t)q0.N + (size_t)i] = m;
                                                         else over = true;
                                                     },
                                                     gath, stg);
            ncount[i] = over ? 1 : 0;
        } else {
            for_each_neighbor<WALK, true, true>(q, i, xi, cell, cs, b.posr, walk_body, NoRec{}, gath, stg);
        }
        const float trL = lxx + lyy + lzz;
        if (ell && q.dim == 2) lyy += 1.0f;  // x-z plane: unit yy entry keeps L invertible
        const float c00 = lyy * lzz - lyz * lyz;
        const float c01 = lxz * lyz - lxy * lzz;
        const float c02 = lxy * lyz - lxz * lyy;
        const float det = lxx * c00 + lxy * c01 + lxz * c02;
        if (delta) {
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
        if constexpr (FEAT) {
            const float ic0 = rsqrtf(q.cs2), hc = q.h * ic0, hr = q.h / q.rho0;
            const float self = q.adh * q.m / xi.w;  // W(0) V_i
            const float wx = kzy - kyz, wy = kxz - kzx, wz = kyx - kxy;
            const float dv = kxx + kyy + kzz, third = dv / (float)q.dim;
            const float sxx = kxx - third, syy = (q.dim == 3) ? kyy - third : 0.0f, szz = kzz - third;
            const float sxy = 0.5f * (kxy + kyx), sxz = 0.5f * (kxz + kzx), syz = 0.5f * (kyz + kzy);
            const float g2 = gx * gx + gy * gy + gz * gz;
            const float v2 = vi.x * vi.x + vi.y * vi.y + vi.z * vi.z;
            float f[ZSPH_NF];
            f[0] = 1.0f - (ws + self);
            f[1] = trL / (float)q.dim;
            f[2] = det;
            f[3] = sqrtf(lxy * lxy + lxz * lxz + lyz * lyz);
            f[4] = gx * hr; f[5] = gy * hr; f[6] = gz * hr;
            f[7] = sqrtf(g2) * hr;
            f[8] = D * hc;
            f[9] = wx * hc; f[10] = wy * hc; f[11] = wz * hc;
            f[12] = sqrtf(wx * wx + wy * wy + wz * wz) * hc;
            f[13] = sqrtf(sxx * sxx + syy * syy + szz * szz + 2.0f * (sxy * sxy + sxz * sxz + syz * syz)) * hc;
            f[14] = vi.x * ic0; f[15] = vi.y * ic0; f[16] = vi.z * ic0;
            f[17] = sqrtf(v2) * ic0;
            f[18] = (xi.w - q.rho0) / q.rho0;
            f[19] = vi.w / (q.rho0 * q.cs2);
            f[20] = (float)nnb;
            f[21] = (ws > 0.0f) ? sqrtf(wv2 / ws) * ic0 : 0.0f;
            f[22] = (ws > 0.0f) ? wdr / (ws * q.rho0) : 0.0f;
            f[23] = nnb ? (float)ndum / (float)nnb : 0.0f;
            float *o = feat + (size_t)i * ZSPH_NF;
    #pragma unroll
            for (int k = 0; k < ZSPH_NF; k += 4) *reinterpret_cast<float4 *>(o + k) = make_float4(f[k], f[k + 1], f[k + 2], f[k + 3]);
        }
    });
}

// Continuity + momentum of fluid particles (delta_ns.cu:204-265; MC form
// simple_ns.cu:123-177).  Pressure and continuity over all neighbours,
// delta term and viscosity over fluid pairs, gravity added.
// MODE 0: plain.  MODE 1 (learned closure): the density-diffusion part of
// drho/dt is scaled by scale[i].x and the viscous part of the acceleration
// (bulk + artificial viscosity) by scale[i].y.  MODE 2: plain, and the parts
// are written to terms[i * 8 ...] = (continuity, diffusion, pressure + g (3),
// viscous (3)) for training.
// At least 7 waves per SIMD in the single-scene plain walks (as k_ldg).
template <int WALK, bool BATCH, int MODE, int DF, int BS>
__global__ __launch_bounds__(BS, ((WALK != WALK_LIST && !BATCH && MODE == 0) ? ZS_GOV_WAVES : 1)) void k_governing(Prm q0, const Prm *__restrict__ sp, Buf b,
                                                        const uint32_t *__restrict__ skey, const int *__restrict__ cs,
                                                        const float4 *__restrict__ gd, float4 *__restrict__ acc,
                                                        const int *__restrict__ nlist, const int *__restrict__ ncount,
                                                        const float2 *__restrict__ scale, float *__restrict__ terms,
                                                        const float4 *__restrict__ xref) {
    stage_stamp(q0, ZSPH_ST_LDG, ZSPH_ST_GOVERNING);
    const int nf = cs[q0.G];
    const int i = pass_particle<BS>(q0, nf);
    if (i >= nf) {
        // dummies and removed particles: 0 (k_governing_dummy, launched after
        // this kernel when there are moving dummies, overwrites theirs)
        if (i < q0.N) acc[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    // the wave's LDS slices for staging its rows' positions and velocities
    constexpr int GS = ZS_GOV_SLOTS + STAGE_PAD;
    __shared__ float4 stage_x[(BS / 64) * GS];
    __shared__ float4 stage_v[(BS / 64) * GS];
    with_scene<BATCH>(q0, sp, b, i, [&](const Prm &q) {
        const bool delta = is_antuono<DF>(q);
        const float4 xi = b.posr[i];
        const float4 vi = b.velp[i];
        const float4 gi = gd[i];
        const float rhoi = xi.w, pi = vi.w;
        const float irhoi = hw_rcp(rhoi);  // 1/(rho_i rho_j) = irhoi irhoj: one reciprocal per pair
        float ax = 0, ay = 0, az = 0, dr = 0;
        float dd = 0, vx = 0, vy = 0, vz = 0;  // diffusion and viscous parts (MODE != 0)
        const float nhh = -0.5f * q.inv_h;       // t = 1 - r/(2h) in one multiply-add
        // gj: gd[j] (read only for fluid j)
        auto body = [&](int j, float dx, float dy, float dz, float r2, float4 xj, float4 vj, float4 gj, auto kind) {
            const float t = fmaf(hw_sqrt(r2), nhh, 1.0f);
            const float F = q.fk * t * t * t;
            const float rhoj = xj.w, pj = vj.w;
            const float irhoj = hw_rcp(rhoj);
            const float Vj = q.m * irhoj;
            const float dvdx = (vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz;
            dr += rhoi * dvdx * F * Vj;
            const float irr = irhoi * irhoj;
            float pt = -(pi + pj) * q.m * F * irr;
            float pv = 0.0f;
            if (fluid_j(kind, j, nf)) {
                const float ire = hw_rcp(r2 + q.eta2);
                if (delta) {
                    const float psi = 2.0f * (rhoi - rhoj) * r2 * ire -
                                      ((gi.x + gj.x) * dx + (gi.y + gj.y) * dy + (gi.z + gj.z) * dz);
                    if (MODE) dd += q.delta_h_cs * F * psi * Vj;
                    else dr += q.delta_h_cs * F * psi * Vj;
                }
                if (MODE) pv = q.h_cs_rho_m * F * (gi.w + gj.w + q.alpha * dvdx * ire) * irr;
                else pt += q.h_cs_rho_m * F * (gi.w + gj.w + q.alpha * dvdx * ire) * irr;
            }
            if (!delta) {
                const float ire = hw_rcp(r2 + q.eta2);
                if (MODE) dd += (rhoi - rhoj) * r2 * F * q.delta_h_cs * ire * Vj;
                else dr += (rhoi - rhoj) * r2 * F * q.delta_h_cs * ire * Vj;
            }
            ax += pt * dx; ay += pt * dy; az += pt * dz;
            if (MODE) { vx += pv * dx; vy += pv * dy; vz += pv * dz; }
        };
        // gathered per neighbour besides its position: velp[j], and gd[j] for fluid j
        auto gath = [&](int j, auto kind) {
            if constexpr (decltype(kind)::value == 1) {
                return ld(b.velp, j);
            } else if constexpr (decltype(kind)::value == 0) {
                return VelGrad{ld(b.velp, j), ld(gd, j)};
            } else {
                VelGrad e;
                e.v = ld(b.velp, j);
                if (j < nf) e.g = ld(gd, j);  // read by the body for fluid neighbours only
                return e;
            }
        };
        auto walk_body = [&](int j, float dx, float dy, float dz, float r2, float4 xj, auto kind, const auto &e) {
            if constexpr (std::is_same_v<std::decay_t<decltype(e)>, float4>) {
                float4 gj;  // not read for a dummy neighbour
                body(j, dx, dy, dz, r2, xj, e, gj, kind);
            } else {
                body(j, dx, dy, dz, r2, xj, e.v, e.g, kind);
            }
        };
        const uint32_t cell = cell_of(q, skey[i]);
        const int n = (WALK == WALK_LIST || WALK == WALK_HITS || WALK == WALK_SKIN) ? ncount[i] : 0;
        LdsStage stg{stage_x + (threadIdx.x >> 6) * GS, stage_v + (threadIdx.x >> 6) * GS, b.velp, ZS_GOV_SLOTS};
        if constexpr (BATCH) stg.tag = b.scene[i];
        // the gathered data of a hit whose velocity is staged: gd[j] for fluid j
        auto gath_v = [&](int j, auto kind, float4 v) {
            if constexpr (decltype(kind)::value == 1) {
                return v;
            } else if constexpr (decltype(kind)::value == 0) {
                return VelGrad{v, ld(gd, j)};
            } else {
                VelGrad e;
                e.v = v;
                if (j < nf) e.g = ld(gd, j);  // read by the body for fluid neighbours only
                return e;
            }
        };
        if (WALK == WALK_SKIN && n == 0) {  // the exact record, rows around the rebuild position
            Prm qr = q;
            qr.trim2 = q.trim2_skin;
            for_each_recorded_at<true, true>(qr, i, xi, xref[i], cell, cs, b.posr,
                                             reinterpret_cast<const uint32_t *>(nlist), (size_t)q0.N, walk_body, gath,
                                             stg, gath_v);
        } else if (WALK == WALK_HITS && n == 0) {
            for_each_recorded<true, true>(q, i, xi, cell, cs, b.posr, reinterpret_cast<const uint32_t *>(nlist),
                                          (size_t)q0.N, walk_body, gath, stg, gath_v);
        } else if (WALK == WALK_LIST && n <= LIST_MAX) {
            // R quads (4R recorded neighbours) per round: the round's list reads
            // and all its gathers are issued before its first interaction, so
            // they are in flight together (this pass is latency-bound; 8 per
            // round measured best).  gd of a dummy is zero (k_ldg writes it), so
            // every gather is unconditional.
            constexpr int R = 2;
            const int4 *list = reinterpret_cast<const int4 *>(nlist);
            for (int k = 0; k < n; k += 4 * R) {
                int js[4 * R];
    #pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int4 jj = (k + 4 * r < n) ? nt_load(&list[list_quad(i, (k >> 2) + r)]) : make_int4(0, 0, 0, 0);
                    js[4 * r] = jj.x; js[4 * r + 1] = jj.y; js[4 * r + 2] = jj.z; js[4 * r + 3] = jj.w;
                }
    #pragma unroll
                for (int e = 1; e < 4 * R; ++e) js[e] = (k + e < n) ? js[e] : js[0];  // past the end: a valid index
                float4 xs[4 * R], vs[4 * R], gs[4 * R];
    #pragma unroll
                for (int e = 0; e < 4 * R; ++e) {
                    xs[e] = ld(b.posr, js[e]);
                    vs[e] = ld(b.velp, js[e]);
                    gs[e] = ld(gd, js[e]);
                }
    #pragma unroll
                for (int e = 0; e < 4 * R; ++e) {
                    if (k + e < n) {
                        const float dx = xi.x - xs[e].x, dy = xi.y - xs[e].y, dz = xi.z - xs[e].z;
                        body(js[e], dx, dy, dz, dist2(dx, dy, dz), xs[e], vs[e], gs[e], AnyJ{});
                    }
                }
            }
        } else {
            for_each_neighbor<WALK_MASK, true, true>(q, i, xi, cell, cs, b.posr, walk_body, NoRec{}, gath);
        }
        if constexpr (MODE == 0) {
            acc[i] = make_float4(ax, ay, az + q.g, dr);
        } else if constexpr (MODE == 1) {
            const float2 sc = scale[i];
            acc[i] = make_float4(ax + sc.y * vx, ay + sc.y * vy, az + q.g + sc.y * vz, dr + sc.x * dd);
        } else {
            float4 *o = reinterpret_cast<float4 *>(terms + (size_t)i * 8);
            o[0] = make_float4(dr, dd, ax, ay);
            o[1] = make_float4(az + q.g, vx, vy, vz);
            acc[i] = make_float4(ax + vx, ay + vy, az + q.g + vz, dr + dd);
        }
    });
}

// Dummies: moving ones (body, piston) get the pressure force of their fluid
// neighbours, without gravity (their sum is the fluid force on the body);
// fixed walls and removed particles get 0.
template <int WALK, bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_governing_dummy(Prm q0, const Prm *__restrict__ sp, Buf b,
        const uint32_t *__restrict__ skey,
                                                              const int *__restrict__ cs, float4 *__restrict__ acc) {
    stage_stamp(q0, ZSPH_ST_GOVERNING, -1);
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
        for_each_neighbor<WALK, true, false>(q, i, xi, cell_of(q, skey[i]), cs, b.posr,
                                       [&](int j, float dx, float dy, float dz, float r2, float4 xj) {
            const float qq = hw_sqrt(r2) * q.inv_h;
            const float t = 1.0f - 0.5f * qq;
            const float F = q.fk * t * t * t;
            const float4 vj = ld(b.velp, j);
            const float rhoj = xj.w;
            const float dvdx = (vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz;
            dr += rhoi * dvdx * F * (q.m * hw_rcp(rhoj));
            const float pt = -(pi + vj.w) * q.m * F * hw_rcp(rhoi * rhoj);
            ax += pt * dx; ay += pt * dy; az += pt * dz;
        });
    }
    acc[i] = make_float4(ax, ay, az, dr);
}

// Midpoint predictor (time_int.cu:3-48), fluid only: save state n, then
// x += dt/2 v, v += dt/2 a, rho = max(rho + dt/2 drho, rho0), p = EOS(rho).
template <bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_predict(Prm q0, const Prm *__restrict__ sp, Buf b,
                                                      const int *__restrict__ cs, const float4 *__restrict__ acc,
                                                      SkinCheck chk) {
    stage_stamp(q0, ZSPH_ST_GOVERNING, -1);  // the integrator follows the continuity+momentum pass
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
    if (chk.xref) skin_flag_wave(chk, skin_moved(q, chk, i, x));
}

// Midpoint corrector (time_int.cu:50-91): x = x^n + dt v*, v = v^n + dt a*,
// rho = max(rho^n + dt drho*, rho0), p = EOS(rho).
template <bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_correct(Prm q0, const Prm *__restrict__ sp, Buf b,
                                                      const int *__restrict__ cs, const float4 *__restrict__ acc,
                                                      SkinCheck chk) {
    stage_stamp(q0, ZSPH_ST_GOVERNING, -1);  // the integrator follows the continuity+momentum pass
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
    if (chk.xref) skin_flag_wave(chk, skin_moved(q, chk, i, x));
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Neighbour count and set digest sum_j mix64(id_j) (mod 2^64) per particle
// over the neighbours' original ids id_j = table[j] + j, so the digest does
// not depend on the particle order; through the same walk as the physics
// kernels.
template <int WALK, bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_nbr_stats(Prm q0, const Prm *__restrict__ sp, Buf b,
        const uint32_t *__restrict__ skey,
                                                        const int *__restrict__ cs, int *__restrict__ count,
                                                        uint64_t *__restrict__ digest) {
    const int i = blockIdx.x * ZS_BLOCK + threadIdx.x;
    if (i >= q0.N) return;
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    const float4 xi = b.posr[i];
    int c = 0;
    uint64_t h = 0;
    for_each_neighbor<WALK, true, true>(q, i, xi, cell_of(q, skey[i]), cs, b.posr,
                                  [&](int j, float, float, float, float, float4) {
                                      ++c;
                                      h += mix64((uint64_t)(b.table[j] + j));
                                  });
    count[i] = c;
    digest[i] = h;
}

// A batch (sp != nullptr) runs the per-scene-parameter instantiation.  The
// wall pass and the moving-dummy pass only ever take the mask walk (the walk
// variants differ in the gradient and continuity+momentum passes alone).
static inline int nblocks_bs(int n, int bs) { return (n + bs - 1) / bs; }

void launch_wall(hipStream_t s, int walk, int block, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                 const int *cs) {
    (void)walk;
    if (!q.N) return;
    with_block(block, [&](auto BSc) {
        constexpr int BS = decltype(BSc)::value;
        if (sp)
            k_wall<WALK_MASK, true, BS><<<nblocks_bs(q.N, BS), BS, 0, s>>>(q, sp, b, skey, cs);
        else
            k_wall<WALK_MASK, false, BS><<<nblocks_bs(q.N, BS), BS, 0, s>>>(q, sp, b, skey, cs);
    });
}
template <int WALK, bool BATCH, int BS>
static void ldg_walk(hipStream_t s, bool feat, const Prm &q, const Prm *sp, const Buf &b,
                     const uint32_t *skey, const int *cs, float4 *gd, int *nlist, int *ncount, float *f,
                     const SkinRec &sr = SkinRec{}) {
    const int grid = pass_grid(q, BS);
    if constexpr (!BATCH) {
        const bool a = q.diffusion == ZSPH_DIFF_ANTUONO;
        if (a)
            k_ldg<WALK, false, false, ZSPH_DIFF_ANTUONO, BS><<<grid, BS, 0, s>>>(q, sp, b, skey, cs, gd, nlist, ncount, f, sr);
        else
            k_ldg<WALK, false, false, ZSPH_DIFF_MC, BS><<<grid, BS, 0, s>>>(q, sp, b, skey, cs, gd, nlist, ncount, f, sr);
    } else {
        k_ldg<WALK, true, false, DIFF_PER_SCENE, BS><<<grid, BS, 0, s>>>(q, sp, b, skey, cs, gd, nlist, ncount, f, sr);
    }
}
// The feature-writing gradient pass of the learned closure (one scene): 256 threads.
template <int WALK>
static void ldg_feat(hipStream_t s, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey, const int *cs,
                     float4 *gd, int *nlist, int *ncount, float *f) {
    const int grid = pass_grid(q, ZS_BLOCK);
    if (q.diffusion == ZSPH_DIFF_ANTUONO)
        k_ldg<WALK, false, true, ZSPH_DIFF_ANTUONO, ZS_BLOCK><<<grid, ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, gd, nlist, ncount, f, SkinRec{});
    else
        k_ldg<WALK, false, true, ZSPH_DIFF_MC, ZS_BLOCK><<<grid, ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, gd, nlist, ncount, f, SkinRec{});
}

void launch_ldg_skin(hipStream_t s, int block, const Prm &q, const Buf &b, const uint32_t *skey, const int *cs,
                     float4 *gd, int *hits, int *ncount, const SkinRec &sr) {
    if (!q.N) return;
    with_block(block, [&](auto BSc) {
        constexpr int BS = decltype(BSc)::value;
        // exactly one of the two forms does the work (the other returns at once)
        ldg_walk<WALK_SKIN_REC, false, BS>(s, false, q, nullptr, b, skey, cs, gd, hits, ncount, nullptr, sr);
        ldg_walk<WALK_SKIN_REPLAY, false, BS>(s, false, q, nullptr, b, skey, cs, gd, hits, ncount, nullptr, sr);
    });
}

void launch_ldg(hipStream_t s, int walk, int block, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                const int *cs, float4 *gd, int *nlist, int *ncount, float *feat) {
    if (!q.N) return;
    if (feat) {
        if (walk == WALK_LIST) ldg_feat<WALK_LIST>(s, q, sp, b, skey, cs, gd, nlist, ncount, feat);
        else if (walk == WALK_HITS) ldg_feat<WALK_HITS>(s, q, sp, b, skey, cs, gd, nlist, ncount, feat);
        else ldg_feat<WALK_MASK>(s, q, sp, b, skey, cs, gd, nlist, ncount, feat);
        return;
    }
    with_block(block, [&](auto BSc) {
        constexpr int BS = decltype(BSc)::value;
        if (walk == WALK_LIST) {
            if (sp) ldg_walk<WALK_LIST, true, BS>(s, false, q, sp, b, skey, cs, gd, nlist, ncount, nullptr);
            else ldg_walk<WALK_LIST, false, BS>(s, false, q, sp, b, skey, cs, gd, nlist, ncount, nullptr);
        } else if (walk == WALK_HITS) {
            if (sp) ldg_walk<WALK_HITS, true, BS>(s, false, q, sp, b, skey, cs, gd, nlist, ncount, nullptr);
            else ldg_walk<WALK_HITS, false, BS>(s, false, q, sp, b, skey, cs, gd, nlist, ncount, nullptr);
        } else {
            if (sp) ldg_walk<WALK_MASK, true, BS>(s, false, q, sp, b, skey, cs, gd, nlist, ncount, nullptr);
            else ldg_walk<WALK_MASK, false, BS>(s, false, q, sp, b, skey, cs, gd, nlist, ncount, nullptr);
        }
    });
}

// Plain continuity+momentum (mode 0) at the tuned block size; the closure
// modes (1, 2: one scene) at 256 threads.
template <int WALK, bool BATCH>
static void gov_walk(hipStream_t s, int mode, int block, const Prm &q, const Prm *sp, const Buf &b,
                     const uint32_t *skey, const int *cs, const float4 *gd, float4 *acc, const int *nlist,
                     const int *ncount, const float2 *scale, float *terms, const float4 *xref = nullptr) {
#define ZS_GOV(MODE_, DF_, BS_)                                                                  \
    k_governing<WALK, BATCH, MODE_, DF_, BS_><<<pass_grid(q, BS_), BS_, 0, s>>>(              \
        q, sp, b, skey, cs, gd, acc, nlist, ncount, scale, terms, xref)
    if constexpr (!BATCH) {
        const bool a = q.diffusion == ZSPH_DIFF_ANTUONO;
        if (mode == 1) {
            if (a) ZS_GOV(1, ZSPH_DIFF_ANTUONO, ZS_BLOCK);
            else ZS_GOV(1, ZSPH_DIFF_MC, ZS_BLOCK);
        } else if (mode == 2) {
            if (a) ZS_GOV(2, ZSPH_DIFF_ANTUONO, ZS_BLOCK);
            else ZS_GOV(2, ZSPH_DIFF_MC, ZS_BLOCK);
        } else {
            with_block(block, [&](auto BSc) {
                constexpr int BS = decltype(BSc)::value;
                if (a) ZS_GOV(0, ZSPH_DIFF_ANTUONO, BS);
                else ZS_GOV(0, ZSPH_DIFF_MC, BS);
            });
        }
    } else {
        with_block(block, [&](auto BSc) {
            constexpr int BS = decltype(BSc)::value;
            ZS_GOV(0, DIFF_PER_SCENE, BS);
        });
    }
#undef ZS_GOV
}

void launch_governing(hipStream_t s, int walk, int block, const Prm &q, const Prm *sp, const Buf &b,
                      const uint32_t *skey, const int *cs, const float4 *gd, float4 *acc, const int *nlist,
                      const int *ncount, int mode, const float2 *scale, float *terms, bool moving,
                      const float4 *xref) {
    if (!q.N) return;
    if (walk == WALK_SKIN) {  // one scene, plain terms (the caller falls back otherwise)
        gov_walk<WALK_SKIN, false>(s, 0, block, q, nullptr, b, skey, cs, gd, acc, nlist, ncount, scale, terms, xref);
    } else if (walk == WALK_LIST) {
        if (sp) gov_walk<WALK_LIST, true>(s, mode, block, q, sp, b, skey, cs, gd, acc, nlist, ncount, scale, terms);
        else gov_walk<WALK_LIST, false>(s, mode, block, q, sp, b, skey, cs, gd, acc, nlist, ncount, scale, terms);
    } else if (walk == WALK_HITS) {
        if (sp) gov_walk<WALK_HITS, true>(s, mode, block, q, sp, b, skey, cs, gd, acc, nlist, ncount, scale, terms);
        else gov_walk<WALK_HITS, false>(s, mode, block, q, sp, b, skey, cs, gd, acc, nlist, ncount, scale, terms);
    } else {
        if (sp) gov_walk<WALK_MASK, true>(s, mode, block, q, sp, b, skey, cs, gd, acc, nlist, ncount, scale, terms);
        else gov_walk<WALK_MASK, false>(s, mode, block, q, sp, b, skey, cs, gd, acc, nlist, ncount, scale, terms);
    }
    if (moving) {
        if (sp)
            k_governing_dummy<WALK_MASK, true><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, acc);
        else
            k_governing_dummy<WALK_MASK, false><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, acc);
    }
}
void launch_predict(hipStream_t s, const Prm &q, const Prm *sp, const Buf &b, const int *cs, const float4 *acc,
                    const SkinCheck &chk) {
    if (!q.N) return;
    if (sp)
        k_predict<true><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, cs, acc, chk);
    else
        k_predict<false><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, cs, acc, chk);
}
void launch_correct(hipStream_t s, const Prm &q, const Prm *sp, const Buf &b, const int *cs, const float4 *acc,
                    const SkinCheck &chk) {
    if (!q.N) return;
    if (sp)
        k_correct<true><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, cs, acc, chk);
    else
        k_correct<false><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, cs, acc, chk);
}
void launch_nbr_stats(hipStream_t s, int walk, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                      const int *cs, int *count, uint64_t *digest) {
    (void)walk;
    if (!q.N) return;
    if (sp)
        k_nbr_stats<WALK_MASK, true><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, count, digest);
    else
        k_nbr_stats<WALK_MASK, false><<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, sp, b, skey, cs, count, digest);
}

}  // namespace zs
