// Run diagnostics: fluid count, kinetic energy, momentum, front position and
// probe pressures.  The reference has none (it prints per-step wall time
// only, deltasph/main.cu:91-95); these are what the long-run accuracy gates
// (mass, wave front, probe pressure) are measured with.
#include "common.hpp"
#include "launch.hpp"

namespace zs {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Per-block partials: part[b*DIAG_D + {count, kin, mx, my, mz}], pmax[b*2 + {front, zmax}].
__global__ __launch_bounds__(ZS_BLOCK) void k_diag(Prm q, Buf b, float front_zmax, double *__restrict__ part,
                                                   float *__restrict__ pmax) {
    __shared__ double sd[DIAG_D][ZS_BLOCK / 64];
    __shared__ float sf[2][ZS_BLOCK / 64];
    const int i = blockIdx.x * ZS_BLOCK + threadIdx.x;
    double v[DIAG_D] = {0.0, 0.0, 0.0, 0.0, 0.0};
    float fr = -INFINITY, zm = -INFINITY;
    if (i < q.N && b.type[i] == ZSPH_FLUID) {
        const float4 x = b.posr[i], u = b.velp[i];
        v[0] = 1.0;
        v[1] = 0.5 * (double)q.m * ((double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z);
        v[2] = (double)q.m * u.x;
        v[3] = (double)q.m * u.y;
        v[4] = (double)q.m * u.z;
        if (x.z < front_zmax) fr = x.x;
        zm = x.z;
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < DIAG_D; ++k) v[k] = wave_sum(v[k]);
    fr = wave_max(fr);
    zm = wave_max(zm);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < DIAG_D; ++k) sd[k][w] = v[k];
        sf[0][w] = fr;
        sf[1][w] = zm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < DIAG_D; ++k) {
            double s = 0.0;
            for (int ww = 0; ww < ZS_BLOCK / 64; ++ww) s += sd[k][ww];
            part[(size_t)blockIdx.x * DIAG_D + k] = s;
        }
        float a = -INFINITY, c = -INFINITY;
        for (int ww = 0; ww < ZS_BLOCK / 64; ++ww) {
            a = fmaxf(a, sf[0][ww]);
            c = fmaxf(c, sf[1][ww]);
        }
        pmax[2 * blockIdx.x] = a;
        pmax[2 * blockIdx.x + 1] = c;
    }
}

// One block per probe point: Shepard average of fluid pressure over every
// particle within the cut-off (a full scan, so it does not depend on the
// cell table being current).
__global__ __launch_bounds__(ZS_BLOCK) void k_probe(Prm q, Buf b, const float *__restrict__ pts,
                                                    float *__restrict__ out) {
    __shared__ float s_pw[ZS_BLOCK / 64], s_w[ZS_BLOCK / 64];
    const float px = pts[3 * blockIdx.x], py = pts[3 * blockIdx.x + 1], pz = pts[3 * blockIdx.x + 2];
    float pw = 0.0f, ws = 0.0f;
    for (int j = threadIdx.x; j < q.N; j += ZS_BLOCK) {
        const float4 x = b.posr[j];
        const float r2 = dist2(px - x.x, py - x.y, pz - x.z);
        if (r2 <= q.rc2 && b.type[j] == ZSPH_FLUID) {
            const float qq = sqrtf(r2) * q.inv_h;
            const float t = 1.0f - 0.5f * qq;
            const float W = q.adh * t * t * t * t * (2.0f * qq + 1.0f);
            pw += b.velp[j].w * W;
            ws += W;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        pw += __shfl_xor(pw, o, 64);
        ws += __shfl_xor(ws, o, 64);
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        s_pw[w] = pw;
        s_w[w] = ws;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.0f, c = 0.0f;
        for (int ww = 0; ww < ZS_BLOCK / 64; ++ww) {
            a += s_pw[ww];
            c += s_w[ww];
        }
        out[blockIdx.x] = (c > 1e-8f) ? a / c : 0.0f;
    }
}

// One block per scene of a batch (or the single scene): the scene's fluid
// is the contiguous range of its cells' fluid segments, its removed
// particles sit in its dummy segments.  part[s*SCENE_D + {count, kinetic,
// mx, my, mz, front, zmax, removed}], summed in a fixed order.
template <bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_scene_diag(Prm q0, const Prm *__restrict__ sp, Buf b,
                                                         const int *__restrict__ cs, float front_zmax,
                                                         double *__restrict__ part) {
    __shared__ double red[SCENE_D][ZS_BLOCK];
    const Prm q = BATCH ? sp[blockIdx.x] : q0;
    const int c0 = q.cell_base, c1 = q.cell_base + q.X * q.Y * q.Z;
    double v[SCENE_D] = {0.0, 0.0, 0.0, 0.0, 0.0, -INFINITY, -INFINITY, 0.0};
    for (int i = cs[c0] + threadIdx.x; i < cs[c1]; i += ZS_BLOCK) {
        const float4 x = b.posr[i], u = b.velp[i];
        v[0] += 1.0;
        v[1] += 0.5 * (double)q.m * ((double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z);
        v[2] += (double)q.m * u.x;
        v[3] += (double)q.m * u.y;
        v[4] += (double)q.m * u.z;
        if (x.z < front_zmax) v[5] = fmax(v[5], (double)x.x);
        v[6] = fmax(v[6], (double)x.z);
    }
    for (int i = cs[q.G + 1 + c0] + threadIdx.x; i < cs[q.G + 1 + c1]; i += ZS_BLOCK)
        if (b.type[i] == ZSPH_KILLED) v[7] += 1.0;
    for (int k = 0; k < SCENE_D; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = ZS_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < SCENE_D; ++k) {
                const double a = red[k][threadIdx.x], c = red[k][threadIdx.x + w];
                red[k][threadIdx.x] = (k == 5 || k == 6) ? fmax(a, c) : a + c;
            }
        __syncthreads();
    }
    if (threadIdx.x < SCENE_D) part[(size_t)blockIdx.x * SCENE_D + threadIdx.x] = red[threadIdx.x][0];
}

void launch_scene_diag(hipStream_t s, const Prm &q, const Prm *sp, int nscenes, const Buf &b, const int *cs,
                       float front_zmax, double *part) {
    if (sp)
        k_scene_diag<true><<<nscenes, ZS_BLOCK, 0, s>>>(q, sp, b, cs, front_zmax, part);
    else
        k_scene_diag<false><<<1, ZS_BLOCK, 0, s>>>(q, sp, b, cs, front_zmax, part);
}

// Shepard filter of the fluid fields at particles rows[t] of a buffer in cell
// order: with w_j = V_j W(r_ij; hc) over fluid j within r <= radius (the
// walk's box is qf.k cells, trimmed at qf.trim2),
//   out[t] = (sum w, <rho>, <v> (3), <a> (3), <drho/dt>).
__global__ __launch_bounds__(ZS_BLOCK) void k_filter(Prm qf, Buf b, const uint32_t *__restrict__ skey,
                                                     const int *__restrict__ cs, const float4 *__restrict__ acc,
                                                     const int *__restrict__ rows, int nq, float hc, float r2max,
                                                     float *__restrict__ out) {
    const int t = blockIdx.x * ZS_BLOCK + threadIdx.x;
    if (t >= nq) return;
    const int i = rows[t];
    const float4 xi = b.posr[i];
    const float ih = 1.0f / hc;
    float sw = 0, sr = 0, svx = 0, svy = 0, svz = 0, sax = 0, say = 0, saz = 0, sd = 0;
    const uint32_t cell = (skey[i] & (qf.flag - 1u)) - (uint32_t)qf.cell_base;
    for_each_row<true, false>(qf, xi, cell, cs, [&](int fb, int fe, int, int, int) {
        for (int j = fb; j < fe; ++j) {
            const float4 xj = b.posr[j];
            const float r2 = dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
            if (r2 > r2max) continue;
            const float q = sqrtf(r2) * ih;
            const float u = fmaxf(1.0f - 0.5f * q, 0.0f);
            const float w = u * u * u * u * (2.0f * q + 1.0f) * qf.m / xj.w;
            const float4 vj = b.velp[j], aj = acc[j];
            sw += w;
            sr += w * xj.w;
            svx += w * vj.x; svy += w * vj.y; svz += w * vj.z;
            sax += w * aj.x; say += w * aj.y; saz += w * aj.z;
            sd += w * aj.w;
        }
    });
    const float inv = sw > 0.0f ? 1.0f / sw : 0.0f;
    float *o = out + (size_t)t * 9;
    o[0] = sw;
    o[1] = sr * inv;
    o[2] = svx * inv; o[3] = svy * inv; o[4] = svz * inv;
    o[5] = sax * inv; o[6] = say * inv; o[7] = saz * inv;
    o[8] = sd * inv;
}

void launch_filter(hipStream_t s, const Prm &qf, const Buf &b, const uint32_t *skey, const int *cs, const float4 *acc,
                   const int *rows, int nq, float hc, float radius, float *out) {
    if (nq > 0)
        k_filter<<<(nq + ZS_BLOCK - 1) / ZS_BLOCK, ZS_BLOCK, 0, s>>>(qf, b, skey, cs, acc, rows, nq, hc,
                                                                     radius * radius, out);
}

int diag_blocks(int n) { return (n + ZS_BLOCK - 1) / ZS_BLOCK; }

void launch_diag(hipStream_t s, const Prm &q, const Buf &b, float front_zmax, double *part, float *pmax) {
    if (q.N) k_diag<<<diag_blocks(q.N), ZS_BLOCK, 0, s>>>(q, b, front_zmax, part, pmax);
}

void launch_probe(hipStream_t s, const Prm &q, const Buf &b, const float *pts, int npts, float *out) {
    if (npts > 0) k_probe<<<npts, ZS_BLOCK, 0, s>>>(q, b, pts, out);
}

}  // namespace zs
