// Moving dummies for gfx950: the floating body (rigid 6-DOF coupling) and the
// piston wavemaker.
//   rigid equations  <- Test_6DOF/equation/acc_alpha.cpp:3-35, time_int.cpp:3-22
//                       (FP64, shared with the host module in host/rigid.hpp)
//   wall acceleration in the dummy pressure <- equation/delta_ns.cu:83 (a_w,
//                       zero there; prescribed here by the body / piston motion)
// Per half step, after the continuity+momentum pass:
//   k_body_reduce   fluid force and torque on the body, FP64 per-block partials
//                   over the dummy segment (fixed block count; each wave sums
//                   its lanes with a wave64 shuffle butterfly, then the block
//                   sums its waves in order: a fixed order throughout);
//   k_body_update   one wave: sums the partials in block order, adds M g,
//                   accelerations, midpoint predictor / corrector of the FP64
//                   state, device time;
//   k_kinematics    places every body / piston dummy at the new time and
//                   stores its velocity and wall acceleration.
// Dummy slots: tposr.xyz holds the body-frame offset (body) or the rest
// position (piston); tvelp.xyz holds the wall acceleration a_w read by the
// dummy-pressure pass.
#include "common.hpp"
#include "launch.hpp"
#include "../host/rigid.hpp"

namespace zs {

using rb::V3;

__device__ __forceinline__ V3 f3(float4 a) { return {(double)a.x, (double)a.y, (double)a.z}; }

__global__ __launch_bounds__(ZS_BLOCK) void k_body_reduce(Prm q, Buf b, const int *__restrict__ cs,
                                                          const float4 *__restrict__ acc,
                                                          const DevBody *__restrict__ body, double *__restrict__ part) {
    stage_stamp(q, ZSPH_ST_GOVERNING, -1);  // the integrator follows the continuity+momentum pass
    __shared__ double red[BODY_D][ZS_BLOCK / 64];
    const int nf = cs[q.G];
    const V3 c = rb::v3(body->s.c);
    const double m = (double)q.m;
    double s[BODY_D] = {0, 0, 0, 0, 0, 0};
    for (int i = nf + blockIdx.x * ZS_BLOCK + threadIdx.x; i < q.N; i += BODY_BLOCKS * ZS_BLOCK) {
        if (b.type[i] != ZSPH_BODY) continue;
        const V3 f = f3(acc[i]) * m;
        const V3 t = rb::cross(f3(b.posr[i]) - c, f);
        s[0] += f.x; s[1] += f.y; s[2] += f.z;
        s[3] += t.x; s[4] += t.y; s[5] += t.z;
    }
    // wave64 butterfly in FP64 (a fixed pattern: bitwise reproducible), then the
    // block's waves in order
#pragma unroll
    for (int k = 0; k < BODY_D; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_xor(s[k], o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < BODY_D; ++k) red[k][w] = s[k];
    __syncthreads();
    if (threadIdx.x < BODY_D) {
        double t = 0.0;
        for (int ww = 0; ww < ZS_BLOCK / 64; ++ww) t += red[threadIdx.x][ww];
        part[blockIdx.x * BODY_D + threadIdx.x] = t;
    }
}

// One wave.  Lane k < 6 sums partial k over the blocks in block order; lane 0
// then integrates.  Predictor: save (state, t), advance by dt/2 with the
// rates of the state.  Corrector: advance the saved state by dt with the
// rates of the mid state.
__global__ __launch_bounds__(64) void k_body_update(Prm q, DevBody *__restrict__ body,
                                                    const double *__restrict__ part, int predictor) {
    __shared__ double ft[BODY_D];
    DevBody &B = *body;
    if (B.has_body && threadIdx.x < BODY_D) {
        double s = 0.0;
        for (int k = 0; k < BODY_BLOCKS; ++k) s += part[k * BODY_D + threadIdx.x];
        ft[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (predictor) B.t0 = B.t;
    const double dt = predictor ? (double)q.half_dt : (double)q.dt;
    B.t = B.t0 + dt;
    if (!B.has_body) return;
    zsph_rigid &s = B.s;
    for (int k = 0; k < 3; ++k) {
        s.F[k] = ft[k];
        s.T[k] = ft[3 + k];
    }
    // accelerations from the fluid force plus the body's weight
    zsph_rigid w = s;
    w.F[2] += s.mass * (double)q.g;
    rb::acc_alpha(w);
    for (int k = 0; k < 3; ++k) {
        s.acc[k] = w.acc[k];
        s.alpha[k] = w.alpha[k];
    }
    if (predictor) B.s0 = s;
    const zsph_rigid rates = s;
    rb::advance(B.s0, rates, dt, s);
}

// Body dummies: r = R r0, x = c + r, v = V + w x r, a_w = A + alpha x r + w x (w x r).
// Piston dummies: x = x0 + d(t) e, v = d'(t) e, a_w = d''(t) e along the axis e.
__global__ __launch_bounds__(ZS_BLOCK) void k_kinematics(Prm q, Buf b, const int *__restrict__ cs,
                                                         const DevBody *__restrict__ body, SkinCheck chk) {
    const int i = cs[q.G] + xcd_block(blocks_of(q.N - cs[q.G])) * ZS_BLOCK + threadIdx.x;
    if (i >= q.N) return;
    const int ti = b.type[i];
    if (ti != ZSPH_BODY && ti != ZSPH_PISTON) return;
    const DevBody &B = *body;
    float4 x = b.posr[i], v = b.velp[i], a = b.tvelp[i];
    const float4 t = b.tposr[i];
    if (ti == ZSPH_BODY) {
        if (!B.has_body) return;
        const zsph_rigid &s = B.s;
        const V3 r = rb::mul(rb::rotation(rb::qget(s)), f3(t));
        const V3 w = rb::v3(s.omega);
        const V3 xp = rb::v3(s.c) + r;
        const V3 vp = rb::v3(s.vel) + rb::cross(w, r);
        const V3 ap = rb::v3(s.acc) + rb::cross(rb::v3(s.alpha), r) + rb::cross(w, rb::cross(w, r));
        x.x = (float)xp.x; x.y = (float)xp.y; x.z = (float)xp.z;
        v.x = (float)vp.x; v.y = (float)vp.y; v.z = (float)vp.z;
        a.x = (float)ap.x; a.y = (float)ap.y; a.z = (float)ap.z;
    } else {
        if (!B.has_piston) return;
        const double ph = B.pw * B.t + B.phase;
        const float d = (float)(B.amp * (sin(ph) - sin(B.phase)));
        const float dv = (float)(B.amp * B.pw * cos(ph));
        const float da = (float)(-B.amp * B.pw * B.pw * sin(ph));
        x.x = t.x; x.y = t.y; x.z = t.z;
        v.x = v.y = v.z = 0.0f;
        a.x = a.y = a.z = 0.0f;
        if (B.axis == 0) { x.x += d; v.x = dv; a.x = da; }
        else if (B.axis == 1) { x.y += d; v.y = dv; a.y = da; }
        else { x.z += d; v.z = dv; a.z = da; }
    }
    b.posr[i] = x;
    b.velp[i] = v;
    b.tvelp[i] = a;
    if (chk.xref && skin_moved(q, chk, i, x)) *chk.flag = 1;  // moving dummies are few: no wave vote
}

static inline int nblocks(int n) { return (n + ZS_BLOCK - 1) / ZS_BLOCK; }

void launch_body_reduce(hipStream_t s, const Prm &q, const Buf &b, const int *cs, const float4 *acc,
                        const DevBody *body, double *part) {
    k_body_reduce<<<BODY_BLOCKS, ZS_BLOCK, 0, s>>>(q, b, cs, acc, body, part);
}
void launch_body_update(hipStream_t s, const Prm &q, DevBody *body, const double *part, bool predictor) {
    k_body_update<<<1, 64, 0, s>>>(q, body, part, predictor ? 1 : 0);
}
void launch_kinematics(hipStream_t s, const Prm &q, const Buf &b, const int *cs, const DevBody *body,
                       const SkinCheck &chk) {
    if (q.N) k_kinematics<<<nblocks(q.N), ZS_BLOCK, 0, s>>>(q, b, cs, body, chk);
}

}  // namespace zs
