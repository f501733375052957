// FP64 rigid-body (6-DOF) mathematics shared by the host C ABI and the
// device coupling kernel.  Equations of the reference's Test_6DOF module:
//   mass properties    Test_6DOF/equation/rigid_info.cpp:3-39
//   accelerations      Test_6DOF/equation/acc_alpha.cpp:3-35
//   explicit Euler     Test_6DOF/equation/time_int.cpp:3-22
//   quaternion algebra Test_6DOF/lib/math.cpp:86-144 (scalar-first Hamilton product, q2t)
// Differences: double precision; a general 3x3 inverse (the reference's
// inv() returns the transpose of the inverse, correct only for symmetric
// matrices, math.cpp:245-271).
#pragma once
#include <cmath>

#include "zsph.h"

#if defined(__HIPCC__)
#define ZS_HD __host__ __device__
#else
#define ZS_HD
#endif

namespace zs {
namespace rb {

struct V3 {
    double x, y, z;
};
ZS_HD inline V3 v3(const double *a) { return {a[0], a[1], a[2]}; }
ZS_HD inline void put(double *a, V3 v) { a[0] = v.x; a[1] = v.y; a[2] = v.z; }
ZS_HD inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
ZS_HD inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
ZS_HD inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
ZS_HD inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct M3 {
    double m[9];  // row-major
};
ZS_HD inline V3 mul(const M3 &A, V3 v) {
    return {A.m[0] * v.x + A.m[1] * v.y + A.m[2] * v.z, A.m[3] * v.x + A.m[4] * v.y + A.m[5] * v.z,
            A.m[6] * v.x + A.m[7] * v.y + A.m[8] * v.z};
}
ZS_HD inline M3 mul(const M3 &A, const M3 &B) {
    M3 C;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C.m[3 * r + c] = A.m[3 * r] * B.m[c] + A.m[3 * r + 1] * B.m[3 + c] + A.m[3 * r + 2] * B.m[6 + c];
    return C;
}
ZS_HD inline M3 transpose(const M3 &A) {
    M3 T;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) T.m[3 * r + c] = A.m[3 * c + r];
    return T;
}
// General inverse by the adjugate; the zero matrix when singular.
ZS_HD inline M3 inverse(const M3 &A) {
    const double *a = A.m;
    M3 R;
    R.m[0] = a[4] * a[8] - a[5] * a[7];
    R.m[1] = a[2] * a[7] - a[1] * a[8];
    R.m[2] = a[1] * a[5] - a[2] * a[4];
    R.m[3] = a[5] * a[6] - a[3] * a[8];
    R.m[4] = a[0] * a[8] - a[2] * a[6];
    R.m[5] = a[2] * a[3] - a[0] * a[5];
    R.m[6] = a[3] * a[7] - a[4] * a[6];
    R.m[7] = a[1] * a[6] - a[0] * a[7];
    R.m[8] = a[0] * a[4] - a[1] * a[3];
    const double det = a[0] * R.m[0] + a[1] * R.m[3] + a[2] * R.m[6];
    const double id = (det != 0.0) ? 1.0 / det : 0.0;
    for (int k = 0; k < 9; ++k) R.m[k] *= id;
    return R;
}

// Quaternions, scalar first (a, b, c, d) = a + bi + cj + dk.
struct Q {
    double a, b, c, d;
};
ZS_HD inline Q qmul(Q p, Q q) {
    return {p.a * q.a - p.b * q.b - p.c * q.c - p.d * q.d, p.b * q.a + p.a * q.b - p.d * q.c + p.c * q.d,
            p.c * q.a + p.d * q.b + p.a * q.c - p.b * q.d, p.d * q.a - p.c * q.b + p.b * q.c + p.a * q.d};
}
ZS_HD inline Q qnormalise(Q q) {
    const double s = 1.0 / sqrt(q.a * q.a + q.b * q.b + q.c * q.c + q.d * q.d);
    return {q.a * s, q.b * s, q.c * s, q.d * s};
}
// Rotation matrix of a unit quaternion (math.cpp:128-144).
ZS_HD inline M3 rotation(Q q) {
    M3 R;
    R.m[0] = 1 - 2 * q.c * q.c - 2 * q.d * q.d;
    R.m[1] = 2 * q.b * q.c - 2 * q.a * q.d;
    R.m[2] = 2 * q.a * q.c + 2 * q.b * q.d;
    R.m[3] = 2 * q.b * q.c + 2 * q.a * q.d;
    R.m[4] = 1 - 2 * q.b * q.b - 2 * q.d * q.d;
    R.m[5] = 2 * q.c * q.d - 2 * q.a * q.b;
    R.m[6] = 2 * q.b * q.d - 2 * q.a * q.c;
    R.m[7] = 2 * q.a * q.b + 2 * q.c * q.d;
    R.m[8] = 1 - 2 * q.b * q.b - 2 * q.c * q.c;
    return R;
}
ZS_HD inline Q qget(const zsph_rigid &r) { return {r.q[0], r.q[1], r.q[2], r.q[3]}; }
ZS_HD inline void qput(zsph_rigid &r, Q q) { r.q[0] = q.a; r.q[1] = q.b; r.q[2] = q.c; r.q[3] = q.d; }

// a = F / M;  alpha = R I^-1 R^T (T - w x (R I R^T w))  (acc_alpha.cpp:5-31)
ZS_HD inline void acc_alpha(zsph_rigid &r) {
    put(r.acc, v3(r.F) * (1.0 / r.mass));
    M3 I;
    for (int k = 0; k < 9; ++k) I.m[k] = r.I[k];
    const M3 R = rotation(qget(r));
    const M3 Rt = transpose(R);
    const M3 Iw = mul(mul(R, I), Rt);
    const M3 Iwinv = mul(mul(R, inverse(I)), Rt);
    const V3 w = v3(r.omega);
    put(r.alpha, mul(Iwinv, v3(r.T) - cross(w, mul(Iw, w))));
}

// Time derivative of the orientation: dq/dt = 1/2 (0, w) (x) q  (time_int.cpp:16)
ZS_HD inline Q qdot(Q q, V3 w) {
    const Q wq = qmul(Q{0.0, w.x, w.y, w.z}, q);
    return {0.5 * wq.a, 0.5 * wq.b, 0.5 * wq.c, 0.5 * wq.d};
}

// Advance (c, v, q, w) from the saved state s0 by dt using the rates of `rates`
// (v, w, a, alpha): the building block of the midpoint predictor-corrector
// the SPH coupling uses.  q is renormalised.
ZS_HD inline void advance(const zsph_rigid &s0, const zsph_rigid &rates, double dt, zsph_rigid &out) {
    put(out.c, v3(s0.c) + v3(rates.vel) * dt);
    put(out.vel, v3(s0.vel) + v3(rates.acc) * dt);
    put(out.omega, v3(s0.omega) + v3(rates.alpha) * dt);
    const Q d = qdot(qget(rates), v3(rates.omega));
    const Q q0 = qget(s0);
    qput(out, qnormalise(Q{q0.a + d.a * dt, q0.b + d.b * dt, q0.c + d.c * dt, q0.d + d.d * dt}));
}

}  // namespace rb
}  // namespace zs
