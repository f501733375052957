// Host C ABI of the FP64 rigid-body module (Test_6DOF equations).
#include "rigid.hpp"

#include "zsph.h"

using namespace zs::rb;

// rigid_info.cpp:3-39
extern "C" int zsph_rigid_mass_properties(const float *pts, int n, double mass, zsph_rigid *r) {
    if (!pts || n <= 0 || !r || !(mass > 0.0)) return ZSPH_E_ARG;
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) c[k] += pts[3 * i + k];
    for (int k = 0; k < 3; ++k) c[k] /= n;
    double I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        const double x = pts[3 * i] - c[0], y = pts[3 * i + 1] - c[1], z = pts[3 * i + 2] - c[2];
        I[0] += y * y + z * z;
        I[4] += x * x + z * z;
        I[8] += x * x + y * y;
        I[1] -= x * y;
        I[2] -= x * z;
        I[5] -= y * z;
    }
    I[3] = I[1];
    I[6] = I[2];
    I[7] = I[5];
    const double s = mass / n;
    for (int k = 0; k < 9; ++k) r->I[k] = I[k] * s;
    for (int k = 0; k < 3; ++k) r->c[k] = c[k];
    r->mass = mass;
    return ZSPH_OK;
}

// acc_alpha.cpp:3-35
extern "C" void zsph_rigid_acc_alpha(zsph_rigid *r) {
    if (r) acc_alpha(*r);
}

// time_int.cpp:3-22 (same order of updates: points with the old c, v, w;
// then q, c, v, w; accelerations reset)
extern "C" void zsph_rigid_euler_step(zsph_rigid *r, float *pts, int n, double dt) {
    if (!r) return;
    const V3 c = v3(r->c), v = v3(r->vel), w = v3(r->omega);
    for (int i = 0; pts && i < n; ++i) {
        const V3 x{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        const V3 xn = x + (v + cross(w, x - c)) * dt;
        pts[3 * i] = (float)xn.x;
        pts[3 * i + 1] = (float)xn.y;
        pts[3 * i + 2] = (float)xn.z;
    }
    const Q q = qget(*r);
    const Q d = qdot(q, w);
    qput(*r, qnormalise(Q{q.a + d.a * dt, q.b + d.b * dt, q.c + d.c * dt, q.d + d.d * dt}));
    put(r->c, c + v * dt);
    put(r->vel, v + v3(r->acc) * dt);
    put(r->omega, w + v3(r->alpha) * dt);
    for (int k = 0; k < 3; ++k) r->acc[k] = r->alpha[k] = 0.0;
}
