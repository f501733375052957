// C entry points over the reference's own Test_6DOF routines (TEST
// INFRASTRUCTURE ONLY).  Compiled, together with the reference sources where
// they lie (Test_6DOF/lib/math.cpp, Test_6DOF/equation/*.cpp), into
// oracle/_ref/libref6dof.so by `make -C oracle ref`; nothing of the reference
// is copied.  The rigid state travels as a flat float array in this order:
//   c(3) F(3) T(3) vel(3) omega(3) acc(3) alpha(3) I(9, row-major) q(4) mass(1)  = 35
#include <vector>

#include "equation.hpp"

namespace {
void unpack(const float *s, rigid_t &r) {
    vector_t *v[7] = {&r.c_pos, &r.F, &r.T, &r.vel, &r.omega, &r.acc, &r.alpha};
    for (int k = 0; k < 7; ++k) *v[k] = {s[3 * k], s[3 * k + 1], s[3 * k + 2]};
    r.I = {s[21], s[22], s[23], s[24], s[25], s[26], s[27], s[28], s[29]};
    r.q = {s[30], s[31], s[32], s[33]};
    r.mass = s[34];
    r.ptc_num = 0;
    r.rigid_id = 0;
}
void pack(const rigid_t &r, float *s) {
    const vector_t *v[7] = {&r.c_pos, &r.F, &r.T, &r.vel, &r.omega, &r.acc, &r.alpha};
    for (int k = 0; k < 7; ++k) { s[3 * k] = v[k]->x; s[3 * k + 1] = v[k]->y; s[3 * k + 2] = v[k]->z; }
    const float I[9] = {r.I.x1, r.I.x2, r.I.x3, r.I.y1, r.I.y2, r.I.y3, r.I.z1, r.I.z2, r.I.z3};
    for (int k = 0; k < 9; ++k) s[21 + k] = I[k];
    s[30] = r.q.a; s[31] = r.q.b; s[32] = r.q.c; s[33] = r.q.d;
    s[34] = r.mass;
}
std::vector<vector_t> points(const float *p, int n) {
    std::vector<vector_t> v(n);
    for (int i = 0; i < n; ++i) v[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    return v;
}
}  // namespace

extern "C" {
void ref6_gravity_center(const float *pts, int n, float *out) {
    const vector_t c = gravity_center(points(pts, n));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
void ref6_inertia_tensor(const float *pts, int n, float *st) {
    rigid_t r;
    unpack(st, r);
    r.I = inertia_tensor(points(pts, n), r);
    pack(r, st);
}
void ref6_cal_acc_alpha(float *st) {
    rigid_t r;
    unpack(st, r);
    cal_acc_alpha(r);
    pack(r, st);
}
// time_int with the reference's fixed dt = 0.001 s (time_int.cpp:5)
void ref6_time_int(float *pts, int n, float *st) {
    rigid_t r;
    unpack(st, r);
    std::vector<vector_t> p = points(pts, n);
    time_int(p, r);
    for (int i = 0; i < n; ++i) { pts[3 * i] = p[i].x; pts[3 * i + 1] = p[i].y; pts[3 * i + 2] = p[i].z; }
    pack(r, st);
}
}
