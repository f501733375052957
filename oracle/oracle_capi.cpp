// C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY): the same
// call sequence as libzsph's C ABI, so tests drive both identically.
// fp64 = 1 builds the double-precision parity oracle, 0 the FP32 baseline.
#include <omp.h>
#include <cstring>
#include <memory>
#include <vector>

#include "zsph.h"
#include "mlp_oracle.hpp"
#include "zsph_oracle.hpp"

namespace {

struct Handle {
    std::unique_ptr<orc::Oracle<double>> d;
    std::unique_ptr<orc::Oracle<float>> f;
};

template <class R, class Fn>
int dispatch(Handle *h, Fn &&fn) {
    if (h->d) return fn(*h->d);
    return fn(*h->f);
}

template <class R>
int upload(orc::Oracle<R> &o, int bi, const zsph_host_state *st) {
    auto &b = o.buf[bi];
    for (int i = 0; i < o.N; ++i) {
        for (int c = 0; c < 3; ++c) {
            b.posr[4 * i + c] = st->pos[3 * i + c];
            b.velp[4 * i + c] = st->vel ? st->vel[3 * i + c] : 0.0f;
        }
        b.posr[4 * i + 3] = st->rhop ? st->rhop[2 * i] : o.p.rho0;
        b.velp[4 * i + 3] = st->rhop ? st->rhop[2 * i + 1] : 0.0f;
        b.type[i] = st->type[i];
        b.table[i] = st->table ? st->table[i] : 0;
    }
    b.tposr = b.posr;
    b.tvelp = b.velp;
    if (st->acc_drhodt)
        for (size_t t = 0; t < o.acc.size(); ++t) o.acc[t] = st->acc_drhodt[t];
    o.hashed = o.sorted = false;
    return ZSPH_OK;
}

template <class R>
int download(orc::Oracle<R> &o, int bi, zsph_host_state *st, int fields) {
    auto &b = o.buf[bi];
    for (int i = 0; i < o.N; ++i) {
        for (int c = 0; c < 3; ++c) {
            if ((fields & ZSPH_F_POS) && st->pos) st->pos[3 * i + c] = (float)b.posr[4 * i + c];
            if ((fields & ZSPH_F_VEL) && st->vel) st->vel[3 * i + c] = (float)b.velp[4 * i + c];
        }
        if ((fields & ZSPH_F_RHOP) && st->rhop) {
            st->rhop[2 * i] = (float)b.posr[4 * i + 3];
            st->rhop[2 * i + 1] = (float)b.velp[4 * i + 3];
        }
        if ((fields & ZSPH_F_ACC) && st->acc_drhodt)
            for (int c = 0; c < 4; ++c) st->acc_drhodt[4 * i + c] = (float)o.acc[4 * i + c];
        if ((fields & ZSPH_F_TYPE) && st->type) st->type[i] = b.type[i];
        if ((fields & ZSPH_F_TABLE) && st->table) st->table[i] = b.table[i];
        if ((fields & ZSPH_F_HASH) && st->hash) st->hash[i] = (int)(o.key[i] & (o.flag - 1u));
    }
    return ZSPH_OK;
}

// Array access in double (which: ZSPH_A_*; ints are widened).
template <class R>
int array_get(orc::Oracle<R> &o, int which, int bi, double *out, size_t count) {
    auto &b = o.buf[bi];
    auto copy = [&](const auto &v) {
        if (v.size() != count) return ZSPH_E_ARG;
        for (size_t t = 0; t < count; ++t) out[t] = (double)v[t];
        return ZSPH_OK;
    };
    switch (which) {
    case ZSPH_A_POSR: return copy(b.posr);
    case ZSPH_A_VELP: return copy(b.velp);
    case ZSPH_A_TPOSR: return copy(b.tposr);
    case ZSPH_A_TVELP: return copy(b.tvelp);
    case ZSPH_A_TYPE: return copy(b.type);
    case ZSPH_A_TABLE: return copy(b.table);
    case ZSPH_A_KEY: return copy(o.key);
    case ZSPH_A_INDEX: return copy(o.idx);
    case ZSPH_A_CELL_START: return copy(o.cs);
    case ZSPH_A_ACC: return copy(o.acc);
    case ZSPH_A_GD: return copy(o.gd);
    case 100: return copy(o.acc_scale);
    case 101: return copy(o.gd_scale);
    case 102: return copy(o.feat);
    case 103: return copy(o.terms);
    case 104: return copy(o.scale);
    default: return ZSPH_E_ARG;
    }
}

template <class R>
int array_set(orc::Oracle<R> &o, int which, int bi, const double *in, size_t count) {
    auto &b = o.buf[bi];
    auto put = [&](auto &v) {
        if (v.size() != count) return ZSPH_E_ARG;
        using T = typename std::decay_t<decltype(v)>::value_type;
        for (size_t t = 0; t < count; ++t) v[t] = (T)in[t];
        return ZSPH_OK;
    };
    switch (which) {
    case ZSPH_A_POSR: return put(b.posr);
    case ZSPH_A_VELP: return put(b.velp);
    case ZSPH_A_TPOSR: return put(b.tposr);
    case ZSPH_A_TVELP: return put(b.tvelp);
    case ZSPH_A_TYPE: return put(b.type);
    case ZSPH_A_TABLE: return put(b.table);
    case ZSPH_A_ACC: return put(o.acc);
    case ZSPH_A_GD: return put(o.gd);
    default: return ZSPH_E_ARG;
    }
}

}  // namespace

extern "C" {

void *orc_create(const zsph_params *p, int fp64) {
    auto *h = new Handle();
    if (fp64) h->d.reset(new orc::Oracle<double>(*p));
    else h->f.reset(new orc::Oracle<float>(*p));
    return h;
}
void orc_destroy(void *h) { delete static_cast<Handle *>(h); }

#define ORC_CALL(expr) \
    dispatch<double>(static_cast<Handle *>(h), [&](auto &o) { expr; return ZSPH_OK; })

int orc_upload(void *h, int b, const zsph_host_state *st) {
    return dispatch<double>(static_cast<Handle *>(h), [&](auto &o) { return upload(o, b, st); });
}
int orc_download(void *h, int b, zsph_host_state *st, int fields) {
    return dispatch<double>(static_cast<Handle *>(h), [&](auto &o) { return download(o, b, st, fields); });
}
int orc_hash(void *h, int b) { return ORC_CALL(o.hash(b)); }
int orc_sort_index(void *h) { return ORC_CALL(o.sort_index()); }
int orc_sort_data(void *h, int from, int to) { return ORC_CALL(o.sort_data(from, to)); }
int orc_boundary(void *h, int b) { return ORC_CALL(o.boundary(b)); }
int orc_wall(void *h, int b) { return ORC_CALL(o.wall(b)); }
int orc_ldg(void *h, int b) { return ORC_CALL(o.ldg(b)); }
int orc_governing(void *h, int b) { return ORC_CALL(o.governing(b)); }
int orc_prediction(void *h, int b) { return ORC_CALL(o.prediction(b)); }
int orc_correction(void *h, int b) { return ORC_CALL(o.correction(b)); }
int orc_step(void *h, int n) { return ORC_CALL(o.step(n)); }
// OpenMP threads of the calling thread's later parallel regions (the CPU
// baseline's one-thread run); returns the previous setting.
int orc_set_threads(int n) {
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}
int orc_killed(void *h, long long *k) { return ORC_CALL(*k = o.killed); }
int orc_array_get(void *h, int which, int b, double *out, size_t count) {
    return dispatch<double>(static_cast<Handle *>(h), [&](auto &o) { return array_get(o, which, b, out, count); });
}
int orc_array_set(void *h, int which, int b, const double *in, size_t count) {
    return dispatch<double>(static_cast<Handle *>(h), [&](auto &o) { return array_set(o, which, b, in, count); });
}
int orc_neighbor_stats(void *h, int b, int brute, int *count, uint64_t *digest) {
    return ORC_CALL(o.nbr_stats(b, brute != 0, count, digest));
}

int orc_diagnostics(void *h, int b, float front_zmax, zsph_diag *d) { return ORC_CALL(o.diagnostics(b, front_zmax, d)); }
int orc_probe_pressure(void *h, int b, const float *pts, int np, double *out) {
    return ORC_CALL(o.probes(b, pts, np, out));
}

// Moving dummies (floating body, piston) on buffer 0; time in seconds.
int orc_set_body(void *h, const zsph_rigid *r) {
    return dispatch<double>(static_cast<Handle *>(h), [&](auto &o) { return o.set_body(*r); });
}
int orc_get_body(void *h, zsph_rigid *r) { return ORC_CALL(*r = o.bs); }
int orc_set_piston(void *h, double amp, double freq, double phase, int axis) {
    return dispatch<double>(static_cast<Handle *>(h), [&](auto &o) { return o.set_piston(amp, freq, phase, axis); });
}
int orc_set_time(void *h, double t) { return ORC_CALL(o.t = o.t0 = t); }
int orc_get_time(void *h, double *t) { return ORC_CALL(*t = o.t); }
int orc_body_reduce(void *h, int b, double *F, double *T) { return ORC_CALL(o.body_force(b, F, T)); }

// The oracle's rigid-body accelerations (gravity g along z) and midpoint
// building block, for tests against the reference's Test_6DOF routines.
void orc_rigid_acc_alpha(zsph_rigid *r, double g) { orc::Oracle<double>::body_accelerations(*r, g); }
void orc_rigid_advance(const zsph_rigid *s0, const zsph_rigid *rates, double dt, zsph_rigid *out) {
    orc::Oracle<double>::body_advance(*s0, *rates, dt, *out);
}

// Closure in the oracle's step (NULL: plain); features of a buffer after ldg.
int orc_set_closure(void *h, const zsph_closure *net, int emulate_bf16) {
    return ORC_CALL(o.closure_on = net != nullptr; if (net) o.closure = *net; o.closure_bf16 = emulate_bf16 != 0);
}
int orc_features(void *h, int b) { return ORC_CALL(o.features(b)); }

// Closure network (emulate_bf16: the device's BF16 roundings).
double orc_mlp_loss_grad(const zsph_closure *net, int emulate_bf16, const float *x, const float *rec, int n,
                         double w_rho, double w_acc, double *grad) {
    return orc::MlpOracle(*net, emulate_bf16 != 0).loss_grad(x, rec, n, w_rho, w_acc, grad);
}
void orc_mlp_infer(const zsph_closure *net, int emulate_bf16, const float *x, int n, double *s) {
    const orc::MlpOracle m(*net, emulate_bf16 != 0);
    for (int p = 0; p < n; ++p) m.infer(x + (size_t)p * ZSPH_NF, s + 2 * (size_t)p);
}
double orc_bf16_round(double x) { return orc::bf16_round(x); }

// Scalar kernels of the oracle, exposed for closed-form tests (a particle-
// free oracle carries only the constants).
static orc::Oracle<double> lite(const zsph_params *p) {
    zsph_params q = *p;
    q.ptc_num = 0;
    q.grid_xdim = q.grid_ydim = q.grid_zdim = 1;
    return orc::Oracle<double>(q);
}
double orc_kernel_W(const zsph_params *p, double r) { return lite(p).kernel_W(r); }
double orc_kernel_F(const zsph_params *p, double r) { return lite(p).kernel_F(r); }
double orc_eos_p(const zsph_params *p, double rho) { return lite(p).eos_p(rho); }
double orc_eos_rho(const zsph_params *p, double pr) { return lite(p).eos_rho(pr); }
int orc_cell(float x, float o, float size, int dim) {
    int c = -1;
    return orc::cell_f32(x, o, size, dim, &c) ? c : -1;
}

}  // extern "C"
