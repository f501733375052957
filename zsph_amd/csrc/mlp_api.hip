// C ABI of the closure network object (zsph_mlp): a device-resident network
// (FP32 master parameters, BF16 packed image, Adam state) and a training set,
// driven by the MFMA kernels in kernels/mlp.hip.  Host code only.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kernels/common.hpp"
#include "kernels/launch.hpp"
#include "zsph.h"

static_assert(zs::MLP_NPARAM == ZSPH_NPARAM, "parameter count");
static_assert(offsetof(zsph_closure, b3) + 2 * sizeof(float) - offsetof(zsph_closure, w0) ==
                  ZSPH_NPARAM * sizeof(float),
              "zsph_closure parameters are contiguous");

struct zsph_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *P = nullptr, *m = nullptr, *v = nullptr;  // master parameters, Adam moments
    float *grad = nullptr;                           // ZSPH_NPARAM + 1 (loss)
    float *part = nullptr;                           // per-block partials
    void *pack = nullptr;                            // BF16 image
    zs::MlpAux *aux = nullptr;
    zs::MlpAux haux{};
    float *feat = nullptr, *rec = nullptr;
    int n = 0;
    int step = 0;  // Adam step count
    std::vector<void *> allocs;
    std::string err;
};

namespace {

int fail(zsph_mlp *m, int code, const char *fmt, ...) {
    if (m) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        m->err = buf;
    }
    return code;
}

#define MLPCHK(m, call)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail(m, ZSPH_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int alloc(zsph_mlp *m, T **p, size_t count) {
    void *d = nullptr;
    const size_t bytes = (count ? count : 1) * sizeof(T);
    if (hipMalloc(&d, bytes) != hipSuccess) return fail(m, ZSPH_E_NOMEM, "hipMalloc(%zu) failed", bytes);
    m->allocs.push_back(d);
    *p = static_cast<T *>(d);
    MLPCHK(m, hipMemsetAsync(d, 0, bytes, m->stream));
    return ZSPH_OK;
}

int upload(zsph_mlp *m, void *dst, const void *src, size_t bytes) {
    MLPCHK(m, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, m->stream));
    MLPCHK(m, hipStreamSynchronize(m->stream));
    return ZSPH_OK;
}

int download(zsph_mlp *m, void *dst, const void *src, size_t bytes) {
    MLPCHK(m, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, m->stream));
    MLPCHK(m, hipStreamSynchronize(m->stream));
    return ZSPH_OK;
}

int range_ok(zsph_mlp *m, int first, int count) {
    if (first < 0 || count < 1 || (long long)first + count > m->n)
        return fail(m, ZSPH_E_ARG, "sample range [%d, %d) outside the training set of %d", first, first + count, m->n);
    return ZSPH_OK;
}

}  // namespace

extern "C" {

int zsph_mlp_create(int device, zsph_mlp **out) {
    if (!out) return ZSPH_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ZSPH_E_NODEV;
    if (device >= ndev) return ZSPH_E_NODEV;
    if (device >= 0 && hipSetDevice(device) != hipSuccess) return ZSPH_E_HIP;
    zsph_mlp *m = new zsph_mlp();
    (void)hipGetDevice(&m->device);
    int r = ZSPH_OK;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) r = ZSPH_E_HIP;
    const size_t np = ZSPH_NPARAM;
    uint16_t *pack = nullptr;
    if (!r) r = alloc(m, &m->P, np);
    if (!r) r = alloc(m, &m->m, np);
    if (!r) r = alloc(m, &m->v, np);
    if (!r) r = alloc(m, &m->grad, np + 1);
    if (!r) r = alloc(m, &m->part, (size_t)256 * (np + 1));
    if (!r) r = alloc(m, &pack, zs::mlp_pack_elems());
    if (!r) r = alloc(m, &m->aux, 1);
    if (!r && zs::mlp_prepare() != hipSuccess) r = fail(m, ZSPH_E_HIP, "MLP kernels: LDS size attribute refused");
    m->pack = pack;
    if (r) {
        zsph_mlp_destroy(m);
        return r;
    }
    *out = m;
    return ZSPH_OK;
}

void zsph_mlp_destroy(zsph_mlp *m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *d : m->allocs) (void)hipFree(d);
    if (m->feat) (void)hipFree(m->feat);
    if (m->rec) (void)hipFree(m->rec);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char *zsph_mlp_last_error(zsph_mlp *m) { return m ? m->err.c_str() : "null network"; }

int zsph_mlp_set(zsph_mlp *m, const zsph_closure *net) {
    if (!m || !net) return ZSPH_E_ARG;
    std::memcpy(m->haux.mean, net->mean, sizeof(net->mean));
    std::memcpy(m->haux.inv_std, net->inv_std, sizeof(net->inv_std));
    std::memcpy(m->haux.lo, net->lo, sizeof(net->lo));
    std::memcpy(m->haux.hi, net->hi, sizeof(net->hi));
    m->haux.out_scale = net->out_scale;
    int r;
    if ((r = upload(m, m->P, net->w0, ZSPH_NPARAM * sizeof(float)))) return r;
    if ((r = upload(m, m->aux, &m->haux, sizeof(m->haux)))) return r;
    MLPCHK(m, hipMemsetAsync(m->m, 0, ZSPH_NPARAM * sizeof(float), m->stream));
    MLPCHK(m, hipMemsetAsync(m->v, 0, ZSPH_NPARAM * sizeof(float), m->stream));
    zs::launch_mlp_pack(m->stream, m->P, m->pack);
    MLPCHK(m, hipGetLastError());
    MLPCHK(m, hipStreamSynchronize(m->stream));
    m->step = 0;
    return ZSPH_OK;
}

int zsph_mlp_get(zsph_mlp *m, zsph_closure *net) {
    if (!m || !net) return ZSPH_E_ARG;
    int r = download(m, net->w0, m->P, ZSPH_NPARAM * sizeof(float));
    if (r) return r;
    std::memcpy(net->mean, m->haux.mean, sizeof(net->mean));
    std::memcpy(net->inv_std, m->haux.inv_std, sizeof(net->inv_std));
    std::memcpy(net->lo, m->haux.lo, sizeof(net->lo));
    std::memcpy(net->hi, m->haux.hi, sizeof(net->hi));
    net->out_scale = m->haux.out_scale;
    return ZSPH_OK;
}

int zsph_mlp_data(zsph_mlp *m, const float *x, const float *rec, int n, float w_rho, float w_acc) {
    if (!m || !x || !rec || n < 1) return ZSPH_E_ARG;
    if (m->feat) (void)hipFree(m->feat);
    if (m->rec) (void)hipFree(m->rec);
    m->feat = m->rec = nullptr;
    m->n = 0;
    MLPCHK(m, hipMalloc(&m->feat, (size_t)n * ZSPH_NF * sizeof(float)));
    MLPCHK(m, hipMalloc(&m->rec, (size_t)n * ZSPH_NT * sizeof(float)));
    int r;
    if ((r = upload(m, m->feat, x, (size_t)n * ZSPH_NF * sizeof(float)))) return r;
    if ((r = upload(m, m->rec, rec, (size_t)n * ZSPH_NT * sizeof(float)))) return r;
    m->haux.w_rho = w_rho;
    m->haux.w_acc = w_acc;
    if ((r = upload(m, m->aux, &m->haux, sizeof(m->haux)))) return r;
    m->n = n;
    return ZSPH_OK;
}

int zsph_mlp_eval(zsph_mlp *m, int first, int count, float *loss, float *grad) {
    if (!m) return ZSPH_E_ARG;
    int r = range_ok(m, first, count);
    if (r) return r;
    zs::launch_mlp_grad(m->stream, m->P, m->pack, m->aux, m->feat + (size_t)first * ZSPH_NF,
                        m->rec + (size_t)first * ZSPH_NT, count, m->part, m->grad);
    MLPCHK(m, hipGetLastError());
    if (grad && (r = download(m, grad, m->grad, ZSPH_NPARAM * sizeof(float)))) return r;
    if (loss && (r = download(m, loss, m->grad + ZSPH_NPARAM, sizeof(float)))) return r;
    MLPCHK(m, hipStreamSynchronize(m->stream));
    return ZSPH_OK;
}

int zsph_mlp_adam(zsph_mlp *m, int first, int count, float lr, float beta1, float beta2, float eps, float *loss) {
    if (!m || !(lr > 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return fail(m, ZSPH_E_ARG, "zsph_mlp_adam: bad hyper-parameters");
    int r = range_ok(m, first, count);
    if (r) return r;
    zs::launch_mlp_grad(m->stream, m->P, m->pack, m->aux, m->feat + (size_t)first * ZSPH_NF,
                        m->rec + (size_t)first * ZSPH_NT, count, m->part, m->grad);
    zs::launch_mlp_adam(m->stream, m->P, m->m, m->v, m->grad, lr, beta1, beta2, eps, ++m->step);
    zs::launch_mlp_pack(m->stream, m->P, m->pack);
    MLPCHK(m, hipGetLastError());
    if (loss && (r = download(m, loss, m->grad + ZSPH_NPARAM, sizeof(float)))) return r;
    return ZSPH_OK;
}

int zsph_mlp_infer(zsph_mlp *m, const float *x, int n, float *s_out) {
    if (!m || !x || !s_out || n < 0) return ZSPH_E_ARG;
    if (n == 0) return ZSPH_OK;
    float *dx = nullptr, *ds = nullptr;
    MLPCHK(m, hipMalloc(&dx, (size_t)n * ZSPH_NF * sizeof(float)));
    if (hipMalloc(&ds, (size_t)n * 2 * sizeof(float)) != hipSuccess) {
        (void)hipFree(dx);
        return fail(m, ZSPH_E_NOMEM, "zsph_mlp_infer: out of memory");
    }
    int r = upload(m, dx, x, (size_t)n * ZSPH_NF * sizeof(float));
    if (!r) {
        zs::launch_mlp_infer(m->stream, m->P, m->pack, m->aux, dx, n, nullptr, reinterpret_cast<float2 *>(ds),
                             zs::mlp_infer_blocks(n));
        if (hipGetLastError() != hipSuccess) r = fail(m, ZSPH_E_HIP, "zsph_mlp_infer: launch failed");
    }
    if (!r) r = download(m, s_out, ds, (size_t)n * 2 * sizeof(float));
    (void)hipFree(dx);
    (void)hipFree(ds);
    return r;
}

}  // extern "C"
