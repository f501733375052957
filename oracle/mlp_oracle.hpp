// CPU oracle of the closure network: TEST INFRASTRUCTURE ONLY.
//
// The per-particle MLP (24 -> 64 -> 64 -> 64 with SiLU -> 2, outputs mapped
// to s = 1 + out_scale tanh(o)), its training loss and the loss gradient, in
// FP64.  `emulate_bf16` reproduces the device's roundings to BF16 (inputs,
// weights, hidden activations and back-propagated deltas as the MFMA kernels
// store them; bias gradients from the unrounded deltas), so device and oracle
// differ only by accumulation order; with it off the arithmetic is exact
// FP64, which the finite-difference tests use to pin the backward pass.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "zsph.h"

namespace orc {

// Round to the nearest BF16 (ties to even), returned as double.
inline double bf16_round(double x) {
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

struct MlpOracle {
    static constexpr int NF = ZSPH_NF, NH = ZSPH_NH, NO = 2, NT = ZSPH_NT;
    std::vector<double> P;  // w0 b0 w1 b1 w2 b2 w3 b3, weights [in][out]
    zsph_closure aux{};
    bool bf = false;

    explicit MlpOracle(const zsph_closure &net, bool emulate_bf16) : bf(emulate_bf16) {
        aux = net;
        P.assign(&net.w0[0], &net.w0[0] + ZSPH_NPARAM);
    }
    double r(double x) const { return bf ? bf16_round(x) : x; }
    static double sig(double z) { return 1.0 / (1.0 + std::exp(-z)); }

    // offsets
    static constexpr int W0 = 0, B0 = NF * NH, W1 = B0 + NH, B1 = W1 + NH * NH, W2 = B1 + NH, B2 = W2 + NH * NH,
                         W3 = B2 + NH, B3 = W3 + NH * NO;

    struct Act {
        double x[NF], z[3][NH], h[3][NH], o[NO];
        bool keep;
    };

    void forward(const float *feat, Act &a) const {
        a.keep = true;
        for (int f = 0; f < NF; ++f) {
            const float v = (feat[f] - aux.mean[f]) * aux.inv_std[f];  // FP32, as the device computes it
            if (!(v >= aux.lo[f] && v <= aux.hi[f])) a.keep = false;
            a.x[f] = r(v);
        }
        const double *in = a.x;
        int nin = NF;
        const int wo[3] = {W0, W1, W2}, bo[3] = {B0, B1, B2};
        for (int L = 0; L < 3; ++L) {
            for (int j = 0; j < NH; ++j) {
                double s = 0.0;
                for (int i = 0; i < nin; ++i) s += in[i] * r(P[wo[L] + i * NH + j]);
                const double z = s + P[bo[L] + j];
                a.z[L][j] = z;
                a.h[L][j] = r(z * sig(z));
            }
            in = a.h[L];
            nin = NH;
        }
        for (int k = 0; k < NO; ++k) {
            double s = 0.0;
            for (int i = 0; i < NH; ++i) s += a.h[2][i] * r(P[W3 + i * NO + k]);
            a.o[k] = s + P[B3 + k];
        }
    }

    // Scale factors with the range fallback.
    void infer(const float *feat, double *s) const {
        Act a;
        forward(feat, a);
        for (int k = 0; k < NO; ++k) s[k] = a.keep ? 1.0 + aux.out_scale * std::tanh(a.o[k]) : 1.0;
    }

    // Mean loss over n samples and (if grad) its gradient w.r.t. P.
    double loss_grad(const float *X, const float *R, int n, double w_rho, double w_acc, double *grad) const {
        std::vector<double> g(ZSPH_NPARAM, 0.0);
        double L = 0.0;
        for (int p = 0; p < n; ++p) {
            Act a;
            forward(X + (size_t)p * NF, a);
            const float *rec = R + (size_t)p * NT;
            double dO[NO];
            for (int k = 0; k < NO; ++k) {
                const double th = std::tanh(a.o[k]);
                const double s = 1.0 + aux.out_scale * th;
                double dls;
                if (k == 0) {
                    const double res = rec[0] + s * rec[1] - rec[2];
                    L += w_rho * res * res;
                    dls = 2.0 * w_rho * res * rec[1];
                } else {
                    double rr[3], dot = 0.0;
                    for (int c = 0; c < 3; ++c) {
                        rr[c] = rec[3 + c] + s * rec[6 + c] - rec[9 + c];
                        L += w_acc * rr[c] * rr[c];
                        dot += rr[c] * rec[6 + c];
                    }
                    dls = 2.0 * w_acc * dot;
                }
                dO[k] = dls * aux.out_scale * (1.0 - th * th);
            }
            if (!grad) continue;
            // output layer
            double d[NH], dn[NH];
            for (int k = 0; k < NO; ++k) {
                g[B3 + k] += dO[k];
                const double dk = r(dO[k]);
                for (int i = 0; i < NH; ++i) g[W3 + i * NO + k] += a.h[2][i] * dk;
            }
            for (int i = 0; i < NH; ++i) {
                double s = 0.0;
                for (int k = 0; k < NO; ++k) s += r(dO[k]) * r(P[W3 + i * NO + k]);
                d[i] = s;
            }
            const int wo[3] = {W0, W1, W2}, bo[3] = {B0, B1, B2};
            for (int L2 = 2; L2 >= 0; --L2) {
                for (int j = 0; j < NH; ++j) {
                    const double z = a.z[L2][j], sg = sig(z);
                    d[j] *= sg * (1.0 + z * (1.0 - sg));
                    g[bo[L2] + j] += d[j];
                }
                const double *in = L2 ? a.h[L2 - 1] : a.x;
                const int nin = L2 ? NH : NF;
                for (int i = 0; i < nin; ++i)
                    for (int j = 0; j < NH; ++j) g[wo[L2] + i * NH + j] += in[i] * r(d[j]);
                if (L2 == 0) break;
                for (int i = 0; i < NH; ++i) {
                    double s = 0.0;
                    for (int j = 0; j < NH; ++j) s += r(d[j]) * r(P[wo[L2] + i * NH + j]);
                    dn[i] = s;
                }
                std::memcpy(d, dn, sizeof(d));
            }
        }
        if (grad)
            for (int k = 0; k < ZSPH_NPARAM; ++k) grad[k] = g[k] / n;
        return L / n;
    }
};

}  // namespace orc
