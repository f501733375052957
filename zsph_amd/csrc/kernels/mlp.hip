// Per-particle closure network on gfx950 MFMA: 24 features -> 64 -> 64 -> 64
// (SiLU) -> 2 outputs, BF16 operands with FP32 accumulation
// (v_mfma_f32_16x16x32_bf16).  A wave processes tiles of 32 samples:
// activations live in LDS as BF16 in both orientations ([sample][neuron]
// feeds the next layer, [neuron][sample] feeds the weight gradient), weights
// are copied once per block from a pre-packed BF16 image.  Training runs one
// wave per block; inference runs INFER_WAVES waves per block, which share the
// weights and each keep their own activations.
//
// MFMA operand layout (16x16x32, lane l, g = l / 16):
//   A[m][k]: m = l % 16, k = 8 g .. 8 g + 7   (8 contiguous BF16)
//   B[k][n]: n = l % 16, k = 8 g .. 8 g + 7   (read from B^T rows)
//   D[m][n]: m = 4 g + i (i = 0..3), n = l % 16
// so every operand fetch is one 16-byte LDS read from a row-major buffer
// with the reduction index contiguous.
//
// Training (TRAIN = true): the loss of the closure,
//   L = sum_p w_rho (cont + s_d diff - t_rho)^2 + w_a |pa + s_v visc - t_a|^2,
//   s = 1 + out_scale tanh(o),
// is back-propagated through the same tiles; per-block FP32 partial
// gradients (accumulated in LDS) are summed in block order by k_mlp_reduce.
#include <algorithm>
#include <climits>

#include "common.hpp"
#include "launch.hpp"

namespace zs {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace mlp {

constexpr int MT = 32;   // samples per tile
constexpr int LDH = 72;  // row stride of 64-wide BF16 rows (36 banks: conflict-free 16-byte reads)
constexpr int LDI = 40;  // row stride of 32-wide BF16 rows (20 banks)

// FP32 master parameters, flat: w0[24][64] b0 w1[64][64] b1 w2[64][64] b2 w3[64][2] b3.
constexpr int P_W0 = 0, P_B0 = P_W0 + MLP_NF * MLP_NH;
constexpr int P_W1 = P_B0 + MLP_NH, P_B1 = P_W1 + MLP_NH * MLP_NH;
constexpr int P_W2 = P_B1 + MLP_NH, P_B2 = P_W2 + MLP_NH * MLP_NH;
constexpr int P_W3 = P_B2 + MLP_NH, P_B3 = P_W3 + MLP_NH * MLP_NO;
static_assert(P_B3 + MLP_NO == MLP_NPARAM, "parameter layout");

// Packed BF16 image (element offsets).  xT = [out][in] (forward B^T),
// x = [in][out] (backward B^T); padding is zero.
constexpr int O_W0T = 0;                       // [64][LDI], in 0..23
constexpr int O_W1T = O_W0T + 64 * LDI;        // [64][LDH]
constexpr int O_W2T = O_W1T + 64 * LDH;        // [64][LDH]
constexpr int O_W3T = O_W2T + 64 * LDH;        // [16][LDH], out 0..1
constexpr int PACK_FWD = O_W3T + 16 * LDH;     // forward part ends here
constexpr int O_W1 = PACK_FWD;                 // [64][LDH]
constexpr int O_W2 = O_W1 + 64 * LDH;          // [64][LDH]
constexpr int O_W3 = O_W2 + 64 * LDH;          // [64][LDI], out 0..1
constexpr int PACK_ALL = O_W3 + 64 * LDI;
static_assert(PACK_FWD % 8 == 0 && O_W1 % 8 == 0 && O_W2 % 8 == 0 && O_W3 % 8 == 0, "16-byte aligned parts");

// LDS activation buffers (BF16 element offsets after the weights).
constexpr int A_X = 0;                  // [32][LDI] normalised features
constexpr int A_H1 = A_X + MT * LDI;    // [32][LDH]
constexpr int A_H2 = A_H1 + MT * LDH;
constexpr int A_H3 = A_H2 + MT * LDH;
constexpr int A_FWD = A_H3 + MT * LDH;  // the forward buffers training keeps
// Inference keeps no activations for a backward pass: the third hidden layer
// reuses the first's buffer (read for the last time by the second layer's
// product), so a block needs only X, H1 and H2.
constexpr int A_INF = A_H2 + MT * LDH;
constexpr int A_XT = A_FWD;             // [32 in][LDI] (32 samples)
constexpr int A_H1T = A_XT + 32 * LDI;  // [64][LDI]
constexpr int A_H2T = A_H1T + 64 * LDI;
constexpr int A_H3T = A_H2T + 64 * LDI;
constexpr int A_DO = A_H3T + 64 * LDI;  // [32][LDI], cols >= 2 zero
constexpr int A_DOT = A_DO + MT * LDI;  // [16][LDI]
constexpr int A_DZ = A_DOT + 16 * LDI;  // [32][LDH]
constexpr int A_DZT = A_DZ + MT * LDH;  // [64][LDI]
constexpr int A_ALL = A_DZT + 64 * LDI;

// Waves per inference block: they share one copy of the weights, each runs
// its own tiles.  73.5 KB of LDS per block, two blocks (eight waves) per CU.
constexpr int INFER_WAVES = 4;

template <bool TRAIN, int NW>
constexpr size_t lds_bytes() {
    return TRAIN ? (size_t)(PACK_ALL + A_ALL) * 2 + (size_t)(MLP_NPARAM + MT + 8) * 4
                 : (size_t)(PACK_FWD + NW * A_INF) * 2 + (size_t)(NW * MT + 8) * 4;
}
constexpr int LDS_PER_CU = 160 * 1024;
constexpr int CUS = 256;  // MI355X
constexpr int INFER_BLOCKS_PER_CU = LDS_PER_CU / (int)lds_bytes<false, INFER_WAVES>();
static_assert(INFER_BLOCKS_PER_CU >= 1, "inference block exceeds the LDS");

__device__ __forceinline__ int lane() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ bf16x8 frag(const bf16 *base, int ld, int r0, int k0) {
    const int l = lane();
    return *reinterpret_cast<const bf16x8 *>(base + (r0 + (l & 15)) * ld + k0 + 8 * (l >> 4));
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// C[mb][nb] = A[32 x 32 KB] * B, with A row-major (lda) and B given as B^T
// row-major (ldb); MB = 2 sample blocks, NB output blocks.
template <int NB, int KB>
__device__ __forceinline__ void gemm(const bf16 *A, int lda, const bf16 *BT, int ldb, f32x4 (&C)[2][NB]) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) acc = mfma(frag(A, lda, 16 * mb, 32 * kb), frag(BT, ldb, 16 * nb, 32 * kb), acc);
            C[mb][nb] = acc;
        }
}

// Weight gradient dW[in][out] += sum over the 32 samples of X[s][in] dZ[s][out]:
// A = X^T ([in][sample], ldx), B^T = dZ^T ([out][sample], ldz); D rows = in.
template <int MB, int NB>
__device__ __forceinline__ void grad_acc(const bf16 *XT, int ldx, const bf16 *DZT, int ldz, float *dW, int n_in,
                                         int n_out) {
    const int l = lane();
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma(frag(XT, ldx, 16 * mb, 0), frag(DZT, ldz, 16 * nb, 0), acc);
            const int col = 16 * nb + (l & 15);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = 16 * mb + 4 * (l >> 4) + i;
                if (row < n_in && col < n_out) dW[row * n_out + col] += acc[i];
            }
        }
}

// Orders a tile's LDS writes before the reads that follow: the whole block
// when it is one wave (training), else the calling wave only (each inference
// wave works on tiles of its own; a wave's LDS operations complete in order).
template <int NW>
__device__ __forceinline__ void tile_sync() {
    if constexpr (NW == 1) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// Store D-layout values v[mb][nb] (rows = samples, cols = neurons) as BF16
// into X ([sample][ld]) and, when XT is given, X^T ([neuron][ldt]).
template <int NB>
__device__ __forceinline__ void store_act(const f32x4 (&v)[2][NB], bf16 *X, int ld, bf16 *XT, int ldt) {
    const int l = lane();
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int col = 16 * nb + (l & 15);
            const int row0 = 16 * mb + 4 * (l >> 4);
            bf16x4 t;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bf16 h = (bf16)v[mb][nb][i];
                X[(row0 + i) * ld + col] = h;
                t[i] = h;
            }
            if (XT) *reinterpret_cast<bf16x4 *>(XT + col * ldt + row0) = t;
        }
}

// Bias + SiLU on a hidden layer's D values; keeps the pre-activation in z.
template <int NB>
__device__ __forceinline__ void bias_silu(f32x4 (&c)[2][NB], f32x4 (&z)[2][NB], const float *__restrict__ b) {
    const int l = lane();
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const float bb = b[16 * nb + (l & 15)];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float zz = c[mb][nb][i] + bb;
                z[mb][nb][i] = zz;
                c[mb][nb][i] = zz * sigm(zz);
            }
    }
}

// d = dH * silu'(z), column sums of d added to db.
template <int NB>
__device__ __forceinline__ void silu_back(f32x4 (&d)[2][NB], const f32x4 (&z)[2][NB], float *db) {
    const int l = lane();
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        float cs = 0.0f;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float zz = z[mb][nb][i];
                const float sg = sigm(zz);
                d[mb][nb][i] *= sg * (1.0f + zz * (1.0f - sg));
                cs += d[mb][nb][i];
            }
        atomicAdd(&db[16 * nb + (l & 15)], cs);
    }
}

}  // namespace mlp

// NW waves per block; wave w of block b takes tiles b NW + w, then every
// (grid NW)-th, of the 32-sample tiles of [0, n) (n = *n_dev when n_dev is
// given: the fluid count of a sorted buffer).
//   feat:   n x MLP_NF raw features (FP32)
//   TRAIN = false: out = n x 2 scale factors (1, 1 where a normalised
//                  feature leaves [lo, hi])
//   TRAIN = true:  rec = n x MLP_NT training records; part = per-block
//                  gradient partials (MLP_NPARAM) + loss partial
template <bool TRAIN, int NW>
__global__ __launch_bounds__(64 * NW) void k_mlp(const float *__restrict__ P, const bf16 *__restrict__ pack,
                                                const MlpAux *__restrict__ aux, const float *__restrict__ feat,
                                                int n, const int *__restrict__ n_dev, float2 *__restrict__ out,
                                                const float *__restrict__ rec, float *__restrict__ part) {
    using namespace mlp;
    static_assert(!TRAIN || NW == 1, "training runs one wave per block");
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int PACK = TRAIN ? PACK_ALL : PACK_FWD;
    constexpr int ACT = TRAIN ? A_ALL : A_INF;  // activations of one wave
    const int tid = (int)threadIdx.x, wv = tid >> 6;
    bf16 *W = reinterpret_cast<bf16 *>(smem);
    bf16 *act = W + PACK + wv * ACT;
    float *dW = reinterpret_cast<float *>(W + PACK + NW * ACT);
    int *keep0 = reinterpret_cast<int *>(dW + (TRAIN ? MLP_NPARAM : 0));
    int *keep = keep0 + wv * MT;  // per sample of the wave's tile: inside the range
    float *lsum = reinterpret_cast<float *>(keep0 + NW * MT);
    const int l = lane();
    if (n_dev) n = *n_dev;

    // weights (16-byte copies, seven loads in flight before their stores),
    // zeroed activations (16-byte stores) and gradient accumulators
    {
        constexpr int STEP = 8 * 64 * NW, PASSES = (PACK + STEP - 1) / STEP, BATCH = 7;
        static_assert(PACK % 8 == 0, "weights copy in 8-element pieces");
        for (int p0 = 0; p0 < PASSES; p0 += BATCH) {
            bf16x8 tmp[BATCH];
#pragma unroll
            for (int k = 0; k < BATCH; ++k) {
                const int e = (p0 + k) * STEP + 8 * tid;
                if (e < PACK) tmp[k] = *reinterpret_cast<const bf16x8 *>(pack + e);
            }
#pragma unroll
            for (int k = 0; k < BATCH; ++k) {
                const int e = (p0 + k) * STEP + 8 * tid;
                if (e < PACK) *reinterpret_cast<bf16x8 *>(W + e) = tmp[k];
            }
        }
    }
    {
        static_assert(ACT % 8 == 0 && PACK % 8 == 0, "activations cleared in 8-element pieces");
        bf16x8 zero;
#pragma unroll
        for (int k = 0; k < 8; ++k) zero[k] = (bf16)0.0f;
        for (int e = 8 * tid; e < NW * ACT; e += 8 * 64 * NW) *reinterpret_cast<bf16x8 *>(W + PACK + e) = zero;
    }
    if (TRAIN)
        for (int e = l; e < MLP_NPARAM; e += 64) dW[e] = 0.0f;
    if (tid == 0) lsum[0] = 0.0f;
    __syncthreads();

    bf16 *X = act + A_X, *H1 = act + A_H1, *H2 = act + A_H2, *H3 = act + (TRAIN ? A_H3 : A_H1);
    bf16 *XT = TRAIN ? act + A_XT : nullptr, *H1T = TRAIN ? act + A_H1T : nullptr;
    bf16 *H2T = TRAIN ? act + A_H2T : nullptr, *H3T = TRAIN ? act + A_H3T : nullptr;
    const float os = aux->out_scale;

    // Inference: lane l holds tile elements e = l + 64 k (k < FL), of feature
    // column (l + 64 k) % MLP_NF, which repeats with period FP in k: the
    // normalisation constants of a lane's FP columns stay in registers.
    // (Training keeps the plain loop: its registers are all in use.)
    constexpr int FL = MT * MLP_NF / 64, FP = 3;
    static_assert(MT * MLP_NF % 64 == 0 && 64 * FP % MLP_NF == 0 && FL % FP == 0, "feature tile layout");
    float fmean[FP], fstd[FP], flo[FP], fhi[FP];
#pragma unroll
    for (int j = 0; j < FP; ++j) {
        const int f = (l + 64 * j) % MLP_NF;
        fmean[j] = aux->mean[f];
        fstd[j] = aux->inv_std[f];
        flo[j] = aux->lo[f];
        fhi[j] = aux->hi[f];
    }
    // Raw features of a tile: a lane's FL loads are issued together, the next
    // tile's while the current one runs through the network.
    float fv[FL];
    auto load_tile = [&](int t) {
#pragma unroll
        for (int k = 0; k < FL; ++k) {
            const int e = l + 64 * k;
            fv[k] = (t + e / MLP_NF < n) ? feat[(size_t)t * MLP_NF + e] : 0.0f;
        }
    };
    const int first = ((int)blockIdx.x * NW + wv) * MT, stride = (int)gridDim.x * NW * MT;
    if (!TRAIN && first < n) load_tile(first);

    for (int t0 = first; t0 < n; t0 += stride) {
        // normalised features of the tile; out-of-range samples flagged
        if (l < MT) keep[l] = 1;
        tile_sync<NW>();
        if constexpr (TRAIN) {
            for (int e = l; e < MT * MLP_NF; e += 64) {
                const int s = e / MLP_NF, f = e - s * MLP_NF;
                float v = 0.0f;
                if (t0 + s < n) {
                    v = (feat[(size_t)(t0 + s) * MLP_NF + f] - aux->mean[f]) * aux->inv_std[f];
                    if (!(v >= aux->lo[f] && v <= aux->hi[f])) keep[s] = 0;
                }
                const bf16 h = (bf16)v;
                X[s * LDI + f] = h;
                XT[f * LDI + s] = h;
            }
        } else {
#pragma unroll
            for (int k = 0; k < FL; ++k) {
                const int e = l + 64 * k, s = e / MLP_NF, f = e - s * MLP_NF;
                float v = 0.0f;
                if (t0 + s < n) {
                    v = (fv[k] - fmean[k % FP]) * fstd[k % FP];
                    if (!(v >= flo[k % FP] && v <= fhi[k % FP])) keep[s] = 0;
                }
                X[s * LDI + f] = (bf16)v;
            }
            if (t0 + stride < n) load_tile(t0 + stride);
        }
        tile_sync<NW>();

        f32x4 c4[2][4], z1[2][4], z2[2][4], z3[2][4];
        gemm<4, 1>(X, LDI, W + O_W0T, LDI, c4);
        bias_silu<4>(c4, z1, P + P_B0);
        store_act<4>(c4, H1, LDH, H1T, LDI);
        tile_sync<NW>();
        gemm<4, 2>(H1, LDH, W + O_W1T, LDH, c4);
        bias_silu<4>(c4, z2, P + P_B1);
        store_act<4>(c4, H2, LDH, H2T, LDI);
        tile_sync<NW>();
        gemm<4, 2>(H2, LDH, W + O_W2T, LDH, c4);
        bias_silu<4>(c4, z3, P + P_B2);
        store_act<4>(c4, H3, LDH, H3T, LDI);
        tile_sync<NW>();
        f32x4 o[2][1];
        gemm<1, 2>(H3, LDH, W + O_W3T, LDH, o);

        const int k = l & 15;  // output index held by this lane
        if constexpr (!TRAIN) {
            if (k < MLP_NO)
#pragma unroll
                for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int s = 16 * mb + 4 * (l >> 4) + i;
                        if (t0 + s >= n) continue;
                        const float sk = keep[s] ? 1.0f + os * tanhf(o[mb][0][i] + P[P_B3 + k]) : 1.0f;
                        reinterpret_cast<float *>(out)[(size_t)(t0 + s) * 2 + k] = sk;
                    }
            tile_sync<NW>();
            continue;
        }

        // ---- training: dL/do per sample, then back through the layers ----
        bf16 *DO = act + A_DO, *DOT = act + A_DOT, *DZ = act + A_DZ, *DZT = act + A_DZT;
        f32x4 d1[2][1];
        float lacc = 0.0f;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int s = 16 * mb + 4 * (l >> 4) + i;
                float dlo = 0.0f;
                if (k < MLP_NO && t0 + s < n) {
                    const float *r = rec + (size_t)(t0 + s) * MLP_NT;
                    const float th = tanhf(o[mb][0][i] + P[P_B3 + k]);
                    const float sk = 1.0f + os * th;
                    float dls;
                    if (k == 0) {
                        const float res = r[0] + sk * r[1] - r[2];
                        lacc += aux->w_rho * res * res;
                        dls = 2.0f * aux->w_rho * res * r[1];
                    } else {
                        const float rx = r[3] + sk * r[6] - r[9];
                        const float ry = r[4] + sk * r[7] - r[10];
                        const float rz = r[5] + sk * r[8] - r[11];
                        lacc += aux->w_acc * (rx * rx + ry * ry + rz * rz);
                        dls = 2.0f * aux->w_acc * (rx * r[6] + ry * r[7] + rz * r[8]);
                    }
                    dlo = dls * os * (1.0f - th * th);
                }
                d1[mb][0][i] = dlo;
            }
        atomicAdd(&lsum[0], lacc);
        // output layer: db3, dW3 = H3^T dO, dH3 = dO W3^T
        {
            float cs = 0.0f;
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int i = 0; i < 4; ++i) cs += d1[mb][0][i];
            if (k < MLP_NO) atomicAdd(&dW[P_B3 + k], cs);
        }
        store_act<1>(d1, DO, LDI, DOT, LDI);
        __syncthreads();
        grad_acc<4, 1>(H3T, LDI, DOT, LDI, dW + P_W3, MLP_NH, MLP_NO);
        f32x4 dh[2][4];
        gemm<4, 1>(DO, LDI, W + O_W3, LDI, dh);
        silu_back<4>(dh, z3, dW + P_B2);
        __syncthreads();
        store_act<4>(dh, DZ, LDH, DZT, LDI);
        __syncthreads();
        // layer 2: dW2 = H2^T dZ3, dH2 = dZ3 W2^T
        grad_acc<4, 4>(H2T, LDI, DZT, LDI, dW + P_W2, MLP_NH, MLP_NH);
        gemm<4, 2>(DZ, LDH, W + O_W2, LDH, dh);
        silu_back<4>(dh, z2, dW + P_B1);
        __syncthreads();
        store_act<4>(dh, DZ, LDH, DZT, LDI);
        __syncthreads();
        // layer 1
        grad_acc<4, 4>(H1T, LDI, DZT, LDI, dW + P_W1, MLP_NH, MLP_NH);
        gemm<4, 2>(DZ, LDH, W + O_W1, LDH, dh);
        silu_back<4>(dh, z1, dW + P_B0);
        __syncthreads();
        store_act<4>(dh, DZ, LDH, DZT, LDI);
        __syncthreads();
        // layer 0: dW0 = X^T dZ1 (inputs 0..23)
        grad_acc<2, 4>(XT, LDI, DZT, LDI, dW + P_W0, MLP_NF, MLP_NH);
        __syncthreads();
    }
    if constexpr (TRAIN) {
        __syncthreads();
        float *pb = part + (size_t)blockIdx.x * (MLP_NPARAM + 1);
        for (int e = l; e < MLP_NPARAM; e += 64) pb[e] = dW[e];
        if (l == 0) pb[MLP_NPARAM] = lsum[0];
    }
}

// grad[p] = scale * sum_b part[b][p] (block order); p = MLP_NPARAM is the loss.
__global__ __launch_bounds__(ZS_BLOCK) void k_mlp_reduce(const float *__restrict__ part, int nblk, float scale,
                                                         float *__restrict__ grad) {
    const int p = blockIdx.x * ZS_BLOCK + threadIdx.x;
    if (p > MLP_NPARAM) return;
    float s = 0.0f;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)b * (MLP_NPARAM + 1) + p];
    grad[p] = s * scale;
}

// Adam on the FP32 master parameters (bias-corrected; step t >= 1).
__global__ __launch_bounds__(ZS_BLOCK) void k_mlp_adam(float *__restrict__ P, float *__restrict__ m,
                                                       float *__restrict__ v, const float *__restrict__ g, float lr,
                                                       float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * ZS_BLOCK + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float gg = g[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gg;
    const float vv = b2 * v[p] + (1.0f - b2) * gg * gg;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// BF16 image of the master parameters in the LDS layout (zero padding).
__global__ __launch_bounds__(ZS_BLOCK) void k_mlp_pack(const float *__restrict__ P, bf16 *__restrict__ pack) {
    using namespace mlp;
    const int e = blockIdx.x * ZS_BLOCK + threadIdx.x;
    if (e >= PACK_ALL) return;
    float v = 0.0f;
    int r, c;
    if (e < O_W1T) {  // W0T [out][in]
        r = e / LDI; c = e - r * LDI;
        if (c < MLP_NF) v = P[P_W0 + c * MLP_NH + r];
    } else if (e < O_W2T) {
        r = (e - O_W1T) / LDH; c = (e - O_W1T) - r * LDH;
        if (c < MLP_NH) v = P[P_W1 + c * MLP_NH + r];
    } else if (e < O_W3T) {
        r = (e - O_W2T) / LDH; c = (e - O_W2T) - r * LDH;
        if (c < MLP_NH) v = P[P_W2 + c * MLP_NH + r];
    } else if (e < O_W1) {  // W3T [out 16][in]
        r = (e - O_W3T) / LDH; c = (e - O_W3T) - r * LDH;
        if (r < MLP_NO && c < MLP_NH) v = P[P_W3 + c * MLP_NO + r];
    } else if (e < O_W2) {  // W1 [in][out]
        r = (e - O_W1) / LDH; c = (e - O_W1) - r * LDH;
        if (c < MLP_NH) v = P[P_W1 + r * MLP_NH + c];
    } else if (e < O_W3) {
        r = (e - O_W2) / LDH; c = (e - O_W2) - r * LDH;
        if (c < MLP_NH) v = P[P_W2 + r * MLP_NH + c];
    } else {  // W3 [in][out 32]
        r = (e - O_W3) / LDI; c = (e - O_W3) - r * LDI;
        if (c < MLP_NO) v = P[P_W3 + r * MLP_NO + c];
    }
    pack[e] = (bf16)v;
}

size_t mlp_pack_elems() { return mlp::PACK_ALL; }

int mlp_blocks(int n) { return std::max(1, std::min(256, (n + mlp::MT - 1) / mlp::MT)); }
// Inference has no cross-block reduction: blocks of INFER_WAVES waves, as
// many as the LDS lets run at once on every CU (one round; each block copies
// the weights once), fewer when the tiles do not fill them.  Each sample's
// result does not depend on which block or wave computes it.
int mlp_infer_blocks(int n) {
    const int tiles = (n + mlp::MT - 1) / mlp::MT;
    const int want = (tiles + mlp::INFER_WAVES - 1) / mlp::INFER_WAVES;
    return std::max(1, std::min(mlp::CUS * mlp::INFER_BLOCKS_PER_CU, want));
}

hipError_t mlp_prepare() {
    using mlp::INFER_WAVES;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mlp<true, 1>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)mlp::lds_bytes<true, 1>());
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mlp<false, INFER_WAVES>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)mlp::lds_bytes<false, INFER_WAVES>());
    return e;
}

void launch_mlp_pack(hipStream_t s, const float *P, void *pack) {
    k_mlp_pack<<<(mlp::PACK_ALL + ZS_BLOCK - 1) / ZS_BLOCK, ZS_BLOCK, 0, s>>>(P, static_cast<bf16 *>(pack));
}

void launch_mlp_infer(hipStream_t s, const float *P, const void *pack, const MlpAux *aux, const float *feat, int n,
                      const int *n_dev, float2 *out, int nblk) {
    using mlp::INFER_WAVES;
    k_mlp<false, INFER_WAVES><<<nblk, 64 * INFER_WAVES, mlp::lds_bytes<false, INFER_WAVES>(), s>>>(
        P, static_cast<const bf16 *>(pack), aux, feat, n, n_dev, out, nullptr, nullptr);
}

void launch_mlp_grad(hipStream_t s, const float *P, const void *pack, const MlpAux *aux, const float *feat,
                     const float *rec, int n, float *part, float *grad) {
    const int nblk = mlp_blocks(n);
    k_mlp<true, 1><<<nblk, 64, mlp::lds_bytes<true, 1>(), s>>>(P, static_cast<const bf16 *>(pack), aux, feat, n,
                                                              nullptr, nullptr, rec, part);
    k_mlp_reduce<<<(MLP_NPARAM + 1 + ZS_BLOCK - 1) / ZS_BLOCK, ZS_BLOCK, 0, s>>>(part, nblk, 1.0f / (float)n, grad);
}

void launch_mlp_adam(hipStream_t s, float *P, float *m, float *v, const float *g, float lr, float b1, float b2,
                     float eps, int t) {
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    k_mlp_adam<<<(MLP_NPARAM + ZS_BLOCK - 1) / ZS_BLOCK, ZS_BLOCK, 0, s>>>(P, m, v, g, lr, b1, b2, eps, c1, c2);
}

}  // namespace zs
