// Host launchers of the gfx950 kernels (one translation unit each).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace zs {

// nbr_build.hip
// sp: per-scene parameter table of a batch (device memory), nullptr for one
// scene; b.scene must then hold each particle's scene.
// gate (every neighbour-build launcher; nullptr: always run): skin mode's
// rebuild flag; the kernels do nothing while *gate == 0.
void launch_hash(hipStream_t s, const Prm &q, const Prm *sp, const Buf &b, uint32_t *key, int *idx, int *killed,
                 const int *gate = nullptr);
int radix_blocks(int n, int items);
// Histogram buffer words for any digit width up to `bits` and any tile.
size_t radix_hist_words(int n, int bits);
// Sorts (key[0], val[0]) over the low `keybits` bits; returns the slot (0/1)
// holding the result, -1 for an unsupported digit width (8..11) or tile
// (items = keys per thread, 8 or 16).  identity_vals: val[0] is known to be
// 0..n-1 and is not read.
int launch_sort(hipStream_t s, int n, int keybits, int bits, int items, uint32_t *key[2], int *val[2],
                uint32_t *hist, bool identity_vals, const int *gate = nullptr);
// cs[0..G]: fluid cell starts; cs[G+1..2G+1]: dummy cell starts (key flag + c).
// scratch: cell_start_scratch(G) ints (tile bounds).  n >= 1.
size_t cell_start_scratch(int G);
void launch_cell_start(hipStream_t s, const uint32_t *skey, int n, int G, uint32_t flag, int *cs, int *scratch,
                       const int *gate = nullptr);
void launch_reorder(hipStream_t s, int n, const int *perm, const Buf &a, const Buf &b, const int *gate = nullptr);
// Skin mode: copy the reordered buffer a into the working buffer b and keep
// the positions in xref (gated); count and clear the rebuild flag.
void launch_skin_commit(hipStream_t s, int n, const Buf &a, const Buf &b, float4 *xref, const int *gate);
void launch_skin_flag(hipStream_t s, int *flag, unsigned long long *rebuilds, int *rebuilt);
// Wave order of the neighbour passes (one scene, gated): units of yt cell rows
// of a plane, tile by tile (k_wave_table).  wtab holds 1 + cap entries, cap =
// wave_table_cap(N, Y, Z, yt) for a grid of at most Y x Z rows.
int wave_table_cap(int N, int Y, int Z, int yt);
void launch_wave_table(hipStream_t s, const Prm &q, const int *cs, int yt, int2 *wtab, int cap, const int *gate);

// physics.hip; walk = WALK_MASK, WALK_HITS or WALK_LIST (common.hpp).
// WALK_LIST: launch_ldg records the neighbour lists (nlist: list_words(N)
// ints, ncount: N ints) that launch_governing then reads.  WALK_HITS: the
// same with the chunk hit masks (nlist: hits_slots(k) x N words; ncount:
// overflow flags).  A caller without a current record passes WALK_MASK to
// launch_governing.
// block: threads per block of the neighbour passes (128, 256 or 512; the
// closure's feature and scaled variants always run 256).
void launch_wall(hipStream_t s, int walk, int block, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                 const int *cs);
// feat (one scene only; may be nullptr): the closure features, ZSPH_NF floats per particle.
void launch_ldg(hipStream_t s, int walk, int block, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                const int *cs, float4 *gd, int *nlist, int *ncount, float *feat);
// mode (one scene only): 0 plain, 1 closure-scaled (scale: 2 floats per
// particle), 2 plain + term decomposition (terms: 8 floats per particle).
// moving: the context has moving dummies (a body or a piston), whose
// pressure force k_governing_dummy computes; fixed walls get zero either way.
// xref (WALK_SKIN only): the rebuild positions its rows are centred on.
void launch_governing(hipStream_t s, int walk, int block, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                      const int *cs, const float4 *gd, float4 *acc, const int *nlist, const int *ncount, int mode,
                      const float2 *scale, float *terms, bool moving,
                      const float4 *xref = nullptr);
// WALK_SKIN's gradient pass (one scene, plain): the recording form on rebuild
// half steps, the replaying form on the others (gated on *sr.rebuilt);
// hits/ncount: the exact record and the per-particle "not recorded" flags.
void launch_ldg_skin(hipStream_t s, int block, const Prm &q, const Buf &b, const uint32_t *skey, const int *cs,
                     float4 *gd, int *hits, int *ncount, const SkinRec &sr);
// chk (skin mode): the displacement test after the update (chk.xref == nullptr: none).
void launch_predict(hipStream_t s, const Prm &q, const Prm *sp, const Buf &b, const int *cs, const float4 *acc,
                    const SkinCheck &chk = SkinCheck{});
void launch_correct(hipStream_t s, const Prm &q, const Prm *sp, const Buf &b, const int *cs, const float4 *acc,
                    const SkinCheck &chk = SkinCheck{});
void launch_nbr_stats(hipStream_t s, int walk, const Prm &q, const Prm *sp, const Buf &b, const uint32_t *skey,
                      const int *cs, int *count, uint64_t *digest);

// diag.hip
constexpr int DIAG_D = 5;  // per-block FP64 partials: count, kinetic, momentum x/y/z
int diag_blocks(int n);
void launch_diag(hipStream_t s, const Prm &q, const Buf &b, float front_zmax, double *part, float *pmax);
void launch_probe(hipStream_t s, const Prm &q, const Buf &b, const float *pts, int npts, float *out);
// Shepard-filtered fluid fields at particles `rows` (cell order; qf: the
// buffer's grid with k and trim2 covering `radius`); out: nq x 9 floats.
void launch_filter(hipStream_t s, const Prm &qf, const Buf &b, const uint32_t *skey, const int *cs, const float4 *acc,
                   const int *rows, int nq, float hc, float radius, float *out);
// Per-scene diagnostics from the cell table of buffer b (one block per scene).
constexpr int SCENE_D = 8;  // count, kinetic, momentum x/y/z, front, zmax, removed
void launch_scene_diag(hipStream_t s, const Prm &q, const Prm *sp, int nscenes, const Buf &b, const int *cs,
                       float front_zmax, double *part);

// body.hip: floating body and piston.  Device-resident FP64 state, so a
// captured step advances it with every replay.
struct DevBody {
    zsph_rigid s;       // state at time t
    zsph_rigid s0;      // state at the start of the step (time t0)
    double t, t0;       // simulated time now and at the start of the step
    double amp, pw, phase;  // piston: d(t) = amp (sin(pw t + phase) - sin(phase)), pw = 2 pi f
    int axis, has_body, has_piston, pad;
};
constexpr int BODY_BLOCKS = 256;  // fixed partial count: the force sum has a fixed order
constexpr int BODY_D = 6;         // F x/y/z, T x/y/z
void launch_body_reduce(hipStream_t s, const Prm &q, const Buf &b, const int *cs, const float4 *acc,
                        const DevBody *body, double *part);
void launch_body_update(hipStream_t s, const Prm &q, DevBody *body, const double *part, bool predictor);
void launch_kinematics(hipStream_t s, const Prm &q, const Buf &b, const int *cs, const DevBody *body,
                       const SkinCheck &chk = SkinCheck{});

// mlp.hip: closure network 24 -> 64 -> 64 -> 64 -> 2 on MFMA (BF16, FP32 accumulate).
constexpr int MLP_NF = ZSPH_NF;          // features
constexpr int MLP_NH = ZSPH_NH;          // hidden width
constexpr int MLP_NO = 2;                // outputs: scale of the density diffusion, of the viscous term
constexpr int MLP_NT = ZSPH_NT;          // training record per sample
constexpr int MLP_NPARAM = MLP_NF * MLP_NH + MLP_NH + 2 * (MLP_NH * MLP_NH + MLP_NH) + MLP_NH * MLP_NO + MLP_NO;
struct MlpAux {
    float mean[MLP_NF], inv_std[MLP_NF], lo[MLP_NF], hi[MLP_NF];
    float out_scale, w_rho, w_acc;
};
size_t mlp_pack_elems();          // BF16 elements of the packed weight image
int mlp_blocks(int n);        // blocks of a training launch over n samples (partials buffer size)
int mlp_infer_blocks(int n);  // blocks for launch_mlp_infer over n samples
hipError_t mlp_prepare();         // allows the kernels' dynamic LDS (call once per device)
void launch_mlp_pack(hipStream_t s, const float *P, void *pack);
// n samples, or *n_dev when n_dev is given; out: n x 2 scale factors.
void launch_mlp_infer(hipStream_t s, const float *P, const void *pack, const MlpAux *aux, const float *feat, int n,
                      const int *n_dev, float2 *out, int nblk);
// grad[0..MLP_NPARAM) = mean gradient over the n samples, grad[MLP_NPARAM] = mean loss;
// part: mlp_blocks(n) x (MLP_NPARAM + 1) floats.
void launch_mlp_grad(hipStream_t s, const float *P, const void *pack, const MlpAux *aux, const float *feat,
                     const float *rec, int n, float *part, float *grad);
void launch_mlp_adam(hipStream_t s, float *P, float *m, float *v, const float *g, float lr, float b1, float b2,
                     float eps, int t);

}  // namespace zs
