// C ABI of libzsph: device context, transfers, the per-stage entry points of
// the reference driver (single_gpu_sph_src/deltasph/main.cu:58-90) and the
// graph-captured full step.  Host code only; kernels live in kernels/.
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "host/host_internal.hpp"
#include "host/rigid.hpp"
#include "kernels/common.hpp"
#include "kernels/launch.hpp"
#include "zsph.h"

// One in-flight asynchronous download (zsph_download_async): a device
// snapshot of the requested arrays and its pinned host copy.
struct zsph_frame {
    zsph_ctx *ctx = nullptr;
    void *dev = nullptr;                     // snapshot: posr, velp, acc (16 B), type, table, key (4 B) per particle
    void *host = nullptr;                    // pinned copy of the same layout
    hipEvent_t snap_done = nullptr, copy_done = nullptr;
    int fields = 0, n = 0;
    uint32_t cell_mask = 0;                  // the cell bits of a sort key (ZSPH_F_HASH)
    std::atomic<int> busy{0};
    unsigned long long seq = 0;              // issue order (the oldest is waited for first)
};

struct zsph_ctx {
    zsph_params p{};
    zs::Prm q{};
    int device = 0;
    hipStream_t stream = nullptr;
    zs::Buf buf[2]{};
    uint32_t *key[2]{};
    int *idx[2]{};
    int sorted_slot = -1;  // slot of (key, idx) holding the sorted result
    int hashed = 0;        // key[0]/idx[0] hold a fresh (unsorted) hash
    int *cell_start = nullptr;
    int *cs_scratch = nullptr;  // tile bounds of the cell-table build
    float4 *acc = nullptr;
    float4 *gd = nullptr;
    uint32_t *hist = nullptr;
    int *killed = nullptr;
    double *diag_part = nullptr;  // per-block diagnostics partials
    float *diag_max = nullptr;
    int keybits = 1;
    int radix_bits = 8;
    int walk = zs::WALK_MASK;  // neighbour-walk variant (tuning)
    int block = ZS_BLOCK;      // threads per block of the neighbour passes (tuning)
    int sort_tile = 16;        // keys per thread of the radix tile (tuning)
    float skin = 0.0f;         // zsph_step's neighbour skin in h (tuning; 0: rebuild every half step)
    zs::Prm qs{};              // the skin-mode neighbour grid (make_skin_prm)
    int *rebuild = nullptr;    // device flag: the next skin-mode half step rebuilds
    unsigned long long *rebuilds = nullptr;  // device count of skin-mode rebuilds
    float4 *xref = nullptr;    // positions at the last skin-mode rebuild
    bool skin_valid = false;   // buffer 0 is as zsph_step left it (or a skin rebuild is flagged)
    bool skin_pending = false; // a rebuild is flagged: the skin structures do not describe buffer 0 yet
    int skin_slot = 0;         // key slot holding the skin grid's sorted keys
    int *nlist = nullptr;      // WALK_LIST neighbour lists (list_words(N) ints) and counts
    int *ncount = nullptr;     // WALK_LIST counts / WALK_HITS overflow flags
    int *hits = nullptr;       // WALK_HITS / WALK_SKIN exact hit masks (hits_slots(k) x N words)
    uint32_t *skin_hits = nullptr;  // WALK_SKIN skin hit masks of the last rebuild (same layout)
    bool list_valid = false;   // the lists / hit masks describe the current sorted buffer
    int tile_rows = 0;         // skin mode's wave order: cell rows per unit (0: cell order; one scene)
    int2 *wtab = nullptr;      // its wave table, rebuilt with the cell table (room for any tile_rows)
    int wt_cap = 0;            // waves of the table for tile_rows
    int nscenes = 1;                  // > 1: a batch of independent scenes
    zs::Prm *scene_prm = nullptr;     // per-scene parameters (device, batches only)
    std::vector<zs::Prm> scene_host;  // the same on the host
    zs::Prm *scene_prm_skin = nullptr;     // skin mode: per-scene skin grids (device, batches only)
    std::vector<zs::Prm> scene_skin_host;  // the same on the host
    std::vector<int> scene_off;       // first particle of each scene in upload order (+ N)
    int cells_buf = -1;               // buffer the cell table describes
    double *scene_part = nullptr;     // per-scene diagnostics (batches only)
    unsigned long long *clock = nullptr;  // per stage {start, ticks, launches} (3 x ZSPH_ST_COUNT words)
    unsigned clock_mask = 0;              // stages (1 << ZSPH_ST_*) bracketed by clock stamps
    bool closure_on = false;      // learned closure: features, network, scaled terms
    float *feat = nullptr;        // N x ZSPH_NF closure features
    float2 *scale = nullptr;      // N scale factors (diffusion, viscous)
    float *terms = nullptr;       // N x 8 term decomposition (zsph_closure_sample)
    float *mlp_P = nullptr;       // network parameters (FP32)
    void *mlp_pack = nullptr;     // their BF16 image
    zs::MlpAux *mlp_aux = nullptr;
    zs::DevBody *body = nullptr;  // moving dummies: FP64 body state, piston, device time
    double *body_part = nullptr;  // BODY_BLOCKS x BODY_D force/torque partials
    bool has_body = false, has_piston = false;
    double t = 0.0, t0 = 0.0;     // simulated time while no dummies move (else on the device)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool use_graph = true;        // zsph_step replays a captured graph (else launches directly)
    std::vector<unsigned char> closure_net;  // the zsph_closure last set (checkpoints carry it)
    float *probe_scratch = nullptr;  // this device's probe_scratch (zsph_probe_pressure_noalloc)
    hipStream_t copy_stream = nullptr;  // device-to-host copies of asynchronous downloads
    zsph_frame frames[ZSPH_FRAME_SLOTS];
    unsigned long long frame_seq = 0;
    std::mutex frame_mu;
    std::condition_variable frame_cv;
    std::string err;
    std::vector<void *> allocs;
};

namespace {

int set_err(zsph_ctx *c, int code, const char *fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return set_err(ctx, ZSPH_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
    } while (0)

template <class T>
int dalloc(zsph_ctx *c, T **p, size_t count) {
    void *d = nullptr;
    const size_t bytes = count ? count * sizeof(T) : sizeof(T);
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess) return set_err(c, ZSPH_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    c->allocs.push_back(d);
    *p = static_cast<T *>(d);
    e = hipMemset(d, 0, bytes);
    if (e != hipSuccess) return set_err(c, ZSPH_E_HIP, "hipMemset failed: %s", hipGetErrorString(e));
    return ZSPH_OK;
}

int check_launch(zsph_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, ZSPH_E_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return ZSPH_OK;
}

bool valid_buf(int b) { return b == 0 || b == 1; }

// Copy on the context's stream and wait for it: ordered with the kernels on
// that (non-blocking) stream, which a plain hipMemcpy is not.
int copy_sync(zsph_ctx *c, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, kind, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ZSPH_OK;
}

// Optional per-stage event timing (zsph_profile_steps).
struct Prof {
    hipStream_t s;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
    hipEvent_t open_ev = nullptr;
    int open_stage = -1;
    void begin(int st) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s);
        open_ev = e;
        open_stage = st;
    }
    void end() {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s);
        ev.push_back({open_stage, {open_ev, e}});
    }
};
struct Scope {
    Prof *p;
    Scope(Prof *pp, int st) : p(pp) { if (p) p->begin(st); }
    ~Scope() { if (p) p->end(); }
};

// The stages of one half step, in the reference's order (main.cu:63-72).
int stage_hash(zsph_ctx *c, int b, Prof *pf = nullptr) {
    {
        Scope sc(pf, ZSPH_ST_HASH);
        zs::launch_hash(c->stream, c->q, c->scene_prm, c->buf[b], c->key[0], c->idx[0], c->killed);
    }
    c->hashed = 1;
    c->sorted_slot = -1;
    return check_launch(c, "hash");
}
int stage_sort(zsph_ctx *c, Prof *pf = nullptr) {
    if (!c->hashed) return set_err(c, ZSPH_E_STATE, "zsph_sort_index called before zsph_hash");
    int slot;
    {
        Scope sc(pf, ZSPH_ST_SORT);
        slot = zs::launch_sort(c->stream, c->q.N, c->keybits, c->radix_bits, c->sort_tile, c->key, c->idx, c->hist,
                               true);
    }
    if (slot < 0) return set_err(c, ZSPH_E_ARG, "unsupported radix width %d / tile %d", c->radix_bits, c->sort_tile);
    c->sorted_slot = slot;
    c->hashed = 0;
    return check_launch(c, "radix sort");
}
int stage_sort_data(zsph_ctx *c, int from, int to, Prof *pf = nullptr) {
    if (c->sorted_slot < 0) return set_err(c, ZSPH_E_STATE, "zsph_sort_data called before zsph_sort_index");
    if (from == to) return set_err(c, ZSPH_E_ARG, "zsph_sort_data needs two different buffers");
    {
        Scope sc(pf, ZSPH_ST_CELLS);
        zs::launch_cell_start(c->stream, c->key[c->sorted_slot], c->q.N, c->q.G, c->q.flag, c->cell_start,
                              c->cs_scratch);
    }
    {
        Scope sc(pf, ZSPH_ST_REORDER);
        zs::launch_reorder(c->stream, c->q.N, c->idx[c->sorted_slot], c->buf[from], c->buf[to]);
    }
    c->list_valid = false;
    c->cells_buf = to;
    return check_launch(c, "sort_data");
}

// The per-scene parameter table of a batch's kernels (nullptr for one
// scene): in skin mode's half step (skin_path) each scene's skin grid.
const zs::Prm *scene_table(const zsph_ctx *c, bool skin_path) {
    return (skin_path && c->scene_prm_skin && c->skin > 0.0f) ? c->scene_prm_skin : c->scene_prm;
}

// The walk a stage runs: WALK_SKIN only on skin mode's plain path (its rows
// live around the rebuild positions); elsewhere (staged reference calls, the
// closure's passes) the hit-mask walk stands in for it.
int eff_walk(const zsph_ctx *c, bool skin_path) {
    if (c->walk == zs::WALK_SKIN && (!skin_path || c->closure_on)) return zs::WALK_HITS;
    return c->walk;
}
// The record a walk's gradient pass writes and its continuity+momentum pass reads.
int *walk_record(zsph_ctx *c) { return (c->walk == zs::WALK_HITS || c->walk == zs::WALK_SKIN) ? c->hits : c->nlist; }
bool walk_records(int walk) { return walk == zs::WALK_LIST || walk == zs::WALK_HITS || walk == zs::WALK_SKIN; }

int stage_boundary(zsph_ctx *c, const zs::Prm &q, int b, Prof *pf = nullptr, bool skin_path = false) {
    if (c->sorted_slot < 0) return set_err(c, ZSPH_E_STATE, "zsph_boundary called before zsph_sort_data");
    const uint32_t *sk = c->key[c->sorted_slot];
    const int walk = eff_walk(c, skin_path);
    {
        Scope sc(pf, ZSPH_ST_WALL);
        zs::launch_wall(c->stream, walk, c->block, q, scene_table(c, skin_path), c->buf[b], sk, c->cell_start);
    }
    {
        Scope sc(pf, ZSPH_ST_LDG);
        if (walk == zs::WALK_SKIN) {
            zs::SkinRec sr;
            sr.xref = c->xref;
            sr.skin = c->skin_hits;
            sr.rebuilt = c->rebuild + 1;
            zs::launch_ldg_skin(c->stream, c->block, q, c->buf[b], sk, c->cell_start, c->gd, c->hits, c->ncount, sr);
        } else {
            zs::launch_ldg(c->stream, walk, c->block, q, scene_table(c, skin_path), c->buf[b], sk, c->cell_start, c->gd,
                           walk_record(c), c->ncount, c->closure_on ? c->feat : nullptr);
        }
    }
    c->list_valid = walk_records(walk);
    return check_launch(c, "boundary");
}
int stage_governing(zsph_ctx *c, const zs::Prm &q, int b, Prof *pf = nullptr, bool skin_path = false) {
    if (c->sorted_slot < 0) return set_err(c, ZSPH_E_STATE, "zsph_governing called before zsph_sort_data");
    {
        Scope sc(pf, ZSPH_ST_GOVERNING);
        // the recording walks need this half step's record; otherwise walk the cells
        const int ew = eff_walk(c, skin_path);
        const int walk = (walk_records(ew) && !c->list_valid) ? zs::WALK_MASK : ew;
        if (c->closure_on)  // scale factors of the fluid particles [0, cs[G])
            zs::launch_mlp_infer(c->stream, c->mlp_P, c->mlp_pack, c->mlp_aux, c->feat, 0, c->cell_start + q.G,
                                 c->scale, zs::mlp_infer_blocks(q.N));
        zs::launch_governing(c->stream, walk, c->block, q, scene_table(c, skin_path), c->buf[b], c->key[c->sorted_slot], c->cell_start,
                             c->gd, c->acc, walk_record(c), c->ncount, c->closure_on ? 1 : 0, c->scale, c->terms,
                             c->has_body || c->has_piston, c->xref);
    }
    return check_launch(c, "governing");
}

// Integrator of one half step on buffer b: the body's force sum and FP64
// update, the fluid's predictor or corrector, then the moving dummies placed
// at the new time.  Without moving dummies only the fluid kernel runs.
int stage_integrate(zsph_ctx *c, const zs::Prm &q, int b, bool predictor, Prof *pf = nullptr,
                    const zs::SkinCheck &chk = zs::SkinCheck{}) {
    Scope sc(pf, ZSPH_ST_INTEGRATE);
    const bool moving = c->has_body || c->has_piston;
    if (c->has_body)
        zs::launch_body_reduce(c->stream, q, c->buf[b], c->cell_start, c->acc, c->body, c->body_part);
    if (moving) zs::launch_body_update(c->stream, q, c->body, c->body_part, predictor);
    const zs::Prm *sp = scene_table(c, chk.xref != nullptr);  // skin mode's half step checks displacements
    if (predictor)
        zs::launch_predict(c->stream, q, sp, c->buf[b], c->cell_start, c->acc, chk);
    else
        zs::launch_correct(c->stream, q, sp, c->buf[b], c->cell_start, c->acc, chk);
    if (moving) zs::launch_kinematics(c->stream, q, c->buf[b], c->cell_start, c->body, chk);
    return check_launch(c, predictor ? "prediction" : "correction");
}

// Host clock of a half step while no dummies move (the device keeps it otherwise).
void host_time(zsph_ctx *c, bool predictor) {
    if (c->has_body || c->has_piston) return;
    if (predictor) c->t0 = c->t;
    c->t = c->t0 + (predictor ? (double)c->q.half_dt : (double)c->q.dt);
}

int half_step(zsph_ctx *c, int from, int to, bool predictor, Prof *pf = nullptr) {
    int r;
    if ((r = stage_hash(c, from, pf))) return r;
    if ((r = stage_sort(c, pf))) return r;
    if ((r = stage_sort_data(c, from, to, pf))) return r;
    if ((r = stage_boundary(c, c->q, to, pf))) return r;
    if ((r = stage_governing(c, c->q, to, pf))) return r;
    return stage_integrate(c, c->q, to, predictor, pf);
}

// Skin mode's half step, always on buffer 0.  The rebuild kernels (hash on
// the skin grid, sort, cell table, reorder into buffer 1 and back with the
// reference positions) run only when the rebuild flag is set; the passes use
// the skin grid; the integrator sets the flag again when a particle has
// moved more than s/2 since the rebuild or left the grid.
// Skin mode's rebuild kernels (hash on the skin grids, sort, cell table,
// reorder into buffer 1 and back with the reference positions), gated on the
// device's rebuild flag, which they then count and clear.
int skin_rebuild(zsph_ctx *c, Prof *pf = nullptr) {
    const zs::Prm &q = c->qs;
    const int *gate = c->rebuild;
    {
        Scope sc(pf, ZSPH_ST_HASH);
        zs::launch_hash(c->stream, q, scene_table(c, true), c->buf[0], c->key[0], c->idx[0], c->killed, gate);
    }
    int slot;
    {
        Scope sc(pf, ZSPH_ST_SORT);
        slot = zs::launch_sort(c->stream, q.N, c->keybits, c->radix_bits, c->sort_tile, c->key, c->idx, c->hist,
                               true, gate);
    }
    if (slot < 0) return set_err(c, ZSPH_E_ARG, "unsupported radix width %d / tile %d", c->radix_bits, c->sort_tile);
    {
        Scope sc(pf, ZSPH_ST_CELLS);
        zs::launch_cell_start(c->stream, c->key[slot], q.N, q.G, q.flag, c->cell_start, c->cs_scratch, gate);
        if (c->tile_rows && c->wtab)
            zs::launch_wave_table(c->stream, q, c->cell_start, c->tile_rows, c->wtab, c->wt_cap, gate);
    }
    {
        Scope sc(pf, ZSPH_ST_REORDER);
        zs::launch_reorder(c->stream, q.N, c->idx[slot], c->buf[0], c->buf[1], gate);
        zs::launch_skin_commit(c->stream, q.N, c->buf[1], c->buf[0], c->xref, gate);
        zs::launch_skin_flag(c->stream, c->rebuild, c->rebuilds, c->rebuild + 1);
    }
    if (int r = check_launch(c, "skin rebuild")) return r;
    c->sorted_slot = slot;
    c->skin_slot = slot;
    c->list_valid = false;
    return ZSPH_OK;
}

int skin_half_step(zsph_ctx *c, bool predictor, Prof *pf = nullptr) {
    zs::Prm q = c->qs;
    if (c->tile_rows && c->wtab) {  // the neighbour passes take the wave table's order
        q.wtab = c->wtab;
        q.wt_cap = c->wt_cap;
        q.wt_rows = c->tile_rows;
    }
    int r;
    if ((r = skin_rebuild(c, pf))) return r;
    if ((r = stage_boundary(c, q, 0, pf, true))) return r;
    if ((r = stage_governing(c, q, 0, pf, true))) return r;
    zs::SkinCheck chk;
    chk.xref = c->xref;
    chk.flag = c->rebuild;
    chk.lim_h = 0.5f * c->skin * 0.999f;
    return stage_integrate(c, q, 0, predictor, pf, chk);
}

// Skin mode: make the next half step rebuild (the state in buffer 0 changed
// outside zsph_step, or its structures describe another grid).
int skin_force_rebuild(zsph_ctx *c) {
    static const int one = 1;
    if (!c->rebuild) return ZSPH_OK;
    HIPCHK(c, hipMemcpyAsync(c->rebuild, &one, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->skin_valid = true;
    c->skin_pending = true;
    return ZSPH_OK;
}

void drop_graph(zsph_ctx *c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
}

// ---- probes without allocation ----------------------------------------------
//
// Staging for zsph_probe_pressure_noalloc: the points (3 floats each) and the
// results (1 float each) of at most ZSPH_PROBE_NOALLOC_MAX probes.  Each GPU
// holds its own copy of this array until the process ends.  In host code the
// name `probe_scratch` is only a host-side stand-in: the device address of a
// GPU's copy comes from hipGetSymbolAddress, resolved once per device (with
// that device current) and kept in probe_scratch_dev.
static __device__ __align__(256) float probe_scratch[4 * ZSPH_PROBE_NOALLOC_MAX];

constexpr int kMaxDevices = 64;
float *probe_scratch_dev[kMaxDevices] = {};
std::mutex probe_scratch_mu;

// Device address of this device's probe_scratch (the caller has made `device`
// current); nullptr with the HIP error in *err on failure.
float *probe_scratch_ptr(int device, hipError_t *err) {
    *err = hipSuccess;
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(probe_scratch_mu);
    if (!probe_scratch_dev[device]) {
        void *ptr = nullptr;
        *err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(probe_scratch));
        if (*err != hipSuccess) return nullptr;
        probe_scratch_dev[device] = static_cast<float *>(ptr);
    }
    return probe_scratch_dev[device];
}

bool valid_params(const zsph_params *p) {
    if (p->ptc_num < 1 || p->ptc_num >= zs::MAX_PARTICLES || p->grid_xdim < 1 || p->grid_ydim < 1 || p->grid_zdim < 1 || !(p->h > 0.0f) ||
        !(p->grid_size > 0.0f) || p->grid_factor < 1 || (p->dim != 2 && p->dim != 3) || !(p->gamma > 0.0f) ||
        !(p->rc2 > 0.0f) || !(p->rho0 > 0.0f) || !(p->cs2 > 0.0f))
        return false;
    // at most 2^29 cells: the cell table's 2G + 2 entries and its indices stay in int
    const long long G = (long long)p->grid_xdim * p->grid_ydim * p->grid_zdim;
    return G <= (1ll << 29) && G - 1 == (long long)p->grid_hash_max;
}

// Context for N particles over G cells (all scenes), c->p / c->q already set.
int create_ctx(zsph_ctx *c, int device, size_t N, size_t G, bool batch) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ZSPH_E_NODEV;
    if (device >= 0) {
        if (device >= ndev) return ZSPH_E_NODEV;
        if (hipSetDevice(device) != hipSuccess) return ZSPH_E_HIP;
    }
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return ZSPH_E_HIP;
    {  // the device address of this device's static probe staging array (no allocation)
        hipError_t e = hipSuccess;
        c->probe_scratch = probe_scratch_ptr(c->device, &e);
        if (!c->probe_scratch) return set_err(c, ZSPH_E_HIP, "probe scratch address: %s", hipGetErrorString(e));
    }
    int r;
    for (int b = 0; b < 2; ++b) {
        // posr carries zs::POS_PAD zeroed spare elements read by the neighbour walks' candidate rounds
        if ((r = dalloc(c, &c->buf[b].posr, N + zs::POS_PAD)) || (r = dalloc(c, &c->buf[b].velp, N)) ||
            (r = dalloc(c, &c->buf[b].tposr, N)) || (r = dalloc(c, &c->buf[b].tvelp, N)) ||
            (r = dalloc(c, &c->buf[b].type, N)) || (r = dalloc(c, &c->buf[b].table, N)))
            return r;
        if (batch && (r = dalloc(c, &c->buf[b].scene, N))) return r;
        if ((r = dalloc(c, &c->key[b], N)) || (r = dalloc(c, &c->idx[b], N))) return r;
    }
    // sort keys: cell bits + the non-fluid flag bit above them
    int kb = 1;
    while ((1u << kb) < c->q.flag) ++kb;
    ++kb;
    c->keybits = kb;
    // fewest passes of at most 11-bit digits, spread evenly (19 bits: 2 x 10)
    const int passes = (kb + 10) / 11;
    c->radix_bits = std::max(8, (kb + passes - 1) / passes);
    if ((r = dalloc(c, &c->cell_start, 2 * (G + 1))) || (r = dalloc(c, &c->cs_scratch, zs::cell_start_scratch((int)G))) || (r = dalloc(c, &c->acc, N)) || (r = dalloc(c, &c->gd, N)) ||
        (r = dalloc(c, &c->hist, zs::radix_hist_words((int)N, 11))) || (r = dalloc(c, &c->killed, 1)) ||
        (r = dalloc(c, &c->diag_part, (size_t)zs::diag_blocks((int)N) * zs::DIAG_D)) ||
        (r = dalloc(c, &c->diag_max, (size_t)zs::diag_blocks((int)N) * 2)) || (r = dalloc(c, &c->body, 1)) ||
        (r = dalloc(c, &c->body_part, (size_t)zs::BODY_BLOCKS * zs::BODY_D)))
        return r;
    const size_t S = (size_t)c->nscenes;
    if ((r = dalloc(c, &c->scene_part, S * zs::SCENE_D))) return r;
    if (batch) {
        if ((r = dalloc(c, &c->scene_prm, S))) return r;
        HIPCHK(c, hipMemcpy(c->scene_prm, c->scene_host.data(), S * sizeof(zs::Prm), hipMemcpyHostToDevice));
    }
    return ZSPH_OK;
}

}  // namespace

extern "C" {

int zsph_create(const zsph_params *p, int device, zsph_ctx **out) {
    if (!p || !out) return ZSPH_E_ARG;
    *out = nullptr;
    if (!valid_params(p)) return ZSPH_E_ARG;
    zsph_ctx *c = new zsph_ctx();
    c->p = *p;
    c->q = zs::make_prm(*p);
    c->scene_off = {0, p->ptc_num};
    c->tile_rows = 16;  // skin mode's wave order (one scene)
    const int r = create_ctx(c, device, (size_t)p->ptc_num, (size_t)c->q.G, false);
    if (r) {
        zsph_destroy(c);
        return r;
    }
    *out = c;
    return ZSPH_OK;
}

int zsph_create_batch(const zsph_params *scenes, int nscenes, int device, zsph_ctx **out) {
    if (!scenes || !out || nscenes < 1) return ZSPH_E_ARG;
    *out = nullptr;
    long long N = 0, G = 0;
    for (int k = 0; k < nscenes; ++k) {
        if (!valid_params(&scenes[k])) return ZSPH_E_ARG;
        N += scenes[k].ptc_num;
        G += (long long)scenes[k].grid_xdim * scenes[k].grid_ydim * scenes[k].grid_zdim;
    }
    if (N >= zs::MAX_PARTICLES || G > (1ll << 29)) return ZSPH_E_ARG;
    zsph_ctx *c = new zsph_ctx();
    c->nscenes = nscenes;
    int cb = 1;
    while (cb < 30 && (1ll << cb) < G) ++cb;
    const uint32_t flag = 1u << cb;
    long long base = 0, off = 0;
    c->scene_off.push_back(0);
    for (int k = 0; k < nscenes; ++k) {
        zs::Prm sq = zs::make_prm(scenes[k]);
        const long long Gk = sq.G;
        sq.G = (int)G;
        sq.flag = flag;
        sq.N = (int)N;
        sq.cell_base = (int)base;
        base += Gk;
        off += scenes[k].ptc_num;
        c->scene_host.push_back(sq);
        c->scene_off.push_back((int)off);
    }
    // context-wide view: scene 0's constants with the batch's sizes
    c->p = scenes[0];
    c->p.ptc_num = (int)N;
    c->q = c->scene_host[0];
    c->q.cell_base = 0;
    const int r = create_ctx(c, device, (size_t)N, (size_t)G, true);
    if (r) {
        zsph_destroy(c);
        return r;
    }
    *out = c;
    return ZSPH_OK;
}

void zsph_destroy(zsph_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    for (zsph_frame &f : c->frames) {
        if (f.snap_done) (void)hipEventDestroy(f.snap_done);
        if (f.copy_done) (void)hipEventDestroy(f.copy_done);
        if (f.host) (void)hipHostFree(f.host);  // f.dev is one of c->allocs
    }
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (void *d : c->allocs) (void)hipFree(d);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int zsph_sync(zsph_ctx *c) {
    if (!c) return ZSPH_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    return ZSPH_OK;
}

const char *zsph_last_error(zsph_ctx *c) { return c ? c->err.c_str() : "null context"; }

int zsph_upload(zsph_ctx *c, int b, const zsph_host_state *st) {
    if (!c || !st || !valid_buf(b) || !st->pos || !st->type) return set_err(c, ZSPH_E_ARG, "zsph_upload: bad argument");
    const int N = c->q.N;
    // A batch's particle belongs to the scene of its original index i + table[i]
    // (the upload order of zsph_create_batch's scenes): a downloaded state in
    // cell order keeps its scenes.  Every original index must lie in [0, N).
    std::vector<int> sid;
    if (c->buf[b].scene) {
        sid.resize(N);
        for (int i = 0; i < N; ++i) {
            const long long o = (long long)i + (st->table ? st->table[i] : 0);
            if (o < 0 || o >= N)
                return set_err(c, ZSPH_E_ARG, "zsph_upload: particle %d has original index %lld outside [0, %d)", i,
                               o, N);
            sid[i] = (int)(std::upper_bound(c->scene_off.begin(), c->scene_off.end(), (int)o) -
                           c->scene_off.begin()) - 1;
        }
    }
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    std::vector<float4> posr(N), velp(N);
    for (int i = 0; i < N; ++i) {
        posr[i] = make_float4(st->pos[3 * i], st->pos[3 * i + 1], st->pos[3 * i + 2],
                              st->rhop ? st->rhop[2 * i] : c->q.rho0);
        velp[i] = st->vel ? make_float4(st->vel[3 * i], st->vel[3 * i + 1], st->vel[3 * i + 2], 0.0f)
                          : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        velp[i].w = st->rhop ? st->rhop[2 * i + 1] : 0.0f;
    }
    const zs::Buf &d = c->buf[b];
    HIPCHK(c, hipMemcpyAsync(d.posr, posr.data(), N * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d.velp, velp.data(), N * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d.tposr, posr.data(), N * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d.tvelp, velp.data(), N * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d.type, st->type, N * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (st->table) {
        HIPCHK(c, hipMemcpyAsync(d.table, st->table, N * sizeof(int), hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(c, hipMemsetAsync(d.table, 0, N * sizeof(int), c->stream));
    }
    if (st->acc_drhodt)
        HIPCHK(c, hipMemcpyAsync(c->acc, st->acc_drhodt, N * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    if (d.scene) HIPCHK(c, hipMemcpyAsync(d.scene, sid.data(), N * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->hashed = 0;
    c->sorted_slot = -1;
    c->cells_buf = -1;
    if (b == 0 && (c->has_body || c->has_piston)) {
        // new particles in buffer 0: the moving-dummy set-up no longer applies
        zs::DevBody hb;
        if (int e_ = copy_sync(c, &hb, c->body, sizeof(hb), hipMemcpyDeviceToHost)) return e_;
        c->t = hb.t;
        c->t0 = hb.t0;
        hb.has_body = hb.has_piston = 0;
        if (int e_ = copy_sync(c, c->body, &hb, sizeof(hb), hipMemcpyHostToDevice)) return e_;
        c->has_body = c->has_piston = false;
        drop_graph(c);
    }
    return ZSPH_OK;
}

int zsph_download(zsph_ctx *c, int b, zsph_host_state *st, int fields) {
    if (!c || !st || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_download: bad argument");
    const int N = c->q.N;
    const zs::Buf &d = c->buf[b];
    std::vector<float4> tmp(N);
    if ((fields & ZSPH_F_POS) || (fields & ZSPH_F_RHOP)) {
        HIPCHK(c, hipMemcpyAsync(tmp.data(), d.posr, N * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < N; ++i) {
            if ((fields & ZSPH_F_POS) && st->pos) {
                st->pos[3 * i] = tmp[i].x; st->pos[3 * i + 1] = tmp[i].y; st->pos[3 * i + 2] = tmp[i].z;
            }
            if ((fields & ZSPH_F_RHOP) && st->rhop) st->rhop[2 * i] = tmp[i].w;
        }
    }
    if ((fields & ZSPH_F_VEL) || (fields & ZSPH_F_RHOP)) {
        HIPCHK(c, hipMemcpyAsync(tmp.data(), d.velp, N * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < N; ++i) {
            if ((fields & ZSPH_F_VEL) && st->vel) {
                st->vel[3 * i] = tmp[i].x; st->vel[3 * i + 1] = tmp[i].y; st->vel[3 * i + 2] = tmp[i].z;
            }
            if ((fields & ZSPH_F_RHOP) && st->rhop) st->rhop[2 * i + 1] = tmp[i].w;
        }
    }
    if ((fields & ZSPH_F_ACC) && st->acc_drhodt)
        HIPCHK(c, hipMemcpyAsync(st->acc_drhodt, c->acc, N * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    if ((fields & ZSPH_F_TYPE) && st->type)
        HIPCHK(c, hipMemcpyAsync(st->type, d.type, N * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if ((fields & ZSPH_F_TABLE) && st->table)
        HIPCHK(c, hipMemcpyAsync(st->table, d.table, N * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if ((fields & ZSPH_F_HASH) && st->hash) {
        const uint32_t *src = c->key[c->sorted_slot >= 0 ? c->sorted_slot : 0];
        HIPCHK(c, hipMemcpyAsync(st->hash, src, N * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((fields & ZSPH_F_HASH) && st->hash) {
        // the cell index only (the sort key's non-fluid flag is internal)
        const int mask = (int)(c->q.flag - 1u);
        for (int i = 0; i < N; ++i) st->hash[i] &= mask;
    }
    return ZSPH_OK;
}

namespace {
constexpr size_t kFrameBytesPerParticle = 3 * 16 + 3 * 4;  // posr, velp, acc; type, table, key

// Allocates the copy stream and the frame slots (snapshot + pinned copy) at first use.
int frames_init(zsph_ctx *c) {
    if (c->copy_stream) return ZSPH_OK;
    const size_t bytes = (size_t)c->q.N * kFrameBytesPerParticle;
    for (int k = 0; k < ZSPH_FRAME_SLOTS; ++k) {
        zsph_frame &f = c->frames[k];
        f.ctx = c;
        char *d = nullptr;
        if (int r = dalloc(c, &d, bytes)) return r;
        f.dev = d;
        if (hipHostMalloc(&f.host, bytes, hipHostMallocDefault) != hipSuccess) {
            f.host = nullptr;
            return set_err(c, ZSPH_E_NOMEM, "zsph_download_async: %zu bytes of pinned host memory", bytes);
        }
        HIPCHK(c, hipEventCreateWithFlags(&f.snap_done, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&f.copy_done, hipEventDisableTiming));
    }
    HIPCHK(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    return ZSPH_OK;
}

// Byte offsets of the arrays inside a frame's snapshot.
struct FrameLayout {
    size_t posr, velp, acc, type, table, key;
    explicit FrameLayout(size_t n)
        : posr(0), velp(16 * n), acc(32 * n), type(48 * n), table(52 * n), key(56 * n) {}
};
}  // namespace

int zsph_download_async(zsph_ctx *c, int b, int fields, zsph_frame **out) {
    if (!c || !out || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_download_async: bad argument");
    *out = nullptr;
    if (int r = frames_init(c)) return r;
    zsph_frame *f = nullptr;
    {  // a free slot, else wait for a writer to release one
        std::unique_lock<std::mutex> lock(c->frame_mu);
        for (;;) {
            for (zsph_frame &x : c->frames)
                if (!x.busy.load() && (!f || x.seq < f->seq)) f = &x;
            if (f) break;
            c->frame_cv.wait(lock);
        }
        f->busy.store(1);
        f->seq = ++c->frame_seq;
    }
    const size_t N = (size_t)c->q.N;
    const FrameLayout L(N);
    const zs::Buf &d = c->buf[b];
    char *dev = static_cast<char *>(f->dev), *host = static_cast<char *>(f->host);
    f->fields = fields;
    f->n = (int)N;
    f->cell_mask = c->q.flag - 1u;
    struct Part { bool on; const void *src; size_t off, bytes; };
    const Part parts[6] = {
        {(fields & (ZSPH_F_POS | ZSPH_F_RHOP)) != 0, d.posr, L.posr, 16 * N},
        {(fields & (ZSPH_F_VEL | ZSPH_F_RHOP)) != 0, d.velp, L.velp, 16 * N},
        {(fields & ZSPH_F_ACC) != 0, c->acc, L.acc, 16 * N},
        {(fields & ZSPH_F_TYPE) != 0, d.type, L.type, 4 * N},
        {(fields & ZSPH_F_TABLE) != 0, d.table, L.table, 4 * N},
        {(fields & ZSPH_F_HASH) != 0, c->key[c->sorted_slot >= 0 ? c->sorted_slot : 0], L.key, 4 * N}};
    hipError_t e = hipSuccess;
    // snapshot in stream order after the issued steps, then the host copy on the copy stream
    for (const Part &p : parts)
        if (p.on && e == hipSuccess) e = hipMemcpyAsync(dev + p.off, p.src, p.bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipEventRecord(f->snap_done, c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->copy_stream, f->snap_done, 0);
    for (const Part &p : parts)
        if (p.on && e == hipSuccess)
            e = hipMemcpyAsync(host + p.off, dev + p.off, p.bytes, hipMemcpyDeviceToHost, c->copy_stream);
    if (e == hipSuccess) e = hipEventRecord(f->copy_done, c->copy_stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(c->copy_stream);
        {
            std::lock_guard<std::mutex> lock(c->frame_mu);
            f->busy.store(0);
        }
        c->frame_cv.notify_all();
        return set_err(c, ZSPH_E_HIP, "zsph_download_async: %s", hipGetErrorString(e));
    }
    *out = f;
    return ZSPH_OK;
}

int zsph_frame_wait(zsph_frame *f, zsph_host_state *st) {
    if (!f || !f->ctx || !f->busy.load() || !st) return ZSPH_E_ARG;
    zsph_ctx *c = f->ctx;
    const hipError_t e = hipEventSynchronize(f->copy_done);
    if (e == hipSuccess) {
        const size_t N = (size_t)f->n;
        const FrameLayout L(N);
        const char *host = static_cast<const char *>(f->host);
        const float4 *posr = reinterpret_cast<const float4 *>(host + L.posr);
        const float4 *velp = reinterpret_cast<const float4 *>(host + L.velp);
        const int fields = f->fields;
        for (size_t i = 0; i < N; ++i) {  // the same unpacking as zsph_download
            if ((fields & ZSPH_F_POS) && st->pos) {
                st->pos[3 * i] = posr[i].x; st->pos[3 * i + 1] = posr[i].y; st->pos[3 * i + 2] = posr[i].z;
            }
            if ((fields & ZSPH_F_VEL) && st->vel) {
                st->vel[3 * i] = velp[i].x; st->vel[3 * i + 1] = velp[i].y; st->vel[3 * i + 2] = velp[i].z;
            }
            if ((fields & ZSPH_F_RHOP) && st->rhop) {
                st->rhop[2 * i] = posr[i].w;
                st->rhop[2 * i + 1] = velp[i].w;
            }
        }
        if ((fields & ZSPH_F_ACC) && st->acc_drhodt) std::memcpy(st->acc_drhodt, host + L.acc, 16 * N);
        if ((fields & ZSPH_F_TYPE) && st->type) std::memcpy(st->type, host + L.type, 4 * N);
        if ((fields & ZSPH_F_TABLE) && st->table) std::memcpy(st->table, host + L.table, 4 * N);
        if ((fields & ZSPH_F_HASH) && st->hash) {
            const uint32_t *key = reinterpret_cast<const uint32_t *>(host + L.key);
            for (size_t i = 0; i < N; ++i) st->hash[i] = (int)(key[i] & f->cell_mask);
        }
    }
    {
        std::lock_guard<std::mutex> lock(c->frame_mu);
        f->busy.store(0);
    }
    c->frame_cv.notify_all();
    return e == hipSuccess ? ZSPH_OK : ZSPH_E_HIP;
}

int zsph_hash(zsph_ctx *c, int b) {
    if (!c || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_hash: bad buffer");
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    return stage_hash(c, b);
}
int zsph_sort_index(zsph_ctx *c) {
    if (!c) return ZSPH_E_ARG;
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    return stage_sort(c);
}
int zsph_sort_data(zsph_ctx *c, int from, int to) {
    if (!c || !valid_buf(from) || !valid_buf(to)) return set_err(c, ZSPH_E_ARG, "zsph_sort_data: bad buffer");
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    return stage_sort_data(c, from, to);
}
int zsph_boundary(zsph_ctx *c, int b) {
    if (!c || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_boundary: bad buffer");
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    return stage_boundary(c, c->q, b);
}
int zsph_governing(zsph_ctx *c, int b) {
    if (!c || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_governing: bad buffer");
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    return stage_governing(c, c->q, b);
}
int zsph_prediction(zsph_ctx *c, int b) {
    if (!c || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_prediction: bad buffer");
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    if (c->sorted_slot < 0) return set_err(c, ZSPH_E_STATE, "zsph_prediction called before zsph_sort_data");
    host_time(c, true);
    return stage_integrate(c, c->q, b, true);
}
int zsph_correction(zsph_ctx *c, int b) {
    if (!c || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_correction: bad buffer");
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    if (c->sorted_slot < 0) return set_err(c, ZSPH_E_STATE, "zsph_correction called before zsph_sort_data");
    host_time(c, false);
    return stage_integrate(c, c->q, b, false);
}

int zsph_step(zsph_ctx *c, int nsteps) {
    if (!c || nsteps < 0) return set_err(c, ZSPH_E_ARG, "zsph_step: bad argument");
    if (nsteps == 0 || c->q.N == 0) return ZSPH_OK;
    const bool skin = c->skin > 0.0f;
    if (skin && !c->skin_valid) {  // the first step after any outside change rebuilds
        if (int r = skin_force_rebuild(c)) return r;
    }
    // one step: two half steps (skin mode: both on buffer 0)
    auto one_step = [&](bool host_clock) {
        if (host_clock) host_time(c, true);
        int r = skin ? skin_half_step(c, true) : half_step(c, 0, 1, true);
        if (host_clock) host_time(c, false);
        if (!r) r = skin ? skin_half_step(c, false) : half_step(c, 1, 0, false);
        return r;
    };
    auto finish = [&]() {
        // after a step the sorted tables describe buffer 0 (skin mode: on the skin grid)
        c->hashed = 0;
        if (skin) {
            c->cells_buf = -1;
            c->sorted_slot = -1;
            c->skin_valid = true;
            c->skin_pending = false;
        } else {
            c->cells_buf = 0;
            c->sorted_slot = (c->keybits + c->radix_bits - 1) / c->radix_bits % 2;
        }
    };
    if (!c->use_graph) {  // the same kernels, launched one by one
        for (int s = 0; s < nsteps; ++s)
            if (int r = one_step(true)) return r;
        finish();
        return ZSPH_OK;
    }
    if (!c->graph_exec) {
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int r = one_step(false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        if (e != hipSuccess) return set_err(c, ZSPH_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
        c->graph = g;
        e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) return set_err(c, ZSPH_E_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    }
    for (int s = 0; s < nsteps; ++s) {
        HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
        host_time(c, true);
        host_time(c, false);
    }
    finish();
    return ZSPH_OK;
}

int zsph_rebuilds(zsph_ctx *c, long long *n) {
    if (!c || !n) return ZSPH_E_ARG;
    *n = 0;
    if (!c->rebuilds) return ZSPH_OK;
    unsigned long long v = 0;
    if (int r = copy_sync(c, &v, c->rebuilds, sizeof(v), hipMemcpyDeviceToHost)) return r;
    *n = (long long)v;
    return ZSPH_OK;
}

namespace {
// Skin-mode neighbour grids for a skin of `skin` h: the context's (qs) and,
// for a batch, each scene's own (make_skin_prm on the scene's grid), laid out
// scene after scene in the key space like the case grids (cell_base), with
// the batch's particle count and sort-key flag (a skin grid has no more cells
// than its case grid, so the keys still fit below the flag).
int set_skin_grids(zsph_ctx *c, float skin) {
    if (c->nscenes == 1) {
        c->qs = zs::make_skin_prm(c->q, skin);
        return ZSPH_OK;
    }
    std::vector<zs::Prm> v;
    long long base = 0;
    for (const zs::Prm &p : c->scene_host) {
        zs::Prm sk = zs::make_skin_prm(p, skin);
        sk.cell_base = (int)base;
        base += sk.G;  // the scene's own skin cells
        v.push_back(sk);
    }
    for (zs::Prm &sk : v) sk.G = (int)base;
    int r;
    if (!c->scene_prm_skin && (r = dalloc(c, &c->scene_prm_skin, v.size()))) return r;
    HIPCHK(c, hipMemcpy(c->scene_prm_skin, v.data(), v.size() * sizeof(zs::Prm), hipMemcpyHostToDevice));
    c->scene_skin_host = v;
    zs::Prm qs = v[0];  // context-wide view, as c->q: scene 0's constants with the batch's sizes
    qs.cell_base = 0;
    qs.clk = c->q.clk;
    qs.clk_mask = c->q.clk_mask;
    c->qs = qs;
    return ZSPH_OK;
}
}  // namespace

int zsph_set_tuning(zsph_ctx *c, const zsph_tuning *t) {
    if (!c || !t) return ZSPH_E_ARG;
    const int block = t->block ? t->block : ZS_BLOCK;
    const int tile = t->sort_tile ? t->sort_tile : 16;
    if (t->walk < ZSPH_WALK_MASK || t->walk > ZSPH_WALK_SKIN || t->radix_bits < 8 || t->radix_bits > 11 ||
        (block != 128 && block != 256 && block != 512) || (tile != 8 && tile != 16) || !(t->skin >= 0.0f) ||
        t->skin > 1.0f)
        return set_err(c, ZSPH_E_ARG,
                       "zsph_set_tuning: walk %d / radix_bits %d / block %d / sort_tile %d / skin %g out of range",
                       t->walk, t->radix_bits, t->block, t->sort_tile, (double)t->skin);
    const int rows = t->tile_rows < 0 ? c->tile_rows : t->tile_rows;
    if (rows > 4096) return set_err(c, ZSPH_E_ARG, "zsph_set_tuning: tile_rows %d out of range", t->tile_rows);
    if (t->walk == ZSPH_WALK_SKIN && c->nscenes > 1)
        return set_err(c, ZSPH_E_ARG, "zsph_set_tuning: the recorded skin walk is for one-scene contexts");
    if (rows > 0 && c->nscenes > 1)
        return set_err(c, ZSPH_E_ARG, "zsph_set_tuning: the tiled wave order is for one-scene contexts");
    if (t->walk == c->walk && t->radix_bits == c->radix_bits && block == c->block && tile == c->sort_tile &&
        t->skin == c->skin && rows == c->tile_rows)
        return ZSPH_OK;
    int r;
    if (t->skin > 0.0f && !c->rebuild &&
        ((r = dalloc(c, &c->rebuild, 2)) || (r = dalloc(c, &c->rebuilds, 1)) ||
         (r = dalloc(c, &c->xref, (size_t)c->q.N))))
        return r;
    // the wave table: room for units of one row (any tile_rows fits); the case's
    // grid bounds the skin grids' rows and planes
    if (t->skin > 0.0f && rows > 0 && !c->wtab &&
        (r = dalloc(c, &c->wtab, (size_t)zs::wave_table_cap(c->q.N, c->q.Y, c->q.Z, 1) + 1)))
        return r;
    if (rows != c->tile_rows) c->skin_valid = false;  // the table is built by the next rebuild
    c->tile_rows = rows;
    c->wt_cap = rows > 0 ? zs::wave_table_cap(c->q.N, c->q.Y, c->q.Z, rows) : 0;
    if (t->skin != c->skin) {
        if (t->skin > 0.0f && (r = set_skin_grids(c, t->skin))) return r;
        c->skin = t->skin;
        c->skin_valid = false;
    }
    if (walk_records(t->walk) && !c->ncount && (r = dalloc(c, &c->ncount, (size_t)c->q.N))) return r;
    if (t->walk == ZSPH_WALK_LIST && !c->nlist && (r = dalloc(c, &c->nlist, zs::list_words(c->q.N)))) return r;
    if ((t->walk == ZSPH_WALK_HITS || t->walk == ZSPH_WALK_SKIN) && !c->hits) {
        int k = c->q.k;  // the widest search of any scene
        for (const zs::Prm &p : c->scene_host) k = std::max(k, p.k);
        if ((r = dalloc(c, &c->hits, (size_t)zs::hits_slots(k) * (size_t)c->q.N))) return r;
    }
    if (t->walk == ZSPH_WALK_SKIN && !c->skin_hits &&
        (r = dalloc(c, &c->skin_hits, (size_t)zs::hits_slots(c->q.k) * (size_t)c->q.N)))
        return r;
    if (t->walk == ZSPH_WALK_SKIN && c->walk != ZSPH_WALK_SKIN) c->skin_valid = false;  // record at the next step
    c->list_valid = false;
    c->walk = t->walk;
    c->radix_bits = t->radix_bits;
    c->block = block;
    c->sort_tile = tile;
    drop_graph(c);  // the captured step embeds the old launches
    c->sorted_slot = -1;
    c->hashed = 0;
    return ZSPH_OK;
}

int zsph_get_tuning(zsph_ctx *c, zsph_tuning *t) {
    if (!c || !t) return ZSPH_E_ARG;
    t->walk = c->walk;
    t->radix_bits = c->radix_bits;
    t->block = c->block;
    t->sort_tile = c->sort_tile;
    t->skin = c->skin;
    t->tile_rows = c->tile_rows;
    return ZSPH_OK;
}

int zsph_autotune(zsph_ctx *c, int steps, zsph_tuning *best, float *best_ms) {
    if (!c || steps < 1) return set_err(c, ZSPH_E_ARG, "zsph_autotune: bad argument");
    const size_t N = (size_t)c->q.N;
    if (N == 0) return ZSPH_OK;
    // save buffer 0 and the removal counter; every trial starts from them
    const zs::Buf &b0 = c->buf[0];
    const int narr = b0.scene ? 7 : 6;
    void *const arrays[7] = {b0.posr, b0.velp, b0.tposr, b0.tvelp, b0.type, b0.table, b0.scene};
    const size_t bytes[7] = {N * 16, N * 16, N * 16, N * 16, N * 4, N * 4, b0.scene ? N * 4 : 0};
    size_t total = 256 + sizeof(zs::DevBody);  // + removal counter (padded) + body state
    for (size_t k : bytes) total += k;
    const double t_keep = c->t, t0_keep = c->t0;
    char *save = nullptr;
    HIPCHK(c, hipMalloc(&save, total));
    auto copy_all = [&](bool to_save) -> hipError_t {
        size_t off = 0;
        hipError_t e = hipSuccess;
        for (int k = 0; k < narr && e == hipSuccess; ++k) {
            e = to_save ? hipMemcpyAsync(save + off, arrays[k], bytes[k], hipMemcpyDeviceToDevice, c->stream)
                        : hipMemcpyAsync(arrays[k], save + off, bytes[k], hipMemcpyDeviceToDevice, c->stream);
            off += bytes[k];
        }
        if (e == hipSuccess)
            e = to_save ? hipMemcpyAsync(save + off, c->killed, 4, hipMemcpyDeviceToDevice, c->stream)
                        : hipMemcpyAsync(c->killed, save + off, 4, hipMemcpyDeviceToDevice, c->stream);
        off += 256;
        if (e == hipSuccess)
            e = to_save ? hipMemcpyAsync(save + off, c->body, sizeof(zs::DevBody), hipMemcpyDeviceToDevice, c->stream)
                        : hipMemcpyAsync(c->body, save + off, sizeof(zs::DevBody), hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        return e;
    };
    hipError_t e = copy_all(true);
    if (e != hipSuccess) {
        (void)hipFree(save);
        return set_err(c, ZSPH_E_HIP, "autotune save: %s", hipGetErrorString(e));
    }
    const int passes = (c->keybits + 10) / 11;  // fewest passes of <= 11-bit digits
    const int even = std::max(8, (c->keybits + passes - 1) / passes);
    zsph_tuning keep{c->walk, c->radix_bits, c->block, c->sort_tile, c->skin, c->tile_rows};
    float best_t = 1e30f;
    int r = ZSPH_OK;
    // times one candidate from the saved state over n steps; keeps it if it is the fastest so far
    auto trial = [&](const zsph_tuning &t, int n) {
        if ((r = zsph_set_tuning(c, &t))) return;
        if ((e = copy_all(false)) != hipSuccess) {
            r = set_err(c, ZSPH_E_HIP, "autotune restore: %s", hipGetErrorString(e));
            return;
        }
        c->skin_valid = false;  // the restored state rebuilds first
        float ms = 0.0f;
        if ((r = zsph_time_steps(c, n, &ms))) return;
        ms *= (float)steps / (float)n;
        if (ms < best_t) {
            best_t = ms;
            keep = t;
        }
    };
    // coordinate search: walk x digit width, then block size, then sort tile
    const zsph_tuning start = keep;
    const int nskin = start.skin > 0.0f ? 4 * steps : steps;  // skin mode: time over more steps
    const int last_walk = (start.skin > 0.0f && c->nscenes == 1) ? ZSPH_WALK_SKIN : ZSPH_WALK_LIST;
    for (int walk = ZSPH_WALK_MASK; walk <= last_walk && !r; ++walk) {
        trial({walk, even, start.block, start.sort_tile, start.skin, start.tile_rows}, nskin);
        if (even != 8 && !r) trial({walk, 8, start.block, start.sort_tile, start.skin, start.tile_rows}, nskin);
    }
    const int blocks[3] = {128, 256, 512};
    zsph_tuning base = keep;
    for (int k = 0; k < 3 && !r; ++k)
        if (blocks[k] != base.block)
            trial({base.walk, base.radix_bits, blocks[k], base.sort_tile, base.skin, base.tile_rows}, nskin);
    base = keep;
    for (int tile = 8; tile <= 16 && !r; tile *= 2)
        if (tile != base.sort_tile)
            trial({base.walk, base.radix_bits, base.block, tile, base.skin, base.tile_rows}, nskin);
    if (start.skin > 0.0f) {
        base = keep;
        const float skins[4] = {0.05f, 0.1f, 0.2f, 0.3f};
        for (int k = 0; k < 4 && !r; ++k)
            if (skins[k] != base.skin)
                trial({base.walk, base.radix_bits, base.block, base.sort_tile, skins[k], base.tile_rows}, nskin);
        if (c->nscenes == 1 && !r) {  // the wave order: cell order against units of 16 rows
            base = keep;
            trial({base.walk, base.radix_bits, base.block, base.sort_tile, base.skin, base.tile_rows ? 0 : 16}, nskin);
        }
    }
    if (!r) r = zsph_set_tuning(c, &keep);
    e = copy_all(false);
    (void)hipFree(save);
    c->t = t_keep;
    c->t0 = t0_keep;
    c->sorted_slot = -1;  // buffer 0 is back to its unsorted starting state
    c->hashed = 0;
    c->cells_buf = -1;
    c->skin_valid = false;
    if (r) return r;
    if (e != hipSuccess) return set_err(c, ZSPH_E_HIP, "autotune restore: %s", hipGetErrorString(e));
    if (best) *best = keep;
    if (best_ms) *best_ms = best_t;
    return ZSPH_OK;
}

int zsph_array_get(zsph_ctx *c, int which, int b, void *host, size_t bytes) {
    if (!c || !host || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_array_get: bad argument");
    const size_t N = (size_t)c->q.N;
    const void *src = nullptr;
    size_t need = 0;
    switch (which) {
    case ZSPH_A_POSR: src = c->buf[b].posr; need = N * 16; break;
    case ZSPH_A_VELP: src = c->buf[b].velp; need = N * 16; break;
    case ZSPH_A_TPOSR: src = c->buf[b].tposr; need = N * 16; break;
    case ZSPH_A_TVELP: src = c->buf[b].tvelp; need = N * 16; break;
    case ZSPH_A_TYPE: src = c->buf[b].type; need = N * 4; break;
    case ZSPH_A_TABLE: src = c->buf[b].table; need = N * 4; break;
    case ZSPH_A_KEY: src = c->key[c->sorted_slot >= 0 ? c->sorted_slot : 0]; need = N * 4; break;
    case ZSPH_A_INDEX: src = c->idx[c->sorted_slot >= 0 ? c->sorted_slot : 0]; need = N * 4; break;
    case ZSPH_A_CELL_START: src = c->cell_start; need = 2 * ((size_t)c->q.G + 1) * 4; break;
    case ZSPH_A_ACC: src = c->acc; need = N * 16; break;
    case ZSPH_A_GD: src = c->gd; need = N * 16; break;
    default: return set_err(c, ZSPH_E_ARG, "zsph_array_get: unknown array %d", which);
    }
    if (bytes != need) return set_err(c, ZSPH_E_ARG, "zsph_array_get: %zu bytes given, %zu needed", bytes, need);
    if (need) HIPCHK(c, hipMemcpyAsync(host, src, need, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ZSPH_OK;
}

int zsph_array_set(zsph_ctx *c, int which, int b, const void *host, size_t bytes) {
    if (!c || !host || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_array_set: bad argument");
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    const size_t N = (size_t)c->q.N;
    void *dst = nullptr;
    size_t need = 0;
    switch (which) {
    case ZSPH_A_POSR: dst = c->buf[b].posr; need = N * 16; break;
    case ZSPH_A_VELP: dst = c->buf[b].velp; need = N * 16; break;
    case ZSPH_A_TPOSR: dst = c->buf[b].tposr; need = N * 16; break;
    case ZSPH_A_TVELP: dst = c->buf[b].tvelp; need = N * 16; break;
    case ZSPH_A_TYPE: dst = c->buf[b].type; need = N * 4; break;
    case ZSPH_A_TABLE: dst = c->buf[b].table; need = N * 4; break;
    case ZSPH_A_ACC: dst = c->acc; need = N * 16; break;
    case ZSPH_A_GD: dst = c->gd; need = N * 16; break;
    default: return set_err(c, ZSPH_E_ARG, "zsph_array_set: array %d is not writable", which);
    }
    if (bytes != need) return set_err(c, ZSPH_E_ARG, "zsph_array_set: %zu bytes given, %zu needed", bytes, need);
    if (need) HIPCHK(c, hipMemcpyAsync(dst, host, need, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ZSPH_OK;
}

int zsph_neighbor_stats(zsph_ctx *c, int b, int *count, uint64_t *digest) {
    if (!c || !valid_buf(b) || !count || !digest) return set_err(c, ZSPH_E_ARG, "zsph_neighbor_stats: bad argument");
    // after skin-mode steps: buffer 0 on the skin grid, the structures of the last rebuild
    const bool skin = c->skin > 0.0f && c->skin_valid && !c->skin_pending && c->sorted_slot < 0 && c->hashed == 0;
    if (skin && b != 0) return set_err(c, ZSPH_E_STATE, "zsph_neighbor_stats: skin-mode state is in buffer 0");
    if (!skin && c->sorted_slot < 0) return set_err(c, ZSPH_E_STATE, "zsph_neighbor_stats called before zsph_sort_data");
    if (skin) {
        // the structures the next half step's passes would use: its rebuild, if one is due
        // (some particle moved more than s/2 in the last half step), runs first
        int r = skin_rebuild(c);
        if (!r && c->walk == ZSPH_WALK_SKIN) r = skin_force_rebuild(c);  // its skin record is of the old order
        c->sorted_slot = -1;
        if (r) return r;
    }
    const size_t N = (size_t)c->q.N;
    int *dc = nullptr;
    uint64_t *dd = nullptr;
    HIPCHK(c, hipMalloc(&dc, N ? N * 4 : 4));
    hipError_t e = hipMalloc(&dd, N ? N * 8 : 8);
    if (e != hipSuccess) {
        (void)hipFree(dc);
        return set_err(c, ZSPH_E_NOMEM, "hipMalloc failed");
    }
    zs::launch_nbr_stats(c->stream, c->walk, skin ? c->qs : c->q, scene_table(c, skin), c->buf[b],
                         c->key[skin ? c->skin_slot : c->sorted_slot], c->cell_start, dc, dd);
    e = hipGetLastError();
    if (e == hipSuccess && N) e = hipMemcpyAsync(count, dc, N * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && N) e = hipMemcpyAsync(digest, dd, N * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dc);
    (void)hipFree(dd);
    if (e != hipSuccess) return set_err(c, ZSPH_E_HIP, "neighbour stats failed: %s", hipGetErrorString(e));
    return ZSPH_OK;
}

int zsph_time_steps(zsph_ctx *c, int nsteps, float *ms) {
    if (!c || !ms || nsteps < 1) return set_err(c, ZSPH_E_ARG, "zsph_time_steps: bad argument");
    int r = zsph_step(c, 1);  // make sure the graph exists before the timed region
    if (r) return r;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipEventRecord(e0, c->stream));
    r = zsph_step(c, nsteps);
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    HIPCHK(c, hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return r;
}

int zsph_profile_steps(zsph_ctx *c, int nsteps, float *stage_ms) {
    if (!c || !stage_ms || nsteps < 1) return set_err(c, ZSPH_E_ARG, "zsph_profile_steps: bad argument");
    Prof pf;
    pf.s = c->stream;
    int r = ZSPH_OK;
    const bool skin = c->skin > 0.0f;
    if (skin && !c->skin_valid) r = skin_force_rebuild(c);
    for (int s = 0; s < nsteps && !r; ++s) {
        host_time(c, true);
        r = skin ? skin_half_step(c, true, &pf) : half_step(c, 0, 1, true, &pf);
        host_time(c, false);
        if (!r) r = skin ? skin_half_step(c, false, &pf) : half_step(c, 1, 0, false, &pf);
    }
    if (skin && !r) {  // as after zsph_step
        c->hashed = 0;
        c->cells_buf = -1;
        c->sorted_slot = -1;
        c->skin_pending = false;
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    for (int k = 0; k < ZSPH_ST_COUNT; ++k) stage_ms[k] = 0.0f;
    for (auto &x : pf.ev) {
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, x.second.first, x.second.second);
        stage_ms[x.first] += t;
        (void)hipEventDestroy(x.second.first);
        (void)hipEventDestroy(x.second.second);
    }
    if (r) return r;
    if (e != hipSuccess) return set_err(c, ZSPH_E_HIP, "profile: %s", hipGetErrorString(e));
    return ZSPH_OK;
}

int zsph_diagnostics(zsph_ctx *c, int b, float front_zmax, zsph_diag *out) {
    if (!c || !out || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_diagnostics: bad argument");
    if (c->nscenes > 1) {  // scenes differ in particle mass: sum the per-scene figures
        std::vector<zsph_diag> per((size_t)c->nscenes);
        int r = zsph_scene_diagnostics(c, b, front_zmax, per.data());
        if (r) return r;
        std::memset(out, 0, sizeof(*out));
        out->front_x = out->zmax_fluid = -INFINITY;
        for (const zsph_diag &d : per) {
            out->n_fluid += d.n_fluid;
            out->mass_fluid += d.mass_fluid;
            out->kinetic += d.kinetic;
            for (int k = 0; k < 3; ++k) out->momentum[k] += d.momentum[k];
            out->front_x = std::max(out->front_x, d.front_x);
            out->zmax_fluid = std::max(out->zmax_fluid, d.zmax_fluid);
        }
        return zsph_killed(c, &out->n_killed);
    }
    std::memset(out, 0, sizeof(*out));
    out->front_x = out->zmax_fluid = -INFINITY;
    const int nb = c->q.N ? zs::diag_blocks(c->q.N) : 0;
    std::vector<double> part((size_t)nb * zs::DIAG_D);
    std::vector<float> mx((size_t)nb * 2);
    zs::launch_diag(c->stream, c->q, c->buf[b], front_zmax, c->diag_part, c->diag_max);
    int r = check_launch(c, "diagnostics");
    if (r) return r;
    if (nb) {
        HIPCHK(c, hipMemcpyAsync(part.data(), c->diag_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(mx.data(), c->diag_max, mx.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(&out->n_killed, c->killed, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double cnt = 0.0;
    for (int k = 0; k < nb; ++k) {
        const double *v = &part[(size_t)k * zs::DIAG_D];
        cnt += v[0];
        out->kinetic += v[1];
        out->momentum[0] += v[2];
        out->momentum[1] += v[3];
        out->momentum[2] += v[4];
        out->front_x = std::max(out->front_x, (double)mx[2 * k]);
        out->zmax_fluid = std::max(out->zmax_fluid, (double)mx[2 * k + 1]);
    }
    out->n_fluid = (int)cnt;
    out->mass_fluid = cnt * (double)c->p.m;
    return ZSPH_OK;
}

int zsph_scene_diagnostics(zsph_ctx *c, int b, float front_zmax, zsph_diag *out) {
    if (!c || !out || !valid_buf(b)) return set_err(c, ZSPH_E_ARG, "zsph_scene_diagnostics: bad argument");
    // after skin-mode steps the cell table is that of the last rebuild, on the
    // skin grids: it still gives each scene's particle ranges in buffer 0
    const bool skin = c->skin > 0.0f && c->skin_valid && !c->skin_pending && c->cells_buf < 0 && c->hashed == 0;
    if (c->cells_buf != b && !(skin && b == 0))
        return set_err(c, ZSPH_E_STATE, "zsph_scene_diagnostics: buffer %d is not in cell order (step or sort first)", b);
    const int S = c->nscenes;
    zs::launch_scene_diag(c->stream, skin ? c->qs : c->q, scene_table(c, skin), S, c->buf[b], c->cell_start,
                          front_zmax, c->scene_part);
    int r = check_launch(c, "scene diagnostics");
    if (r) return r;
    std::vector<double> part((size_t)S * zs::SCENE_D);
    if ((r = copy_sync(c, part.data(), c->scene_part, part.size() * sizeof(double), hipMemcpyDeviceToHost))) return r;
    for (int k = 0; k < S; ++k) {
        const double *v = &part[(size_t)k * zs::SCENE_D];
        const float m = S > 1 ? c->scene_host[k].m : c->q.m;
        zsph_diag &d = out[k];
        d.n_fluid = (int)v[0];
        d.n_killed = (int)v[7];
        d.mass_fluid = v[0] * (double)m;
        d.kinetic = v[1];
        d.momentum[0] = v[2];
        d.momentum[1] = v[3];
        d.momentum[2] = v[4];
        d.front_x = v[5];
        d.zmax_fluid = v[6];
    }
    return ZSPH_OK;
}

int zsph_probe_pressure(zsph_ctx *c, int b, const float *pts, int npts, float *p_out) {
    if (!c || !valid_buf(b) || npts < 0 || (npts && (!pts || !p_out)))
        return set_err(c, ZSPH_E_ARG, "zsph_probe_pressure: bad argument");
    if (c->nscenes > 1) return set_err(c, ZSPH_E_ARG, "zsph_probe_pressure: one scene only (probe a batch's scenes by their own contexts)");
    if (npts == 0) return ZSPH_OK;
    float *d = nullptr;
    HIPCHK(c, hipMalloc(&d, sizeof(float) * 4 * (size_t)npts));
    hipError_t e = hipMemcpyAsync(d, pts, sizeof(float) * 3 * (size_t)npts, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        zs::launch_probe(c->stream, c->q, c->buf[b], d, npts, d + 3 * (size_t)npts);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(p_out, d + 3 * (size_t)npts, sizeof(float) * npts, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return set_err(c, ZSPH_E_HIP, "probe pressure failed: %s", hipGetErrorString(e));
    return ZSPH_OK;
}

int zsph_probe_pressure_noalloc(zsph_ctx *c, int b, const float *pts, int npts, float *p_out) {
    if (!c || !valid_buf(b) || npts < 0 || (npts && (!pts || !p_out)))
        return set_err(c, ZSPH_E_ARG, "zsph_probe_pressure_noalloc: bad argument");
    if (npts > ZSPH_PROBE_NOALLOC_MAX)
        return set_err(c, ZSPH_E_ARG, "zsph_probe_pressure_noalloc: %d points, at most %d (use zsph_probe_pressure)",
                       npts, ZSPH_PROBE_NOALLOC_MAX);
    if (c->nscenes > 1)
        return set_err(c, ZSPH_E_ARG, "zsph_probe_pressure_noalloc: one scene only");
    if (npts == 0) return ZSPH_OK;
    float *d = c->probe_scratch;  // resolved for this context's device at zsph_create
    // points in [0, 3 npts), results in [3 npts, 4 npts): the kernel writes every result
    HIPCHK(c, hipMemcpyAsync(d, pts, sizeof(float) * 3 * (size_t)npts, hipMemcpyHostToDevice, c->stream));
    zs::launch_probe(c->stream, c->q, c->buf[b], d, npts, d + 3 * (size_t)npts);
    if (int r = check_launch(c, "probe pressure")) return r;
    HIPCHK(c, hipMemcpyAsync(p_out, d + 3 * (size_t)npts, sizeof(float) * npts, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ZSPH_OK;
}

// ---- learned closure ------------------------------------------------------

int zsph_set_closure(zsph_ctx *c, const zsph_closure *net) {
    if (!c) return ZSPH_E_ARG;
    if (c->nscenes > 1) return set_err(c, ZSPH_E_ARG, "zsph_set_closure: one scene only");
    drop_graph(c);  // the captured step has the other governing variant
    if (!net) {
        c->closure_on = false;
        c->closure_net.clear();
        return ZSPH_OK;
    }
    const size_t N = (size_t)c->q.N;
    int r;
    // feat may already exist (zsph_closure_sample allocates it): allocate only what is missing
    if (!c->feat && (r = dalloc(c, &c->feat, N * ZSPH_NF))) return r;
    if (!c->scale && (r = dalloc(c, &c->scale, N))) return r;
    if (!c->mlp_P) {
        uint16_t *pack = nullptr;
        if ((r = dalloc(c, &c->mlp_P, (size_t)ZSPH_NPARAM)) || (r = dalloc(c, &pack, zs::mlp_pack_elems())) ||
            (r = dalloc(c, &c->mlp_aux, 1)))
            return r;
        c->mlp_pack = pack;
        HIPCHK(c, zs::mlp_prepare());
    }
    zs::MlpAux aux{};
    std::memcpy(aux.mean, net->mean, sizeof(aux.mean));
    std::memcpy(aux.inv_std, net->inv_std, sizeof(aux.inv_std));
    std::memcpy(aux.lo, net->lo, sizeof(aux.lo));
    std::memcpy(aux.hi, net->hi, sizeof(aux.hi));
    aux.out_scale = net->out_scale;
    if ((r = copy_sync(c, c->mlp_P, net->w0, ZSPH_NPARAM * sizeof(float), hipMemcpyHostToDevice))) return r;
    if ((r = copy_sync(c, c->mlp_aux, &aux, sizeof(aux), hipMemcpyHostToDevice))) return r;
    zs::launch_mlp_pack(c->stream, c->mlp_P, c->mlp_pack);
    if ((r = check_launch(c, "closure pack"))) return r;
    const unsigned char *raw = reinterpret_cast<const unsigned char *>(net);
    c->closure_net.assign(raw, raw + sizeof(zsph_closure));
    c->closure_on = true;
    return ZSPH_OK;
}

int zsph_closure_sample(zsph_ctx *c, int b, float *features, float *terms) {
    if (!c || !valid_buf(b) || !features || !terms) return set_err(c, ZSPH_E_ARG, "zsph_closure_sample: bad argument");
    if (c->nscenes > 1) return set_err(c, ZSPH_E_ARG, "zsph_closure_sample: one scene only");
    if (c->cells_buf != b)
        return set_err(c, ZSPH_E_STATE, "zsph_closure_sample: buffer %d is not in cell order (sort into it first)", b);
    const size_t N = (size_t)c->q.N;
    int r;
    if (!c->feat && (r = dalloc(c, &c->feat, N * ZSPH_NF))) return r;
    if (!c->terms && (r = dalloc(c, &c->terms, N * 8))) return r;
    const uint32_t *sk = c->key[c->sorted_slot];
    // zero rows for the dummies, which the kernels do not write
    HIPCHK(c, hipMemsetAsync(c->feat, 0, N * ZSPH_NF * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(c->terms, 0, N * 8 * sizeof(float), c->stream));
    const int walk = ZSPH_WALK_MASK;  // every walk visits the same neighbours in the same order
    zs::launch_wall(c->stream, walk, c->block, c->q, nullptr, c->buf[b], sk, c->cell_start);
    zs::launch_ldg(c->stream, walk, c->block, c->q, nullptr, c->buf[b], sk, c->cell_start, c->gd, c->nlist, c->ncount,
                   c->feat);
    zs::launch_governing(c->stream, walk, c->block, c->q, nullptr, c->buf[b], sk, c->cell_start, c->gd, c->acc, c->nlist,
                         c->ncount, 2, c->scale, c->terms, c->has_body || c->has_piston);
    if ((r = check_launch(c, "closure sample"))) return r;
    c->list_valid = false;
    if ((r = copy_sync(c, features, c->feat, N * ZSPH_NF * sizeof(float), hipMemcpyDeviceToHost))) return r;
    return copy_sync(c, terms, c->terms, N * 8 * sizeof(float), hipMemcpyDeviceToHost);
}

int zsph_filter_fields(zsph_ctx *c, int b, float radius, float hc, const int *rows, int nq, float *out) {
    if (!c || !valid_buf(b) || nq < 0 || (nq && (!rows || !out)) || !(radius > 0.0f) || !(hc > 0.0f))
        return set_err(c, ZSPH_E_ARG, "zsph_filter_fields: bad argument");
    if (c->nscenes > 1) return set_err(c, ZSPH_E_ARG, "zsph_filter_fields: one scene only");
    if (c->cells_buf != b)
        return set_err(c, ZSPH_E_STATE, "zsph_filter_fields: buffer %d is not in cell order", b);
    for (int t = 0; t < nq; ++t)
        if (rows[t] < 0 || rows[t] >= c->q.N) return set_err(c, ZSPH_E_ARG, "zsph_filter_fields: row out of range");
    if (nq == 0) return ZSPH_OK;
    zs::Prm qf = c->q;
    const float rc = radius * qf.inv_cell;
    qf.k = (int)std::ceil(rc);
    qf.trim2 = rc * rc * 1.002f;
    int *drows = nullptr;
    float *dout = nullptr;
    HIPCHK(c, hipMalloc(&drows, sizeof(int) * (size_t)nq));
    hipError_t e = hipMalloc(&dout, sizeof(float) * 9 * (size_t)nq);
    if (e == hipSuccess) e = hipMemcpyAsync(drows, rows, sizeof(int) * (size_t)nq, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        zs::launch_filter(c->stream, qf, c->buf[b], c->key[c->sorted_slot], c->cell_start, c->acc, drows, nq, hc,
                          radius, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, sizeof(float) * 9 * (size_t)nq, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(drows);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return set_err(c, ZSPH_E_HIP, "zsph_filter_fields: %s", hipGetErrorString(e));
    return ZSPH_OK;
}

int zsph_set_graph(zsph_ctx *c, int on) {
    if (!c) return ZSPH_E_ARG;
    c->use_graph = on != 0;
    drop_graph(c);
    return ZSPH_OK;
}

namespace {
bool clockable(int st) { return st == ZSPH_ST_WALL || st == ZSPH_ST_LDG || st == ZSPH_ST_GOVERNING; }
}  // namespace

int zsph_clock_stage(zsph_ctx *c, int stage, int on) {
    if (!c) return ZSPH_E_ARG;
    if (!clockable(stage)) return set_err(c, ZSPH_E_ARG, "zsph_clock_stage: stage %d has no clock (wall, gradient "
                                                        "or continuity+momentum pass only)", stage);
    int r;
    if (!c->clock && (r = dalloc(c, &c->clock, 3 * ZSPH_ST_COUNT))) return r;
    const unsigned bit = 1u << stage;
    const unsigned mask = on ? (c->clock_mask | bit) : (c->clock_mask & ~bit);
    if (mask != c->clock_mask) drop_graph(c);  // the captured step's kernel arguments change
    c->clock_mask = mask;
    // the kernels stamp through their parameter block (zs::stage_stamp)
    for (zs::Prm *q : {&c->q, &c->qs}) {
        q->clk = mask ? c->clock : nullptr;
        q->clk_mask = mask;
    }
    return zsph_clock_read_stage(c, stage, nullptr, nullptr);
}

int zsph_clock_read_stage(zsph_ctx *c, int stage, double *ms, long long *launches) {
    if (!c) return ZSPH_E_ARG;
    if (!clockable(stage)) return set_err(c, ZSPH_E_ARG, "zsph_clock_read_stage: stage %d has no clock", stage);
    if (!c->clock) return set_err(c, ZSPH_E_STATE, "zsph_clock_read_stage: clock not enabled");
    unsigned long long v[3] = {0, 0, 0};
    int r = copy_sync(c, v, c->clock + 3 * stage, sizeof(v), hipMemcpyDeviceToHost);
    if (r) return r;
    int khz = 0;
    HIPCHK(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    if (ms) *ms = khz > 0 ? (double)v[1] / (double)khz : 0.0;
    if (launches) *launches = (long long)v[2];
    const unsigned long long zero[3] = {0, 0, 0};
    return copy_sync(c, c->clock + 3 * stage, zero, sizeof(zero), hipMemcpyHostToDevice);
}

int zsph_clock_governing(zsph_ctx *c, int on) { return zsph_clock_stage(c, ZSPH_ST_GOVERNING, on); }

int zsph_clock_read(zsph_ctx *c, double *ms, long long *launches) {
    return zsph_clock_read_stage(c, ZSPH_ST_GOVERNING, ms, launches);
}

int zsph_killed(zsph_ctx *c, int *killed) {
    if (!c || !killed) return ZSPH_E_ARG;
    HIPCHK(c, hipMemcpyAsync(killed, c->killed, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ZSPH_OK;
}

// ---- moving dummies --------------------------------------------------------

namespace {

// Device body record as the host sees it; the time comes from the host clock
// while nothing moves.
int body_fetch(zsph_ctx *c, zs::DevBody *hb) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (int e_ = copy_sync(c, hb, c->body, sizeof(*hb), hipMemcpyDeviceToHost)) return e_;
    if (!c->has_body && !c->has_piston) {
        hb->t = c->t;
        hb->t0 = c->t0;
    }
    return ZSPH_OK;
}

// Buffer 0's particle arrays on the host, edited by the set-up calls.
struct HostArrays {
    std::vector<float4> posr, velp, tposr, tvelp;
    std::vector<int> type;
};
int fetch_arrays(zsph_ctx *c, HostArrays &h) {
    const size_t N = (size_t)c->q.N;
    const zs::Buf &d = c->buf[0];
    h.posr.resize(N); h.velp.resize(N); h.tposr.resize(N); h.tvelp.resize(N); h.type.resize(N);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (int e_ = copy_sync(c, h.posr.data(), d.posr, N * sizeof(float4), hipMemcpyDeviceToHost)) return e_;
    if (int e_ = copy_sync(c, h.velp.data(), d.velp, N * sizeof(float4), hipMemcpyDeviceToHost)) return e_;
    if (int e_ = copy_sync(c, h.tposr.data(), d.tposr, N * sizeof(float4), hipMemcpyDeviceToHost)) return e_;
    if (int e_ = copy_sync(c, h.tvelp.data(), d.tvelp, N * sizeof(float4), hipMemcpyDeviceToHost)) return e_;
    if (int e_ = copy_sync(c, h.type.data(), d.type, N * sizeof(int), hipMemcpyDeviceToHost)) return e_;
    return ZSPH_OK;
}
int store_arrays(zsph_ctx *c, const HostArrays &h, const zs::DevBody &hb) {
    const size_t N = (size_t)c->q.N;
    const zs::Buf &d = c->buf[0];
    if (int e_ = copy_sync(c, d.velp, h.velp.data(), N * sizeof(float4), hipMemcpyHostToDevice)) return e_;
    if (int e_ = copy_sync(c, d.tposr, h.tposr.data(), N * sizeof(float4), hipMemcpyHostToDevice)) return e_;
    if (int e_ = copy_sync(c, d.tvelp, h.tvelp.data(), N * sizeof(float4), hipMemcpyHostToDevice)) return e_;
    if (int e_ = copy_sync(c, c->body, &hb, sizeof(hb), hipMemcpyHostToDevice)) return e_;
    drop_graph(c);  // the captured step does not contain the moving-dummy kernels
    c->skin_valid = false;
    return ZSPH_OK;
}

}  // namespace

int zsph_set_body(zsph_ctx *c, const zsph_rigid *r) {
    using namespace zs::rb;
    if (!c || !r) return set_err(c, ZSPH_E_ARG, "zsph_set_body: null argument");
    if (c->nscenes > 1) return set_err(c, ZSPH_E_ARG, "zsph_set_body: one scene only");
    const double qn = std::sqrt(r->q[0] * r->q[0] + r->q[1] * r->q[1] + r->q[2] * r->q[2] + r->q[3] * r->q[3]);
    M3 I;
    for (int k = 0; k < 9; ++k) I.m[k] = r->I[k];
    const double detI = I.m[0] * (I.m[4] * I.m[8] - I.m[5] * I.m[7]) - I.m[1] * (I.m[3] * I.m[8] - I.m[5] * I.m[6]) +
                        I.m[2] * (I.m[3] * I.m[7] - I.m[4] * I.m[6]);
    if (!(r->mass > 0.0) || !(qn > 0.0) || !std::isfinite(qn) || !(std::fabs(detI) > 0.0) || !std::isfinite(detI))
        return set_err(c, ZSPH_E_ARG, "zsph_set_body: needs mass > 0, a non-zero quaternion and an invertible I");
    zs::DevBody hb;
    HostArrays h;
    int e;
    if ((e = body_fetch(c, &hb)) || (e = fetch_arrays(c, h))) return e;
    hb.s = *r;
    for (int k = 0; k < 4; ++k) hb.s.q[k] = r->q[k] / qn;
    const M3 R = rotation(qget(hb.s));
    const M3 Rt = transpose(R);
    const V3 cc = v3(hb.s.c), V = v3(hb.s.vel), w = v3(hb.s.omega), A = v3(hb.s.acc), al = v3(hb.s.alpha);
    int nbody = 0;
    for (size_t i = 0; i < h.type.size(); ++i) {
        if (h.type[i] != ZSPH_BODY) continue;
        ++nbody;
        const V3 d = V3{h.posr[i].x, h.posr[i].y, h.posr[i].z} - cc;
        const V3 r0 = mul(Rt, d);
        const V3 v = V + cross(w, d);
        const V3 a = A + cross(al, d) + cross(w, cross(w, d));
        h.tposr[i] = make_float4((float)r0.x, (float)r0.y, (float)r0.z, 0.0f);
        h.velp[i].x = (float)v.x; h.velp[i].y = (float)v.y; h.velp[i].z = (float)v.z;
        h.tvelp[i].x = (float)a.x; h.tvelp[i].y = (float)a.y; h.tvelp[i].z = (float)a.z;
    }
    if (nbody == 0) return set_err(c, ZSPH_E_ARG, "zsph_set_body: buffer 0 holds no type-2 (body) particles");
    hb.s0 = hb.s;
    hb.has_body = 1;
    if ((e = store_arrays(c, h, hb))) return e;
    c->has_body = true;
    return ZSPH_OK;
}

int zsph_get_body(zsph_ctx *c, zsph_rigid *r) {
    if (!c || !r) return set_err(c, ZSPH_E_ARG, "zsph_get_body: null argument");
    if (!c->has_body) return set_err(c, ZSPH_E_STATE, "zsph_get_body: no body (zsph_set_body)");
    zs::DevBody hb;
    int e = body_fetch(c, &hb);
    if (e) return e;
    *r = hb.s;
    return ZSPH_OK;
}

int zsph_set_piston(zsph_ctx *c, const zsph_piston *p) {
    if (!c || !p) return set_err(c, ZSPH_E_ARG, "zsph_set_piston: null argument");
    if (c->nscenes > 1) return set_err(c, ZSPH_E_ARG, "zsph_set_piston: one scene only");
    if (p->axis < 0 || p->axis > 2 || !(p->frequency >= 0.0) || !std::isfinite(p->frequency) ||
        !std::isfinite(p->amplitude) || !std::isfinite(p->phase))
        return set_err(c, ZSPH_E_ARG, "zsph_set_piston: axis 0..2, finite amplitude/phase, frequency >= 0");
    zs::DevBody hb;
    HostArrays h;
    int e;
    if ((e = body_fetch(c, &hb)) || (e = fetch_arrays(c, h))) return e;
    hb.amp = p->amplitude;
    hb.pw = 2.0 * M_PI * p->frequency;
    hb.phase = p->phase;
    hb.axis = p->axis;
    const double ph = hb.pw * hb.t + hb.phase;
    const double d = hb.amp * (std::sin(ph) - std::sin(hb.phase));
    const double dv = hb.amp * hb.pw * std::cos(ph), da = -hb.amp * hb.pw * hb.pw * std::sin(ph);
    int np = 0;
    for (size_t i = 0; i < h.type.size(); ++i) {
        if (h.type[i] != ZSPH_PISTON) continue;
        ++np;
        float rest[3] = {h.posr[i].x, h.posr[i].y, h.posr[i].z};
        float v[3] = {0.0f, 0.0f, 0.0f}, a[3] = {0.0f, 0.0f, 0.0f};
        rest[p->axis] = (float)((double)rest[p->axis] - d);
        v[p->axis] = (float)dv;
        a[p->axis] = (float)da;
        h.tposr[i] = make_float4(rest[0], rest[1], rest[2], 0.0f);
        h.velp[i].x = v[0]; h.velp[i].y = v[1]; h.velp[i].z = v[2];
        h.tvelp[i].x = a[0]; h.tvelp[i].y = a[1]; h.tvelp[i].z = a[2];
    }
    if (np == 0) return set_err(c, ZSPH_E_ARG, "zsph_set_piston: buffer 0 holds no type-3 (piston) particles");
    hb.has_piston = 1;
    if ((e = store_arrays(c, h, hb))) return e;
    c->has_piston = true;
    return ZSPH_OK;
}

int zsph_set_time(zsph_ctx *c, double t) {
    if (!c || !std::isfinite(t)) return set_err(c, ZSPH_E_ARG, "zsph_set_time: bad argument");
    c->t = c->t0 = t;
    if (c->has_body || c->has_piston) {
        zs::DevBody hb;
        int e = body_fetch(c, &hb);
        if (e) return e;
        hb.t = hb.t0 = t;
        if (int e_ = copy_sync(c, c->body, &hb, sizeof(hb), hipMemcpyHostToDevice)) return e_;
    }
    return ZSPH_OK;
}

int zsph_get_time(zsph_ctx *c, double *t) {
    if (!c || !t) return set_err(c, ZSPH_E_ARG, "zsph_get_time: null argument");
    zs::DevBody hb;
    int e = body_fetch(c, &hb);
    if (e) return e;
    *t = hb.t;
    return ZSPH_OK;
}

int zsph_body_reduce(zsph_ctx *c, int b, double F[3], double T[3]) {
    if (!c || !valid_buf(b) || !F || !T) return set_err(c, ZSPH_E_ARG, "zsph_body_reduce: bad argument");
    if (!c->has_body) return set_err(c, ZSPH_E_STATE, "zsph_body_reduce: no body (zsph_set_body)");
    if (c->sorted_slot < 0) return set_err(c, ZSPH_E_STATE, "zsph_body_reduce called before zsph_sort_data");
    zs::launch_body_reduce(c->stream, c->q, c->buf[b], c->cell_start, c->acc, c->body, c->body_part);
    int e = check_launch(c, "body reduce");
    if (e) return e;
    std::vector<double> part((size_t)zs::BODY_BLOCKS * zs::BODY_D);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->body_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double s[zs::BODY_D] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < zs::BODY_BLOCKS; ++k)
        for (int d = 0; d < zs::BODY_D; ++d) s[d] += part[(size_t)k * zs::BODY_D + d];
    for (int d = 0; d < 3; ++d) {
        F[d] = s[d];
        T[d] = s[3 + d];
    }
    return ZSPH_OK;
}

// ---- checkpoint / restart ------------------------------------------------

}  // extern "C"

namespace {

// FNV-1a 64 of the parameter blocks (4-byte fields and the two 8-byte
// pointers, no padding), without the stage clocks and the wave order
// (measurement and scheduling state, not parameters): a restart must use the
// parameters the checkpoint was written with.
static_assert(sizeof(zs::Prm) == offsetof(zs::Prm, wtab) + sizeof(void *) &&
                  offsetof(zs::Prm, wtab) == offsetof(zs::Prm, clk) + sizeof(void *),
              "Prm has no padding");
uint64_t params_hash(const zsph_ctx *c) {
    zs::Prm q = c->q;
    q.clk = nullptr;
    q.clk_mask = 0;
    q.wtab = nullptr;
    q.wt_cap = q.wt_rows = 0;
    uint64_t h = zsph_fnv1a64(&q, sizeof(q));
    for (const zs::Prm &p : c->scene_host) h = zsph_fnv1a64(&p, sizeof(p), h);
    return h;
}

// The sections of this context's checkpoint, in file order (sizes only).
// "body" is the FP64 rigid-body / piston block in the library's own layout,
// tied to the format version.
std::vector<CkptSection> ckpt_layout(const zsph_ctx *c, bool closure) {
    const size_t N = (size_t)c->q.N;
    std::vector<CkptSection> v;
    auto add = [&](const char *name, size_t bytes) {
        CkptSection s;
        s.name = name;
        s.bytes = bytes;
        v.push_back(s);
    };
    add("posr", N * sizeof(float4));
    add("velp", N * sizeof(float4));
    add("tposr", N * sizeof(float4));
    add("tvelp", N * sizeof(float4));
    add("type", N * sizeof(int));
    add("table", N * sizeof(int));
    if (c->buf[0].scene) add("scene", N * sizeof(int));
    add("body", sizeof(zs::DevBody));
    if (closure) add("closure", sizeof(zsph_closure));
    return v;
}

// Device address of a particle-array section of buffer 0 (null for the others).
void *ckpt_device(zsph_ctx *c, const std::string &name) {
    const zs::Buf &d = c->buf[0];
    if (name == "posr") return d.posr;
    if (name == "velp") return d.velp;
    if (name == "tposr") return d.tposr;
    if (name == "tvelp") return d.tvelp;
    if (name == "type") return d.type;
    if (name == "table") return d.table;
    if (name == "scene") return d.scene;
    return nullptr;
}

}  // namespace

extern "C" {

int zsph_checkpoint_write(zsph_ctx *c, const char *path, long long step) {
    if (!c || !path || !*path || step < 0) return set_err(c, ZSPH_E_ARG, "zsph_checkpoint_write: bad argument");
    c->skin_valid = false;  // buffer 0 changes (or, for a checkpoint, the next step rebuilds)
    std::vector<CkptSection> secs = ckpt_layout(c, c->closure_on);
    size_t total = 0;
    for (const CkptSection &s : secs) total += s.bytes;
    std::vector<char> host(total);
    zs::DevBody hb;
    int killed = 0;
    int e;
    if ((e = body_fetch(c, &hb))) return e;  // waits for the stream
    if ((e = copy_sync(c, &killed, c->killed, sizeof(int), hipMemcpyDeviceToHost))) return e;
    size_t off = 0;
    for (CkptSection &s : secs) {
        char *dst = host.data() + off;
        if (void *src = ckpt_device(c, s.name)) {
            HIPCHK(c, hipMemcpyAsync(dst, src, s.bytes, hipMemcpyDeviceToHost, c->stream));
        } else if (s.name == "body") {
            std::memcpy(dst, &hb, sizeof(hb));
        } else {  // closure
            std::memcpy(dst, c->closure_net.data(), sizeof(zsph_closure));
        }
        s.data = dst;
        off += s.bytes;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    CkptMeta m;
    m.step = step;
    m.time = hb.t;
    m.t0 = hb.t0;
    m.ptc_num = c->q.N;
    m.cells = c->q.G;
    m.scenes = c->nscenes;
    m.dim = c->q.dim;
    m.grid[0] = c->q.X;
    m.grid[1] = c->q.Y;
    m.grid[2] = c->q.Z;
    m.killed = killed;
    m.has_body = c->has_body ? 1 : 0;
    m.has_piston = c->has_piston ? 1 : 0;
    m.closure = c->closure_on ? 1 : 0;
    m.params = params_hash(c);
    std::string err;
    if ((e = zsph_ckpt_write(path, m, secs, &err))) return set_err(c, e, "%s", err.c_str());
    return ZSPH_OK;
}

int zsph_checkpoint_read(zsph_ctx *c, const char *path, long long *step) {
    if (!c || !path || !*path) return set_err(c, ZSPH_E_ARG, "zsph_checkpoint_read: bad argument");
    CkptMeta m;
    std::vector<CkptSection> secs;
    std::vector<char> payload;
    std::string err;
    int e = zsph_ckpt_read(path, &m, &secs, &payload, &err);
    if (e) return set_err(c, e, "%s", err.c_str());
    if (m.ptc_num != c->q.N || m.cells != c->q.G || m.scenes != c->nscenes || m.dim != c->q.dim ||
        m.grid[0] != c->q.X || m.grid[1] != c->q.Y || m.grid[2] != c->q.Z)
        return set_err(c, ZSPH_E_ARG,
                       "%s: checkpoint of %d particles, %d cells, %d scene(s), %d-D; this context has %d, %d, %d, %d-D",
                       path, m.ptc_num, m.cells, m.scenes, m.dim, c->q.N, c->q.G, c->nscenes, c->q.dim);
    if (m.params != params_hash(c))
        return set_err(c, ZSPH_E_ARG, "%s: written with other parameters (parameter hash differs)", path);
    const std::vector<CkptSection> want = ckpt_layout(c, m.closure != 0);
    bool same = want.size() == secs.size();
    for (size_t k = 0; same && k < want.size(); ++k) same = want[k].name == secs[k].name && want[k].bytes == secs[k].bytes;
    if (!same) return set_err(c, ZSPH_E_ARG, "%s: sections differ from this context's layout", path);
    zs::DevBody hb;
    size_t off = 0, body_off = 0, closure_off = 0;
    for (const CkptSection &s : secs) {
        if (s.name == "body") body_off = off;
        if (s.name == "closure") closure_off = off;
        off += s.bytes;
    }
    std::memcpy(&hb, payload.data() + body_off, sizeof(hb));
    if ((hb.has_body != 0) != (m.has_body != 0) || (hb.has_piston != 0) != (m.has_piston != 0) ||
        ((m.has_body || m.has_piston) && c->nscenes > 1))
        return set_err(c, ZSPH_E_ARG, "%s: inconsistent moving-dummy flags", path);
    // validated: from here on the context takes the checkpoint's state
    if (m.closure) {
        zsph_closure net;
        std::memcpy(&net, payload.data() + closure_off, sizeof(net));
        if ((e = zsph_set_closure(c, &net))) return e;
    } else if (c->closure_on && (e = zsph_set_closure(c, nullptr))) {
        return e;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    off = 0;
    for (const CkptSection &s : secs) {
        if (void *dst = ckpt_device(c, s.name))
            HIPCHK(c, hipMemcpyAsync(dst, payload.data() + off, s.bytes, hipMemcpyHostToDevice, c->stream));
        off += s.bytes;
    }
    HIPCHK(c, hipMemcpyAsync(c->body, &hb, sizeof(hb), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->killed, &m.killed, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->has_body = m.has_body != 0;
    c->has_piston = m.has_piston != 0;
    c->t = m.time;
    c->t0 = m.t0;
    c->sorted_slot = -1;  // buffer 0 holds a state between steps, not yet hashed
    c->hashed = 0;
    c->cells_buf = -1;
    c->list_valid = false;
    c->skin_valid = false;
    drop_graph(c);  // the captured step may lack (or hold) the moving-dummy kernels
    if (step) *step = m.step;
    return ZSPH_OK;
}

}  // extern "C"
