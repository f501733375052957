// Neighbour build for gfx950: cell hash, stable LSD radix sort of
// (hash, index) with wave64 ballot ranking, cell-start table, and the
// cell-order gather of the particle buffer.
//   hash      <- single_gpu_sph_src/lib/hash.cu:3-26   (cuda_ptc_hash)
//   sort      <- single_gpu_sph_src/lib/sort.cu:3-6    (thrust::sort_by_key)
//   cells +
//   reorder   <- single_gpu_sph_src/lib/sort.cu:8-84   (cuda_sort_data)
#include <algorithm>
#include <climits>

#include "common.hpp"
#include "launch.hpp"

namespace zs {

// ---------------------------------------------------------------------------
// Cell hash.  H = (cz*Y + cy)*X + cx; a particle whose cell lies outside the
// grid on any axis is removed: type 100, parked at the grid origin with
// v = 0, (rho, p) = (rho0, 0) and H = 0 (hash.cu:16-23).  The sort key adds
// the non-fluid flag above the cell bits, so fluid sorts first.
// ---------------------------------------------------------------------------
// Cell coordinate on the neighbour grid when it differs from the case's grid
// (skin mode), clamped to the grid: a particle inside the case's grid lies
// inside the neighbour grid, so the clamp only absorbs a last-bit rounding of
// the quotient at the far end.
__device__ __forceinline__ int skin_axis(float x, float o, float size, int dim) {
    const float t = floorf((float)((double)(x - o) / (double)size));
    return min(max((int)t, 0), dim - 1);
}

// gate (skin mode; nullptr otherwise): the kernel does nothing unless *gate
// is set (the step's rebuild flag, written by the previous integrator).
__device__ __forceinline__ bool gated_off(const int *gate) { return gate && *gate == 0; }

// One particle's cell key (k_hash's body).
template <bool BATCH>
__device__ __forceinline__ void hash_one(const Prm &q0, const Prm *__restrict__ sp, const Buf &b,
                                         uint32_t *__restrict__ key, int *__restrict__ idx, int *__restrict__ killed,
                                         int i) {
    const Prm q = scene_prm<BATCH>(q0, sp, b, i);
    const float4 x = b.posr[i];
    int cx, cy, cz;
    const bool in = cell_axis(x.x, q.gx0, q.kill_size, q.kill_X, &cx) &&
                    cell_axis(x.y, q.gy0, q.kill_size, q.kill_Y, &cy) &&
                    cell_axis(x.z, q.gz0, q.kill_size, q.kill_Z, &cz);
    uint32_t h = 0;
    int t;
    if (in) {
        if (q.kill_size != q.grid_size) {  // skin mode: the key on the neighbour grid
            cx = skin_axis(x.x, q.gx0, q.grid_size, q.X);
            cy = skin_axis(x.y, q.gy0, q.grid_size, q.Y);
            cz = skin_axis(x.z, q.gz0, q.grid_size, q.Z);
        }
        h = (uint32_t)((cz * q.Y + cy) * q.X + cx);
        t = b.type[i];
    } else {
        t = ZSPH_KILLED;
        b.type[i] = ZSPH_KILLED;
        b.posr[i] = make_float4(q.gx0, q.gy0, q.gz0, q.rho0);
        b.velp[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.tvelp[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        atomicAdd(killed, 1);
    }
    // a removed particle stays in its own scene's cell 0 (dummy segment)
    key[i] = ((uint32_t)q.cell_base + h) | (t == ZSPH_FLUID ? 0u : q.flag);
    idx[i] = i;
}

// Gated (skin mode) launches cap the grid so that a skipped rebuild costs a
// small launch; each thread then takes every (grid)-th particle.
template <bool BATCH>
__global__ __launch_bounds__(ZS_BLOCK) void k_hash(Prm q0, const Prm *__restrict__ sp, Buf b,
                                                   uint32_t *__restrict__ key, int *__restrict__ idx,
                                                   int *__restrict__ killed, const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    for (int i = xcd_block(gridDim.x) * ZS_BLOCK + threadIdx.x; i < q0.N; i += gridDim.x * ZS_BLOCK)
        hash_one<BATCH>(q0, sp, b, key, idx, killed, i);
}

// ---------------------------------------------------------------------------
// Radix sort.  Tile = 256 threads x ITEMS keys (ITEMS = 8 or 16, a tuning
// parameter: the scatter stages its whole tile in LDS, 8 bytes per key);
// wave w of a block owns the contiguous keys [tile + w*64*ITEMS, +64*ITEMS)
// and walks them 64 at a time, so (wave, item, lane) order equals input order
// and the scatter is stable.  Equal digits inside a wave are found with BITS
// ballots.
// ---------------------------------------------------------------------------
constexpr int RS_ITEMS_MIN = 8;   // smallest tile: sizes the histogram buffer
constexpr int RS_WAVES = ZS_BLOCK / 64;

__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Mask of the lanes (among `valid`) whose BITS-bit digit equals this lane's.
template <int BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
    uint64_t m = valid;
#pragma unroll
    for (int bit = 0; bit < BITS; ++bit) {
        const bool set = (d >> bit) & 1u;
        const uint64_t bal = __ballot(set);
        m &= set ? bal : ~bal;
    }
    return m;
}

template <int BITS, int RS_ITEMS>
__global__ __launch_bounds__(ZS_BLOCK) void k_radix_hist(const uint32_t *__restrict__ kin, int n, int shift,
                                                         uint32_t *__restrict__ hist, int nblk,
                                                         const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    constexpr int RS_TILE = ZS_BLOCK * RS_ITEMS;
    constexpr int R = 1 << BITS;
    __shared__ uint32_t h[R];
    for (int d = threadIdx.x; d < R; d += ZS_BLOCK) h[d] = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int base = blockIdx.x * RS_TILE + w * 64 * RS_ITEMS;
    for (int it = 0; it < RS_ITEMS; ++it) {
        const int e = base + it * 64 + lane;
        const bool valid = e < n;
        const uint32_t d = valid ? (kin[e] >> shift) & (R - 1) : 0u;
        const uint64_t m = match_digit<BITS>(d, __ballot(valid));
        if (valid && lane_rank(m) == 0) atomicAdd(&h[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    for (int d = threadIdx.x; d < R; d += ZS_BLOCK) hist[d * nblk + blockIdx.x] = h[d];
}

// Row scan: block d turns hist[d][0..nblk) into its exclusive prefix (the
// offset of each tile among digit-d keys) and writes the digit total.
__global__ __launch_bounds__(ZS_BLOCK) void k_radix_rowscan(uint32_t *__restrict__ hist, int nblk,
                                                            uint32_t *__restrict__ total,
                                                            const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    __shared__ uint32_t part[ZS_BLOCK];
    uint32_t *row = hist + (size_t)blockIdx.x * nblk;
    const int t = threadIdx.x;
    const int chunk = (nblk + ZS_BLOCK - 1) / ZS_BLOCK;
    const int b = min(t * chunk, nblk), e = min(b + chunk, nblk);
    uint32_t s = 0;
    for (int i = b; i < e; ++i) s += row[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < ZS_BLOCK; off <<= 1) {
        const uint32_t v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int i = b; i < e; ++i) {
        const uint32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (t == ZS_BLOCK - 1) total[blockIdx.x] = part[t];
}

template <int BITS, int RS_ITEMS>
__global__ __launch_bounds__(ZS_BLOCK) void k_radix_scatter(const uint32_t *__restrict__ kin,
                                                            const int *__restrict__ vin,
                                                            uint32_t *__restrict__ kout,
                                                            int *__restrict__ vout, int n, int shift,
                                                            const uint32_t *__restrict__ scanned, int nblk,
                                                            const uint32_t *__restrict__ total,
                                                            const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    constexpr int RS_TILE = ZS_BLOCK * RS_ITEMS;
    constexpr int R = 1 << BITS;
    constexpr int PER = (R + ZS_BLOCK - 1) / ZS_BLOCK;
    __shared__ uint32_t cnt[RS_WAVES][R];
    __shared__ uint32_t goff[R];
    __shared__ uint32_t part[ZS_BLOCK];
    __shared__ uint32_t dstart[R];     // start of each digit's run in the tile
    __shared__ uint32_t sk[RS_TILE];   // the tile in digit order
    __shared__ int sv[RS_TILE];
    // digit base = exclusive scan of the digit totals; thread t owns digits [t*PER, t*PER+PER)
    {
        const int t = threadIdx.x;
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int d = t * PER + k;
            if (d < R) s += total[d];
        }
        part[t] = s;
        __syncthreads();
        for (int off = 1; off < ZS_BLOCK; off <<= 1) {
            const uint32_t v = (t >= off) ? part[t - off] : 0u;
            __syncthreads();
            part[t] += v;
            __syncthreads();
        }
        uint32_t run = part[t] - s;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int d = t * PER + k;
            if (d < R) {
                goff[d] = run + scanned[(size_t)d * nblk + blockIdx.x];
                run += total[d];
#pragma unroll
                for (int w = 0; w < RS_WAVES; ++w) cnt[w][d] = 0;
            }
        }
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int base = blockIdx.x * RS_TILE + w * 64 * RS_ITEMS;
    uint32_t k[RS_ITEMS];
    int v[RS_ITEMS];
    uint32_t loc[RS_ITEMS];
#pragma unroll
    for (int it = 0; it < RS_ITEMS; ++it) {
        const int e = base + it * 64 + lane;
        const bool valid = e < n;
        k[it] = valid ? kin[e] : 0u;
        v[it] = valid ? (vin ? vin[e] : e) : 0;
    }
#pragma unroll
    for (int it = 0; it < RS_ITEMS; ++it) {
        const int e = base + it * 64 + lane;
        const bool valid = e < n;
        const uint32_t d = (k[it] >> shift) & (R - 1);
        const uint64_t m = match_digit<BITS>(d, __ballot(valid));
        const uint32_t r = lane_rank(m);
        const uint32_t before = valid ? cnt[w][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && r == (uint32_t)__popcll(m) - 1u) cnt[w][d] = before + (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
        loc[it] = before + r;
    }
    __syncthreads();
    // per digit: the waves' exclusive prefix (cnt) and the tile's count
    for (int d = threadIdx.x; d < R; d += ZS_BLOCK) {
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ++ww) {
            const uint32_t c = cnt[ww][d];
            cnt[ww][d] = run;
            run += c;
        }
        dstart[d] = run;
    }
    __syncthreads();
    // exclusive scan of the tile's digit counts; thread t owns digits [t*PER, t*PER+PER)
    {
        const int t = threadIdx.x;
        uint32_t s = 0;
#pragma unroll
        for (int k2 = 0; k2 < PER; ++k2) {
            const int d = t * PER + k2;
            if (d < R) s += dstart[d];
        }
        part[t] = s;
        __syncthreads();
        for (int off = 1; off < ZS_BLOCK; off <<= 1) {
            const uint32_t v2 = (t >= off) ? part[t - off] : 0u;
            __syncthreads();
            part[t] += v2;
            __syncthreads();
        }
        uint32_t run = part[t] - s;
#pragma unroll
        for (int k2 = 0; k2 < PER; ++k2) {
            const int d = t * PER + k2;
            if (d < R) {
                const uint32_t c = dstart[d];
                dstart[d] = run;
                run += c;
            }
        }
    }
    __syncthreads();
    // the tile in digit order in LDS (within a digit: wave, item, lane order)
#pragma unroll
    for (int it = 0; it < RS_ITEMS; ++it) {
        const int e = base + it * 64 + lane;
        if (e < n) {
            const uint32_t d = (k[it] >> shift) & (R - 1);
            const uint32_t p = dstart[d] + cnt[w][d] + loc[it];
            sk[p] = k[it];
            sv[p] = v[it];
        }
    }
    __syncthreads();
    // consecutive positions of one digit's run go to consecutive outputs
    const int valid_n = min(RS_TILE, n - (int)blockIdx.x * RS_TILE);
    for (int p = threadIdx.x; p < valid_n; p += ZS_BLOCK) {
        const uint32_t kk = sk[p];
        const uint32_t d = (kk >> shift) & (R - 1);
        const uint32_t dst = goff[d] + ((uint32_t)p - dstart[d]);
        kout[dst] = kk;
        vout[dst] = sv[p];
    }
}

// ---------------------------------------------------------------------------
// Cell table: entry e < G+1 is the first sorted position with key >= e (the
// fluid particles of cell c are [cs[c], cs[c+1]), so cs[G] = fluid count);
// entry G+1+c is the first position with key >= flag + c (the dummies of
// cell c; cs[2G+1] = N).  Empty cells have equal bounds.  Every entry is
// rewritten on every call.
// ---------------------------------------------------------------------------
// Two passes.  The key of entry e, ek(e) = e for e <= G and flag + (e - G - 1)
// above, is non-decreasing in e (flag >= G), so cs[e] is the lower bound of
// ek(e) in the sorted keys, i.e. the first particle whose entry index
// ent(key) (key for fluid, G + 1 + key - flag for the others) is >= e.
// k_cs_bounds: bound[t] = the first particle whose entry lies in tile t or
// later (tiles of CS_TILE entries, t = 0..T): the particle where the entry
// tile changes writes the bounds of the tiles it skips, so each is written
// once.  k_cell_start: a block owns one tile; its entries' lower bounds lie in
// [bound[t], bound[t+1]]; the keys between are staged in LDS and each entry
// is a binary search there (in global memory when they do not fit, e.g.
// hundreds of particles in one cell).
constexpr int CS_ITEMS = 4;
constexpr int CS_TILE = ZS_BLOCK * CS_ITEMS;
constexpr int CS_KEYS = 4096;  // staged keys per block (16 KB of LDS)

__device__ __forceinline__ uint32_t cs_entry_key(int e, int G, uint32_t flag) {
    return (e <= G) ? (uint32_t)e : flag + (uint32_t)(e - G - 1);
}
__device__ __forceinline__ int cs_tile_of(uint32_t key, int G, uint32_t flag) {
    return ((key < flag) ? (int)key : G + 1 + (int)(key - flag)) / CS_TILE;
}

// Lower bound of k in a[lo, hi) by one thread.
template <class A>
__device__ __forceinline__ int lower_bound_in(const A *a, int lo, int hi, uint32_t k) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// bound[ta..tb] = v for every lane's (ta, tb, v), the wave writing 64 tiles
// at a time (a particle before a run of empty cells skips many tiles).
__device__ __forceinline__ void wave_fill(int *__restrict__ bound, int ta, int tb, int v) {
    const int lane = (int)threadIdx.x & 63;
    uint64_t m = __ballot(tb >= ta);
    while (m) {
        const int k = __builtin_ctzll(m);
        m &= m - 1u;
        const int a = __shfl(ta, k), b = __shfl(tb, k), w = __shfl(v, k);
        for (int t = a + lane; t <= b; t += 64) bound[t] = w;
    }
}

__global__ __launch_bounds__(ZS_BLOCK) void k_cs_bounds(const uint32_t *__restrict__ skey, int n, int G,
                                                        uint32_t flag, int T, int *__restrict__ bound,
                                                        const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    // the block's base is uniform, so every lane of a wave makes the same trips
    for (int base = blockIdx.x * ZS_BLOCK; base < n; base += gridDim.x * ZS_BLOCK) {
        const int i = base + (int)threadIdx.x;
        int ta = 0, tb = -1, ua = 0, ub = -1;
        if (i < n) {
            const int t = cs_tile_of(skey[i], G, flag);
            ta = (i == 0) ? 0 : cs_tile_of(skey[i - 1], G, flag) + 1;
            tb = t;
            if (i == n - 1) {  // the tiles after the last particle's
                ua = t + 1;
                ub = T;
            }
        }
        wave_fill(bound, ta, tb, i);  // every lane of the wave takes part
        wave_fill(bound, ua, ub, n);
    }
}

__global__ __launch_bounds__(ZS_BLOCK) void k_cell_start(const uint32_t *__restrict__ skey, int G, uint32_t flag,
                                                         const int *__restrict__ bound, int *__restrict__ cs,
                                                         const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    __shared__ uint32_t keys[CS_KEYS];
    const int E = 2 * G + 2;
    const int e0 = blockIdx.x * CS_TILE;
    const int e1 = min(e0 + CS_TILE, E);
    // every entry of the tile has its lower bound in [lo, lo + len]
    const int lo = bound[blockIdx.x], len = bound[blockIdx.x + 1] - lo;
    const bool staged = len <= CS_KEYS;
    if (staged)
        for (int j = (int)threadIdx.x; j < len; j += ZS_BLOCK) keys[j] = skey[lo + j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CS_ITEMS; ++j) {
        const int e = e0 + j * ZS_BLOCK + (int)threadIdx.x;
        if (e < e1) {
            const uint32_t k = cs_entry_key(e, G, flag);
            cs[e] = staged ? lo + lower_bound_in(keys, 0, len, k) : lower_bound_in(skey, lo, lo + len, k);
        }
    }
}

// Gather buffer `a` into cell order in buffer `b`; table'[k] = table[p] + p - k
// keeps table + index equal to the original particle id (sort.cu:107).
__global__ __launch_bounds__(ZS_BLOCK) void k_reorder(int n, const int *__restrict__ perm, Buf a, Buf b,
                                                      const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    for (int k = xcd_block(gridDim.x) * ZS_BLOCK + threadIdx.x; k < n; k += gridDim.x * ZS_BLOCK) {
        const int p = perm[k];
        b.posr[k] = a.posr[p];
        b.velp[k] = a.velp[p];
        b.tposr[k] = a.tposr[p];
        b.tvelp[k] = a.tvelp[p];
        b.type[k] = a.type[p];
        b.table[k] = a.table[p] + p - k;
        if (a.scene) b.scene[k] = a.scene[p];
    }
}

// Skin mode, end of a rebuild: the reordered state in buffer `a` becomes the
// working buffer `b` (the step's passes always run on one buffer), and each
// particle's position is kept as its reference position for the displacement
// test of the integrators.
__global__ __launch_bounds__(ZS_BLOCK) void k_skin_commit(int n, Buf a, Buf b, float4 *__restrict__ xref,
                                                          const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    for (int k = xcd_block(gridDim.x) * ZS_BLOCK + threadIdx.x; k < n; k += gridDim.x * ZS_BLOCK) {
        const float4 x = a.posr[k];
        b.posr[k] = x;
        xref[k] = x;
        b.velp[k] = a.velp[k];
        b.tposr[k] = a.tposr[k];
        b.tvelp[k] = a.tvelp[k];
        b.type[k] = a.type[k];
        b.table[k] = a.table[k];
        if (a.scene) b.scene[k] = a.scene[k];
    }
}

// Skin mode, after the rebuild kernels of a half step: count a rebuild that
// happened, tell the passes whether it did, and clear the flag (the
// integrator sets it again when needed).
__global__ void k_skin_flag(int *__restrict__ flag, unsigned long long *__restrict__ rebuilds,
                            int *__restrict__ rebuilt) {
    const int f = *flag;
    if (f) ++*rebuilds;
    *rebuilt = f;  // read by the passes of this half step (WALK_SKIN)
    *flag = 0;
}

// Wave order of the gradient and continuity+momentum passes (one scene).  In
// cell order a sweep over whole planes keeps about five planes of neighbour
// rows in use, more than one XCD's L2 holds.  Here the fluid goes in units of
// yt consecutive cell rows of one plane (rows [t yt, (t + 1) yt) of plane z:
// one contiguous range of the fluid), tile by tile and, inside a tile, plane
// by plane, so the waves in flight on an XCD read a compact block of rows.
// Each unit's fluid makes waves of up to 64 consecutive particles; the other
// particles [nf, N) follow in waves of 64.  wtab[0] = (fluid waves, all
// waves); wave w is wtab[1 + w] = (first, end); the entries up to cap are
// filled with empty waves.  Every particle lies in exactly one wave, and a
// particle's own arithmetic does not depend on its wave, so the passes'
// results do not depend on the order.  One block of 1024 threads.
__global__ __launch_bounds__(1024) void k_wave_table(Prm q, const int *__restrict__ cs, int yt,
                                                     int2 *__restrict__ wtab, int cap,
                                                     const int *__restrict__ gate) {
    if (gated_off(gate)) return;
    __shared__ int wsum[1024 / 64];
    const int tid = (int)threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int T = (q.Y + yt - 1) / yt;
    const int U = T * q.Z;
    const int nf = cs[q.G];
    int base = 0;  // waves of the units before this round
    for (int u0 = 0; u0 < U; u0 += 1024) {
        const int u = u0 + tid;
        int a = 0, e = 0, nw = 0;
        if (u < U) {
            const int t = u / q.Z, z = u - t * q.Z;
            const int y0 = t * yt, y1 = min(y0 + yt, q.Y);
            a = cs[z * q.XY + y0 * q.X];
            e = cs[z * q.XY + y1 * q.X];
            nw = (e - a + 63) >> 6;
        }
        int s = nw;  // inclusive scan over the wave, then over the block's waves
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(s, o, 64);
            if (lane >= o) s += v;
        }
        if (lane == 63) wsum[wv] = s;
        __syncthreads();
        int pre = 0, total = 0;
        for (int k = 0; k < 1024 / 64; ++k) {
            pre += (k < wv) ? wsum[k] : 0;
            total += wsum[k];
        }
        const int off = base + pre + s - nw;
        for (int k = 0; k < nw; ++k) wtab[1 + off + k] = make_int2(a + 64 * k, min(a + 64 * (k + 1), e));
        base += total;
        __syncthreads();  // wsum is written again next round
    }
    const int nd = (q.N - nf + 63) >> 6;
    for (int k = tid; k < nd; k += 1024) wtab[1 + base + k] = make_int2(nf + 64 * k, min(nf + 64 * (k + 1), q.N));
    for (int w = base + nd + tid; w < cap; w += 1024) wtab[1 + w] = make_int2(q.N, q.N);
    if (tid == 0) wtab[0] = make_int2(base, base + nd);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

int wave_table_cap(int N, int Y, int Z, int yt) { return (N + 63) / 64 + ((Y + yt - 1) / yt) * Z + 2; }

void launch_wave_table(hipStream_t s, const Prm &q, const int *cs, int yt, int2 *wtab, int cap, const int *gate) {
    k_wave_table<<<1, 1024, 0, s>>>(q, cs, yt, wtab, cap, gate);
}
static inline int nblocks(int n) { return (n + ZS_BLOCK - 1) / ZS_BLOCK; }
// Gated (skin mode) launches of the per-particle rebuild kernels use at most
// ZS_GATED_BLOCKS blocks (8 per CU), so that a half step without a rebuild
// costs a small launch; ungated launches take one block per ZS_BLOCK
// particles.
#ifndef ZS_GATED_BLOCKS
#define ZS_GATED_BLOCKS 2048
#endif
static inline int gated_blocks(int n, const int *gate) {
    const int b = nblocks(n);
    return (gate && b > ZS_GATED_BLOCKS) ? ZS_GATED_BLOCKS : b;
}

void launch_hash(hipStream_t s, const Prm &q, const Prm *sp, const Buf &b, uint32_t *key, int *idx, int *killed,
                 const int *gate) {
    if (q.N == 0) return;
    if (sp)
        k_hash<true><<<gated_blocks(q.N, gate), ZS_BLOCK, 0, s>>>(q, sp, b, key, idx, killed, gate);
    else
        k_hash<false><<<gated_blocks(q.N, gate), ZS_BLOCK, 0, s>>>(q, sp, b, key, idx, killed, gate);
}

int radix_blocks(int n, int items) { return (n + ZS_BLOCK * items - 1) / (ZS_BLOCK * items); }

size_t radix_hist_words(int n, int bits) {
    return ((size_t)1 << bits) * ((size_t)radix_blocks(n, RS_ITEMS_MIN) + 1);
}

template <int BITS, int ITEMS>
static int sort_passes(hipStream_t s, int n, int keybits, uint32_t *key[2], int *val[2], uint32_t *hist,
                       bool identity_vals, const int *gate) {
    const int nblk = radix_blocks(n, ITEMS);
    const int R = 1 << BITS;
    uint32_t *total = hist + (size_t)R * nblk;
    int src = 0;
    for (int shift = 0; shift < keybits; shift += BITS) {
        k_radix_hist<BITS, ITEMS><<<nblk, ZS_BLOCK, 0, s>>>(key[src], n, shift, hist, nblk, gate);
        k_radix_rowscan<<<R, ZS_BLOCK, 0, s>>>(hist, nblk, total, gate);
        k_radix_scatter<BITS, ITEMS><<<nblk, ZS_BLOCK, 0, s>>>(key[src], (identity_vals && shift == 0) ? nullptr : val[src],
                                                        key[1 - src], val[1 - src], n, shift, hist, nblk, total, gate);
        src = 1 - src;
    }
    return src;
}

template <int ITEMS>
static int sort_bits(hipStream_t s, int n, int keybits, int bits, uint32_t *key[2], int *val[2], uint32_t *hist,
                     bool identity_vals, const int *gate) {
    switch (bits) {
    case 8: return sort_passes<8, ITEMS>(s, n, keybits, key, val, hist, identity_vals, gate);
    case 9: return sort_passes<9, ITEMS>(s, n, keybits, key, val, hist, identity_vals, gate);
    case 10: return sort_passes<10, ITEMS>(s, n, keybits, key, val, hist, identity_vals, gate);
    case 11: return sort_passes<11, ITEMS>(s, n, keybits, key, val, hist, identity_vals, gate);
    default: return -1;
    }
}

int launch_sort(hipStream_t s, int n, int keybits, int bits, int items, uint32_t *key[2], int *val[2],
                uint32_t *hist, bool identity_vals, const int *gate) {
    if (n == 0) return 0;
    if (items == 8) return sort_bits<8>(s, n, keybits, bits, key, val, hist, identity_vals, gate);
    if (items == 16) return sort_bits<16>(s, n, keybits, bits, key, val, hist, identity_vals, gate);
    return -1;
}

static int cs_tiles(int G) { return (2 * G + 2 + CS_TILE - 1) / CS_TILE; }
size_t cell_start_scratch(int G) { return (size_t)cs_tiles(G) + 1; }

void launch_cell_start(hipStream_t s, const uint32_t *skey, int n, int G, uint32_t flag, int *cs, int *scratch,
                       const int *gate) {
    const int T = cs_tiles(G);
    k_cs_bounds<<<gated_blocks(n, gate), ZS_BLOCK, 0, s>>>(skey, n, G, flag, T, scratch, gate);
    k_cell_start<<<T, ZS_BLOCK, 0, s>>>(skey, G, flag, scratch, cs, gate);
}

void launch_reorder(hipStream_t s, int n, const int *perm, const Buf &a, const Buf &b, const int *gate) {
    if (n == 0) return;
    k_reorder<<<gated_blocks(n, gate), ZS_BLOCK, 0, s>>>(n, perm, a, b, gate);
}

void launch_skin_commit(hipStream_t s, int n, const Buf &a, const Buf &b, float4 *xref, const int *gate) {
    if (n == 0) return;
    k_skin_commit<<<gated_blocks(n, gate), ZS_BLOCK, 0, s>>>(n, a, b, xref, gate);
}

void launch_skin_flag(hipStream_t s, int *flag, unsigned long long *rebuilds, int *rebuilt) {
    k_skin_flag<<<1, 1, 0, s>>>(flag, rebuilds, rebuilt);
}

}  // namespace zs
