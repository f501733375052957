// Device-side definitions shared by the gfx950 kernels of libzsph.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <type_traits>

#include "zsph.h"

namespace zs {

// Kernel view of zsph_params plus derived constants, passed by value.
struct Prm {
    float half_dt, dt, h, inv_h, g, m, rho0, adh, h2, eta2, cs2, alpha, delta_h_cs, h_cs_rho_m;
    float fk;        // -5 adh / h^2 : F(q) = fk (1 - q/2)^3
    float rc2;       // neighbour iff r^2 <= rc2
    float gamma, B;  // EOS: p = B ((rho/rho0)^gamma - 1), B = rho0 c^2 / gamma
    float inv_gamma;
    float grid_size;
    float inv_cell;  // 1 / grid_size
    float trim2;     // (cut-off in cell units)^2, padded: row-trimming radius
    float trim_pad;  // extra half-width (cell units) kept around each trimmed row
    float gx0, gy0, gz0;
    int k;           // cells searched each side
    int X, Y, Z, XY;
    int G;           // number of cells (all scenes of a batch)
    int cell_base;   // first cell of this scene in the key space (0 for one scene)
    uint32_t flag;   // sort-key bit of non-fluid particles (1 << cell bits)
    int N;
    int dim;
    int diffusion;
    // The grid whose outside removes particles (the case's grid).  It equals the
    // neighbour grid above except in skin mode, where the neighbour grid has
    // larger cells (make_skin_prm) and particles are still removed exactly where
    // the case's grid ends (lib/hash.cu:16-23).
    float kill_size;
    int kill_X, kill_Y, kill_Z;
    // Skin mode: the rows of the recorded skin walk (WALK_SKIN) are trimmed to
    // rc + s around a particle's rebuild position (trim2_skin, cell units,
    // padded like trim2), and its rebuild records every candidate within
    // rc + s (rcs2 = (rc + s)^2); 0 outside skin mode.
    float trim2_skin, rcs2;
    // Stage clocks of the neighbour passes (clk nullptr: off): three words per
    // stage ZSPH_ST_* (open time, ticks, launches) for the stages whose bit is
    // set in clk_mask; see stage_stamp.
    unsigned clk_mask;
    // Wave order of the gradient and continuity+momentum passes (one scene):
    // wtab[w] = (first, end) of the particles of wave w, for wt_cap waves, in
    // units of wt_rows cell rows (k_wave_table); wtab nullptr: wave w of the
    // XCD-dealt blocks takes particles [64 w, 64 w + 64).  (An even number of
    // 4-byte fields before the pointers: no padding.)
    int wt_cap, wt_rows;
    unsigned long long *clk;
    const int2 *wtab;
};

// Host-side construction (also used by the C ABI layer).
inline Prm make_prm(const zsph_params &p) {
    Prm q{};
    q.half_dt = p.half_dt;
    q.dt = p.half_dt * 2.0f;
    q.h = p.h;
    q.inv_h = 1.0f / p.h;
    q.g = p.g;
    q.m = p.m;
    q.rho0 = p.rho0;
    q.adh = p.adh;
    q.h2 = p.h2;
    q.eta2 = p.eta2;
    q.cs2 = p.cs2;
    q.alpha = p.alpha;
    q.delta_h_cs = p.delta_h_cs;
    q.h_cs_rho_m = p.h_cs_rho_m;
    q.fk = -5.0f * p.adh / p.h2;
    q.rc2 = p.rc2;
    q.gamma = p.gamma;
    q.B = p.rho0 * p.cs2 / p.gamma;
    q.inv_gamma = 1.0f / p.gamma;
    q.grid_size = p.grid_size;
    q.inv_cell = 1.0f / p.grid_size;
    // Cut-off radius in cell units, padded by 0.1 % plus 1e-3 cells: far
    // above the FP32 rounding of r^2 and of the cell-unit coordinates, so
    // trimming never drops a pair the r^2 <= rc2 predicate accepts.
    q.trim2 = p.rc2 * q.inv_cell * q.inv_cell * 1.002f;
    q.trim_pad = 1e-3f;
    q.gx0 = p.grid_xmin;
    q.gy0 = p.grid_ymin;
    q.gz0 = p.grid_zmin;
    q.k = p.grid_factor;
    q.X = p.grid_xdim;
    q.Y = p.grid_ydim;
    q.Z = p.grid_zdim;
    q.XY = p.grid_xdim * p.grid_ydim;
    q.G = p.grid_xdim * p.grid_ydim * p.grid_zdim;
    int cb = 1;
    while (cb < 30 && (1 << cb) < q.G) ++cb;
    q.flag = 1u << cb;
    q.N = p.ptc_num;
    q.dim = p.dim;
    q.diffusion = p.diffusion;
    q.kill_size = q.grid_size;
    q.kill_X = q.X;
    q.kill_Y = q.Y;
    q.kill_Z = q.Z;
    q.trim2_skin = 0.0f;
    q.rcs2 = 0.0f;
    return q;
}

// Skin-mode neighbour grid for a skin of `skin` h (> 0): cells of edge
// e = 1.001 (rc + s) / k with s = skin h, so the (2k+1)^3 cell box around a
// particle's cell at the last rebuild reaches rc + s; rows are trimmed to
// rc + s/2 around the particle's current position.  While no particle has
// moved more than s/2 since the rebuild, every pair within rc has each
// partner inside a trimmed row of the other's box, so the r^2 <= rc2
// predicate alone still decides the neighbour set.  The grid has the same
// origin and covers the case's grid (X' = floor(X size / e) + 1 <= X cells
// per axis, so G' <= G and the flag bit and cell-table sizes still hold);
// removal stays on the case's grid (kill_*).
inline Prm make_skin_prm(const Prm &q, float skin) {
    Prm s = q;
    const double rc = std::sqrt((double)q.rc2);
    const double sk = (double)skin * (double)q.h;
    const double e = 1.001 * (rc + sk) / (double)q.k;
    s.grid_size = (float)e;
    s.inv_cell = (float)(1.0 / e);
    s.X = (int)std::floor((double)q.X * (double)q.grid_size / e) + 1;
    s.Y = (int)std::floor((double)q.Y * (double)q.grid_size / e) + 1;
    s.Z = (int)std::floor((double)q.Z * (double)q.grid_size / e) + 1;
    if (q.dim == 2) s.Y = q.Y;  // the x-z plane keeps its single row of cells in y
    s.XY = s.X * s.Y;
    s.G = s.XY * s.Z;
    const double rt = (rc + 0.5 * sk) / e;  // trimming radius in cell units
    s.trim2 = (float)(rt * rt * 1.002);
    const double rs = (rc + sk) / e;  // the recorded skin walk's rows
    s.trim2_skin = (float)(rs * rs * 1.002);
    s.rcs2 = (float)((rc + sk) * (rc + sk));
    return s;
}

// One particle buffer (the reference's gpu_ptc_t, data_structure.cuh:158-168,
// with (pos,rho) and (vel,p) packed into float4 for 16-byte coalesced loads).
struct Buf {
    float4 *posr;   // x, y, z, rho
    float4 *velp;   // vx, vy, vz, p
    float4 *tposr;  // saved x, rho at the start of the step
    float4 *tvelp;  // saved v, p
    int *type;
    int *table;
    int *scene;     // scene of each particle (batches only; nullptr for one scene)
};

// The parameters of particle i's scene: the kernel argument for one scene,
// the scene's entry of the parameter table for a batch.
template <bool BATCH>
__device__ __forceinline__ Prm scene_prm(const Prm &q0, const Prm *__restrict__ sp, const Buf &b, int i) {
    if constexpr (BATCH) return sp[b.scene[i]];
    else return q0;
}

// Runs f(q) with the parameters of particle i's scene as a wave-uniform
// value: q0 itself for one scene; in a batch, once per distinct scene among
// the wave's active lanes (scenes are contiguous in cell order, so nearly
// every wave has one) with only that scene's lanes active, the scene's
// parameters read once for the wave instead of per lane (they stay out of
// the vector registers the neighbour walks need).
template <bool BATCH, class F>
__device__ __forceinline__ void with_scene(const Prm &q0, const Prm *__restrict__ sp, const Buf &b, int i, F &&f) {
    if constexpr (!BATCH) {
        f(q0);
    } else {
        const int sc = b.scene[i];
        uint64_t todo = __ballot(1);
        while (todo) {
            const int s = __builtin_amdgcn_readlane(sc, (int)__builtin_ctzll(todo));
            const bool mine = sc == s;
            todo &= ~__ballot(mine);
            if (mine) {
                // the scene index restated as wave-uniform where it is used, so the
                // parameter block is read with scalar loads into scalar registers
                const Prm q = sp[__builtin_amdgcn_readfirstlane(s)];
                f(q);
            }
        }
    }
}

#define ZS_BLOCK 256

// Blocks are dispatched round-robin over the 8 XCDs; remap blocks
// [0, nblocks) among themselves so each XCD (own L2) works on one contiguous
// range of them, and leave blocks at or beyond nblocks where they are.
// nblocks counts the blocks that hold the launch's work: a pass over the
// fluid alone (the first cs[G] particles) gives the remap only the fluid's
// blocks, so that all eight XCDs share that work (remapping the whole grid
// would leave the XCDs whose ranges hold no fluid idle).  Measured on the
// neighbour passes: contiguous ranges beat round-robin blocks and chunks of
// 8 to 128 blocks dealt in turn.
__device__ __forceinline__ int xcd_block(int nblocks) {
    const int b = blockIdx.x;
    if (b >= nblocks) return b;
    const int per = nblocks >> 3, rem = nblocks & 7;
    const int x = b & 7, local = b >> 3;
    return x * per + (x < rem ? x : rem) + local;
}
// Blocks of ZS_BLOCK particles that cover n particles.
__device__ __forceinline__ int blocks_of(int n) { return (n + ZS_BLOCK - 1) / ZS_BLOCK; }
// Blocks of BS particles that cover n particles (the neighbour passes take their
// block size as a tuning parameter: 128, 256 or 512 threads).
template <int BS>
__device__ __forceinline__ int blocks_of_bs(int n) { return (n + BS - 1) / BS; }
// The particle of this thread in a gradient or continuity+momentum pass of
// BS-thread blocks (nf = cs[G]): with a wave table (q.wtab, k_wave_table)
// lane l of the block's wave in that order, the fluid's blocks dealt to the
// XCDs in contiguous ranges; otherwise particle (block) * BS + thread with the
// fluid's blocks dealt the same way.  q.N for a lane without a particle.
template <int BS>
__device__ __forceinline__ int pass_particle(const Prm &q, int nf) {
    if (q.wtab) {
        constexpr int WPB = BS / 64;
        const int2 h = q.wtab[0];  // (fluid waves, all waves)
        const int w = xcd_block((h.x + WPB - 1) / WPB) * WPB + (int)(threadIdx.x >> 6);
        if (w >= h.y) return q.N;
        const int2 e = q.wtab[1 + w];
        const int i = e.x + (int)(threadIdx.x & 63);
        return i < e.y ? i : q.N;
    }
    return xcd_block(blocks_of_bs<BS>(nf)) * BS + (int)threadIdx.x;
}
// Blocks of a gradient or continuity+momentum pass of BS-thread blocks.
inline int pass_grid(const Prm &q, int bs) {
    return q.wtab ? (q.wt_cap + bs / 64 - 1) / (bs / 64) : (q.N + bs - 1) / bs;
}
// Runs f(std::integral_constant<int, BS>) for the tuned block size bs of the
// neighbour passes (any other value runs ZS_BLOCK).
template <class F>
inline void with_block(int bs, F &&f) {
    if (bs == 128) f(std::integral_constant<int, 128>{});
    else if (bs == 512) f(std::integral_constant<int, 512>{});
    else f(std::integral_constant<int, ZS_BLOCK>{});
}

// Cell coordinate along one axis: floor of the FP32 quotient (x - o) / size
// rounded to nearest; returns false outside [0, dim).  NaN positions fail.
// The quotient is the FP64 quotient rounded to FP32: for FP32 operands that
// double rounding equals the correctly rounded FP32 quotient (53 >= 2*24 + 2
// bits, Figueroa 1995), and it does not depend on the build's FP32 division
// mode (this library is built with approximate FP32 division, which puts a
// point whose quotient rounds to an integer k in cell k - 1).
__host__ __device__ inline bool cell_axis(float x, float o, float size, int dim, int *c) {
    const float t = floorf((float)((double)(x - o) / (double)size));
    if (!(t >= 0.0f && t < (float)dim)) return false;
    *c = (int)t;
    return true;
}

// WALK_SKIN's extra inputs of the gradient pass: the rebuild positions, the
// skin hit masks of the last rebuild (slot-major like the exact record) and
// the flag telling whether this half step rebuilt (gates the two forms).
struct SkinRec {
    const float4 *xref = nullptr;
    uint32_t *skin = nullptr;
    const int *rebuilt = nullptr;
};

// Skin mode's displacement test of the integrators (xref == nullptr: off):
// a particle that has moved more than s/2 from its position at the last
// rebuild, or has left the case's grid (to be removed by the next hash),
// sets *flag so that the next half step rebuilds the neighbour structures.
struct SkinCheck {
    const float4 *xref = nullptr;
    int *flag = nullptr;
    float lim_h = 0.0f;  // s/2 in units of the particle's scene h, shrunk by 0.1 %
};

// EOS p(rho).  gamma == 1 is the reference's linear form c^2 (rho - rho0)
// (lib/time_int.cu:39-40); gamma == 7 uses the expanded (1+e)^7 - 1 so the
// small-compression regime keeps full relative precision.
__host__ __device__ inline float eos_p(const Prm &q, float rho) {
    if (q.gamma == 1.0f) return q.cs2 * (rho - q.rho0);
    const float e = (rho - q.rho0) / q.rho0;
    if (q.gamma == 7.0f) {
        const float poly = 7.0f + e * (21.0f + e * (35.0f + e * (35.0f + e * (21.0f + e * (7.0f + e)))));
        return q.B * e * poly;
    }
    return q.B * expm1f(q.gamma * log1pf(e));
}

// Inverse EOS rho(p) for dummy particles (delta_ns.cu:129 for gamma == 1).
__host__ __device__ inline float eos_rho(const Prm &q, float p) {
    if (q.gamma == 1.0f) return p / q.cs2 + q.rho0;
    return q.rho0 + q.rho0 * expm1f(q.inv_gamma * log1pf(p / q.B));
}

// Element j of a particle array, addressed by a 32-bit byte offset from the
// wave-uniform base: a scalar-base vector load with no 64-bit address
// arithmetic.  Needs j * sizeof(T) < 2^32, i.e. N < 2^28 for 16-byte
// elements (zsph_create refuses larger N).
template <class T>
__device__ __forceinline__ T ld(const T *__restrict__ base, int j) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (uint32_t)j * (uint32_t)sizeof(T));
}
// x, y, z of a float4 particle array (12-byte load).
__device__ __forceinline__ float3 ld_xyz(const float4 *__restrict__ base, int j) {
    return *reinterpret_cast<const float3 *>(reinterpret_cast<const char *>(base) + (uint32_t)j * 16u);
}
// Spare elements after the N particles of every position array: the mask
// walk's candidate rounds read up to three past a row's end.
constexpr int POS_PAD = 64;
// N + POS_PAD elements of 16 bytes stay below 2^32 bytes.
constexpr int MAX_PARTICLES = (1 << 28) - POS_PAD;

// Hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32: about 1 ulp),
// without the denormal range scaling of the library forms; the operands in
// the particle-pair bodies (densities, r^2 + eta^2, r^2 of distinct
// particles) are never denormal.
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// r^2 in one fixed FP32 form (shared with the oracle): fma(dz,dz,fma(dy,dy,dx*dx)).
__host__ __device__ inline float dist2(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// True when particle i at x needs a rebuild (see SkinCheck); NaN positions do.
__device__ __forceinline__ bool skin_moved(const Prm &q, const SkinCheck &k, int i, float4 x) {
    const float4 r = k.xref[i];
    const float lim = k.lim_h * q.h;  // the scene's own s/2 (a batch's scenes differ in h)
    if (!(dist2(x.x - r.x, x.y - r.y, x.z - r.z) <= lim * lim)) return true;
    int c;
    return !(cell_axis(x.x, q.gx0, q.kill_size, q.kill_X, &c) && cell_axis(x.y, q.gy0, q.kill_size, q.kill_Y, &c) &&
             cell_axis(x.z, q.gz0, q.kill_size, q.kill_Z, &c));
}

// Stage clock stamp, taken by the first thread of the first block of a kernel
// (blocks are dispatched in index order, so block 0 starts when the kernel
// does, right after the previous kernel on the stream has finished): closes
// the interval of stage close_st opened by an earlier kernel (its ticks and
// launch count grow; a stage not open is left alone) and opens stage
// open_st (-1: none).  A stage's interval thus runs from the start of its
// first kernel to the start of the kernel after its last one: the kernels'
// time plus one launch gap, with no clock kernels of its own in the stream.
__device__ __forceinline__ void stage_stamp(const Prm &q, int close_st, int open_st) {
    if (!q.clk || blockIdx.x != 0 || threadIdx.x != 0) return;
    const unsigned long long now = (unsigned long long)wall_clock64();
    if (close_st >= 0 && (q.clk_mask >> close_st & 1u)) {
        unsigned long long *w = q.clk + 3 * close_st;
        const unsigned long long t0 = w[0];
        if (t0) {
            w[1] += now - t0;
            w[2] += 1ull;
            w[0] = 0ull;
        }
    }
    if (open_st >= 0 && (q.clk_mask >> open_st & 1u)) q.clk[3 * open_st] = now;
}

// One store of the flag per wave whose lanes found a particle that moved.
__device__ __forceinline__ void skin_flag_wave(const SkinCheck &k, bool moved) {
    const uint64_t m = __ballot(moved);
    if (m && (int)(threadIdx.x & 63) == __builtin_ctzll(m)) *k.flag = 1;
}

// Decode a linear cell index into the particle's cell and the x range of
// its search box (clamped to the grid).  A box row (cz + dz, cy + dy) with
// |dz|, |dy| <= k lies outside the clamped box exactly when it lies outside
// the grid, so the rows are checked against the grid's bounds instead of
// keeping the box's y and z limits.
struct CellBox {
    int xlo, xhi;
    int cy, cz;  // the particle's own cell row
};
__device__ __forceinline__ CellBox cell_box(const Prm &q, uint32_t key) {
    const int cz = (int)(key / (uint32_t)q.XY);
    const int rem = (int)key - cz * q.XY;
    const int cy = rem / q.X;
    const int cx = rem - cy * q.X;
    CellBox b;
    b.xlo = max(cx - q.k, 0); b.xhi = min(cx + q.k, q.X - 1);
    b.cy = cy;
    b.cz = cz;
    return b;
}

// Sort keys put every fluid particle before every other particle: key =
// cell | (type != fluid ? flag : 0), flag = 1 << cell bits.  In a batch the
// cell is global: scene s owns cells [cell_base_s, cell_base_s + G_s), so a
// scene's fluid (and its dummies) stay contiguous after the sort.  The cell table
// cs has 2 (G + 1) entries: fluid particles of cell c are [cs[c], cs[c+1]),
// the others (dummies, removed) are [cs[G+1+c], cs[G+1+c+1]).  Fluid
// particles therefore occupy [0, cs[G]) and dummies [cs[G], N).

// Candidate ranges of one row of the search box: fluid particles [fb, fe),
// the others (dummies, removed) [db, de); empty when the row lies outside
// the grid or beyond the cut-off.
struct RowRange {
    int fb = 0, fe = 0, db = 0, de = 0;
};
struct NoPre {
    __device__ __forceinline__ void operator()(int, int) const {}
};

// Row (dz, dy) of the (2k+1)^3 cell box around xi's cell (offsets from that
// cell), trimmed to the x-cells that can hold a point within the cut-off of
// xi.  The trimming works in cell units with a margin, so it only ever keeps
// extra candidates; the r^2 predicate alone decides the neighbour set.
// pre(r, s) runs when the row lies within the cut-off distance, before its
// cell-table loads are issued (a caller's own loads for the row then travel
// together with them).
template <bool FLUID_J, bool DUMMY_J, class P>
__device__ __forceinline__ RowRange box_row(const Prm &q, const CellBox &b, float fx, float fy, float fz,
                                            const int *__restrict__ cs, int dz, int dy, int r, int s, P &pre) {
    RowRange w;
    const int cz = b.cz + dz, cy = b.cy + dy;
    if ((unsigned)cz >= (unsigned)q.Z || (unsigned)cy >= (unsigned)q.Y) return w;
    const float dzc = fmaxf(0.0f, fmaxf((float)cz - fz, fz - (float)(cz + 1)));
    const float dyc = fmaxf(0.0f, fmaxf((float)cy - fy, fy - (float)(cy + 1)));
    const float remz = q.trim2 - dzc * dzc;
    const float rem = remz - dyc * dyc;
    if (rem < 0.0f) return w;
    pre(r, s);
    // the hardware square root (about 1 ulp) is far inside the trimming pad
    const float wd = hw_sqrt(rem) + q.trim_pad;
    const int xa = max(b.xlo, (int)floorf(fx - wd));
    const int xb = min(b.xhi, (int)floorf(fx + wd));
    if (xa > xb) return w;
    const int row = q.cell_base + cz * q.XY + cy * q.X;
    if (FLUID_J) { w.fb = ld(cs, row + xa); w.fe = ld(cs, row + xb + 1); }
    if (DUMMY_J) { w.db = ld(cs, q.G + 1 + row + xa); w.de = ld(cs, q.G + 1 + row + xb + 1); }
    return w;
}

// Calls g(A, rA, B, rB) for the rows of the box in point-mirror pairs: row
// (dz, dy) with row (-dz, -dy), for dz < 0 (every dy) and for dz = 0, dy < 0,
// then the centre row (0, 0) alone (B empty, rB = -1).  r = (dz + k)(2k + 1)
// + dy + k is a row's place in the (2k+1)^2 rows of the box; the mirror of r
// is (2k+1)^2 - 1 - r.  A particle near the top of its cell has more
// neighbours in the rows above and fewer in the rows below, so the hit counts
// of a mirror pair vary less across the lanes of a wave than those of either
// row: the walks visit a pair's hits as one sequence.  pre(r, s): see box_row
// (s = 0 for A, 1 for B).
template <bool FLUID_J, bool DUMMY_J, class G, class P = NoPre>
__device__ __forceinline__ void for_each_row_pair(const Prm &q, float4 xi, uint32_t cell, const int *__restrict__ cs,
                                                  G &&g, P &&pre = P{}) {
    const CellBox b = cell_box(q, cell);
    const float fx = (xi.x - q.gx0) * q.inv_cell;
    const float fy = (xi.y - q.gy0) * q.inv_cell;
    const float fz = (xi.z - q.gz0) * q.inv_cell;
    const int K = 2 * q.k + 1, last = K * K - 1;
    for (int dz = -q.k; dz <= 0; ++dz) {
        const int dyend = (dz < 0) ? q.k : -1;
        for (int dy = -q.k; dy <= dyend; ++dy) {
            const int r = (dz + q.k) * K + dy + q.k;
            const RowRange A = box_row<FLUID_J, DUMMY_J>(q, b, fx, fy, fz, cs, dz, dy, r, 0, pre);
            const RowRange B = box_row<FLUID_J, DUMMY_J>(q, b, fx, fy, fz, cs, -dz, -dy, last - r, 1, pre);
            g(A, r, B, last - r);
        }
    }
    const RowRange C = box_row<FLUID_J, DUMMY_J>(q, b, fx, fy, fz, cs, 0, 0, last / 2, 0, pre);
    g(C, last / 2, RowRange{}, -1);
}

// Calls g(fb, fe, db, de, r) for every row of the box, in the order of
// for_each_row_pair (A then B of each pair); rows beyond the cut-off come
// with empty ranges.
template <bool FLUID_J, bool DUMMY_J, class G>
__device__ __forceinline__ void for_each_row(const Prm &q, float4 xi, uint32_t cell, const int *__restrict__ cs,
                                             G &&g) {
    for_each_row_pair<FLUID_J, DUMMY_J>(q, xi, cell, cs, [&](const RowRange &A, int rA, const RowRange &B, int rB) {
        g(A.fb, A.fe, A.db, A.de, rA);
        if (rB >= 0) g(B.fb, B.fe, B.db, B.de, rB);
    });
}

// Neighbour-walk variants (chosen per context by measurement):
//  WALK_MASK:  each row's candidates (fluid range, then the others) are
//              tested 32 at a time into a per-lane hit mask, then the hits
//              are visited (a point-mirror row pair's as one sequence when
//              both rows fit one chunk); the wave runs the interaction body
//              once per hit of its busiest lane in that sequence.
//  WALK_HITS:  the gradient pass runs the mask walk and records every row
//              chunk's hit mask; the continuity+momentum pass of the same
//              half step walks the same rows and visits the recorded hits
//              without testing candidates.  Particles with a row of more
//              than 64 candidates take the mask walk there.
//  WALK_LIST:  the gradient pass runs the mask walk and records each fluid
//              particle's neighbours, in visiting order, in a global list; the continuity+momentum pass of the same half
//              step runs its interactions from the list with every lane
//              busy and tests no candidates.  Particles with more than
//              LIST_MAX neighbours take the mask walk there (same order,
//              same results).
constexpr int WALK_MASK = 0;
constexpr int WALK_HITS = 1;
constexpr int WALK_LIST = 2;
// WALK_SKIN (skin mode only): the gradient pass of a rebuild half step tests
// each particle's candidates against (rc + s)^2 from its rebuild position,
// records those skin hit masks, and visits them with the exact predicate;
// between rebuilds it replays the skin masks instead of testing.  Either way
// it records the exact hit masks, which the continuity+momentum pass replays
// as in WALK_HITS.  All its rows lie around the rebuild position, trimmed to
// rc + s.  The two gradient-pass forms are separate launches gated on
// whether this half step rebuilt (WALK_SKIN_REC, WALK_SKIN_REPLAY).
constexpr int WALK_SKIN = 3;
constexpr int WALK_SKIN_REC = 4;
constexpr int WALK_SKIN_REPLAY = 5;
constexpr int LIST_MAX = 128;
// List layout: 4-entry pieces (16 bytes), interleaved per wave: quad q
// (entries 4q .. 4q+3) of particle i is at int4 index
// ((i / 64) * LIST_MAX / 4 + q) * 64 + i % 64, so the 64 lanes of a wave
// reading their q-th quads read 1 KB contiguously.
__host__ __device__ inline size_t list_quad(int i, int q) {
    return ((size_t)(i >> 6) * (LIST_MAX / 4) + (size_t)q) * 64 + (size_t)(i & 63);
}
inline size_t list_words(int n) { return ((size_t)(n + 63) >> 6) * 64 * LIST_MAX; }
// WALK_HITS record: 32-bit hit masks, two chunk slots per row of the box,
// slot-major over the particles (hits[slot * N + i]), so a wave's lanes
// storing or loading the same slot touch 256 contiguous bytes.
inline int hits_slots(int k) { return 2 * (2 * k + 1) * (2 * k + 1); }
// Kind of a visited neighbour j, known from the part of its row it came from:
// the visits of one-chunk row pairs pass FluidJ or DummyJ, the chunked rows
// AnyJ (then j < cs[G] tells).
template <int K>
struct JKind {
    static constexpr int value = K;
};
using FluidJ = JKind<0>;
using DummyJ = JKind<1>;
using AnyJ = JKind<2>;
// Whether neighbour j of kind K is a fluid particle (nf = cs[G]).
template <class K>
__device__ __forceinline__ bool fluid_j(K, int j, int nf) {
    if constexpr (K::value == 0) return true;
    else if constexpr (K::value == 1) return false;
    else return j < nf;
}
// Per-neighbour data a walk gathers besides the position: gath(j, kind)
// returns it (NoGather: nothing); its loads are issued with the position's.
struct Empty {};
struct VelGrad {  // velp[j] and gd[j]
    float4 v, g;
};
struct NoGather {
    template <class K>
    __device__ __forceinline__ Empty operator()(int, K) const { return {}; }
};
// Body call: f(j, dx, dy, dz, r2, xj, kind, ext), f(j, dx, dy, dz, r2, xj,
// kind) or f(j, dx, dy, dz, r2, xj), whichever the body takes.
template <class F, class K, class E>
__device__ __forceinline__ decltype(auto) call_body(F &f, int j, float dx, float dy, float dz, float r2, float4 xj,
                                                    K k, const E &e) {
    if constexpr (std::is_invocable_v<F &, int, float, float, float, float, float4, K, const E &>)
        return f(j, dx, dy, dz, r2, xj, k, e);
    else if constexpr (std::is_invocable_v<F &, int, float, float, float, float, float4, K>)
        return f(j, dx, dy, dz, r2, xj, k);
    else
        return f(j, dx, dy, dz, r2, xj);
}
// A hit's index, position and gathered data.
template <class E>
struct Hit {
    int j;
    float4 x;
    E e;
};
// The hit j of kind K: position and gathered data, loads issued together.
template <class G, class K>
__device__ __forceinline__ auto load_hit(const float4 *__restrict__ posr, G &gath, int j, K k) {
    using E = decltype(gath(j, k));
    return Hit<E>{j, ld(posr, j), gath(j, k)};
}
// Keeps a loaded value live at this point.  A body that tests its hit before
// using it (the skin walk's exact predicate) reads the density and the
// gathered record only inside the test's branch, and the compiler then issues
// those loads inside it: a second L2 round trip per hit after the one the
// test waits for.  Holding them before the test issues every load of the hit
// together.
__device__ __forceinline__ void hold(float v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void hold_xyz(const float4 &v) {
    hold(v.x); hold(v.y); hold(v.z);
}
__device__ __forceinline__ void hold_xyz(const Empty &) {}
// Runs the body on hit h of particle i at xi.
template <class F, class K, class E>
__device__ __forceinline__ decltype(auto) run_hit(F &f, float4 xi, const Hit<E> &h, K k) {
    const float dx = xi.x - h.x.x, dy = xi.y - h.x.y, dz = xi.z - h.x.z;
    return call_body(f, h.j, dx, dy, dz, dist2(dx, dy, dz), h.x, k, h.e);
}

// Bits [0, n) of m, and m shifted down by n (n <= 32).
__device__ __forceinline__ uint32_t low_bits(uint32_t m, int n) { return n >= 32 ? m : m & ((1u << n) - 1u); }
__device__ __forceinline__ uint32_t high_bits(uint32_t m, int n) { return n >= 32 ? 0u : m >> n; }

// Visits the hits of a pair of one-chunk rows A, B (at most 32 candidates
// each) as one sequence: A's hits, then B's, each row's in candidate order.
// Bit u of mA is candidate u of row A: the fluid particle A.fb + u for
// u < nA = A.fe - A.fb, else the other particle A.db + u - nA; the same for
// B.  For each hit, in that order, use(load(j, kind), kind, side, u) runs
// (side 0 for row A, 1 for B; u = the hit's bit in its row's mask).  When
// no lane of the wave has a non-fluid candidate in either row and every
// lane's hits fit in 32 bits with B's placed above A's candidates, the
// sequence runs on that 32-bit mask with kind FluidJ; otherwise on the
// 64-bit mask (mA | mB << 32) with kind AnyJ.  Both give the same order.
template <bool FLUID_J, bool DUMMY_J, class L, class U>
__device__ __forceinline__ void visit_rows(uint32_t mA, uint32_t mB, const RowRange &A, const RowRange &B, L &&load,
                                           U &&use) {
    const int nA = A.fe - A.fb, nB = B.fe - B.fb;
    const int bw = mB ? 32 - __clz(mB) : 0;  // B's highest hit + 1
    const bool fluid_only = !DUMMY_J || (A.de == A.db && B.de == B.db);
    if (FLUID_J && __all(fluid_only && nA + bw <= 32)) {
        uint32_t m = mA | (mB ? mB << nA : 0u);
        const int ob = B.fb - nA;
        while (m) {
            const int t = __builtin_ctz(m);
            m &= m - 1u;
            const bool hi = t >= nA;
            use(load((hi ? ob : A.fb) + t, FluidJ{}), FluidJ{}, (int)hi, hi ? t - nA : t);
        }
        return;
    }
    const int dA = A.db - nA, dB = B.db - nB;
    uint64_t m = (uint64_t)mA | ((uint64_t)mB << 32);
    while (m) {
        const int t = __builtin_ctzll(m);
        m &= m - 1u;
        const bool hi = t >= 32;
        const int u = t & 31;
        const int j = ((u < (hi ? nB : nA)) ? (hi ? B.fb : A.fb) : (hi ? dB : dA)) + u;
        use(load(j, AnyJ{}), AnyJ{}, (int)hi, u);
    }
}

// visit_rows with the positions read from the staged rows: the hit of row A's
// fluid range at u is LDS slot fA + u, of its other range slot dA + u (the
// same for B).  The hit's gathered data is gath(j, kind), or, when the slice
// also holds the records of src2, gath_v(j, kind, staged record).  Same order
// and kinds as visit_rows.
template <bool FLUID_J, bool DUMMY_J, class Stg, class G, class GV, class U>
__device__ __forceinline__ void visit_rows_lds(uint32_t mA, uint32_t mB, const RowRange &A, const RowRange &B,
                                               const Stg &stg, int fA, int dA_, int fB, int dB_, G &gath, GV &gath_v,
                                               U &&use) {
    const int nA = A.fe - A.fb, nB = B.fe - B.fb;
    auto hit = [&](int j, int slot, auto kind) {
        if constexpr (std::is_invocable_v<GV &, int, decltype(kind), float4>) {
            using E = decltype(gath_v(j, kind, float4{}));
            return Hit<E>{j, stg.lds[slot], gath_v(j, kind, stg.lds2[slot])};
        } else {
            using E = decltype(gath(j, kind));
            return Hit<E>{j, stg.lds[slot], gath(j, kind)};
        }
    };
    const int bw = mB ? 32 - __clz(mB) : 0;  // B's highest hit + 1
    const bool fluid_only = !DUMMY_J || (A.de == A.db && B.de == B.db);
    if (FLUID_J && __all(fluid_only && nA + bw <= 32)) {
        uint32_t m = mA | (mB ? mB << nA : 0u);
        const int ob = B.fb - nA, sob = fB - nA;
        while (m) {
            const int t = __builtin_ctz(m);
            m &= m - 1u;
            const bool hi = t >= nA;
            use(hit((hi ? ob : A.fb) + t, (hi ? sob : fA) + t, FluidJ{}), FluidJ{}, (int)hi, hi ? t - nA : t);
        }
        return;
    }
    const int dA = A.db - nA, dB = B.db - nB;
    const int sdA = dA_ - nA, sdB = dB_ - nB;
    uint64_t m = (uint64_t)mA | ((uint64_t)mB << 32);
    while (m) {
        const int t = __builtin_ctzll(m);
        m &= m - 1u;
        const bool hi = t >= 32;
        const int u = t & 31;
        const bool fl = u < (hi ? nB : nA);
        const int j = (fl ? (hi ? B.fb : A.fb) : (hi ? dB : dA)) + u;
        const int slot = (fl ? (hi ? fB : fA) : (hi ? sdB : sdA)) + u;
        use(hit(j, slot, AnyJ{}), AnyJ{}, (int)hi, u);
    }
}

// No-op recorder of the mask walk's chunk masks.
struct NoRec {
    __device__ __forceinline__ void operator()(int, int, uint32_t) const {}
};

// LDS staging of a wave's candidate rows.  The lanes of a wave are
// consecutive particles in cell order, so when every active lane lies in one
// cell row (the same y and z cell), the candidates of all lanes in the box
// row at offset (dz, dy) lie in that row's cells [cx_lo - k, cx_hi + k]
// (cx_lo, cx_hi: the first and last active lanes' x cells): one contiguous
// range of the fluid segment and one of the others.  The wave copies those
// ranges of both rows of a pair into its LDS slice with coalesced loads
// (their bounds are read once per wave from the cell table), and each lane
// tests its own candidates there instead of gathering them from L2, then
// reads its hits' positions there too.  The staged records are the positions
// themselves, so the hit masks, and everything after them, are those of the
// unstaged walk.
// Records per wave: the four ranges of a row pair; a larger pair is gathered
// from L2.  Staging a very wide pair costs more than gathering its hits, so
// the best cap is below the LDS limit (about 350 at 7 waves per SIMD).
// Gradient pass per launch, developed flows: dam break 0.376 / 0.372 / 0.367
// / 0.374 / 0.377 ms at 256 / 288 / 316 / 332 / 352 records, wave tank 2.42 /
// 2.38 / 2.37 ms at 256 / 316 / 352.
#ifndef ZS_STAGE_SLOTS
#define ZS_STAGE_SLOTS 316
#endif
constexpr int STAGE_SLOTS = ZS_STAGE_SLOTS;
constexpr int STAGE_PAD = 4;      // a test round reads up to three records past a row's end
struct NoStage {
    float4 *lds = nullptr;  // unused
    float4 *lds2 = nullptr;
    const float4 *src2 = nullptr;
    int cap = 0;
};
// A wave's LDS slice: cap + STAGE_PAD position records (lds) and, when lds2
// is set, the same records of the array src2 (the continuity+momentum pass
// stages velocities and pressures with the positions).  tag: the lane's
// scene in a batch (0 for one scene); a wave stages only when its lanes share
// one scene (and so one grid), as they nearly always do: scenes are
// contiguous in cell order.
struct LdsStage {
    float4 *lds;
    float4 *lds2 = nullptr;
    const float4 *src2 = nullptr;
    int cap = STAGE_SLOTS;
    int tag = 0;
};
// The wave-uniform cell row of the walk (ok: every active lane lies in it).
struct WaveRow {
    bool ok;
    int cz, cy, x0, x1;  // the row's z and y cell; the first and last lanes' x cells
};
__device__ __forceinline__ WaveRow wave_row(const Prm &q, uint32_t cell, int tag) {
    const uint64_t ex = __builtin_amdgcn_read_exec();
    const int last = 63 - __builtin_clzll(ex);
    const uint32_t c0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)cell);
    const uint32_t c1 = (uint32_t)__builtin_amdgcn_readlane((int)cell, last);
    // scenes are non-decreasing along the sorted particles: equal tags at the
    // first and last active lanes put the whole wave in one scene
    const bool one_scene = __builtin_amdgcn_readfirstlane(tag) == __builtin_amdgcn_readlane(tag, last);
    const int r0 = (int)(c0 / (uint32_t)q.X), r1 = (int)(c1 / (uint32_t)q.X);
    WaveRow w;
    w.ok = one_scene && r0 == r1;
    w.x0 = (int)c0 - r0 * q.X;
    w.x1 = (int)c1 - r1 * q.X;
    w.cz = r0 / q.Y;
    w.cy = r0 - w.cz * q.Y;
    return w;
}
// Staged ranges of a row pair: segment starts (particle index) and LDS slots.
struct StagedPair {
    bool ok;
    int sAf, sAd, sBf, sBd;  // first particle of each staged range
    int oAf, oAd, oBf, oBd;  // its first LDS slot
};
// The uniform fluid and other ranges [s, e) of box row (dz, dy) of the wave's row.
template <bool DUMMY_J>
__device__ __forceinline__ void stage_range(const Prm &q, const WaveRow &w, const int *__restrict__ cs, int dz, int dy,
                                            int &sf, int &ef, int &sd, int &ed) {
    sf = ef = sd = ed = 0;
    const int z = w.cz + dz, y = w.cy + dy;
    if (z < 0 || z >= q.Z || y < 0 || y >= q.Y) return;
    const int xa = max(w.x0 - q.k, 0), xb = min(w.x1 + q.k, q.X - 1);
    const int row = q.cell_base + z * q.XY + y * q.X;
    sf = __builtin_amdgcn_readfirstlane(cs[row + xa]);
    ef = __builtin_amdgcn_readfirstlane(cs[row + xb + 1]);
    if (DUMMY_J) {
        sd = __builtin_amdgcn_readfirstlane(cs[q.G + 1 + row + xa]);
        ed = __builtin_amdgcn_readfirstlane(cs[q.G + 1 + row + xb + 1]);
    }
}
// Copies the ranges of box rows (dz, dy) (when useA) and (-dz, -dy) (when
// useB) into the wave's LDS slice; ok = false (nothing copied) when they
// exceed STAGE_SLOTS.  Every active lane of the wave takes part.
template <bool DUMMY_J>
__device__ __forceinline__ StagedPair stage_pair(const Prm &q, const WaveRow &w, const int *__restrict__ cs,
                                                 const float4 *__restrict__ posr, const LdsStage &stg, int dz,
                                                 int dy, bool useA, bool useB) {
    StagedPair p;
    int eAf, eAd, eBf, eBd;
    p.sAf = p.sAd = p.sBf = p.sBd = 0;
    eAf = eAd = eBf = eBd = 0;
    if (useA) stage_range<DUMMY_J>(q, w, cs, dz, dy, p.sAf, eAf, p.sAd, eAd);
    if (useB) stage_range<DUMMY_J>(q, w, cs, -dz, -dy, p.sBf, eBf, p.sBd, eBd);
    const int lAf = eAf - p.sAf, lAd = eAd - p.sAd, lBf = eBf - p.sBf, lBd = eBd - p.sBd;
    p.oAf = 0;
    p.oAd = lAf;
    p.oBf = lAf + lAd;
    p.oBd = p.oBf + lBf;
    const int total = p.oBd + lBd;
    p.ok = total <= stg.cap;
    if (!p.ok) return p;
    const uint64_t ex = __builtin_amdgcn_read_exec();
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ex >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ex, 0u));
    const int nact = __builtin_popcountll(ex);
    // the previous pair's reads of the slice come before these writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int t = rank; t < total; t += nact) {
        const int src = (t < p.oAd) ? p.sAf + t
                        : (t < p.oBf) ? p.sAd + (t - p.oAd)
                        : (t < p.oBd) ? p.sBf + (t - p.oBf)
                                      : p.sBd + (t - p.oBd);
        stg.lds[t] = ld(posr, src);
        if (stg.lds2) stg.lds2[t] = ld(stg.src2, src);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return p;
}
// Hit mask of the cnt (<= 32) consecutive staged records from slot s against
// c: bit u set when record s + u lies within r2max of c.  A round past cnt
// reads up to three records that follow (STAGE_PAD) and the final mask
// clears their bits.
__device__ __forceinline__ uint32_t test_run_lds(const float4 *lds, int s, int cnt, float4 c, float r2max) {
    uint32_t m = 0;
    const float4 *p = lds + s;
    for (int t = 0; t < cnt; t += 4, p += 4) {
        float3 xs[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) xs[e] = make_float3(p[e].x, p[e].y, p[e].z);
        uint32_t nib = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) nib |= (dist2(c.x - xs[e].x, c.y - xs[e].y, c.z - xs[e].z) <= r2max) ? (1u << e) : 0u;
        m |= nib << t;  // t <= 28
    }
    return (cnt < 32) ? m & ((1u << cnt) - 1u) : m;
}
// Hit mask of a row (at most 32 candidates) from its staged ranges: fluid
// candidates from slot so + (w.fb - sf), the others from sd's.
template <bool FLUID_J, bool DUMMY_J>
__device__ __forceinline__ uint32_t test_row_lds(const float4 *lds, const RowRange &w, int sf, int of, int sd, int od,
                                                 float4 c, float r2max) {
    const int nfl = w.fe - w.fb;
    uint32_t m = FLUID_J ? test_run_lds(lds, of + (w.fb - sf), nfl, c, r2max) : 0u;
    if (DUMMY_J && w.de > w.db) m |= test_run_lds(lds, od + (w.db - sd), w.de - w.db, c, r2max) << nfl;
    return m;
}

// Visit every neighbour j != i (FP32 r^2 <= rc2) of particle i at xi whose
// segment is enabled (FLUID_J: j < cs[G]; DUMMY_J: j >= cs[G]).
// f(j, dx, dy, dz, r2, xj[, kind]) with x_ij = xi - xj (kind: FluidJ, DummyJ
// or AnyJ).  Every walk variant visits the neighbours in this order: row
// pairs in for_each_row_pair order, row A before row B, each row's
// candidates (fluid range, then the others) in ascending order.  A row's
// candidates are tested into a per-lane hit mask; when both rows of a pair
// have at most 32 candidates their hits are visited as one sequence
// (visit_rows), else row by row, 32 candidates at a time.  The wave runs f
// once per hit of its busiest lane in each sequence.  rec(r, c, m) receives
// the hit mask m of chunk c of box row r; gath(j, kind) the data gathered
// with each hit's position (passed to f as ext).
template <int WALK, bool FLUID_J, bool DUMMY_J, class F, class Rec = NoRec, class Gath = NoGather,
          class Stg = NoStage>
__device__ __forceinline__ void for_each_neighbor(const Prm &q, int i, float4 xi, uint32_t cell,
                                                  const int *__restrict__ cs, const float4 *__restrict__ posr, F &&f,
                                                  Rec &&rec = Rec{}, Gath &&gath = Gath{}, Stg stg = Stg{}) {
    constexpr bool STAGED = std::is_same_v<Stg, LdsStage>;
    WaveRow wr{};
    if constexpr (STAGED) wr = wave_row(q, cell, stg.tag);
    auto load = [&](int j, auto kind) { return load_hit(posr, gath, j, kind); };
    auto use = [&](const auto &h, auto kind, int, int) { run_hit(f, xi, h, kind); };
    auto visit = [&](int j, auto kind) { use(load(j, kind), kind, 0, 0); };
    // Hit mask of candidates [base, base + cnt) (cnt <= 32) of a row whose
    // candidate u is off(u) + u, off(u) = fb for u < nfl, else dbo = db - nfl
    // (rows of more than 32 candidates); a round past cnt reads up to three
    // particles that follow (POS_PAD) and the final mask clears their bits.
    auto test_chunk = [&](int base, int cnt, int fb, int nfl, int dbo) {
        uint32_t m = 0;
        for (int t = 0; t < cnt; t += 4) {
            float3 xs[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int u = base + t + e;
                xs[e] = ld_xyz(posr, ((u < nfl) ? fb : dbo) + u);
            }
            __builtin_amdgcn_sched_barrier(0);
            uint32_t nib = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                nib |= (dist2(xi.x - xs[e].x, xi.y - xs[e].y, xi.z - xs[e].z) <= q.rc2) ? (1u << e) : 0u;
            m |= nib << t;  // t <= 28
        }
        return (cnt < 32) ? m & ((1u << cnt) - 1u) : m;
    };
    // i's place among a row's candidates (-1: not in the row)
    auto self_place = [&](const RowRange &w) {
        return (FLUID_J && i >= w.fb && i < w.fe)   ? i - w.fb
               : (DUMMY_J && i >= w.db && i < w.de) ? (w.fe - w.fb) + (i - w.db)
                                                    : -1;
    };
    // one row, 32 candidates at a time
    auto row_chunks = [&](const RowRange &w, int r) {
        const int nfl = w.fe - w.fb;
        const int total = nfl + (w.de - w.db);
        const int dbo = w.db - nfl;
        const int us = self_place(w);
        for (int base = 0, c = 0; base < total; base += 32, ++c) {
            uint32_t m = test_chunk(base, min(32, total - base), w.fb, nfl, dbo);
            if ((uint32_t)(us - base) < 32u) m &= ~(1u << (us - base));
            rec(r, c, m);
            while (m) {
                const int u = base + __builtin_ctz(m);
                m &= m - 1u;
                visit(((u < nfl) ? w.fb : dbo) + u, AnyJ{});
            }
        }
    };
    for_each_row_pair<FLUID_J, DUMMY_J>(q, xi, cell, cs, [&](const RowRange &A, int rA, const RowRange &B, int rB) {
        const int tA = (A.fe - A.fb) + (A.de - A.db);
        const int tB = (B.fe - B.fb) + (B.de - B.db);
        StagedPair sp{};
        if constexpr (STAGED) {
            // wave-uniform: the pair's box-row offsets, and whether any lane has candidates
            const int K = 2 * q.k + 1;
            const bool useA = __any(tA > 0 && tA <= 32 && tB <= 32);
            const bool useB = __any(tB > 0 && tA <= 32 && tB <= 32);
            if (wr.ok && (useA || useB))
                sp = stage_pair<DUMMY_J>(q, wr, cs, posr, stg, rA / K - q.k, rA % K - q.k, useA, useB);
        }
        if (tA <= 32 && tB <= 32) {
            uint32_t mA = 0, mB = 0;
            if (tA) {
                if (STAGED && sp.ok) mA = test_row_lds<FLUID_J, DUMMY_J>(stg.lds, A, sp.sAf, sp.oAf, sp.sAd, sp.oAd, xi, q.rc2);
                else mA = test_chunk(0, tA, A.fb, A.fe - A.fb, A.db - (A.fe - A.fb));
                const int us = self_place(A);
                if ((uint32_t)us < 32u) mA &= ~(1u << us);
                rec(rA, 0, mA);
            }
            if (tB) {
                if (STAGED && sp.ok) mB = test_row_lds<FLUID_J, DUMMY_J>(stg.lds, B, sp.sBf, sp.oBf, sp.sBd, sp.oBd, xi, q.rc2);
                else mB = test_chunk(0, tB, B.fb, B.fe - B.fb, B.db - (B.fe - B.fb));
                const int us = self_place(B);
                if ((uint32_t)us < 32u) mB &= ~(1u << us);
                rec(rB, 0, mB);
            }
            if (STAGED && sp.ok) {  // the hits' positions from the staged rows, their other data gathered
                NoGather none;
                visit_rows_lds<FLUID_J, DUMMY_J>(mA, mB, A, B, stg, sp.oAf + (A.fb - sp.sAf), sp.oAd + (A.db - sp.sAd),
                                                 sp.oBf + (B.fb - sp.sBf), sp.oBd + (B.db - sp.sBd), gath, none, use);
            } else {
                visit_rows<FLUID_J, DUMMY_J>(mA, mB, A, B, load, use);
            }
        } else {
            row_chunks(A, rA);
            if (rB >= 0) row_chunks(B, rB);
        }
    });
}

// WALK_HITS, replay side: the walk of for_each_neighbor with each chunk's
// hit mask read from the record (hits[(2 r + c) n + i], written by the
// gradient pass of the same half step on the same buffer, so the rows and
// their candidate ranges are the same) instead of tested, in the same order
// (row pairs merged the same way).  Every row of the particle must have at
// most 64 candidates.
template <bool FLUID_J, bool DUMMY_J, class F, class Gath = NoGather, class Stg = NoStage, class GV = NoGather>
__device__ __forceinline__ void for_each_recorded_at(const Prm &q, int i, float4 xi, float4 xc, uint32_t cell,
                                                     const int *__restrict__ cs, const float4 *__restrict__ posr,
                                                     const uint32_t *__restrict__ hits, size_t n, F &&f,
                                                     Gath &&gath = Gath{}, Stg stg = Stg{}, GV &&gath_v = GV{});
template <bool FLUID_J, bool DUMMY_J, class F, class Gath = NoGather, class Stg = NoStage, class GV = NoGather>
__device__ __forceinline__ void for_each_recorded(const Prm &q, int i, float4 xi, uint32_t cell,
                                                  const int *__restrict__ cs, const float4 *__restrict__ posr,
                                                  const uint32_t *__restrict__ hits, size_t n, F &&f,
                                                  Gath &&gath = Gath{}, Stg stg = Stg{}, GV &&gath_v = GV{}) {
    for_each_recorded_at<FLUID_J, DUMMY_J>(q, i, xi, xi, cell, cs, posr, hits, n, f, gath, stg, gath_v);
}
// The same with the rows trimmed around xc (WALK_SKIN: the rebuild position)
// while the pair geometry uses the current position xi.
template <bool FLUID_J, bool DUMMY_J, class F, class Gath, class Stg, class GV>
__device__ __forceinline__ void for_each_recorded_at(const Prm &q, int i, float4 xi, float4 xc, uint32_t cell,
                                                     const int *__restrict__ cs, const float4 *__restrict__ posr,
                                                     const uint32_t *__restrict__ hits, size_t n, F &&f,
                                                     Gath &&gath, Stg stg, GV &&gath_v) {
    constexpr bool STAGED = std::is_same_v<Stg, LdsStage>;
    WaveRow wr{};
    if constexpr (STAGED) wr = wave_row(q, cell, stg.tag);
    auto load = [&](int j, auto kind) { return load_hit(posr, gath, j, kind); };
    auto use = [&](const auto &h, auto kind, int, int) { run_hit(f, xi, h, kind); };
    auto visit = [&](int j, auto kind) { use(load(j, kind), kind, 0, 0); };
    // one row, chunk by chunk (m0: its first chunk's mask)
    auto row_chunks = [&](const RowRange &w, int r, uint32_t m0) {
        const int nfl = w.fe - w.fb;
        const int total = nfl + (w.de - w.db);
        for (int base = 0, c = 0; base < total; base += 32, ++c) {
            uint32_t m = c ? hits[(size_t)(2 * r + c) * n + (size_t)i] : m0;
            while (m) {
                const int u = base + __builtin_ctz(m);
                m &= m - 1u;
                visit(((u < nfl) ? w.fb : w.db - nfl) + u, AnyJ{});
            }
        }
    };
    // the first masks of the pair's rows, loaded with their cell-table entries
    uint32_t mA = 0, mB = 0;
    for_each_row_pair<FLUID_J, DUMMY_J>(
        q, xc, cell, cs,
        [&](const RowRange &A, int rA, const RowRange &B, int rB) {
            const int tA = (A.fe - A.fb) + (A.de - A.db);
            const int tB = (B.fe - B.fb) + (B.de - B.db);
            // a row without candidates has no record (its load, if any, is stale)
            const uint32_t a = tA ? mA : 0u, b = tB ? mB : 0u;
            const bool fast = tA <= 32 && tB <= 32;
            StagedPair sp{};
            if constexpr (STAGED) {  // wave-uniform, before any lane leaves
                const int K = 2 * q.k + 1;
                const bool useA = __any(fast && a != 0u), useB = __any(fast && b != 0u);
                if (wr.ok && (useA || useB))
                    sp = stage_pair<DUMMY_J>(q, wr, cs, posr, stg, rA / K - q.k, rA % K - q.k, useA, useB);
            }
            if (fast) {
                if ((a | b) == 0u) return;  // no neighbour in either row
                if (STAGED && sp.ok)
                    visit_rows_lds<FLUID_J, DUMMY_J>(a, b, A, B, stg, sp.oAf + (A.fb - sp.sAf), sp.oAd + (A.db - sp.sAd),
                                                     sp.oBf + (B.fb - sp.sBf), sp.oBd + (B.db - sp.sBd), gath, gath_v,
                                                     use);
                else
                    visit_rows<FLUID_J, DUMMY_J>(a, b, A, B, load, use);
            } else {
                row_chunks(A, rA, a);
                if (rB >= 0) row_chunks(B, rB, b);
            }
        },
        [&](int r, int side) {
            const uint32_t m = hits[(size_t)(2 * r) * n + (size_t)i];
            if (side) mB = m;
            else mA = m;
        });
}

// WALK_SKIN's gradient-pass walk of particle i (current position xi, rebuild
// position xc): the rows of the box around its rebuild cell trimmed to
// rc + s around xc (qr.trim2 = trim2_skin).  REPLAY = false (the rebuild
// half step): each chunk's candidates are tested against qr.rcs2 from xc and
// the skin mask goes to rec_skin(r, c, m); REPLAY = true: the skin mask is
// read from skin[(2 r + c) n + i] (every row of the particle within 64
// candidates).  Every skin hit is visited: f(j, dx, dy, dz, r2, xj[, kind])
// returns whether j is a neighbour (the caller applies r^2 <= rc2), and the chunk's
// exact mask goes to rec_exact(r, c, m).  Row pairs and chunks are visited
// as in for_each_neighbor, so the exact record replays in the same order.
template <bool REPLAY, class F, class RS, class RE, class Gath = NoGather>
__device__ __forceinline__ void skin_walk(const Prm &qr, int i, float4 xi, float4 xc, uint32_t cell,
                                          const int *__restrict__ cs, const float4 *__restrict__ posr,
                                          const uint32_t *__restrict__ skin, size_t n, F &&f, RS &&rec_skin,
                                          RE &&rec_exact, Gath &&gath = Gath{}) {
    auto load = [&](int j, auto kind) { return load_hit(posr, gath, j, kind); };
    auto visit = [&](int j, auto kind) -> bool { return run_hit(f, xi, load(j, kind), kind); };
    // skin mask of candidates [base, base + cnt) (cnt <= 32), tested from xc
    // (rows of more than 32 candidates)
    auto test_chunk = [&](int base, int cnt, int fb, int nfl, int dbo) {
        uint32_t m = 0;
        for (int t = 0; t < cnt; t += 4) {
            float3 xs[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int u = base + t + e;
                xs[e] = ld_xyz(posr, ((u < nfl) ? fb : dbo) + u);
            }
            __builtin_amdgcn_sched_barrier(0);
            uint32_t nib = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                nib |= (dist2(xc.x - xs[e].x, xc.y - xs[e].y, xc.z - xs[e].z) <= qr.rcs2) ? (1u << e) : 0u;
            m |= nib << t;
        }
        return (cnt < 32) ? m & ((1u << cnt) - 1u) : m;
    };
    auto self_place = [&](const RowRange &w) {
        return (i >= w.fb && i < w.fe) ? i - w.fb : (i >= w.db && i < w.de) ? (w.fe - w.fb) + (i - w.db) : -1;
    };
    auto row_chunks = [&](const RowRange &w, int r, uint32_t m0) {
        const int nfl = w.fe - w.fb;
        const int total = nfl + (w.de - w.db);
        const int dbo = w.db - nfl;
        const int us = self_place(w);
        for (int base = 0, c = 0; base < total; base += 32, ++c) {
            uint32_t m;
            if constexpr (REPLAY) {
                m = c ? skin[(size_t)(2 * r + c) * n + (size_t)i] : m0;
            } else {
                m = test_chunk(base, min(32, total - base), w.fb, nfl, dbo);
                if ((uint32_t)(us - base) < 32u) m &= ~(1u << (us - base));
                rec_skin(r, c, m);
            }
            uint32_t e = 0;
            while (m) {
                const int b = __builtin_ctz(m);
                m &= m - 1u;
                const int u = base + b;
                if (visit(((u < nfl) ? w.fb : dbo) + u, AnyJ{})) e |= 1u << b;
            }
            rec_exact(r, c, e);
        }
    };
    uint32_t pA = 0, pB = 0;  // REPLAY: the first skin masks of the pair's rows
    for_each_row_pair<true, true>(
        qr, xc, cell, cs,
        [&](const RowRange &A, int rA, const RowRange &B, int rB) {
            const int tA = (A.fe - A.fb) + (A.de - A.db);
            const int tB = (B.fe - B.fb) + (B.de - B.db);
            uint32_t mA = 0, mB = 0;
            if constexpr (REPLAY) {  // a row without candidates has no record
                mA = tA ? pA : 0u;
                mB = tB ? pB : 0u;
            }
            if (tA <= 32 && tB <= 32) {
                if constexpr (!REPLAY) {
                    if (tA) {
                        mA = test_chunk(0, tA, A.fb, A.fe - A.fb, A.db - (A.fe - A.fb));
                        const int us = self_place(A);
                        if ((uint32_t)us < 32u) mA &= ~(1u << us);
                        rec_skin(rA, 0, mA);
                    }
                    if (tB) {
                        mB = test_chunk(0, tB, B.fb, B.fe - B.fb, B.db - (B.fe - B.fb));
                        const int us = self_place(B);
                        if ((uint32_t)us < 32u) mB &= ~(1u << us);
                        rec_skin(rB, 0, mB);
                    }
                }
                uint32_t eA = 0, eB = 0;  // the exact masks, in the rows' candidate bits
                visit_rows<true, true>(mA, mB, A, B, load, [&](const auto &h, auto kind, int side, int u) {
                    if (run_hit(f, xi, h, kind)) {
                        if (side) eB |= 1u << u;
                        else eA |= 1u << u;
                    }
                });
                if (tA) rec_exact(rA, 0, eA);
                if (tB) rec_exact(rB, 0, eB);
            } else {
                row_chunks(A, rA, mA);
                if (rB >= 0) row_chunks(B, rB, mB);
            }
        },
        [&](int r, int side) {
            if constexpr (REPLAY) {
                const uint32_t m = skin[(size_t)(2 * r) * n + (size_t)i];
                if (side) pB = m;
                else pA = m;
            }
        });
}

// Nontemporal (streaming) 16-byte store and load.
typedef int i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void nt_store(int4 *p, int4 v) {
    const i32x4 x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<i32x4 *>(p));
}
__device__ __forceinline__ int4 nt_load(const int4 *p) {
    const i32x4 x = __builtin_nontemporal_load(reinterpret_cast<const i32x4 *>(p));
    return make_int4(x.x, x.y, x.z, x.w);
}

// WALK_LIST, recording side: the neighbours of particle i, in the order
// put() receives them, into its list quads.  Four entries are staged in
// registers and written as one 16-byte store; entries past LIST_MAX are
// counted but not stored.  The list is written once and read once, by the
// next kernel: the stores (and its loads) are nontemporal, keeping the
// neighbour gathers' lines in L2.
struct ListWriter {
    int4 *list;
    int i;
    int cnt = 0;
    int s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    __device__ ListWriter(int *l, int particle) : list(reinterpret_cast<int4 *>(l)), i(particle) {}
    __device__ __forceinline__ void put(int j) {
        s0 = s1; s1 = s2; s2 = s3; s3 = j;
        ++cnt;
        if ((cnt & 3) == 0 && cnt <= LIST_MAX) nt_store(&list[list_quad(i, (cnt >> 2) - 1)], make_int4(s0, s1, s2, s3));
    }
    // Stores the last cnt % 4 entries (the rest of their quad is never
    // read) and returns the neighbour count.
    __device__ __forceinline__ int finish() {
        const int r = cnt & 3;
        if (r && cnt <= LIST_MAX) {
            const int4 v = (r == 1) ? make_int4(s3, 0, 0, 0)
                         : (r == 2) ? make_int4(s2, s3, 0, 0)
                                    : make_int4(s1, s2, s3, 0);
            nt_store(&list[list_quad(i, cnt >> 2)], v);
        }
        return cnt;
    }
};

}  // namespace zs
