/*
 * zsph.h - C ABI of the MI355X (gfx950) delta-SPH step library (libzsph.so).
 *
 * Every entry point replaces one host function or kernel launch of the
 * reference single-GPU driver (single_gpu_sph_src/deltasph/main.cu:58-90);
 * the reference symbol is cited on each declaration.  The ABI uses plain C
 * types only: host arrays are caller-owned, device memory is owned by an
 * opaque context, every call returns 0 on success or a negative ZSPH_E* code
 * (the library never exits the process; zsph_last_error() has the text).
 *
 * Host array layout (the reference's cpu_output_t, data_structure.cuh:146-156):
 *   pos  float[3N]  (x,y,z)        vel  float[3N]
 *   rhop float[2N]  (rho,p)        acc_drhodt float[4N] (ax,ay,az,drho/dt)
 *   type int[N]  (1 fluid, 0 fixed dummy wall, 2 floating-body dummy,
 *                 3 moving-wall dummy (piston), 100 removed)
 *   table int[N] (original id - current index, sort.cu:107)
 *   hash  int[N]
 */
#ifndef ZSPH_H
#define ZSPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZSPH_OK 0
#define ZSPH_E_ARG (-1)       /* invalid argument / inconsistent parameters */
#define ZSPH_E_HIP (-2)       /* HIP runtime error                          */
#define ZSPH_E_NOMEM (-3)     /* device or host allocation failed           */
#define ZSPH_E_DT (-4)        /* dt above the stability bound (check.cu:3-18) */
#define ZSPH_E_IO (-5)        /* file could not be read / written / parsed  */
#define ZSPH_E_STATE (-6)     /* call out of order (e.g. sort before hash)  */
#define ZSPH_E_NODEV (-7)     /* no HIP device                              */

#define ZSPH_PATH_MAX 256

/* Particle types (the reference only distinguishes type==1 from the rest). */
#define ZSPH_FLUID 1
#define ZSPH_WALL 0
#define ZSPH_BODY 2
#define ZSPH_PISTON 3
#define ZSPH_KILLED 100

/* Density-diffusion form. */
#define ZSPH_DIFF_ANTUONO 0 /* equation/delta_ns.cu (deltasph executable)  */
#define ZSPH_DIFF_MC 1      /* equation/simple_ns.cu (simplesph executable) */

/*
 * Case file contents: the reference's cpu_json_t (data_structure.cuh:6-51,
 * keys read at io/read_json.cu:12-49) plus optional extension keys.
 */
typedef struct zsph_json {
    float g, rho0, dx, h_factor, r_factor, eta_factor, cs_factor, delta, alpha;
    float xmin, xmax, ymin, ymax, zmin, zmax;
    float grid_layer_factor, grid_size_factor;
    float dt;
    int start_step, current_step, end_step, output_step;
    int ptc_num, water_ptc_num, air_ptc_num, wall_ptc_num, rigid_ptc_num;
    int thread_num, gpu_id;
    char input_path[ZSPH_PATH_MAX];
    char output_path[ZSPH_PATH_MAX];
    char git_hash[64];
    /* extensions; absent keys keep the reference behaviour */
    int dim;         /* 3 (reference) or 2 (x-z plane, y == 0)                  */
    float gamma;     /* EOS exponent; 1 = linear p = c^2 (rho - rho0) (reference) */
    float c0;        /* > 0: explicit sound speed, else cs_factor*sqrt(|g|(zmax-zmin)) */
    float cfl;       /* > 0: dt = cfl * h / c0, else "dt" key                     */
    int diffusion;   /* ZSPH_DIFF_ANTUONO or ZSPH_DIFF_MC                          */
    int output_fields; /* ZSPH_F_* mask written by the VTK writer                  */
    int binary;      /* 1 = big-endian binary legacy VTK (ZSPH_BINARY), 0 = ASCII */
} zsph_json;

/* Derived host parameters: the reference's cpu_param_t (data_structure.cuh:53-105). */
typedef struct zsph_cpu_param {
    float g, m, rho0, dx, h, r, eta, cs, adh, delta, alpha;
    float dt;
    int start_step, current_step, end_step, output_step;
    float xmin, xmax, ymin, ymax, zmin, zmax;
    int grid_factor;
    float grid_size;
    float grid_xmin, grid_ymin, grid_zmin;
    float grid_xmax, grid_ymax, grid_zmax;
    int grid_xdim, grid_ydim, grid_zdim;
    int grid_num, grid_hash_min, grid_hash_max;
    int ptc_num, water_ptc_num, air_ptc_num, wall_ptc_num, rigid_ptc_num;
    int thread_num, gpu_id;
    char input_path[ZSPH_PATH_MAX];
    char output_path[ZSPH_PATH_MAX];
    char git_hash[64];
    int dim;
    float gamma;
    int diffusion;
    int output_fields;
    int binary;
} zsph_cpu_param;

/*
 * Device parameter block: a superset of the reference's gpu_param_t
 * (data_structure.cuh:107-134), which the reference keeps in __constant__
 * memory (deltasph/main.cu:47).  Here it is copied into the context and
 * passed to every kernel by value.
 */
typedef struct zsph_params {
    float half_dt, h, g, m, rho0, adh, h2, eta2, cs2, alpha, delta_h_cs, h_cs_rho_m;
    float grid_size;
    int grid_factor;
    float grid_xmin, grid_ymin, grid_zmin;
    int grid_xdim, grid_ydim, grid_zdim;
    int grid_hash_min, grid_hash_max;
    int ptc_num;
    /* extensions */
    int dim;        /* 2 or 3                                               */
    float gamma;    /* EOS exponent (1 = reference linear EOS)             */
    int diffusion;  /* ZSPH_DIFF_*                                          */
    float rc2;      /* FP32 cut-off (2h)^2 = 4*h2: neighbour iff r^2 <= rc2 */
} zsph_params;

/* Host particle arrays (any pointer may be NULL to skip that field). */
typedef struct zsph_host_state {
    float *pos;
    float *vel;
    float *rhop;
    float *acc_drhodt;
    int *type;
    int *table;
    int *hash;
} zsph_host_state;

/* Field mask for zsph_download / the VTK writer (replaces ZSPH_OUTPUT_*). */
#define ZSPH_F_POS 1
#define ZSPH_F_VEL 2
#define ZSPH_F_RHOP 4
#define ZSPH_F_ACC 8
#define ZSPH_F_TYPE 16
#define ZSPH_F_TABLE 32
#define ZSPH_F_HASH 64
#define ZSPH_F_ALL 127

typedef struct zsph_ctx zsph_ctx;

/* ------------------------------------------------------------------------ */
/* Host set-up (CPU only)                                                    */
/* ------------------------------------------------------------------------ */

/* Fill *j with the defaults of every optional key (dim 3, gamma 1, ...). */
void zsph_json_defaults(zsph_json *j);

/* io/read_json.cu:5-51 read_json.  Reads a flat JSON object. */
int zsph_read_json(const char *path, zsph_json *out);

/* lib/init.cu:153-219 set_cpu_param + lib/check.cu:3-18 check_dt
 * (returns ZSPH_E_DT instead of exit(1)). */
int zsph_set_cpu_param(zsph_cpu_param *out, const zsph_json *in);

/* lib/init.cu:221-251 set_gpu_param. */
void zsph_set_gpu_param(zsph_params *out, const zsph_cpu_param *in);

/*
 * Dam-break case generator (the reference reads input.vtk instead,
 * io/read_vkt.cu:3-92; no case files ship with it).  Lattice fluid column
 * [0,cw]x[0,wy]x[0,ch] at (i+1/2)dx in a tank [0,lx]x[0,wy]x[0,lz] with
 * `layers` dummy layers on the floor and the four side walls (dim 2: x-z
 * plane, floor and two walls).  Hydrostatic initial density through the EOS.
 * Writes into *j the domain, counts and dt derived from the geometry.  Call
 * with st == NULL to get the particle count in *n_out, then again with
 * arrays of that size.
 */
typedef struct zsph_dambreak {
    int dim;
    float dx;
    float lx, wy, lz;   /* tank interior                              */
    float cw, ch;       /* water column width (x) and height (z)      */
    int layers;         /* dummy wall layers (3 in the configs)       */
    float wall_height;  /* side walls up to this z (<= 0: lz)          */
} zsph_dambreak;
int zsph_case_dambreak(const zsph_dambreak *c, zsph_json *j, zsph_host_state *st, int *n_out);

/* Legacy VTK (io/write_vtk.cu:29-208 write_vtk, io/read_vkt.cu:3-92 read_vtk).
 * The writer emits every field in `fields`; binary = big-endian. The reader
 * accepts float or double arrays, BINARY or ASCII, fields p, rho, vel, acc,
 * type, table (missing p/rho/vel/acc default to 0/rho0/0/0 like the reference). */
int zsph_write_vtk(const char *path, const zsph_host_state *st, int n, int fields, int binary);
int zsph_read_vtk(const char *path, zsph_host_state *st, int n, float rho0);
int zsph_vtk_count(const char *path, int *n_out);

/* ------------------------------------------------------------------------ */
/* Rigid body (6-DOF), FP64                                                  */
/* ------------------------------------------------------------------------ */

/* The reference's rigid_t (Test_6DOF/header/data_structure.hpp:36-50) in
 * double precision.  q is scalar-first; I is the inertia tensor about the
 * centre of mass in the body frame (row-major); F, T are the applied force
 * and torque about c in the world frame. */
typedef struct zsph_rigid {
    double c[3], vel[3], omega[3], q[4];
    double F[3], T[3], acc[3], alpha[3];
    double I[9];
    double mass;
} zsph_rigid;

/* Test_6DOF/equation/rigid_info.cpp:3-39 gravity_center + inertia_tensor:
 * c = mean of the n points, I = (mass/n) sum(|r|^2 1 - r r^T); sets r->c,
 * r->I, r->mass, leaves the rest untouched. */
int zsph_rigid_mass_properties(const float *pts, int n, double mass, zsph_rigid *r);

/* Test_6DOF/equation/acc_alpha.cpp:3-35 cal_acc_alpha: a = F/M,
 * alpha = R I^-1 R^T (T - w x (R I R^T w)). */
void zsph_rigid_acc_alpha(zsph_rigid *r);

/* Test_6DOF/equation/time_int.cpp:3-22 time_int: explicit Euler of the
 * points (x += (v + w x (x - c)) dt), then q += 1/2 (w (x) q) dt with
 * renormalisation, c += v dt, v += a dt, w += alpha dt; acc and alpha are
 * reset to 0.  pts may be NULL (n = 0). */
void zsph_rigid_euler_step(zsph_rigid *r, float *pts, int n, double dt);

/* ------------------------------------------------------------------------ */
/* Device context                                                            */
/* ------------------------------------------------------------------------ */

/* lib/init.cu:38-115 alloc_gpu_ptc_data x2 + alloc_gpu_tmp_data, plus the
 * cudaMemcpyToSymbol(par/tmp_data) of deltasph/main.cu:47-48 and the device
 * checks of lib/check.cu:19-59.  device < 0 keeps the current device.
 * ptc_num must be at least 1 and below 2^28 - 64 (particle arrays, with 64
 * spare elements, are addressed by 32-bit byte offsets), as must the total of
 * a batch; every scene of a batch holds at least one particle.  The grid may
 * have at most 2^29 cells (all scenes of a batch together).  Other sizes
 * return ZSPH_E_ARG before any device call. */
int zsph_create(const zsph_params *p, int device, zsph_ctx **out);

/* lib/init.cu:117-134 delete_gpu_*_data. */
void zsph_destroy(zsph_ctx *ctx);

/* lib/check.cu:96-105 check_gerr, without the exit: synchronise the
 * context's stream and return the first pending error. */
int zsph_sync(zsph_ctx *ctx);
const char *zsph_last_error(zsph_ctx *ctx);

/* lib/init.cu:136-151 cpu_to_gpu: pos, vel, rhop, type, table into buffer
 * `buf` (0 = "old", 1 = "new").  Saved predictor state is cleared.  In a
 * batch, a particle belongs to the scene of its original index i + table[i]
 * (table NULL: i), counted in zsph_create_batch's scene order, so a state
 * downloaded in cell order uploads with its scenes; an original index
 * outside [0, N) is ZSPH_E_ARG. */
int zsph_upload(zsph_ctx *ctx, int buf, const zsph_host_state *st);

/* lib/gmemcpy.cu:3-36 gpu_to_cpu, with the ZSPH_OUTPUT_* macros replaced by
 * the `fields` mask (ZSPH_F_HASH reads the sorted hash of the last sort). */
int zsph_download(zsph_ctx *ctx, int buf, zsph_host_state *st, int fields);

/* gpu_to_cpu without the solver stall (lib/gmemcpy.cu:3-36 feeding the
 * writer threads of io/mul_thread.cu:50-80).  zsph_download_async snapshots
 * the `fields` of buffer `buf` on the device, in stream order after the
 * steps already issued (a device-to-device copy, ~0.05 ms per million
 * particles), and copies the snapshot into pinned host memory on a separate
 * copy stream, so later zsph_step calls run while it crosses PCIe; it
 * returns at once with a frame handle.  zsph_frame_wait blocks until that
 * copy is done, unpacks it into `st` exactly as zsph_download would (same
 * bits) and releases the frame; it may be called from another thread (a
 * writer), once per frame.  A context holds ZSPH_FRAME_SLOTS frames in
 * flight; zsph_download_async waits for the oldest when all are taken.  The
 * snapshot and pinned buffers are allocated at the first call (ZSPH_E_NOMEM
 * if that fails) and freed by zsph_destroy, which must come after every
 * frame has been waited for. */
#define ZSPH_FRAME_SLOTS 3
typedef struct zsph_frame zsph_frame;
int zsph_download_async(zsph_ctx *ctx, int buf, int fields, zsph_frame **frame);
int zsph_frame_wait(zsph_frame *frame, zsph_host_state *st);

/* lib/hash.cu:3-26 cuda_ptc_hash: cell hash of buffer `buf`, index = i;
 * particles outside the grid are removed (type 100, parked at the grid origin). */
int zsph_hash(zsph_ctx *ctx, int buf);

/* lib/sort.cu:3-6 cuda_sort_index (thrust::sort_by_key): stable LSD radix
 * sort of (hash, index) on the device. */
int zsph_sort_index(zsph_ctx *ctx);

/* lib/sort.cu:8-84 cuda_sort_data: cell start/end tables of the sorted
 * hashes, and the gather of buffer `from` into buffer `to` in cell order. */
int zsph_sort_data(zsph_ctx *ctx, int from, int to);

/* equation/delta_ns.cu:5-177 and equation/simple_ns.cu:3-97
 * cuda_boundary_ns: dummy-wall pressure extrapolation, then (delta form)
 * the renormalised density gradient and velocity divergence; (MC form) the
 * velocity divergence only. */
int zsph_boundary(zsph_ctx *ctx, int buf);

/* equation/delta_ns.cu:179-267 and equation/simple_ns.cu:99-180
 * cuda_govering_ns: continuity + momentum -> acc_drhodt. */
int zsph_governing(zsph_ctx *ctx, int buf);

/* lib/time_int.cu:3-48 cuda_prediction and :50-91 cuda_correction. */
int zsph_prediction(zsph_ctx *ctx, int buf);
int zsph_correction(zsph_ctx *ctx, int buf);

/* deltasph/main.cu:58-90: `nsteps` full predictor-corrector steps starting
 * and ending in buffer 0, replayed from a captured HIP graph (no host
 * synchronisation inside; call zsph_sync to wait). */
int zsph_step(zsph_ctx *ctx, int nsteps);

/* ------------------------------------------------------------------------ */
/* Diagnostics and raw access (no counterpart in the reference)             */
/* ------------------------------------------------------------------------ */

/* Internal device arrays, for tests and checkpoints. Element counts:
 * POSR/VELP/TPOSR/TVELP/ACC/GD: 4N floats; TYPE/TABLE/KEY/INDEX: N ints;
 * CELL_START: grid_num+1 ints. */
#define ZSPH_A_POSR 0       /* (x,y,z,rho)      per buffer */
#define ZSPH_A_VELP 1       /* (vx,vy,vz,p)     per buffer */
#define ZSPH_A_TPOSR 2      /* saved (x,rho)^n  per buffer */
#define ZSPH_A_TVELP 3      /* saved (v,p)^n    per buffer */
#define ZSPH_A_TYPE 4       /* per buffer                  */
#define ZSPH_A_TABLE 5      /* per buffer                  */
#define ZSPH_A_KEY 6        /* cell hash (sorted after zsph_sort_index) */
#define ZSPH_A_INDEX 7      /* sort permutation            */
#define ZSPH_A_CELL_START 8 /* first sorted index of each cell; [G] = N */
#define ZSPH_A_ACC 9        /* (ax,ay,az,drho/dt)          */
#define ZSPH_A_GD 10        /* (grad rho^L, div v)         */
int zsph_array_get(zsph_ctx *ctx, int which, int buf, void *host, size_t bytes);
int zsph_array_set(zsph_ctx *ctx, int which, int buf, const void *host, size_t bytes);

/* Per-particle neighbour count and an order-independent 64-bit digest of
 * the neighbour-index set of buffer `buf` (after zsph_sort_data), using the
 * same cell walk and r^2 <= rc2 predicate as the physics kernels.  After
 * skin-mode steps (buffer 0): the structures the next half step's passes
 * would use, so a rebuild that is due runs first (buffer 0 is then in the
 * rebuild's order, as after that half step's rebuild). */
int zsph_neighbor_stats(zsph_ctx *ctx, int buf, int *count, uint64_t *digest);

/* Timing on the context's stream with HIP events (no counterpart in the
 * reference, which prints host wall time per half step, main.cu:91-95).
 * zsph_time_steps: `nsteps` graph-replayed steps bracketed by two events;
 * *ms = elapsed milliseconds.  zsph_profile_steps: `nsteps` steps issued
 * stage by stage (same kernels, no graph) with an event pair around every
 * stage; stage_ms[ZSPH_ST_*] accumulates milliseconds over all half steps. */
#define ZSPH_ST_HASH 0
#define ZSPH_ST_SORT 1
#define ZSPH_ST_CELLS 2
#define ZSPH_ST_REORDER 3
#define ZSPH_ST_WALL 4
#define ZSPH_ST_LDG 5
#define ZSPH_ST_GOVERNING 6
#define ZSPH_ST_INTEGRATE 7
#define ZSPH_ST_COUNT 8
int zsph_time_steps(zsph_ctx *ctx, int nsteps, float *ms);
int zsph_profile_steps(zsph_ctx *ctx, int nsteps, float *stage_ms);

/* Run diagnostics of buffer `buf` (no counterpart in the reference, which
 * only prints timings).  Sums are accumulated in FP64 per block and added
 * in a fixed order, so repeated calls give identical results. */
typedef struct zsph_diag {
    int n_fluid;        /* particles of type 1                              */
    int n_killed;       /* removed since zsph_create                        */
    double mass_fluid;  /* n_fluid * m                                      */
    double front_x;     /* max x over fluid particles with z < front_zmax   */
    double zmax_fluid;  /* max z over fluid particles                       */
    double kinetic;     /* sum 1/2 m |v|^2 over fluid                       */
    double momentum[3]; /* sum m v over fluid                               */
} zsph_diag;
int zsph_diagnostics(zsph_ctx *ctx, int buf, float front_zmax, zsph_diag *out);

/* Fluid pressure at `npts` points (xyz triples): Shepard average
 * sum p_j W / sum W over fluid particles within 2h; 0 where there are none.
 * Buffer `buf` must be in cell order (after zsph_sort_data or zsph_step). */
int zsph_probe_pressure(zsph_ctx *ctx, int buf, const float *pts, int npts, float *p_out);

/* zsph_probe_pressure without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight output loops, harnesses that forbid hipMalloc).  The points
 * and results are staged in a fixed library-owned device array of
 * ZSPH_PROBE_NOALLOC_MAX probes per GPU (64 KiB, held until the process
 * ends); more points are ZSPH_E_ARG.  Every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (other threads, other contexts)
 * race; calls from one thread are ordered.  The kernel writes every result,
 * so nothing in the array needs clearing.  Same results as
 * zsph_probe_pressure, bit for bit. */
#define ZSPH_PROBE_NOALLOC_MAX 4096
int zsph_probe_pressure_noalloc(zsph_ctx *ctx, int buf, const float *pts, int npts, float *p_out);

/* ------------------------------------------------------------------------ */
/* Batches of independent scenes (no counterpart in the reference, which    */
/* runs one case per process).                                              */
/* ------------------------------------------------------------------------ */

/* One context for `nscenes` independent scenes, each with its own
 * parameters (spacing, h, c0, dt, grid; scenes[k].ptc_num particles).  The
 * scenes share the particle buffers and every launch: the cell key space is
 * the concatenation of the scenes' grids, so each scene keeps its own cell
 * grid and its particles never see another scene's.  Upload the particles
 * scene by scene (scene 0's first); every stage, zsph_step and the
 * download then work on the whole batch, and each scene's results are
 * bitwise those of the same scene in a context of its own.  Body, piston
 * and probe calls are for single-scene contexts.  zsph_get_time reports
 * scene 0's clock; scene k's is its step count times its own dt. */
int zsph_create_batch(const zsph_params *scenes, int nscenes, int device, zsph_ctx **ctx);

/* Diagnostics of each scene (out[nscenes]; one entry for a single-scene
 * context) from the cell table of buffer `buf`, which must be in cell order
 * (after zsph_sort_data into it, or zsph_step for buffer 0).  For a batch,
 * zsph_diagnostics returns the sums over scenes. */
int zsph_scene_diagnostics(zsph_ctx *ctx, int buf, float front_zmax, zsph_diag *out);

/* ------------------------------------------------------------------------ */
/* Moving dummies: floating body (6-DOF coupling) and piston wavemaker.     */
/* The reference has the rigid-body equations only as a standalone test    */
/* (Test_6DOF/) and zero wall-acceleration placeholders in the dummy        */
/* extrapolation (equation/delta_ns.cu:83); the coupling here is new.       */
/* ------------------------------------------------------------------------ */

/* Make the type-2 (ZSPH_BODY) particles of buffer 0 one rigid body with
 * state *r (c, vel, omega, q, I about c in the body frame, mass).  Their
 * body-frame offsets are taken from their current positions.  Each half step
 * then: sums the fluid's force and torque on the body particles (their
 * governing accelerations, gravity excluded), adds M g, integrates the FP64
 * state with the fluid's midpoint predictor-corrector, and moves the body
 * particles rigidly (x = c + R r0, v = V + w x R r0, wall acceleration
 * A + alpha x R r0 + w x (w x R r0) for the dummy pressure). */
int zsph_set_body(zsph_ctx *ctx, const zsph_rigid *r);
/* Current body state; F, T are the fluid force and torque of the last half
 * step (without gravity), acc and alpha the resulting accelerations. */
int zsph_get_body(zsph_ctx *ctx, zsph_rigid *r);

/* Piston: type-3 (ZSPH_PISTON) particles of buffer 0 move rigidly along
 * `axis` (0 x, 1 y, 2 z) with displacement
 *   d(t) = amplitude (sin(2 pi frequency t + phase) - sin(phase)),
 * x(t) = x0 + d(t) e, where x0 is chosen so that x(t_call) is the position at
 * the call (t_call = the simulated time, zsph_get_time).  Velocity d'(t) e;
 * the wall acceleration d''(t) e enters the dummy pressure.
 * zsph_upload into buffer 0 ends both the body and the piston set-up (new
 * particles); call zsph_set_body / zsph_set_piston again after it. */
typedef struct zsph_piston {
    double amplitude, frequency, phase;
    int axis;
} zsph_piston;
int zsph_set_piston(zsph_ctx *ctx, const zsph_piston *p);

/* Simulated time (s); zsph_step advances it by dt per step. */
int zsph_set_time(zsph_ctx *ctx, double t);
int zsph_get_time(zsph_ctx *ctx, double *t);

/* Force and torque (about the body centre c) of the fluid on the body
 * particles of buffer `buf`, from the last zsph_governing, FP64 sums of
 * FP32 terms: the body-force reduction on its own. */
int zsph_body_reduce(zsph_ctx *ctx, int buf, double F[3], double T[3]);

/* Wave-tank case: fluid of depth `depth` in [0,lx]x[0,wy], a piston of
 * `layers` dummy layers behind x = 0 over the tank height `lz`, fixed floor
 * and walls elsewhere, and a floating box [box_lo, box_hi] (a `layers`-thick
 * shell of type-2 particles on the fluid's lattice, no fluid inside).  The
 * box is snapped to the nearest lattice faces (multiples of dx) and the
 * snapped extents are written back to *c, so the caller's mass properties
 * match the particle set.  Same two-call protocol as zsph_case_dambreak. */
typedef struct zsph_wavetank {
    float dx;
    float lx, wy, lz, depth;
    float box_lo[3], box_hi[3];  /* box_hi <= box_lo on any axis: no box */
    int layers;
    float piston_travel;         /* floor extends this far behind the piston */
} zsph_wavetank;
int zsph_case_wavetank(zsph_wavetank *c, zsph_json *j, zsph_host_state *st, int *n_out);

/* ------------------------------------------------------------------------ */
/* Learned closure (no counterpart in the reference): a per-particle MLP    */
/* (ZSPH_NF features -> 64 -> 64 -> 64, SiLU -> 2) whose outputs scale the  */
/* density-diffusion and viscous terms, run on MFMA in BF16 with FP32       */
/* accumulation, trained by the library's own backward kernels.             */
/* ------------------------------------------------------------------------ */
#define ZSPH_NF 24  /* features per fluid particle                          */
#define ZSPH_NH 64  /* hidden width                                         */
#define ZSPH_NT 12  /* training record: cont, diff, t_rho, pa[3], visc[3], t_a[3] */
#define ZSPH_NPARAM (ZSPH_NF * ZSPH_NH + ZSPH_NH + 2 * (ZSPH_NH * ZSPH_NH + ZSPH_NH) + ZSPH_NH * 2 + 2)

/* Network and its input statistics.  Weights are [in][out] row-major; the
 * arrays w0 .. b3 are contiguous (ZSPH_NPARAM floats, the gradient layout).
 * Input x is normalised as (x - mean) * inv_std; a particle with any
 * normalised feature outside [lo, hi] (the training range) keeps the plain
 * terms.  Outputs o give the scale factors s = 1 + out_scale tanh(o). */
typedef struct zsph_closure {
    float w0[ZSPH_NF * ZSPH_NH], b0[ZSPH_NH];
    float w1[ZSPH_NH * ZSPH_NH], b1[ZSPH_NH];
    float w2[ZSPH_NH * ZSPH_NH], b2[ZSPH_NH];
    float w3[ZSPH_NH * 2], b3[2];
    float mean[ZSPH_NF], inv_std[ZSPH_NF];
    float lo[ZSPH_NF], hi[ZSPH_NF];
    float out_scale;
} zsph_closure;

/* A device-resident network with a training set, for training and checks.
 * The training loss over samples p is
 *   L = mean_p [ w_rho (cont + s_d diff - t_rho)^2 + w_acc |pa + s_v visc - t_a|^2 ]
 * with the record of sample p = (cont, diff, t_rho, pa[3], visc[3], t_a[3]):
 * the continuity and density-diffusion parts of drho/dt, its target, the
 * pressure + gravity and viscous parts of the acceleration, its target. */
typedef struct zsph_mlp zsph_mlp;
int zsph_mlp_create(int device, zsph_mlp **out);
void zsph_mlp_destroy(zsph_mlp *m);
const char *zsph_mlp_last_error(zsph_mlp *m);
/* Load the network (resets the optimiser state). */
int zsph_mlp_set(zsph_mlp *m, const zsph_closure *net);
int zsph_mlp_get(zsph_mlp *m, zsph_closure *net);
/* Training set: x = n x ZSPH_NF raw features, rec = n x ZSPH_NT records. */
int zsph_mlp_data(zsph_mlp *m, const float *x, const float *rec, int n, float w_rho, float w_acc);
/* Mean loss and its gradient (ZSPH_NPARAM floats, may be NULL) over
 * samples [first, first + count) of the training set. */
int zsph_mlp_eval(zsph_mlp *m, int first, int count, float *loss, float *grad);
/* One Adam step on samples [first, first + count); *loss (may be NULL) is
 * the loss before the step. */
int zsph_mlp_adam(zsph_mlp *m, int first, int count, float lr, float beta1, float beta2, float eps, float *loss);
/* Scale factors (n x 2) of n feature rows, with the range fallback. */
int zsph_mlp_infer(zsph_mlp *m, const float *x, int n, float *s_out);

/* Use the network in the step of a single-scene context (NULL: plain
 * terms again).  Each half step then computes the features of every fluid
 * particle in the gradient pass, runs the network on them, and scales the
 * density-diffusion part of drho/dt by s_d and the viscous part of the
 * acceleration (bulk and artificial viscosity) by s_v, particle by
 * particle; out-of-range particles keep the plain terms. */
int zsph_set_closure(zsph_ctx *ctx, const zsph_closure *net);

/* Training data of the state in buffer `buf` (in cell order): runs the wall,
 * gradient and continuity+momentum passes (plain terms) and returns, per
 * particle in the buffer's order, the ZSPH_NF features (features[N][NF]) and
 * the term decomposition terms[N][8] = (continuity, density diffusion,
 * pressure + gravity (3), viscous (3)); dummy rows are zero. */
int zsph_closure_sample(zsph_ctx *ctx, int buf, float *features, float *terms);

/* Coarse-graining filter (training data of the closure): Shepard averages
 * of the fluid fields around particles rows[0..nq) of buffer `buf` (in cell
 * order, accelerations from the last zsph_governing on it), with weights
 * V_j W(r; hc) over fluid particles within `radius` (<= 2 hc):
 * out[t * 9 ...] = (sum of weights, rho, v (3), a (3), drho/dt). */
int zsph_filter_fields(zsph_ctx *ctx, int buf, float radius, float hc, const int *rows, int nq, float *out);

/* Launch tuning (no counterpart in the reference, which launches 256-thread
 * blocks, deltasph/main.cu:55-56).  Results do not depend on it: all
 * neighbour walks visit neighbours in the same order, and the stable sort
 * defines one order whatever its digit width. */
#define ZSPH_WALK_MASK 0  /* per-row hit bitmask, then the hits            */
#define ZSPH_WALK_HITS 1  /* gradient pass records every row chunk's hit
                           * mask; the continuity+momentum pass visits the
                           * recorded hits without testing (4 B per row
                           * chunk: 200 B of device memory per particle at
                           * grid_factor 2)                                */
#define ZSPH_WALK_LIST 2  /* gradient pass records neighbour lists (in walk
                           * order) that the governing pass reuses; takes
                           * 512 B of device memory per particle          */
#define ZSPH_WALK_SKIN 3  /* skin mode (skin > 0): each rebuild records
                           * every particle's candidates within 2h + s as
                           * hit masks; until the next rebuild the gradient
                           * pass replays them with the exact predicate
                           * instead of testing, and records the exact
                           * masks the governing pass replays (another
                           * 200 B per particle); elsewhere (skin 0, staged
                           * calls, the closure) it runs as ZSPH_WALK_HITS.
                           * Its rows lie around the rebuild positions, so
                           * its neighbour order, and FP32 summation order,
                           * differs from the other walks'.  One-scene
                           * contexts only (ZSPH_E_ARG for a batch)        */
typedef struct zsph_tuning {
    int walk;        /* ZSPH_WALK_*                                          */
    int radix_bits;  /* sort digit width, 8..11                              */
    int block;       /* threads per block of the three neighbour passes:
                      * 128, 256 or 512 (0: 256; the closure's feature and
                      * scaled passes always run 256)                        */
    int sort_tile;   /* keys per thread of the radix tile, 8 or 16 (0: 16);
                      * the scatter stages the whole tile in LDS (8 B per
                      * key: 16 or 32 KB)                                    */
    float skin;      /* zsph_step's neighbour skin s in units of h, 0..1 (0:
                      * the reference schedule: hash, sort, cell table and
                      * reorder every half step).  s > 0: zsph_step keeps the
                      * cell structures of its last rebuild, on a grid of
                      * cells (2h + s)/grid_factor with rows trimmed to
                      * 2h + s/2, and rebuilds them only after some particle
                      * has moved more than s/2 since (or left the grid).  In
                      * a batch, s is in each scene's own h and each scene has
                      * its own skin grid; any scene's particle triggers the
                      * rebuild of all.  Neighbour sets are the same
                      * (r^2 <= (2h)^2 decides every pair); the particle
                      * order, and with it the FP32 summation order, follows
                      * the rebuilds.                                        */
    int tile_rows;   /* skin mode, one scene: the order in which the gradient
                      * and continuity+momentum passes take their waves.
                      * > 0: units of tile_rows cell rows of one plane, tile
                      * by tile and plane by plane inside a tile, so the
                      * neighbour rows in use on an XCD fit its L2; 0: cell
                      * order.  Each particle's results are the same in any
                      * order (bit for bit).  Default 16 (-1 in
                      * zsph_set_tuning: keep the current value)           */
} zsph_tuning;
int zsph_set_tuning(zsph_ctx *ctx, const zsph_tuning *t);
int zsph_get_tuning(zsph_ctx *ctx, zsph_tuning *t);
/* Rebuilds of the neighbour structures done by zsph_step in skin mode since
 * zsph_create (no counterpart in the reference, which rebuilds every half
 * step). */
int zsph_rebuilds(zsph_ctx *ctx, long long *n);
/* Times tuning candidates over `steps` graph-replayed steps each, every one
 * starting from a saved copy of buffer 0, restores that state, and keeps the
 * fastest; *best and *best_ms (may be NULL) receive it.  The search is by
 * coordinates: walk x digit width first (at the current block and tile),
 * then the block size, then the sort tile, each from the best so far; in
 * skin mode (skin > 0 set beforehand) then skins of 0.05, 0.1, 0.2 and 0.3 h, over
 * four times `steps` (a skin's saving shows over several steps). */
int zsph_autotune(zsph_ctx *ctx, int steps, zsph_tuning *best, float *best_ms);

/* zsph_step launches each step's kernels directly instead of replaying the
 * captured graph (on = 0); the results are the same.  For tools that cannot
 * follow graph launches (kernel tracers), and for comparison. */
int zsph_set_graph(zsph_ctx *ctx, int on);

/* Kernel clock of the continuity+momentum kernel (no counterpart in the
 * reference): with it on, every launch of that kernel -- in zsph_step's
 * captured graph too -- is timed on the device's constant-rate wall clock, so
 * its average duration over a long graph-replayed run can be measured without
 * a profiler.  The kernels stamp the clock themselves (the first thread of the
 * first block of the pass and of the kernel that follows it), so no kernel is
 * added to the stream; an interval includes one launch gap.  With the learned
 * closure on, the gradient pass's interval also holds the network's
 * inference, which runs between the two passes.  zsph_clock_read returns the
 * accumulated milliseconds and launch count and resets them (either pointer
 * may be NULL); zsph_clock_governing also resets. */
int zsph_clock_governing(zsph_ctx *ctx, int on);
int zsph_clock_read(zsph_ctx *ctx, double *ms, long long *launches);
/* The same clock for any of the three neighbour passes (no counterpart in the
 * reference): stage = ZSPH_ST_WALL, ZSPH_ST_LDG (the gradient pass) or
 * ZSPH_ST_GOVERNING; several stages may be clocked together, each with its
 * own accumulator.  zsph_clock_governing / zsph_clock_read are the
 * ZSPH_ST_GOVERNING forms.  Other stages are ZSPH_E_ARG. */
int zsph_clock_stage(zsph_ctx *ctx, int stage, int on);
int zsph_clock_read_stage(zsph_ctx *ctx, int stage, double *ms, long long *launches);

/* Number of particles removed by zsph_hash since zsph_create. */
int zsph_killed(zsph_ctx *ctx, int *killed);

/* ---- Checkpoint / restart ------------------------------------------------
 * Replaces the restart the reference names but does not implement: its case
 * file carries start_step / current_step (io/read_json.cu:32-33) and its loop
 * starts at start_step (deltasph/main.cu:58), but write_json
 * (io/write_json.cu:3-12) is never called and reading its own VTK output back
 * loses every field (io/read_vkt.cu:26-40 against io/write_vtk.cu:66-127).
 *
 * A checkpoint holds the complete state between steps: buffer 0's particle
 * arrays in their current order (position, velocity, density/pressure, the
 * saved and moving-dummy arrays, type, table, scene), the removal count, the
 * simulated time, the FP64 body / piston state and the closure network.  A
 * context created with the same parameters and restored from it continues
 * bit for bit as the run that wrote it.
 *
 * File: a JSON object of top-level scalars ("format": "zsph-checkpoint",
 * "version": 1, "step", "time", sizes, flags, "params" = FNV-1a 64 of the
 * parameter blocks, "sections" = comma-separated names, "<name>_bytes",
 * "payload_bytes", "checksum" = FNV-1a 64 of the payload, both as 16 hex
 * digits) padded with spaces to 4096 bytes, the last one '\n'; then the
 * sections' raw little-endian bytes in the listed order.  Written to
 * <path>.tmp and renamed into place.
 *
 * zsph_checkpoint_write: `step` is the caller's step count (the loop index to
 * continue from); call it between steps (after zsph_step or zsph_upload).
 * zsph_checkpoint_read: the file must match the context (particle and cell
 * counts, scenes, dimension and the parameter hash; else ZSPH_E_ARG and the
 * context is unchanged); a truncated or corrupt file is ZSPH_E_IO.  Restores
 * the body / piston / closure set-up with the state; *step (may be NULL)
 * receives the stored step.  zsph_checkpoint_inspect reads the header
 * without a context (host only) and, with verify != 0, checks the payload
 * against its checksum. */
typedef struct zsph_checkpoint_info {
    long long step;
    double time;
    int ptc_num, cells, scenes, dim;
    int grid[3];
    int killed;
    int has_body, has_piston, closure;
    long long payload_bytes;
} zsph_checkpoint_info;
int zsph_checkpoint_write(zsph_ctx *ctx, const char *path, long long step);
int zsph_checkpoint_read(zsph_ctx *ctx, const char *path, long long *step);
int zsph_checkpoint_inspect(const char *path, int verify, zsph_checkpoint_info *info);

/* sizeof of the ABI structs, so bindings can check their mirrors:
 * 0 zsph_json, 1 zsph_cpu_param, 2 zsph_params, 3 zsph_host_state, 4 zsph_dambreak,
 * 5 zsph_rigid, 6 zsph_diag, 7 zsph_tuning, 8 zsph_piston, 9 zsph_wavetank, 10 zsph_closure,
 * 11 zsph_checkpoint_info. */
size_t zsph_sizeof(int which);

#ifdef __cplusplus
}
#endif
#endif /* ZSPH_H */
