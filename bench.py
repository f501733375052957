"""Benchmark: particle-steps/s of the delta-SPH step.

Default workload (--case dambreak3d): the plain 3-D dam break, tank
3.22 x 1.0 x 1.0 m, water column 1.2 x 1.0 x 0.55 m, dx 0.008 m (1,293,750
fluid particles on a lattice + 3 dummy layers on the floor and the four walls,
1,853,844 particles in all), Wendland C2 with h = 1.3 dx, Tait EOS gamma 7,
rho0 1000, c0 = 10 sqrt(2 g H), delta 0.1, alpha 0.02, CFL 0.2, FP32.
One step = one full predictor-corrector step (two half steps: hash, radix
sort, cell table + reorder, wall pass, gradient pass, continuity+momentum,
integrator) for every particle; particles are resident in HBM before timing.
The default timed run is the configured 10,000 steps (t = 0 to 0.63 s, through
the column's collapse and the run-up to the far wall), so the figure is the
whole run's throughput, not the compact initial state's.  A shorter --steps
times a window of the developed flow instead: it starts at the tuning step
(--start-at), not at t = 0.  The autotuner always runs on the state after a
fixed number of steps (--tune-at: 3000 for the dam break), whatever --steps is.

Besides the whole-step figure the line reports, device-clocked over the timed
region, the gradient pass (the dominant kernel) as `roofline` and the
continuity+momentum pass as `roofline_governing`: algorithmic bytes per launch
over the average launch time, against 8 TB/s.  Neither pass is HBM-bound (the
gradient pass is VALU-bound: candidate tests and pair arithmetic; the
continuity+momentum pass also waits on neighbour gathers), so each also carries
`valu_frac`: its recorded SQ_INSTS_VALU over this run's launch time, against
the VALU issue peak of one wave64 FP32 instruction per SIMD every 2 cycles at
2.4 GHz.  That count and the measured HBM traffic are recorded from rocprofv3
counter runs of the same workload and tuning (`RECORDED`, with the profile
file named), not read in the run.
The CPU baseline is the FP32 oracle on the host's cores: median of 5 timings
after 2 warm-up steps with all threads, and a one-thread sample.

Every case runs zsph_step in skin mode by default (--skin 0.1, tuned among
0.05 to 0.3 h; in the ensemble each scene's skin is in its own h): the
neighbour structures are rebuilt only after a particle has moved more than half
the skin since the last rebuild, instead of every half step; neighbour sets are
unchanged (the r^2 <= (2h)^2 predicate decides every pair).  --skin 0 times
the reference schedule.  In skin mode a one-scene case's neighbour passes take
their waves in units of 16 cell rows of a plane, tile by tile, for L2 locality
(--tile-rows; 0: cell order; results are the same bit for bit).

--case wavetank3d: the floating-body tank, 4.0 x 1.0 x 1.2 m filled to 0.6 m
at dx 0.006 m (11.1 M fluid, 12.5 M particles), piston 0.025 m amplitude at
1 Hz, a 0.3 x 0.3 x 0.2 m box of density 500 coupled through its FP64 rigid
state (force/torque reduction, body update and dummy kinematics every half step).

--case ensemble2d: 512 independent 2-D dam breaks (column width U(0.2,0.6),
height U(0.3,0.8), dx U(0.004,0.01) per scene; 3.8 M particles) advanced
together in one context, each on its own cell grid with its own dt.

--closure NPZ (one-scene cases): every step runs with the learned closure
network stored in NPZ (e.g. zsph_amd/data/closure_dambreak3d.npz, trained for
the coarse dam break --dx 0.016): features in the gradient pass, MFMA
inference over the fluid, scaled diffusion and viscous terms.  The CPU
baseline then runs the oracle with the same network.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--case dambreak3d|wavetank3d|ensemble2d]
                    [--dx DX] [--closure NPZ] [--no-cpu-baseline]

Prints one JSON line.  Only one GPU is used: --gpus above 1 exits.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# VALU issue capacity: 256 CUs x 4 SIMDs, one wave64 FP32 instruction per SIMD every 2 cycles at
# 2.4 GHz = 1.23e12 wave instructions/s, the vector FP32 peak of 157.3 TFLOP/s counted in FMAs.  A
# probe of independent FMA chains reaches 9.3e11/s at the clock the chip holds under load (about
# 2.0 GHz); the counter file of the neighbour passes records it.
SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 2, 2.4e9

# Recorded hardware counters of the neighbour passes per launch (rocprofv3 --pmc, one counter per
# pass; the table's profile file gives the commands): HBM traffic = FETCH_SIZE x 2 + WRITE_SIZE x 1
# (KiB, factors from a device copy of known size) and SQ_INSTS_VALU (wave instructions), measured on the
# developed flow (from step 3000).  Keyed by (case, dx, grid_size_factor, walk, skin, block,
# tile_rows); other workloads and tunings report none.
_PROFILES = ("profiles/counters_neighbour_kernels_dambreak3d.md and "
             "profiles/traffic_neighbour_kernels_dambreak3d_dx0.008.md")
RECORDED = {
    ("dambreak3d", 0.008, 2.0, 1, 0.05, 128, 0): {
        "profile": _PROFILES,
        "k_ldg": {"traffic": (139411.0 * 2.0 + 171444.0) * 1024.0, "valu_insts": 2.50797e8},
        "k_governing": {"traffic": (355986.0 * 2.0 + 28966.3) * 1024.0, "valu_insts": 1.92998e8},
    },
    ("dambreak3d", 0.008, 2.0, 1, 0.1, 128, 0): {
        "profile": _PROFILES,
        "k_ldg": {"traffic": (136702.0 * 2.0 + 170848.0) * 1024.0, "valu_insts": 2.57490e8},
        "k_governing": {"traffic": (348649.0 * 2.0 + 28966.3) * 1024.0, "valu_insts": 1.95183e8},
    },
    ("dambreak3d", 0.008, 2.0, 1, 0.2, 128, 0): {
        "profile": _PROFILES,
        "k_ldg": {"traffic": (131766.0 * 2.0 + 170162.0) * 1024.0, "valu_insts": 2.67643e8},
        "k_governing": {"traffic": (337919.0 * 2.0 + 28966.3) * 1024.0, "valu_insts": 1.96918e8},
    },
    ("dambreak3d", 0.008, 2.0, 1, 0.3, 128, 0): {
        "profile": _PROFILES,
        "k_ldg": {"traffic": (129085.0 * 2.0 + 168821.0) * 1024.0, "valu_insts": 2.75235e8},
        "k_governing": {"traffic": (329900.0 * 2.0 + 28966.3) * 1024.0, "valu_insts": 1.96657e8},
    },
    # the default wave order: units of 16 cell rows of a plane, tile by tile
    ("dambreak3d", 0.008, 2.0, 1, 0.05, 128, 16): {
        "profile": _PROFILES,
        "k_ldg": {"traffic": (93810.5 * 2.0 + 187465.0) * 1024.0, "valu_insts": 2.46819e+08},
        "k_governing": {"traffic": (193329.0 * 2.0 + 29546.5) * 1024.0, "valu_insts": 1.85636e+08},
    },
    ("dambreak3d", 0.008, 2.0, 1, 0.1, 128, 16): {
        "profile": _PROFILES,
        "k_ldg": {"traffic": (87634.9 * 2.0 + 186941.0) * 1024.0, "valu_insts": 2.52817e+08},
        "k_governing": {"traffic": (181999.0 * 2.0 + 29555.9) * 1024.0, "valu_insts": 1.8741e+08},
    },
    ("dambreak3d", 0.008, 2.0, 1, 0.2, 128, 16): {
        "profile": _PROFILES,
        "k_ldg": {"traffic": (94168.3 * 2.0 + 185524.0) * 1024.0, "valu_insts": 2.62347e+08},
        "k_governing": {"traffic": (183317.0 * 2.0 + 29539.5) * 1024.0, "valu_insts": 1.88895e+08},
    },
    ("dambreak3d", 0.008, 2.0, 1, 0.3, 128, 16): {
        "profile": _PROFILES,
        "k_ldg": {"traffic": (100933.0 * 2.0 + 183734.0) * 1024.0, "valu_insts": 2.69057e+08},
        "k_governing": {"traffic": (183096.0 * 2.0 + 29546.1) * 1024.0, "valu_insts": 1.87955e+08},
    },
    # the recorded skin walk's gradient pass is two gated launches per half step
    # (recording and replaying forms): their per-launch figures are summed
    ("wavetank3d", 0.006, 2.0, 3, 0.05, 256, 0): {
        "profile": "profiles/kernel_stats_wavetank3d_dx0.006.md",
        "k_ldg": {"traffic": ((258166.0 + 2756080.0) * 2.0 + (202550.0 + 1459580.0)) * 1024.0,
                  "valu_insts": 1.10794e8 + 1.04900e9},
        "k_governing": {"traffic": (3853700.0 * 2.0 + 195312.0) * 1024.0, "valu_insts": 1.31611e9},
    },
    ("wavetank3d", 0.006, 2.0, 1, 0.05, 256, 16): {
        "profile": "profiles/kernel_stats_wavetank3d_dx0.006.md",
        "k_ldg": {"traffic": (1.19967e+06 * 2.0 + 1.45045e+06) * 1024.0, "valu_insts": 1.6886e+09},
        "k_governing": {"traffic": (2.22977e+06 * 2.0 + 197816) * 1024.0, "valu_insts": 1.29785e+09},
    },
    ("wavetank3d", 0.006, 2.0, 3, 0.05, 256, 16): {
        "profile": "profiles/kernel_stats_wavetank3d_dx0.006.md",
        "k_ldg": {"traffic": ((186650.0 + 1848610.0) * 2.0 + (203892.0 + 1462990.0)) * 1024.0,
                  "valu_insts": 1.11066e8 + 1.05222e9},
        "k_governing": {"traffic": (2364560.0 * 2.0 + 197813.0) * 1024.0, "valu_insts": 1.32075e9},
    },
}


def step_bytes(n, G, n_fluid):
    """Algorithmic HBM bytes per particle-step of the step's kernel schedule:
    640 + 48 G/N + 16 f_wall + 204 f_fluid (compulsory traffic, neighbour
    reads served by caches, sort at its single-pass bound)."""
    ff = n_fluid / n
    return 640.0 + 48.0 * G / n + 16.0 * (1.0 - ff) + 204.0 * ff


def ldg_bytes(n, n_fluid, G):
    """Algorithmic bytes of one gradient-pass launch: each fluid particle reads (x, rho) 16,
    (v, p) 16 and its sort key 4 and writes (grad rho^L, div v) 16 = 52 B; each other particle
    writes its zero record, 16 B; plus one pass over the cell table's 2 (G + 1) entries."""
    return 52.0 * n_fluid + 16.0 * (n - n_fluid) + 8.0 * (G + 1)


def governing_bytes(n, n_fluid, G):
    """Algorithmic bytes of one continuity+momentum launch: each fluid particle reads (x, rho)
    16, (v, p) 16, (grad rho^L, div v) 16 and its key 4 and writes (a, drho/dt) 16 = 68 B; each
    other particle writes its zero record, 16 B; plus one pass over the cell table."""
    return 68.0 * n_fluid + 16.0 * (n - n_fluid) + 8.0 * (G + 1)


def kernel_roofline(name, stage_ms, launches, nbytes, rec, bound):
    """Roofline entry of one neighbour pass from its device-clocked average launch time."""
    achieved = nbytes / (stage_ms * 1e-3) / 1e9
    out = {"bound": bound, "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": nbytes,
           "avg_launch_ms": stage_ms, "launches": launches,
           "timing": "device wall clock over every launch of the timed region, stamped by the kernels' first blocks"}
    if rec and name in rec:
        r = rec[name]
        out["traffic"] = r["traffic"]
        # VALU issue: the recorded instruction count over this run's launch time
        out["valu_frac"] = r["valu_insts"] * VALU_CYCLES / (SIMDS * CLOCK_HZ * stage_ms * 1e-3)
        out["recorded"] = (f"traffic and SQ_INSTS_VALU per launch recorded in {rec['profile']} "
                           "(rocprofv3 --pmc of the same workload and tuning)")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            names = [l.split(":", 1)[1].strip() for l in f if l.startswith("model name")]
            f.seek(0)
            sockets = {l.split(":", 1)[1].strip() for l in f if l.startswith("physical id")}
        return (names[0] if names else "unknown"), max(1, len(sockets))
    except OSError:
        return "unknown", 1


def cpu_baseline(scenes, steps=1, repeats=5, warmup=2, single=(1, 3), setup=None, what="the same case"):
    """FP32 CPU oracle (OpenMP over particles) on this host's cores: after `warmup` steps,
    the median of `repeats` timings of `steps` steps, with all threads; then the same with
    one thread, (steps, repeats) = `single` after one warm-up step (a bounded sample:
    a one-thread step of the full case takes about as long as the whole threaded sample)."""
    import statistics

    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def run(nthreads, k, reps, warm):
        prev = oracle.set_threads(nthreads)
        rates = []
        try:
            objs = []
            for case in scenes:
                o = oracle.Oracle(case.params, fp64=False)
                o.upload(case.host)
                if setup:
                    setup(o)
                o.step(warm)
                objs.append(o)
            for _ in range(reps):
                work, elapsed = 0, 0.0
                for case, o in zip(scenes, objs):
                    t0 = time.perf_counter()
                    o.step(k)
                    elapsed += time.perf_counter() - t0
                    work += case.n * k
                rates.append(work / elapsed)
            for o in objs:
                o.close()
        finally:
            oracle.set_threads(prev)
        return statistics.median(rates), rates

    model, sockets = cpu_model()
    value, rates = run(threads, steps, repeats, warmup)
    print(f"[bench] CPU baseline {threads} threads: {value:.4g} particle-steps/s", file=sys.stderr, flush=True)
    one, one_rates = run(1, single[0], single[1], 1)
    print(f"[bench] CPU baseline 1 thread: {one:.4g} particle-steps/s", file=sys.stderr, flush=True)
    n = sum(c.n for c in scenes)
    return {"value": value, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"median of {repeats} timings of {steps} step(s) after {warmup} warm-up steps of {what} "
                      f"({n} particles), FP32 oracle, OpenMP {threads} threads",
            "repeats": rates, "cpu_model": model, "sockets": sockets, "nproc": os.cpu_count(),
            "single_thread": {"value": one, "cores": 1, "repeats": one_rates,
                              "sample": f"median of {single[1]} timings of {single[0]} step(s) after 1 warm-up "
                                        f"step, one OpenMP thread"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the configured run, 10000 for the 3-D dam break)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--event-steps", type=int, default=200, help="steps timed again between HIP events")
    ap.add_argument("--chunk", type=int, default=100, help="graph replays between host syncs in the timed region")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="launch the step's kernels directly (for kernel tracers that cannot follow graph launches)")
    ap.add_argument("--case", choices=("dambreak3d", "wavetank3d", "ensemble2d"), default="dambreak3d")
    ap.add_argument("--scenes", type=int, default=512, help="ensemble2d: number of scenes")
    ap.add_argument("--dx", type=float, default=None, help="particle spacing (default 0.008 / 0.006)")
    ap.add_argument("--gsf", type=float, default=2.0, help="grid_size_factor: cell edge = 2h / gsf")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--walk", type=int, default=None, help="force a neighbour walk (0 mask, 1 hit masks, 2 lists, 3 recorded skin walk)")
    ap.add_argument("--radix-bits", type=int, default=None, help="force the sort's digit width (8 to 11 bits)")
    ap.add_argument("--block", type=int, default=None, help="force the neighbour passes' block size (128/256/512)")
    ap.add_argument("--sort-tile", type=int, default=None, help="force the radix tile's keys per thread (8/16)")
    ap.add_argument("--tile-rows", type=int, default=None,
                    help="force skin mode's wave order: cell rows per unit (0: cell order; default 16)")
    ap.add_argument("--skin", type=float, default=None,
                    help="neighbour skin in units of h (0: rebuild every half step, the reference schedule; > 0: "
                         "rebuild only after a particle moved more than skin/2, the autotuner then also tries "
                         "skins of 0.05, 0.1, 0.2 and 0.3; default 0.1)")
    ap.add_argument("--closure", default=None, metavar="NPZ",
                    help="run every step with this learned closure network (one-scene cases)")
    ap.add_argument("--tune-at", type=int, default=None,
                    help="autotune on the developed state after this many steps of the run (default: 3000 for "
                         "the 3-D dam break, 1000 for the ensemble, 300 for the wave tank, whatever --steps is)")
    ap.add_argument("--state-out", default=None, metavar="PATH",
                    help="write the state at the start of the timed window to this checkpoint file")
    ap.add_argument("--state-in", default=None, metavar="PATH",
                    help="start the timed window from this checkpoint (written by --state-out with the same "
                         "case) instead of stepping there; needs a fixed tuning (--walk etc. or --no-autotune)")
    ap.add_argument("--start-at", type=int, default=None,
                    help="first step of the timed window (default: 0 when the timed steps cover the configured "
                         "run, else the tuning step: a short window is timed in the developed flow)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the step is single-GPU.",
              file=sys.stderr)
        sys.exit(2)

    import zsph_amd as z
    from zsph_amd import cases

    floating = args.case == "wavetank3d"
    if floating:
        args.dx = args.dx or 0.006
        case = cases.wavetank3d(dx=args.dx, grid_size_factor=args.gsf)
        body = cases.box_body(*case.box, cases.BOX_DENSITY)

        def setup(x):  # the same body and piston on the solver and on the CPU baseline
            x.set_body(body)
            x.set_piston(cases.PISTON_AMPLITUDE, cases.PISTON_FREQUENCY, 0.0, 0)
        workload = (f"3-D wave tank 4.0x1.0x1.2 m filled to 0.6 m, dx {args.dx} m, {case.n} particles "
                    f"({case.json.water_ptc_num} fluid), piston {cases.PISTON_AMPLITUDE} m at "
                    f"{cases.PISTON_FREQUENCY} Hz, floating box {case.box} of density {cases.BOX_DENSITY} "
                    f"(FP64 6-DOF coupling), delta-SPH Antuono, Tait gamma 7, CFL 0.2, cell edge 2h/{args.gsf:g}")
        metric = "particle-steps/s, 3D floating-body wave tank"
    elif args.case == "ensemble2d":
        case = cases.ensemble2d(n_scenes=args.scenes, grid_size_factor=args.gsf)
        setup = None
        workload = (f"{case.nscenes} independent 2-D dam breaks in one context (tank 1.61x1.0 m; column width "
                    f"U(0.2,0.6), height U(0.3,0.8), dx U(0.004,0.01) per scene, SplitMix64 seeds), {case.n} "
                    f"particles, per-scene grids and dt, delta-SPH Antuono, Tait gamma 7, CFL 0.2, "
                    f"cell edge 2h/{args.gsf:g}")
        metric = f"particle-steps/s, {case.nscenes}-scene 2D ensemble"
    else:
        args.dx = args.dx or 0.008
        case = cases.dambreak3d(dx=args.dx, grid_size_factor=args.gsf)
        setup = None
        workload = (f"3-D dam break 3.22x1.0x1.0 m, column 1.2x1.0x0.55 m, dx {args.dx} m, "
                    f"{case.n} particles ({case.json.water_ptc_num} fluid), delta-SPH Antuono, "
                    f"Tait gamma 7, CFL 0.2, cell edge 2h/{args.gsf:g}")
        metric = "particle-steps/s, 1.3M 3D dam-break"
    if args.closure:
        if isinstance(case, z.Batch):
            print("bench.py: --closure needs a one-scene case", file=sys.stderr)
            sys.exit(2)
        from zsph_amd import closure
        net = closure.load_network(args.closure)
        case_setup = setup

        def setup(x):  # solver and CPU baseline alike
            if case_setup:
                case_setup(x)
            x.set_closure(net)
        workload += f", learned closure {os.path.basename(args.closure)}"
        metric += " (learned closure)"
        # the network's accuracy against the fine plain run, as its evaluation report recorded
        # it (python -m zsph_amd.train_closure --noise-floor writes <network>_eval.json)
        accuracy = None
        report = os.path.splitext(args.closure)[0] + "_eval.json"
        if os.path.exists(report):
            with open(report) as f:
                ev = json.load(f).get("evaluation", {})
            if "coarse_learned" in ev:
                lr, pl = ev["coarse_learned"], ev["coarse_plain"]
                accuracy = {"recorded_in": os.path.relpath(report, ROOT), "t_end_s": ev["t_end"],
                            "probe_z_m": ev["probe_z"],
                            "front_err": lr["front_err_max"], "front_err_plain_coarse": pl["front_err_max"],
                            "smoothed_peak_p_err": lr["smoothed_peak_p_err"],
                            "smoothed_peak_p_err_plain_coarse": pl["smoothed_peak_p_err"],
                            "stepping_time_ratio_fine_over_coarse": lr["speedup"]}
                if "reordered_sums" in ev:
                    accuracy["fine_run_noise_floor_p_err"] = ev["reordered_sums"]["fine_vs_fine"]["smoothed_peak_p_err"]
    run_len = {"dambreak3d": 10000, "wavetank3d": 1000, "ensemble2d": 3000}[args.case]
    if args.steps is None:  # the configured run length (the ensemble and the tank: a bounded sample)
        args.steps = run_len
    if args.tune_at is None:
        args.tune_at = {"dambreak3d": 3000, "wavetank3d": 300, "ensemble2d": 1000}[args.case]
    if args.start_at is None:
        args.start_at = 0 if args.steps >= run_len else args.tune_at
    batch = isinstance(case, z.Batch)
    n_fluid = int((case.host.type == z.FLUID).sum())
    G = case.grid_num if batch else case.cpu.grid_num
    print(f"[bench] {args.case}: N={case.n} (fluid {n_fluid}), {G} cells", file=sys.stderr, flush=True)
    s = z.Solver(case, device=0)
    # --no-graph: every step of the run launches its kernels directly, the tuning and
    # positioning steps before the timed window included (kernel tracers follow them)
    s.set_graph(args.graph)
    s.upload(case.host)
    if setup:
        setup(s)
    if args.skin is None:
        args.skin = 0.1
    if args.skin > 0:
        s.set_skin(args.skin)
    # pick the neighbour-walk variant and sort digit width on this case
    if any(v is not None for v in (args.walk, args.radix_bits, args.block, args.sort_tile, args.tile_rows)):
        t = s.get_tuning()
        s.set_tuning(t["walk"] if args.walk is None else args.walk,
                     t["radix_bits"] if args.radix_bits is None else args.radix_bits,
                     args.block, args.sort_tile, tile_rows=args.tile_rows)
        args.autotune = False
    import tempfile
    if args.state_in:
        if args.autotune:
            print("bench.py: --state-in needs a fixed tuning (--walk / --no-autotune)", file=sys.stderr)
            sys.exit(2)
        info = z.checkpoint_info(args.state_in)
        s.restore(args.state_in)
        args.start_at = int(info["step"])
        tune = s.get_tuning()
    elif args.autotune or args.start_at > 0:
        # The walks' relative speed changes as the flow develops, so the candidates are
        # timed on a developed state (a fixed step of the run, not a fraction of --steps):
        # checkpoint the initial state, step there, autotune (it restores the state it
        # timed from), then either restart from the initial state bit for bit or time the
        # window from there (all outside the timed region).
        with tempfile.TemporaryDirectory() as tmp:
            ckpt = os.path.join(tmp, "initial.bin")
            s.checkpoint(ckpt, 0)
            s.step(args.tune_at)
            s.sync()
            if args.autotune:
                tune = s.autotune(steps=3)
                tune["tuned_at_step"] = args.tune_at
            else:
                tune = s.get_tuning()
            if args.start_at != args.tune_at:
                s.restore(ckpt)
                if args.start_at:
                    s.step(args.start_at)
    else:
        tune = s.get_tuning()
    if args.state_out:
        s.checkpoint(args.state_out, args.start_at)
    print(f"[bench] tuning: {tune}; timed window starts at step {args.start_at}", file=sys.stderr, flush=True)
    # device clock around every launch of the three neighbour passes, inside the captured step
    for st in (z.ST_WALL, z.ST_LDG, z.ST_GOVERNING):
        s.clock_stage(st, True)
    s.step(args.warmup)
    s.sync()
    print(f"[bench] warm-up {args.warmup} steps done", file=sys.stderr, flush=True)

    # timed region: K graph-replayed steps, synchronised on both sides
    s.sync()
    for st in (z.ST_WALL, z.ST_LDG, z.ST_GOVERNING):
        s.clock_read_stage(st)  # reset
    rebuilds0 = s.rebuilds()
    t0 = time.perf_counter()
    # replayed in chunks with a host sync between them: bounds the launches in
    # flight (profilers that trace every dispatch keep a record per launch)
    last_note = t0
    for k0 in range(0, args.steps, args.chunk):
        s.step(min(args.chunk, args.steps - k0))
        s.sync()
        now = time.perf_counter()
        if now - last_note > 30.0:  # a progress line about every 30 s of a long timed region
            print(f"[bench] {k0 + min(args.chunk, args.steps - k0)} of {args.steps} timed steps done",
                  file=sys.stderr, flush=True)
            last_note = now
    wall = time.perf_counter() - t0
    rebuilds = s.rebuilds() - rebuilds0
    clocks = {st: s.clock_read_stage(st) for st in (z.ST_WALL, z.ST_LDG, z.ST_GOVERNING)}
    for st in (z.ST_WALL, z.ST_LDG, z.ST_GOVERNING):
        s.clock_stage(st, False)
    # a sample of further steps between HIP events on the solver's stream (cross-check; 0: none)
    ev_ms = s.time_steps(args.event_steps) if args.event_steps > 0 else None
    if ev_ms is not None:
        print(f"[bench] {args.steps} steps: host {wall * 1e3 / args.steps:.3f} ms/step; next {args.event_steps} "
              f"steps between events {ev_ms / args.event_steps:.3f} ms/step", file=sys.stderr, flush=True)

    # per-kernel times (HIP events around every stage on the same stream)
    prof = s.profile_steps(args.profile_steps)
    killed = s.killed()
    sim_t = s.time()
    extra = {}
    if floating:
        b = s.get_body()
        extra = {"body_c": list(b.c[:]), "body_q": list(b.q[:])}
    s.close()

    ms_step = wall * 1e3 / args.steps
    value = case.n * args.steps / wall
    b_step = step_bytes(case.n, G, n_fluid)
    rec = None
    if not args.closure:
        key = (args.case, args.dx, args.gsf, tune["walk"], round(tune.get("skin", 0.0), 3))
        rows = tune.get("tile_rows", 0) if tune.get("skin", 0.0) > 0 else 0
        rec = RECORDED.get(key + (tune.get("block"), rows))
        if rec is None:  # the same walk, skin and wave order recorded at another block size
            other = [k for k in RECORDED if k[:5] == key and k[6] == rows]
            if other:
                rec = dict(RECORDED[other[0]])
                rec["profile"] += (f" (recorded with {other[0][5]}-thread blocks; this run's "
                                   f"{tune.get('block')}-thread blocks do the same per-particle work)")
    avg = {st: (ms / max(n, 1), n) for st, (ms, n) in clocks.items()}
    # the dominant kernel: the gradient pass (VALU-bound: candidate tests and pair arithmetic)
    roof = kernel_roofline("k_ldg", *avg[z.ST_LDG], ldg_bytes(case.n, n_fluid, G), rec, "valu")
    roof_gov = kernel_roofline("k_governing", *avg[z.ST_GOVERNING], governing_bytes(case.n, n_fluid, G), rec,
                               "valu")
    out = {
        "metric": metric,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (lattice case generated in-process, no input files)",
        "config": {"workload": workload,
                   **({"closure_accuracy": accuracy} if args.closure and accuracy else {}),
                   "particles": case.n, "fluid_particles": n_fluid, "cells": G,
                   "fluid_particle_steps_per_s": n_fluid * args.steps / wall,
                   "timed_steps": [args.start_at, args.start_at + args.steps], "run_steps": run_len,
                   "event_ms_per_step": ev_ms / args.event_steps if ev_ms is not None else None, "event_steps": args.event_steps,
                   "sim_time_s": sim_t, "removed_particles": killed,
                   "tuning": tune, "graph": args.graph,
                   "neighbour_rebuilds": (rebuilds if tune.get("skin", 0.0) > 0 else 2 * args.steps),
                   **extra},
        "roofline": roof,
        "roofline_governing": roof_gov,
        "wall_pass_ms_per_launch": avg[z.ST_WALL][0],
        "step_roofline": {"bytes_per_particle_step": b_step, "achieved": b_step * value / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b_step * value / 1e9 / HBM_PEAK_GBS},
        "stage_ms_per_step": {k: v / args.profile_steps for k, v in prof.items()},
    }
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (FP32 oracle)...", file=sys.stderr, flush=True)
        if batch:  # every 16th scene: a bounded, size-representative sample
            out["cpu_baseline"] = cpu_baseline(case.cases[::16], steps=10, single=(2, 3),
                                               what=f"every 16th scene of the {case.nscenes}-scene ensemble")
        elif floating:  # 12.5 M particles: one step per timing, fewer repeats
            out["cpu_baseline"] = cpu_baseline([case], steps=1, repeats=3, warmup=1, single=(1, 1), setup=setup,
                                               what=f"the same {case.n}-particle case")
        else:
            out["cpu_baseline"] = cpu_baseline([case], steps=1, setup=setup,
                                               what=f"the same {case.n}-particle case")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
