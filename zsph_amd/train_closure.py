"""Train and evaluate the learned closure on one GPU.

    python -m zsph_amd.train_closure --out net.npz --report report.json

1. Training data: fine plain 3-D dam breaks (tank 3.22 x 1.0 x 1.0 m, dx 0.008)
   of seeded column sizes, coarse-grained to dx 0.016 at evenly spaced times
   through the first far-wall impact (zsph_amd.closure.collect).
2. Training on the device with the library's MFMA forward/backward kernels
   (zsph_amd.closure.train); the network is saved to --out.
3. Evaluation on the held-out column 1.2 x 0.55 m: the fine plain run against
   the coarse plain and coarse learned runs -- wave-front position before the
   far-wall impact, peak far-wall probe pressures, and wall-clock time.
   --noise-floor: the fine and the learned coarse runs again in skin mode
   (skin 0.1 h), the same physics with the FP32 neighbour sums in another
   order, each compared with its first run: how much of a probe's error is
   round-off amplified by the impact rather than resolution.
"""
import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True, help="trained network (.npz)")
    ap.add_argument("--net", default=None, help="evaluate this saved network instead of training one")
    ap.add_argument("--report", required=True, help="results (.json)")
    ap.add_argument("--dx", type=float, default=0.008, help="fine spacing; coarse = 2 dx")
    ap.add_argument("--columns", type=int, default=4, help="seeded training columns")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--t-train", type=float, default=0.8,
                    help="simulated time of each training run (s): through the first far-wall impact (near 0.62 s)")
    ap.add_argument("--snaps", type=int, default=16)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--t-eval", type=float, default=0.9,
                    help="simulated time of the evaluation runs (s); the first far-wall impact is near 0.62 s")
    ap.add_argument("--skip-eval", action="store_true")
    ap.add_argument("--noise-floor", action="store_true",
                    help="also rerun the fine and the learned coarse case in skin mode (another summation order)")
    args = ap.parse_args()

    import numpy as np
    from zsph_amd import closure, cases

    if args.net:
        net = closure.load_network(args.net)
        report = {"network": os.path.basename(args.net)}
        log(f"[closure] evaluating {args.net}")
    else:
        columns = [closure.column_of_seed(args.seed + k) for k in range(args.columns)]
        log(f"[closure] training columns (w, h): {columns}")
        X, rec = closure.collect(columns, dx_fine=args.dx, t_end=args.t_train, n_snap=args.snaps, log=log)
        log(f"[closure] {len(X)} samples")
        net, report = closure.train(X, rec, iters=args.iters, log=log)
        closure.save_network(net, args.out, columns=columns, dx=args.dx, t_train=args.t_train, snaps=args.snaps,
                             iters=args.iters, **report)
        report.update({"network": os.path.basename(args.out), "t_train": args.t_train, "snaps": args.snaps,
                       "iters": args.iters, "columns": columns, "dx_fine": args.dx, "dx_coarse": 2 * args.dx})

    if not args.skip_eval:
        fine = cases.dambreak3d(dx=args.dx, grid_size_factor=2.0)
        coarse = cases.dambreak3d(dx=2 * args.dx, grid_size_factor=2.0)
        xw = closure.far_wall(fine)
        log(f"[closure] evaluation: fine N={fine.n}, coarse N={coarse.n}, far wall at x={xw:.3f}")
        hf = closure.run_history(fine, args.t_eval, log=log)
        log(f"[closure] fine plain: {hf['steps']} steps in {hf['wall_s']:.1f} s")
        hp = closure.run_history(coarse, args.t_eval, log=log)
        hl = closure.run_history(coarse, args.t_eval, net=net, log=log)
        plain = closure.compare(hf, hp, xw)
        learned = closure.compare(hf, hl, xw)
        for name, r in (("plain", plain), ("learned", learned)):
            log(f"[closure] coarse {name:8s} front err {r['front_err_max']:.4f}, smoothed peak p err "
                f"{np.round(r['smoothed_peak_p_err'], 3).tolist()}, raw peak p err "
                f"{np.round(r['peak_p_err'], 3).tolist()}, {r['speedup']:.1f}x less stepping time")
        report["evaluation"] = {"column": [1.2, 0.55], "t_end": args.t_eval, "far_wall_x": xw,
                                "probe_z": list(closure.FAR_WALL_PROBES_Z), "fine_step_s": hf["step_s"],
                                "fine_ms_per_step": hf["ms_per_step"], "fine_steps": hf["steps"],
                                "coarse_plain_ms_per_step": hp["ms_per_step"],
                                "coarse_learned_ms_per_step": hl["ms_per_step"],
                                "coarse_plain": plain, "coarse_learned": learned}
        if args.noise_floor:
            hf2 = closure.run_history(fine, args.t_eval, log=log, skin=0.1)
            hl2 = closure.run_history(coarse, args.t_eval, net=net, log=log, skin=0.1)
            reordered = {"fine_vs_fine": closure.compare(hf, hf2, xw),
                         "learned_vs_learned": closure.compare(hl, hl2, xw),
                         "learned_reordered_vs_fine": closure.compare(hf, hl2, xw)}
            for name, r in reordered.items():
                log(f"[closure] reordered sums, {name}: front {r['front_err_max']:.4f}, smoothed peak p "
                    f"{np.round(r['smoothed_peak_p_err'], 3).tolist()}")
            report["evaluation"]["reordered_sums"] = reordered
        np.savez(os.path.splitext(args.report)[0] + "_histories.npz",
                 t_f=hf["t"], front_f=hf["front"], probes_f=hf["probes"],
                 t_c=hp["t"], front_p=hp["front"], probes_p=hp["probes"],
                 front_l=hl["front"], probes_l=hl["probes"])
    with open(args.report, "w") as f:
        json.dump(report, f, indent=1)
    log(f"[closure] report written to {args.report}")


if __name__ == "__main__":
    main()
