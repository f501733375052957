"""Calibration copy for the HBM traffic counters: `reps` copies of a `mb`-MiB
FP32 buffer into another (each copy reads and writes exactly mb MiB), so
FETCH_SIZE / WRITE_SIZE per copy dispatch can be divided by the known bytes.

    rocprofv3 --pmc FETCH_SIZE -- python3 zsph_amd/copy_probe.py [--mb 512] [--reps 5]
"""
import argparse


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mb", type=int, default=512)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    import torch
    n = a.mb * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {a.reps} x {a.mb} MiB", flush=True)


if __name__ == "__main__":
    main()
