"""HBM copy of known size, used to calibrate the FETCH_SIZE / WRITE_SIZE
hardware counters: under ``rocprofv3 --pmc`` the counters this copy reports,
divided by its known bytes (``--mib`` MiB read, the same written), give the
factor that turns the solver kernel's counters into bytes.

    python -m lassie_amd.hbm_probe --mib 256
"""
import argparse
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    import torch
    n = args.mib * (1 << 20) // 8
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(args.reps):
        t0 = time.perf_counter()
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        print(f"copy {args.mib} MiB read + {args.mib} MiB written: {dt * 1e3:.3f} ms, "
              f"{2 * args.mib * (1 << 20) / dt / 1e9:.1f} GB/s", flush=True)


if __name__ == "__main__":
    main()
