"""Plain device copy of a known size, for calibrating rocprofv3's FETCH_SIZE /
WRITE_SIZE counters against bytes actually moved:

    rocprofv3 --pmc FETCH_SIZE -- python3 calibrate_copy.py --mib 1024

Each launch copies N MiB once (read N MiB, write N MiB); the counter values per
dispatch divided by N MiB give the factor applied to the photon kernel's
counters.
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(a.reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {a.mib} MiB x {a.reps}", flush=True)


if __name__ == "__main__":
    main()
