"""Copy kernel of known size, the calibration of HBM traffic counters.

    python -m sddc_amd.calib [--mib 2048] [--reps 3]

copies a device buffer of the given size to another one `reps` times.
Under `rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE) each copy dispatch should
read (write) exactly that many bytes; the ratio of the known size to the
counter reading is the factor by which the stage kernel's counter
readings are scaled.
"""
from __future__ import annotations

import argparse
import sys


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=2048)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    import torch
    n = a.mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(a.reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {a.mib} MiB x {a.reps}: {float(y[-1])}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
