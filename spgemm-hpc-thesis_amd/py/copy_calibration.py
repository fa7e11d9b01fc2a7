"""Plain device-to-device copy of a known size, for calibrating the HBM byte
counters (FETCH_SIZE / WRITE_SIZE) that the traffic figures are derived from.

    rocprofv3 --pmc FETCH_SIZE -- python copy_calibration.py [MiB]

Runs in its own process with torch only (torch carries its own HIP runtime,
so it never shares a process with libspgemm_amd.so).  Each copy reads and
writes exactly MiB mebibytes.
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x3", flush=True)


if __name__ == "__main__":
    main()
