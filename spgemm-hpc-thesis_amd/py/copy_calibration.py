"""Plain device-to-device copy of a known size, for calibrating the HBM byte
counters (FETCH_SIZE / WRITE_SIZE) that the traffic figures are derived from.

    rocprofv3 --pmc FETCH_SIZE -- python copy_calibration.py [MiB]
    python copy_calibration.py [MiB] time    # also: the copy's own rate

With `time`, ten more copies are timed with HIP events and the median rate
(bytes read + bytes written per second) is printed: the streaming rate the
device reaches on a plain copy, beside which a kernel's counter traffic rate
can be read.

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
    if len(sys.argv) > 2 and sys.argv[2] == "time":
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ms = []
        for _ in range(10):
            ev[0].record()
            y.copy_(x)
            ev[1].record()
            torch.cuda.synchronize()
            ms.append(ev[0].elapsed_time(ev[1]))
        ms.sort()
        med = (ms[4] + ms[5]) / 2
        gbs = 2 * n * 8 / (med * 1e-3) / 1e9
        print(f"copy of {mib} MiB: median {med:.3f} ms, {gbs:.0f} GB/s read + write", flush=True)


if __name__ == "__main__":
    main()
