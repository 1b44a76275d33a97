"""Device copy of a known size, for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings.

    python -m pbi4_amd.calibrate [MiB] [repeats]

Each repeat is one device-to-device copy kernel that reads and writes exactly MiB mebibytes.
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}")


if __name__ == "__main__":
    main()
