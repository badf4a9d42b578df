"""Plain device copy of a known size, to calibrate HBM byte counters.

``python -m lbm_isca_amd.copy_probe [MiB] [reps]`` copies a float64 tensor of
MiB mebibytes ``reps`` times with torch (one read and one write of MiB each).
"""

import sys


def main():
    import torch
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}", flush=True)


if __name__ == "__main__":
    main()
