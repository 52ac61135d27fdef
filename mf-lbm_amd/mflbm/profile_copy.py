"""Calibration workload for hardware byte counters: a plain device-to-device copy of a known size.

Under ``rocprofv3 --pmc FETCH_SIZE`` (or ``WRITE_SIZE``) the copy kernel should read (write) exactly
``--mib`` MiB per dispatch; the ratio between the counter and that size is the factor used to turn the
counters of the solver's kernels into HBM bytes.
"""
import argparse

import torch


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    # the achievable read + write bandwidth of a streaming copy (the practical ceiling of the solver's kernels)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(a.reps):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / a.reps
    print(f"copied {a.mib} MiB x {a.reps}: {2.0 * a.mib * (1 << 20) / (ms * 1e-3) / 1e9:.1f} GB/s read + write")


if __name__ == "__main__":
    main()
