"""Plain device copy of known size, for calibrating HBM traffic counters.

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o run -- python3 -m getm_amd.hbm_copy --mib 1024

Runs ``--reps`` launches of one elementwise kernel (dst = src * 1) over a
float32 buffer of ``--mib`` MiB: each launch reads exactly that many bytes and
writes as many.  The ratio of the FETCH_SIZE / WRITE_SIZE counter values
(KiB per dispatch) to these byte counts is the factor by which the same
counters are scaled for the transport kernels.
"""
from __future__ import annotations

import argparse
import sys


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    import torch

    n = args.mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda:0", dtype=torch.float32)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(args.reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(f"[hbm_copy] {args.reps} launches, each reading and writing {n * 4} bytes ({args.mib} MiB)",
          file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
