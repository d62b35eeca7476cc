"""Calibration kernel for HBM-traffic counters: one plain device-to-device
copy of a known size (default 1 GiB read + 1 GiB written), repeated.  Its
FETCH_SIZE / WRITE_SIZE per dispatch, divided into the known bytes, gives the
factor by which the same counters are scaled for the library's kernels.

    rocprofv3 --pmc FETCH_SIZE -- python -m oops_amd.pmc_copy
"""
import sys

import torch


def main(nbytes=1 << 30, reps=3):
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {nbytes} bytes x {reps}", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1 << 30)
