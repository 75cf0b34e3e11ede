"""HBM calibration for rocprofv3 counter passes: one plain device copy of a known size.

Run under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE); the counters' value for
this copy, divided by the bytes it moves, is the scale applied to the engine kernel's
counters.
"""
import sys

import torch


def main():
    nbytes = int(sys.argv[1]) if len(sys.argv) > 1 else 1 << 30
    x = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda").fill_(1)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {nbytes} bytes 3 times ({nbytes * 3 / 1024:.0f} KiB read, same written)")


if __name__ == "__main__":
    main()
