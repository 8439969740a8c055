"""A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters.

rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE) -- python3 safire_amd/prof_copy.py reports, for the
one elementwise kernel it launches, the kilobytes the counter sees for reading and writing
exactly BYTES bytes each; the ratio to BYTES / 1024 converts the counter of any other kernel
into bytes.
"""
from __future__ import annotations

import sys

BYTES = 256 * 1024 * 1024


def main() -> int:
    import torch
    x = torch.ones(BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    torch.add(x, 0.0, out=y)          # one kernel: reads BYTES, writes BYTES
    torch.cuda.synchronize()
    print(f"copy kernel: read {BYTES} B, wrote {BYTES} B ({BYTES // 1024} KiB each)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
