"""Condenses rocprofv3 output into short markdown tables.

    python safire_amd/prof_summary.py STATS_CSV [TOP]
        kernel_stats.csv (from --kernel-trace --stats --output-format csv): kernel, calls, mean
        and total duration, share of GPU time.

    python safire_amd/prof_summary.py --traffic FETCH_DIR WRITE_DIR CAL_FETCH_DIR CAL_WRITE_DIR [TOP]
        counter_collection.csv of a --pmc FETCH_SIZE run and of a --pmc WRITE_SIZE run (KiB per
        dispatch), each scaled by what the same counter reads for the plain copy kernel of
        prof_copy.py (BYTES read and written): HBM bytes per launch for every kernel.
"""
from __future__ import annotations

import collections
import csv
import glob
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from prof_copy import BYTES  # noqa: E402


def short_name(name: str) -> str:
    name = name.replace("afq::(anonymous namespace)::", "").replace("(anonymous namespace)::", "")
    name = name.replace("afq::", "").replace("void ", "")
    name = re.sub(r"\(.*", "", name)                    # drop the parameter list
    return name.strip()


def summarize(path: str, top: int = 20) -> str:
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| kernel | calls | mean (us) | total (ms) | share |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append(f"| `{short_name(r['Name'])}` | {r['Calls']} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                   f"{float(r['Percentage']):.1f}% |")
    return "\n".join(out)


def counter_means(directory: str, counter: str) -> dict:
    """{kernel: (mean value per dispatch, dispatches)} of one counter in a --pmc run."""
    path = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)[0]
    total = collections.defaultdict(float)
    disp = collections.defaultdict(set)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        k = short_name(r.get("Kernel_Name") or r.get("Kernel-Name"))
        total[k] += float(r["Counter_Value"])
        disp[k].add(r.get("Dispatch_Id") or r.get("Dispatch-Id"))
    return {k: (total[k] / len(disp[k]), len(disp[k])) for k in total}


def copy_factor(directory: str, counter: str) -> float:
    """Bytes per counter unit: BYTES over what the counter reads for the copy kernel (the
    largest dispatch of the calibration run)."""
    kib = max(v for v, _ in counter_means(directory, counter).values())
    return BYTES / kib


def traffic(fetch_dir: str, write_dir: str, cal_fetch: str, cal_write: str, top: int = 12) -> str:
    fetch = counter_means(fetch_dir, "FETCH_SIZE")
    write = counter_means(write_dir, "WRITE_SIZE")
    ff, fw = copy_factor(cal_fetch, "FETCH_SIZE"), copy_factor(cal_write, "WRITE_SIZE")
    out = [f"Copy calibration: FETCH_SIZE x {ff / 1024:.3f} KiB per KiB counted, WRITE_SIZE x "
           f"{fw / 1024:.3f} (a {BYTES >> 20} MiB read + write copy).", "",
           "| kernel | dispatches | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM read (MB) | "
           "HBM write (MB) | total (MB) |", "|---|---:|---:|---:|---:|---:|---:|"]
    rows = []
    for k in set(fetch) & set(write):
        if not k.startswith("k_"):
            continue          # the library's kernels only (not torch's, in the bench's checks)
        rd = fetch[k][0] * ff / 1e6
        wr = write[k][0] * fw / 1e6
        rows.append((rd + wr, k, fetch[k], write[k], rd, wr))
    rows.sort(reverse=True)
    for tot, k, f, w, rd, wr in rows[:top]:
        out.append(f"| `{k}` | {f[1]} | {f[0]:.1f} | {w[0]:.1f} | {rd:.2f} | {wr:.2f} | "
                   f"{tot:.2f} |")
    return "\n".join(out)


if __name__ == "__main__":
    if sys.argv[1] == "--traffic":
        a = sys.argv[2:]
        print(traffic(a[0], a[1], a[2], a[3], int(a[4]) if len(a) > 4 else 12))
    else:
        print(summarize(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 20))
