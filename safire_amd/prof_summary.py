"""Condenses a rocprofv3 kernel_stats.csv (from --kernel-trace --stats --output-format csv) into
a short markdown table: kernel, calls, mean and total duration, share of GPU time."""
from __future__ import annotations

import csv
import re
import sys


def short_name(name: str) -> str:
    name = name.replace("afq::(anonymous namespace)::", "").replace("void ", "")
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


if __name__ == "__main__":
    print(summarize(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 20))
