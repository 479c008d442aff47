"""Short markdown summaries of rocprofv3 CSV output (``--output-format csv``).

    python prof_summary.py stats <prefix>_kernel_stats.csv [--top N]
    python prof_summary.py pmc <prefix>_counter_collection.csv [--match REGEX]

``stats`` prints the per-kernel table of ``--kernel-trace --stats`` (calls, total,
average, share); ``pmc`` prints, per kernel, the number of dispatches and the mean
per dispatch of every counter collected with ``--pmc``.
"""
from __future__ import annotations

import argparse
import csv
import re
import sys
from collections import defaultdict


def _col(row: dict, *names):
    low = {k.lower(): k for k in row}
    for n in names:
        if n.lower() in low:
            return row[low[n.lower()]]
    raise KeyError(f"none of {names} in columns {list(row)}")


def short_name(name: str) -> str:
    name = re.sub(r"\(.*", "", name).replace("void ", "").strip()
    return name


def stats(path: str, top: int) -> str:
    rows = list(csv.DictReader(open(path)))
    out = ["| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        name = short_name(_col(r, "Name", "KernelName", "Kernel_Name"))
        calls = int(_col(r, "Calls"))
        tot = float(_col(r, "TotalDurationNs"))
        avg = float(_col(r, "AverageNs"))
        mn = float(_col(r, "MinNs"))
        mx = float(_col(r, "MaxNs"))
        pct = float(_col(r, "Percentage"))
        out.append(f"| `{name}` | {calls} | {tot / 1e6:.3f} | {avg / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} | {pct:.2f} |")
    return "\n".join(out)


def pmc(path: str, match: str | None) -> str:
    rows = list(csv.DictReader(open(path)))
    acc: dict = defaultdict(lambda: defaultdict(float))
    disp: dict = defaultdict(set)
    counters: list = []
    for r in rows:
        name = short_name(_col(r, "Kernel_Name", "KernelName", "Name"))
        if match and not re.search(match, name):
            continue
        cn = _col(r, "Counter_Name", "CounterName")
        cv = float(_col(r, "Counter_Value", "CounterValue"))
        did = _col(r, "Dispatch_Id", "DispatchId", "Correlation_Id")
        acc[name][cn] += cv
        disp[name].add(did)
        if cn not in counters:
            counters.append(cn)
    head = "| kernel | dispatches | " + " | ".join(f"{c} (mean)" for c in counters) + " |"
    out = [head, "|" + "---|" * (2 + len(counters))]
    for name in sorted(acc, key=lambda n: -len(disp[n])):
        n = len(disp[name])
        vals = " | ".join(f"{acc[name][c] / n:.6g}" for c in counters)
        out.append(f"| `{name}` | {n} | {vals} |")
    return "\n".join(out)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("mode", choices=["stats", "pmc"])
    ap.add_argument("csv")
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--match", default=None)
    a = ap.parse_args(argv)
    print(stats(a.csv, a.top) if a.mode == "stats" else pmc(a.csv, a.match))
    return 0


if __name__ == "__main__":
    sys.exit(main())
