"""Summarise rocprofv3 CSV output of a benchmark run.

    python -m cast_simulator_amd.rocprof_summary kernel_trace.csv [--last N]
    python -m cast_simulator_amd.rocprof_summary counter_collection.csv --counters

Kernel-trace mode prints per-kernel dispatch statistics (count, mean, min,
max, total) over all dispatches, and, with --last N, over the last N
dispatches of gate kernels only (the timed step of a bench run, which comes
after its autotuning passes).  Counter mode prints the per-kernel mean of
each collected counter over dispatches.
"""
from __future__ import annotations

import argparse
import csv
from collections import OrderedDict, defaultdict

GATE_KERNELS = ("k_tile", "k_mfma", "k_dense", "k_monomial")


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "").replace("cast_hip::", "")
    if name.startswith("void "):
        name = name[5:]
    return name.split("(")[0]


def _table(rows, title):
    out = [f"### {title}", "", "| kernel | calls | mean ms | min ms | max ms | total ms |",
           "|---|---:|---:|---:|---:|---:|"]
    for name, ts in sorted(rows.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| `{name}` | {len(ts)} | {sum(ts) / len(ts):.4f} | {min(ts):.4f} | "
                   f"{max(ts):.4f} | {sum(ts):.2f} |")
    return "\n".join(out) + "\n"


def kernel_trace(path: str, last: int | None) -> str:
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    allk = defaultdict(list)
    gate = []
    for r in rows:
        name = _short(r["Kernel_Name"])
        ms = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        allk[name].append(ms)
        if name.startswith(GATE_KERNELS):
            gate.append((name, ms))
    text = _table(allk, f"All dispatches ({len(rows)})")
    if last:
        sel = OrderedDict()
        for name, ms in gate[-last:]:
            sel.setdefault(name, []).append(ms)
        text += "\n" + _table(sel, f"Last {last} gate-kernel dispatches (the timed step)")
        total = sum(ms for _, ms in gate[-last:])
        text += f"\nSum of the last {last} gate-kernel durations: {total:.2f} ms\n"
    return text


def counters(path: str) -> str:
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    acc = defaultdict(lambda: defaultdict(list))
    for r in rows:
        name = _short(r["Kernel_Name"])
        acc[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = ["| kernel | counter | dispatches | mean per dispatch |", "|---|---|---:|---:|"]
    for name, cs in acc.items():
        for c, v in cs.items():
            out.append(f"| `{name}` | {c} | {len(v)} | {sum(v) / len(v):.1f} |")
    return "\n".join(out) + "\n"


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("csv")
    ap.add_argument("--last", type=int, default=None)
    ap.add_argument("--counters", action="store_true")
    a = ap.parse_args()
    print(counters(a.csv) if a.counters else kernel_trace(a.csv, a.last))


if __name__ == "__main__":
    main()
