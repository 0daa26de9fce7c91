"""Summarise rocprofv3 CSV output of a benchmark run.

    python -m cast_simulator_amd.rocprof_summary kernel_trace.csv [--last N] [--precision f64|f32]
    python -m cast_simulator_amd.rocprof_summary counter_collection.csv --counters
    python -m cast_simulator_amd.rocprof_summary FETCH.csv --traffic WRITE.csv \
        --copy-bytes B --last N --algorithmic A [--kernels k_traj]

Kernel-trace mode prints per-kernel dispatch statistics (count, mean, min,
max, total) over all dispatches, and, with --last N, over the last N
dispatches of gate kernels only (the timed step of a bench run, which comes
after its autotuning passes).  Counter mode prints the per-kernel mean of
each collected counter over dispatches.

Traffic mode turns one FETCH_SIZE run and one WRITE_SIZE run (KiB per
dispatch) of the same program into HBM bytes per launch of each gate kernel
over the last N gate dispatches, calibrated by what the same counters read
for the device-to-device copy of B bytes the program makes first (the
`--copy-probe-qubits` option of bench.py): factor = counter bytes / B.
--kernels replaces the gate-kernel name prefixes (k_tile, k_mfma, k_dense,
k_monomial) with others, e.g. k_traj for the batched-trajectory launch.
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


def _precision_of(name: str):
    """'f64' or 'f32' from a gate-kernel symbol (the split-product kernel is
    f32 only), or None."""
    if "<double" in name:
        return "f64"
    if "<float" in name or name.startswith("k_mfma_split"):
        return "f32"
    return None


def kernel_trace(path: str, last: int | None, precision: str | None = None) -> str:
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    allk = defaultdict(list)
    gate = []
    for r in rows:
        name = _short(r["Kernel_Name"])
        ms = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        allk[name].append(ms)
        if name.startswith(GATE_KERNELS) and (precision is None or
                                              _precision_of(name) == precision):
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


def _gate_and_copy(path: str, prefixes=GATE_KERNELS):
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    copy = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
    gate = [(_short(r["Kernel_Name"]), float(r["Counter_Value"])) for r in rows
            if _short(r["Kernel_Name"]).startswith(prefixes)]
    return max(copy) if copy else None, gate


def traffic(fetch_csv: str, write_csv: str, copy_bytes: int, last: int, algorithmic: float,
            prefixes=GATE_KERNELS):
    import json

    fcopy, fgate = _gate_and_copy(fetch_csv, prefixes)
    wcopy, wgate = _gate_and_copy(write_csv, prefixes)
    if not fcopy or not wcopy:
        raise SystemExit("no copy-probe dispatch found in the counter files")
    ff, wf = fcopy * 1024 / copy_bytes, wcopy * 1024 / copy_bytes
    # Per kernel symbol, each run on its own: the autotuner may settle a gate
    # on a different family in the two runs, so dispatches are not paired.
    def means(sel, factor):
        acc = defaultdict(list)
        for name, v in sel:
            acc[name].append(v * 1024 / factor)
        return {n: sum(v) / len(v) for n, v in acc.items()}, {n: len(v) for n, v in acc.items()}

    rd, rn = means(fgate[-last:], ff)
    wr, wn = means(wgate[-last:], wf)
    both = [n for n in rd if n in wr]
    out = {
        "calibration": {"copy_bytes_each_way": copy_bytes, "fetch_raw_kib": fcopy,
                        "write_raw_kib": wcopy, "fetch_factor": ff, "write_factor": wf},
        "algorithmic_bytes_per_launch": algorithmic,
        "hbm_bytes_per_launch": {n: rd[n] + wr[n] for n in both},
        "read_bytes_per_launch": {n: rd[n] for n in both},
        "write_bytes_per_launch": {n: wr[n] for n in both},
        "launches_fetch_run": {n: rn[n] for n in both},
        "launches_write_run": {n: wn[n] for n in both},
    }
    return json.dumps(out, indent=1)


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("csv")
    ap.add_argument("--last", type=int, default=None)
    ap.add_argument("--counters", action="store_true")
    ap.add_argument("--traffic", metavar="WRITE_CSV", default=None)
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--algorithmic", type=float, default=0.0)
    ap.add_argument("--kernels", nargs="+", default=None,
                    help="kernel-name prefixes for traffic mode (default: the gate kernels)")
    ap.add_argument("--precision", choices=["f64", "f32"], default=None,
                    help="kernel-trace mode: --last counts only gate kernels of this precision")
    a = ap.parse_args()
    if a.traffic:
        print(traffic(a.csv, a.traffic, a.copy_bytes, a.last or 0, a.algorithmic,
                      tuple(a.kernels) if a.kernels else GATE_KERNELS))
    else:
        print(counters(a.csv) if a.counters else kernel_trace(a.csv, a.last, a.precision))


if __name__ == "__main__":
    main()
