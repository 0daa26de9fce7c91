"""Benchmark OpenQASM circuits under selectable fusion strategies on the
gfx950 backend (the counterpart of the reference's ``src/bin/bench.rs``).

    python -m cast_simulator_amd.bench_qasm FILE.qasm [...]
        [--fusion none,size-only,hw-adaptive] [--profile PROFILE.json]
        [--max-size 6] [--bench-budget 5] [--precision f32|f64]
        [--force-dense] [--fusion-log] [--kernel-stats]

For every file and fusion mode: parse (``parse_qasm``), build the circuit
graph, fuse a copy (none = ``size_only(1)``, size-only = ``size_only(max)``,
hw-adaptive = the hardware-adaptive cost model with the MI355X profile), then
time repeated executions with the adaptive loop; one row per (file, mode)
with gates, rows (depth), compile time and the per-execution GPU time (the
sum of the kernels' event times, as in src/simulator/mod.rs:295-307).
"""
from __future__ import annotations

import argparse
import math
import sys
from pathlib import Path

PROFILE_DIR = Path(__file__).resolve().parent / "profiles"
MODES = ("none", "size-only", "hw-adaptive")


def fmt_duration(s: float) -> str:
    if s >= 1.0:
        return f"{s:.3f} s"
    if s >= 1e-3:
        return f"{s * 1e3:.3f} ms"
    return f"{s * 1e6:.3f} us"


def fusion_config(cs, mode: str, profile, max_size: int, ztol: float):
    if mode == "none":
        return cs.FusionConfig.size_only(1)
    if mode == "size-only":
        return cs.FusionConfig.size_only(max_size)
    return cs.FusionConfig.hardware_adaptive(profile, max_size, ztol)


def bench_file(cs, path: Path, modes, profile, args) -> list[dict]:
    circuit = cs.parse_qasm(path.read_text())
    graph = cs.CircuitGraph.from_qasm_circuit(circuit)
    spec = cs.HipKernelGenSpec.f64() if args.precision == "f64" else cs.HipKernelGenSpec.f32()
    if args.force_dense:
        spec.ztol = 0.0
        spec.otol = 0.0
    rows = []
    for mode in modes:
        g = cs.CircuitGraph.from_qasm_circuit(circuit)
        cfg = fusion_config(cs, mode, profile, args.max_size, spec.ztol)
        if args.fusion_log and mode == "hw-adaptive":
            print(cs.optimize_with_log(g, cfg).summary(), file=sys.stderr)
        else:
            cs.optimize(g, cfg)
        sim = cs.Simulator(spec)
        r = sim.bench(g, args.bench_budget)
        samples = r["samples_s"]
        mean = sum(samples) / len(samples)
        sd = math.sqrt(sum((x - mean) ** 2 for x in samples) / max(1, len(samples) - 1))
        if args.kernel_stats:
            mgr = sim.manager()
            for kid in range(mgr.n_kernels()):
                info = mgr.plan_info(kid)
                print(f"  {kid:>4}  {info.k:>2}q  variant {info.variant}  "
                      f"nnz {info.scalar_nnz}", file=sys.stderr)
        rows.append({"file": path.stem, "fusion": mode, "gates": r["n_gates"],
                     "rows": g.n_rows(), "qubits": graph.n_qubits(),
                     "compile_s": r["compile_s"], "mean_s": mean, "stddev_s": sd,
                     "iters": len(samples)})
    return rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("files", nargs="+", type=Path)
    ap.add_argument("--fusion", default=",".join(MODES))
    ap.add_argument("--profile", type=Path, default=None)
    ap.add_argument("--max-size", type=int, default=6)
    ap.add_argument("--bench-budget", type=float, default=5.0)
    ap.add_argument("--force-dense", action="store_true")
    ap.add_argument("--precision", choices=["f32", "f64"], default="f64")
    ap.add_argument("--fusion-log", action="store_true")
    ap.add_argument("--kernel-stats", action="store_true")
    args = ap.parse_args(argv)
    modes = [m for m in args.fusion.split(",") if m]
    for m in modes:
        if m not in MODES:
            ap.error(f"unknown fusion mode {m!r} (choose from {', '.join(MODES)})")
    if not modes:
        ap.error("at least one --fusion mode must be given")

    import cast_simulator_amd as cs

    ppath = args.profile or PROFILE_DIR / f"mi355x_{args.precision}.json"
    print(f"Loading profile from {ppath}", file=sys.stderr)
    profile = cs.HardwareProfile.load(str(ppath))

    print(f"{'Circuit':<16} {'Fusion':<14} {'Gates':>7} {'Depth':>6} {'Compile':>10} "
          f"{'Exec time':>28}")
    print("-" * 86)
    for path in args.files:
        for r in bench_file(cs, path, modes, profile, args):
            t = f"{fmt_duration(r['mean_s'])} +/- {fmt_duration(r['stddev_s'])} ({r['iters']})"
            print(f"{r['file']:<16} {r['fusion']:<14} {r['gates']:>7} {r['rows']:>6} "
                  f"{fmt_duration(r['compile_s']):>10} {t:>28}", flush=True)
        print()
    return 0


if __name__ == "__main__":
    sys.exit(main())
