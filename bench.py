"""Benchmark: HBM GB/s per fused gate on a 30-qubit f64 statevector.

Workload (one GPU): the 30-qubit random-layer circuit (depth 20; each layer =
Haar SU(2) on every qubit, then CZ on a seeded random perfect matching),
fused to k <= 5 dense gates by the hardware-adaptive cost model, applied to
one resident 16 GiB f64 statevector.  One "step" = one pass of the whole
fused circuit.  value = fused gates x algorithmic bytes per gate / wall time
of the timed steps, where the algorithmic bytes of a gate on n qubits are
2 * 2^n * 2 * 8 (one read and one write of every amplitude).

Also reported: the dominant kernel's HBM roofline fraction from HIP events on
the launch stream, the f32 figure of the same circuit, gate-updates/s, and
the CPU baseline (the oracle's SIMD split-layout path on this host's cores,
timed on a bounded sample of the same fused gates).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits 30]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from collections import defaultdict
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "HBM GB/s per fused gate at 30q f64; batched circuits/s + p99 at fixed fidelity"
HBM_PEAK_GBS = 8000.0
SEED = 20260101
KERNEL = {1: "k_dense", 2: "k_monomial", 3: "k_tile", 4: "k_mfma", 5: "k_tile(swizzled)",
          6: "k_tile(3 products)", 7: "k_tile(3 products, swizzled)", 8: "k_tile(xcd order)",
          9: "k_tile(3 products, xcd order)", 10: "k_mfma(f32 products)",
          11: "k_mfma(xcd order)", 12: "k_mfma(swizzled)", 13: "k_tile(xcd order, swizzled)",
          14: "k_tile(3 products, xcd order, swizzled)", 15: "k_mfma(xcd order, swizzled)"}


def mfma_symbol(f: int, ctype: str, k: int) -> str:
    """Profiler symbol of a matrix-core or monomial launch: f32 k = 5, 6 run
    the bf16 split-product kernel unless the f32-product kernel (code 10) was
    chosen; codes 11, 12 and 15 are the default kernel in XCD-blocked,
    swizzled and both tile orders (same symbol)."""
    if f in (4, 11, 12, 15) and ctype == "float" and k >= 5:
        return f"k_mfma_split<{k}>"
    if f in (4, 10, 11, 12, 15):
        return f"k_mfma6<{ctype}>" if k == 6 else f"k_mfma<{ctype}, {k}>"
    return f"{KERNEL[f]}<{ctype}, {k}>"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def hw_profile(cs, precision: str):
    """Cost-model profile: the measured MI355X refit when present, else the
    datasheet roofline (8 TB/s; FP64 78.6 TF, FP32 VALU 78.6 TF without
    packed FP32)."""
    path = ROOT / "cast_simulator_amd" / "profiles" / f"mi355x_{precision}.json"
    if path.exists():
        return cs.HardwareProfile.load(str(path)), str(path.relative_to(ROOT))
    cfg = cs.ProfileConfig()
    cfg.name, cfg.arch, cfg.precision, cfg.n_qubits = "MI355X (datasheet)", "gfx950", precision, 30
    sb = 8 if precision == "f64" else 4
    # reference GFLOP/s unit = half the true flops (src/profile.rs:436)
    peak_ref_gflops = 78.6e12 / 2 / 1e9
    return cs.HardwareProfile.from_roofline(cfg, HBM_PEAK_GBS * 1e9 / 2**30, peak_ref_gflops, sb), \
        "datasheet"


def fused_circuit(cs, n: int, precision: str, max_size: int):
    prof, src = hw_profile(cs, precision)
    ztol = 1e-12 if precision == "f64" else 1e-6
    g = cs.CircuitGraph.from_circuit(cs.random_layer_circuit(n, 20, SEED))
    cs.optimize(g, cs.FusionConfig.hardware_adaptive(prof, max_size, ztol))
    return g.gates_in_row_order(), prof, src


def run_gpu(cs, n: int, precision: str, steps: int, warmup: int, max_size: int):
    gates, prof, src = fused_circuit(cs, n, precision, max_size)
    spec = cs.HipKernelGenSpec.f64() if precision == "f64" else cs.HipKernelGenSpec.f32()
    mgr = cs.HipKernelManager(spec)
    ids = [mgr.generate(g) for g in gates]
    info = {i: mgr.plan_info(i) for i in set(ids)}
    fam = {i: mgr.family(i, n) for i in set(ids)}
    sv = cs.HipStatevector(n, cs.Precision.F64 if precision == "f64" else cs.Precision.F32)
    sv.zero()
    # Setup (untimed, like the reference's compile phase): per-gate kernel
    # autotuning on the resident statevector, three timings per surviving
    # candidate (two gave 5721-5724 GB/s, three 5729-5730 on the same box).
    t_tune = time.perf_counter()
    mgr.set_autotune(True, 22, 3)
    rounds = mgr.tune(ids, sv, 48)
    t_tune = time.perf_counter() - t_tune
    fam = {i: mgr.family(i, n) for i in set(ids)}
    log(f"  autotune: {rounds} passes, {t_tune:.1f} s")
    for s in range(warmup):
        for i in ids:
            mgr.apply(i, sv)
        mgr.sync(False)
        log(f"  warmup step {s + 1}/{warmup} done")
    mgr.sync(False)
    ctype = "double" if precision == "f64" else "float"

    def symbol(kid):
        """Kernel symbol as the profiler names it."""
        k, f = info[kid].k, fam[kid]
        if f in (3, 5, 6, 7, 8, 9, 13, 14):
            low = sum(1 for q in info[kid].qubits if q < 6)
            return (f"k_tile3<{ctype}, {k}, {low}>" if f in (6, 7, 9, 14)
                    else f"k_tile<{ctype}, {k}, {low}>")
        if f == 1:  # dense: the store mode is a template argument
            nts = precision == "f64" and min(info[kid].qubits) >= 3
            return f"k_dense<{ctype}, {k}, {'true' if nts else 'false'}>"
        return mfma_symbol(f, ctype, k)

    per_kernel = defaultdict(list)
    t0 = time.perf_counter()
    for s in range(steps):
        for i in ids:
            mgr.apply(i, sv)
        kernels, _ = mgr.sync(True)
        for kid, t in kernels:
            per_kernel[symbol(kid)].append(t)
        log(f"  timed step {s + 1}/{steps} done")
    wall = time.perf_counter() - t0
    scalar = 8 if precision == "f64" else 4
    bytes_per_gate = 2 * (1 << n) * 2 * scalar
    # Streaming ceiling on the same statevector: the in-place scale kernel
    # (16-byte loads and stores, no gather) reads and writes every amplitude
    # once, the same bytes as a gate; best of 5 synchronous calls.
    from cast_simulator_amd import abi

    pcode = abi.F64 if precision == "f64" else abi.F32
    abi.scale(sv.dptr(), n, pcode, 1.0)
    ceil_t = []
    for _ in range(5):
        c0 = time.perf_counter()
        abi.scale(sv.dptr(), n, pcode, 1.0)
        ceil_t.append(time.perf_counter() - c0)
    ceiling_gbs = bytes_per_gate / min(ceil_t) / 1e9
    n_gates = len(ids)
    gbs = n_gates * steps * bytes_per_gate / wall / 1e9
    # dominant kernel = the kernel symbol with the most GPU time
    sym_dom, times = max(per_kernel.items(), key=lambda kv: sum(kv[1]))
    avg = sum(times) / len(times)
    achieved = bytes_per_gate / avg / 1e9
    norm = sv.norm_squared()
    hist = defaultdict(int)
    fam_hist = defaultdict(int)
    for i in ids:
        hist[info[i].k] += 1
        fam_hist[KERNEL[fam[i]]] += 1
    return {
        "autotune": {"passes": rounds, "seconds": round(t_tune, 2),
                     "kernels_chosen": dict(sorted(fam_hist.items()))},
        "mgr": mgr,
        "ids": ids,
        "gates": gates,
        "n_gates": n_gates,
        "k_hist": dict(sorted(hist.items())),
        "wall_s": wall,
        "gbs_per_gate": gbs,
        "ms_per_step": wall / steps * 1e3,
        "gate_updates_per_s": n_gates * steps / wall,
        "dominant": {"kernel": sym_dom, "launches": len(times), "avg_ms": avg * 1e3,
                     "achieved_gbs": achieved},
        "per_k_gbs": {s: bytes_per_gate / (sum(t) / len(t)) / 1e9
                      for s, t in sorted(per_kernel.items())},
        "norm_after": norm,
        "stream_ceiling_gbs": ceiling_gbs,
        "profile": src,
        "crossover_ai": prof.crossover_ai,
    }


def fidelity_f32_vs_f64(cs, r64, r32, n: int):
    """One pass of the headline circuit at f64 and at f32 from |0...0>, with
    the kernels the timed runs used (each precision fused by its own cost
    model), both statevectors resident (16 + 8 GiB at 30 qubits): fidelity
    |<f32|f64>|^2 / (<f32|f32><f64|f64>) and the norms, reduced in f64 on the
    device."""
    s64 = cs.HipStatevector(n, cs.Precision.F64)
    s32 = cs.HipStatevector(n, cs.Precision.F32)
    for sv, r in ((s64, r64), (s32, r32)):
        sv.zero()
        for i in r["ids"]:
            r["mgr"].apply(i, sv)
        r["mgr"].sync(False)
    out = {"fidelity_vs_f64": s64.fidelity(s32),
           "norm_error_f32": abs(s32.norm_squared() - 1.0),
           "norm_error_f64": abs(s64.norm_squared() - 1.0)}
    del s64, s32
    return out


def run_qft20(cs, n: int = 20, reps: int = 20):
    """Config 1: QFT with final swaps on 20 qubits (220 gates), size-only
    fusion <= 2 (Phase 1 only), on a seeded complex Gaussian f64 state.  GPU:
    the fused list on the device, kernel time from events and wall time per
    pass, over `reps` passes after one untimed pass whose result is kept for
    the parity check against the CPU baseline (cpu_qft20)."""
    import numpy as np

    g = cs.CircuitGraph.from_circuit(cs.QuantumCircuit.qft(n, True))
    cs.optimize(g, cs.FusionConfig.size_only(2))
    gates = g.gates_in_row_order()
    init = np.asarray(cs.random_statevector(n, SEED))
    mgr = cs.HipKernelManager(cs.HipKernelGenSpec.f64())
    ids = [mgr.generate(x) for x in gates]
    sv = cs.HipStatevector(n, cs.Precision.F64)
    sv.upload(init)
    for i in ids:
        mgr.apply(i, sv)
    mgr.sync(False)
    first = np.asarray(sv.download())
    gpu = 0.0
    t0 = time.perf_counter()
    for _ in range(reps):
        for i in ids:
            mgr.apply(i, sv)
        kernels, _ = mgr.sync(True)
        gpu += sum(t for _, t in kernels)
    wall = time.perf_counter() - t0
    b = 2 * (1 << n) * 2 * 8
    hist = defaultdict(int)
    for x in gates:
        hist[x.n_qubits()] += 1
    return {
        "workload": f"{n}-qubit QFT with swaps (220 gates), size-only fusion <= 2, seeded complex "
                    f"Gaussian f64 state (seed {SEED}), one resident statevector (16 MiB)",
        "fused_gates": len(ids), "k_histogram": dict(sorted(hist.items())),
        "gpu_kernel_ms_per_pass": round(gpu / reps * 1e3, 3),
        "gpu_wall_ms_per_pass": round(wall / reps * 1e3, 3),
        "gbs_per_gate_kernel": round(len(ids) * b * reps / gpu / 1e9, 1),
        "passes": reps,
        "_gates": gates, "_init": init, "_first": first,
    }


def cpu_qft20(q, threads: int, min_s: float = 2.0):
    """CPU leg of config 1: the oracle's SIMD split-layout port (a
    restatement of the reference CPU path) over the whole fused list on
    `threads` threads, repeated until `min_s` seconds; its first pass is
    compared amplitude by amplitude with the GPU's."""
    import numpy as np

    import oracle

    gates, init, first = q["_gates"], q["_init"], q["_first"]
    n = (init.size // 2).bit_length() - 1
    sv = oracle.SimdStatevector(n, f32=False)
    sv.load(init)
    mats = [(np.asarray(x.matrix()), list(x.qubits())) for x in gates]
    for m, qs in mats:
        sv.apply(m, qs, threads)
    ref = sv.interleaved()
    za, zb = ref[0::2] + 1j * ref[1::2], first[0::2] + 1j * first[1::2]
    rel = float(np.linalg.norm(za - zb) / np.linalg.norm(za))
    passes, t0 = 0, time.perf_counter()
    while True:
        for m, qs in mats:
            sv.apply(m, qs, threads)
        passes += 1
        if time.perf_counter() - t0 >= min_s:
            break
    dt = (time.perf_counter() - t0) / passes
    return {"ms_per_pass": round(dt * 1e3, 3), "threads": threads, "passes": passes,
            "gpu_vs_cpu_rel_l2": rel, "max_abs_diff": float(np.max(np.abs(za - zb)))}


def run_dense_large(cs, n64: int = 33, n32: int = 34, depth: int = 12):
    """Large dense fusion: the random-layer circuit (depth 12) fused
    aggressively (size-only, k <= 6) on a 33-qubit f64 (128 GiB) and a
    34-qubit f32 (128 GiB) statevector; per kernel group GB/s and
    algorithmic TFLOP/s (8 K flops per amplitude for a dense k-qubit gate),
    the f32-vs-f64 fidelity at 33 qubits (both resident: 64 + 128 GiB, inner
    product on the device).  Kernel choice: the automatic family rule (the
    per-gate autotuner would cost minutes of extra passes at this size)."""
    def fused(n):
        g = cs.CircuitGraph.from_circuit(cs.random_layer_circuit(n, depth, SEED))
        cs.optimize(g, cs.FusionConfig.size_only(6))
        return g.gates_in_row_order()

    def run(gates, n, prec, keep=False):
        spec = cs.HipKernelGenSpec.f64() if prec == "f64" else cs.HipKernelGenSpec.f32()
        mgr = cs.HipKernelManager(spec)
        mgr.set_autotune(False)
        ids = [mgr.generate(g) for g in gates]
        info = {i: mgr.plan_info(i) for i in set(ids)}
        fam = {i: mgr.family(i, n) for i in set(ids)}
        sv = cs.HipStatevector(n, cs.Precision.F64 if prec == "f64" else cs.Precision.F32)
        sv.zero()
        t0 = time.perf_counter()
        for i in ids:
            mgr.apply(i, sv)
        kernels, _ = mgr.sync(True)
        wall = time.perf_counter() - t0
        sb = 8 if prec == "f64" else 4
        B = 2 * (1 << n) * 2 * sb
        groups = defaultdict(list)
        ctype = "double" if prec == "f64" else "float"
        for kid, t in kernels:
            k = info[kid].k
            dense = info[kid].scalar_nnz == 2 * (1 << k) ** 2
            f = fam[kid]
            if f in (4, 10, 11, 12, 15):
                kern = mfma_symbol(f, ctype, k)
            elif f == 3 and (k == 4 or (k == 5 and prec == "f32")):
                # automatic wave tiles take the three-product form here
                kern = f"k_tile3<{prec}, k={k}>"
            else:
                kern = f"{KERNEL[f]}<{prec}, k={k}>"
            groups[(kern, k, dense)].append(t)
        rows = []
        for (kern, k, dense), ts in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
            avg = sum(ts) / len(ts)
            row = {"kernel": kern + ("" if dense else " (sparse)"), "k": k,
                   "launches": len(ts), "avg_ms": round(avg * 1e3, 3),
                   "gbs": round(B / avg / 1e9, 1)}
            if dense:
                row["tflops"] = round(8 * (1 << k) * (1 << n) / avg / 1e12, 2)
            rows.append(row)
        hist = defaultdict(int)
        for i in ids:
            hist[info[i].k] += 1
        res = {"qubits": n, "precision": prec, "fused_gates": len(ids),
               "k_histogram": dict(sorted(hist.items())), "wall_s": round(wall, 2),
               "gbs_per_gate": round(len(ids) * B / wall / 1e9, 1),
               "norm_error": abs(sv.norm_squared() - 1.0), "kernels": rows}
        return res, (sv if keep else None)

    log(f"[dense-large] {n64}q f64, depth {depth}, fused k<=6")
    g64 = fused(n64)
    r64, sv64 = run(g64, n64, "f64", keep=True)
    log(f"  {r64['fused_gates']} gates in {r64['wall_s']} s, {r64['gbs_per_gate']} GB/s per gate")
    log(f"[dense-large] {n64}q f32 of the same circuit, fidelity against f64")
    r32s, sv32 = run(g64, n64, "f32", keep=True)
    fid = sv64.fidelity(sv32)
    del sv32, sv64
    log(f"  fidelity {fid:.12f}")
    log(f"[dense-large] {n32}q f32, depth {depth}, fused k<=6")
    r32, _ = run(fused(n32), n32, "f32")
    log(f"  {r32['fused_gates']} gates in {r32['wall_s']} s, {r32['gbs_per_gate']} GB/s per gate")

    def roof(r, prec):
        k6 = [x for x in r["kernels"] if x["k"] == 6 and "tflops" in x]
        if not k6:
            return None
        top = max(k6, key=lambda x: x["launches"] * x["avg_ms"])
        if "split" in top["kernel"]:
            # bf16 split products: 4 real products x 6 partial products on
            # the bf16 matrix cores = 6x the algorithmic flops; the kernel is
            # HBM-bound (bf16 peak 2.5 PFLOP/s dense)
            return {"bound": "hbm", "kernel": top["kernel"], "achieved": top["gbs"],
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(top["gbs"] / HBM_PEAK_GBS, 4),
                    "algorithmic_tflops": top["tflops"],
                    "bf16_mfma_tflops": round(6 * top["tflops"], 1),
                    "bf16_mfma_frac": round(6 * top["tflops"] / 2500.0, 4), "traffic": None}
        peak = 78.6 if prec == "f64" else 157.3
        return {"bound": "mfma", "kernel": top["kernel"], "achieved": top["tflops"],
                "peak": peak, "unit": "TFLOP/s", "frac": round(top["tflops"] / peak, 4),
                "gbs": top["gbs"], "traffic": None}
    return {
        "workload": f"{n64}-qubit f64 and {n32}-qubit f32 random-layer circuit, depth {depth}, "
                    f"seed {SEED}, size-only fusion k<=6, automatic kernel family; fidelity of "
                    f"the {n64}-qubit f32 state against the f64 one (both resident)",
        "f64": r64, "f32": r32, "f32_same_circuit_as_f64": r32s,
        "fidelity_f32_vs_f64": fid,
        "roofline": roof(r64, "f64"), "roofline_f32": roof(r32, "f32"),
        "flop_convention": "tflops = 8 K 2^n true real flops per dense k-qubit gate; the "
                           "reference's GFLOP/s unit (src/profile.rs:436) is half of it",
    }


def dense_large_line(d):
    """The --workload dense-large JSON line around run_dense_large's result."""
    return {
        "metric": "config 3: per-gate GB/s and algorithmic TFLOP/s, 33q f64 / 34q f32, k<=6",
        "value": d["f64"]["gbs_per_gate"], "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(d["f64"]["wall_s"] * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": d["workload"]}, **{k: v for k, v in d.items() if k != "workload"},
    }


def run_marginals(cs, n: int = 30, reps: int = 3):
    """Device marginal probabilities on an n-qubit f64 statevector: one read of
    every amplitude (2^n * 16 bytes) per call; only 2^m doubles reach the
    host.  The reference downloads the whole statevector instead
    (src/simulator/mod.rs:282-291)."""
    from cast_simulator_amd import abi

    sv = cs.HipStatevector(n, cs.Precision.F64)
    sv.zero()
    nbytes = (1 << n) * 16
    rows = []
    for measured in ([0], [n - 1, n // 2, 3], list(range(10)), list(range(n - 10, n)),
                     list(range(0, n, 2))[:16]):
        sv.marginal_probabilities(measured)  # warm
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            p = sv.marginal_probabilities(measured)
            ts.append(time.perf_counter() - t0)
        t = min(ts)
        rows.append({"measured": len(measured), "qubits": measured[:4] + (["..."] if len(measured) > 4 else []),
                     "ms": round(t * 1e3, 3), "gbs": round(nbytes / t / 1e9, 1),
                     "sum": float(p.sum())})
    del sv
    abi.stream_sync(None)
    return {"workload": f"{n}-qubit f64 statevector (16 GiB), marginals over m measured qubits, "
                        "synchronous call including the 2^m result copy",
            "bytes_per_call": nbytes, "calls": rows,
            "best_hbm_frac": round(max(r["gbs"] for r in rows) / HBM_PEAK_GBS, 4)}


def run_density_matrix(cs, n: int = 15, depth: int = 4, p: float = 1e-3, steps: int = 2):
    """Density-matrix mode: the random-layer circuit on n physical qubits with
    a depolarizing channel after every single-qubit gate, each gate lifted to
    its superoperator (2k qubits) and applied to the 2n-qubit vectorised rho
    (n = 15: 30 virtual qubits, 16 GiB f64).  Per-gate GB/s from the kernel
    event times, and the trace of rho after the run.

    Two byte counts: algorithmic (the whole vector read and written, as for a
    statevector gate) and moved.  The monomial kernel reads and writes only
    the amplitudes of rows that differ from the identity row (n_active of K
    per combo), so its moved bytes are n_active / K of the algorithmic ones;
    every other kernel moves the whole vector.  Rates over 8 TB/s are
    possible only in the algorithmic column."""
    import numpy as np
    from cast_simulator_amd import abi

    circ = cs.random_layer_circuit(n, depth, SEED)
    gates = [g.with_depolarizing(p) if g.n_qubits() == 1 else g for g in circ.gates()]
    lifted = [g.to_density_matrix_gate(n) for g in gates]
    mgr = cs.HipKernelManager(cs.HipKernelGenSpec.f64())
    ids = [mgr.generate(g) for g in lifted]
    rho = cs.HipStatevector(2 * n, cs.Precision.F64)
    rho.zero()
    mgr.tune(ids, rho, 16)
    fam = {i: mgr.family(i, 2 * n) for i in set(ids)}
    kq = {i: mgr.plan_info(i).k for i in set(ids)}
    nbytes = 2 * (1 << (2 * n)) * 16

    def active_rows(m):
        K = m.shape[0]
        eye = np.eye(K)
        return sum(1 for j in range(K) if not np.array_equal(m[j], eye[j]))

    moved = {}
    for i, g in zip(ids, lifted):
        if i not in moved:
            m = np.asarray(g.matrix())
            moved[i] = nbytes * (active_rows(m) / m.shape[0] if fam[i] == 2 else 1.0)
    rho.zero()
    gpu = 0.0
    per_kernel = defaultdict(list)
    moved_total = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        for i in ids:
            mgr.apply(i, rho)
        kernels, _ = mgr.sync(True)
        gpu += sum(t for _, t in kernels)
        for kid, t in kernels:
            per_kernel[f"{KERNEL[fam[kid]]} k={kq[kid]}"].append((t, moved[kid]))
            moved_total += moved[kid]
    wall = time.perf_counter() - t0
    trace = float(abi.dm_populations(rho.dptr(), n, abi.F64).sum())
    hist = defaultdict(int)
    for g in lifted:
        hist[g.n_qubits()] += 1
    return {"workload": f"{n}-qubit random-layer circuit, depth {depth}, depolarizing p={p} "
                        f"after every single-qubit gate; density matrix as a {2 * n}-qubit f64 "
                        "vector (16 GiB); superoperator gates",
            "gates": len(ids), "superop_k_histogram": dict(sorted(hist.items())),
            "steps": steps,
            "algorithmic_gbs_per_gate": round(len(ids) * steps * nbytes / gpu / 1e9, 1),
            "moved_gbs": round(moved_total / gpu / 1e9, 1),
            "moved_hbm_frac": round(moved_total / gpu / 1e9 / HBM_PEAK_GBS, 4),
            "wall_algorithmic_gbs_per_gate": round(len(ids) * steps * nbytes / wall / 1e9, 1),
            "per_kernel": {k: {"launches": len(v),
                               "algorithmic_gbs": round(len(v) * nbytes / sum(t for t, _ in v)
                                                        / 1e9, 1),
                               "moved_gbs": round(sum(b for _, b in v) / sum(t for t, _ in v)
                                                  / 1e9, 1)}
                           for k, v in sorted(per_kernel.items())},
            "trace_after": trace}


def run_trajectories(cs, n: int, count: int, tune_sample: int, plain_sample: int):
    """Batched noisy trajectories (one persistent launch) against the plain
    per-circuit f64 GPU path on the same trajectories.

    Workload: the n-qubit random-layer circuit of 40 layers (seed SEED); each
    trajectory runs a depth drawn uniformly from [10, 40] and gets a
    depolarizing Pauli with p = 1e-3 after every gate on each touched qubit;
    f32 amplitudes, f64 norms and <Z_i Z_i+1> sums.  circuits/s counts the
    whole batch: host planning, clean checkpoints and the launch."""
    import numpy as np

    D, p = 40, 1e-3
    lc = cs.LayeredCircuit.random_layers(n, D, SEED)
    pairs = [q for i in range(n - 1) for q in (i, i + 1)]
    noise = cs.NoiseSpec(p, 10, D, SEED)
    S = cs.sample_trajectories(lc, noise, count)
    n_err = sum(len(s.errors) for s in S)
    eng = cs.TrajectoryEngine(lc, cs.Precision.F32, pairs)
    t0 = time.perf_counter()
    best, trials = eng.autotune(S[:tune_sample], cs.TrajectoryTuning(4, 6, 256, 2, 0))
    t_tune = time.perf_counter() - t0
    log(f"  trajectory autotune: {len(trials)} trials on {tune_sample} trajectories, "
        f"{t_tune:.1f} s -> {best!r}")
    r = eng.run(S, best)
    sv_bytes = (2 << n) * 4
    # algorithmic bytes: each fused gate reads and writes one statevector,
    # each task's observation reads it once more
    alg = r.n_gate_passes * 2 * sv_bytes + r.n_tasks * sv_bytes
    zz = np.array(r.traj_zz)
    log(f"  trajectories: {count} in {r.total_s:.2f} s ({count / r.total_s:.1f} circuits/s), "
        f"gpu {r.gpu_s:.2f} s, {r.n_tasks} tasks, {r.n_gate_passes} gate passes")

    # Plain per-circuit f64 GPU path on the first trajectories: explicit gate
    # list (Paulis inserted), fused to k <= 5 by the hardware-adaptive model,
    # one statevector each; its kernel time from events and its end-to-end
    # time (fusion + plans + kernels).  The batched f32 states of the same
    # trajectories are dumped 8 at a time and compared with the f64 ones.
    prof, _ = hw_profile(cs, "f64")
    sim = cs.Simulator.f64()
    sim.manager().set_autotune(False)
    t_plain = gpu_plain = 0.0
    plain_passes = 0
    fids = []
    sub = S[:plain_sample]
    for b0 in range(0, len(sub), 8):
        part = sub[b0:b0 + 8]
        rd = eng.run(part, best, True)
        dump = rd.dump().reshape(rd.n_tasks, -1)
        for t, s in enumerate(part):
            tp = time.perf_counter()
            g = cs.CircuitGraph.from_circuit(cs.trajectory_circuit(lc, s))
            cs.optimize(g, cs.FusionConfig.hardware_adaptive(prof, 5, 1e-12))
            sv = cs.HipStatevector(n, cs.Precision.F64)
            sv.zero()
            kernels, _ = sim.run(g, sv)
            t_plain += time.perf_counter() - tp
            gpu_plain += sum(x for _, x in kernels)
            plain_passes += len(kernels)
            x = np.asarray(sv.download())
            y = dump[rd.task_of[t]]
            zx = x[0::2] + 1j * x[1::2]
            zy = y[0::2] + 1j * y[1::2]
            fids.append(abs(np.vdot(zx, zy)) ** 2 / (np.vdot(zx, zx).real * np.vdot(zy, zy).real))
        del dump, rd
    plain_cps = len(sub) / t_plain
    return {
        "workload": f"{n}-qubit random-layer circuit, 40 layers, seed {SEED}; {count} trajectories, "
                    f"depth ~ U[10, 40], depolarizing p={p} after every gate per touched qubit "
                    f"({n_err} Pauli errors sampled); f32 amplitudes, f64 norm and <Z_i Z_i+1>",
        "circuits_per_s": round(count / r.total_s, 2),
        "gpu_circuits_per_s": round(count / r.gpu_s, 2),
        "total_s": round(r.total_s, 3),
        "plan_s": round(r.plan_s, 3),
        "checkpoint_s": round(r.checkpoint_s, 3),
        "gpu_s": round(r.gpu_s, 3),
        "tasks": r.n_tasks,
        "shared_clean_trajectories": r.n_shared_clean,
        "gate_passes": r.n_gate_passes,
        "dirty_matrices": r.n_dirty_matrices,
        "slots": r.n_slots,
        "tuning": {"chunk_layers": best.chunk_layers, "fusion_max": best.fusion_max,
                   "block": best.block, "slots_per_cu": best.slots_per_cu,
                   "waves_per_simd": best.waves_per_simd,
                   "k5_k6_products": "f32" if best.mfma_f32 else "bf16 split"},
        "autotune": {"trials": len(trials), "sample": tune_sample, "seconds": round(t_tune, 1)},
        "kernel_gbs": round(alg / r.gpu_s / 1e9, 1),
        "kernel_hbm_frac": round(alg / r.gpu_s / 1e9 / HBM_PEAK_GBS, 4),
        "mean_zz": r.mean_zz,
        "mean_zz_stderr": float(zz.std(ddof=1) / np.sqrt(len(zz))),
        "norm_sq_range": [r.min_norm_sq, r.max_norm_sq],
        "plain_f64_circuits_per_s": round(plain_cps, 2),
        "plain_sample": len(sub),
        "speedup_vs_plain_f64": round(count / r.total_s / plain_cps, 2),
        "speedup_split": {
            "kernel_s_per_trajectory_plain_f64": gpu_plain / len(sub),
            "kernel_s_per_trajectory_batched_f32": r.gpu_s / count,
            "device_speedup": round(gpu_plain / len(sub) / (r.gpu_s / count), 2),
            "gate_passes_per_trajectory_plain": round(plain_passes / len(sub), 2),
            "gate_passes_per_trajectory_batched": round(r.n_gate_passes / count, 2),
            "tasks_per_trajectory": round(r.n_tasks / count, 4),
            "host_s_per_trajectory_plain": (t_plain - gpu_plain) / len(sub),
            "host_s_per_trajectory_batched": (r.plan_s + r.checkpoint_s) / count,
            "note": "plain = per-circuit f64 path on the first sample trajectories (their own "
                    "gate lists fused to k<=5); batched = the whole launch divided by the "
                    "trajectory count; device_speedup = kernel time per trajectory, plain over "
                    "batched: it includes the shared clean prefixes (gate passes per "
                    "trajectory) and f32 vs f64 bytes",
        },
        "fidelity_vs_f64_min": float(min(fids)),
        "fidelity_vs_f64_median": float(np.median(fids)),
        "fidelity_sample": len(fids),
    }


def run_serving(cs, count: int, rate: float, capacity_sample: int):
    """Continuous-batching serving of random-layer circuits.

    Workload: seeded Poisson arrivals at `rate` req/s; n ~ U{12..28}, depth ~
    U{10..50}, f32 or f64 with probability 1/2 each; statevectors resident
    on the device from admission to result.  Latency = arrival -> result
    (f64 norm and <Z_i Z_i+1> on the host).  Capacity = the first
    `capacity_sample` requests all queued at t = 0, served with overlap,
    against the plain per-circuit path: the same circuits in f64, one at a
    time on one stream."""
    import numpy as np

    reqs = cs.poisson_requests(count, rate, 12, 28, 10, 50, 0.5, SEED)
    streams, threads = 4, 16

    def serve(rs, n_streams, realtime, shadow=False, batch=True):
        cfg = cs.ServeConfig()
        cfg.streams, cfg.compile_threads, cfg.realtime = n_streams, threads, realtime
        cfg.f64_shadow = shadow
        if not batch:  # every request on the per-request path
            cfg.batch_max_qubits = 0
        return cs.Server(cfg).run(rs)

    res, st = serve(reqs, streams, True)
    log(f"  serving: {count} requests at {rate}/s: {st['throughput']:.1f}/s, "
        f"p50 {st['p50_s'] * 1e3:.1f} ms, p99 {st['p99_s'] * 1e3:.1f} ms")
    sub = reqs[:capacity_sample]
    _, cap = serve(sub, streams, False)
    _, cap_nb = serve(sub, streams, False, batch=False)
    # Plain per-circuit path: the same circuits in f64, one at a time on one
    # stream, compiled ahead by the same host threads.
    _, pl = serve([r.with_precision(cs.Precision.F64) for r in sub], 1, False, batch=False)
    log(f"  capacity: {cap['throughput']:.1f} circuits/s mixed ({cap_nb['throughput']:.1f} "
        f"without the batch lane), plain f64 serial {pl['throughput']:.1f} circuits/s")
    # Latency below saturation: a second stream at 70 % of the measured
    # capacity (the offered load at `rate` exceeds what the HBM can serve).
    sus_rate = round(0.7 * cap["throughput"], 1)
    sus_reqs = cs.poisson_requests(min(300, count), sus_rate, 12, 28, 10, 50, 0.5, SEED + 1)
    _, sus = serve(sus_reqs, streams, True)
    _, sus_nb = serve(sus_reqs, streams, True, batch=False)
    log(f"  at {sus_rate}/s: p50 {sus['p50_s'] * 1e3:.1f} ms, p99 {sus['p99_s'] * 1e3:.1f} ms "
        f"(without the batch lane: p50 {sus_nb['p50_s'] * 1e3:.1f} ms, "
        f"p99 {sus_nb['p99_s'] * 1e3:.1f} ms)")
    # Accuracy pass: every request again, each f32 one with its f64 replay on
    # the device (checking mode, not timed).
    shadow, _ = serve(reqs, streams, False, shadow=True)
    fid = [o.fidelity_vs_f64 for o in shadow if o.precision == cs.Precision.F32]
    same = all(np.array_equal(np.asarray(a.zz), np.asarray(b.zz)) and a.norm_sq == b.norm_sq
               for a, b in zip(sorted(res, key=lambda o: o.id), sorted(shadow, key=lambda o: o.id)))
    log(f"  f32 requests: fidelity vs f64 min {min(fid):.12f} over {len(fid)}")
    norms = [abs(o.norm_sq - 1.0) for o in res]
    return {
        "workload": f"{count} requests, Poisson arrivals at {rate}/s (seed {SEED}), "
                    "random-layer circuits n ~ U{12..28}, depth ~ U{10..50}, f32/f64 50/50, "
                    "fused to k<=6, resident statevectors, admission by free HBM, 4 streams; "
                    "requests of <= 18 qubits in batches of up to 32 (one launch per gate "
                    "step and size)",
        "circuits_per_s": round(st["throughput"], 2),
        "p50_ms": round(st["p50_s"] * 1e3, 2),
        "p99_ms": round(st["p99_s"] * 1e3, 2),
        "mean_ms": round(st["mean_s"] * 1e3, 2),
        "queue_p99_ms": round(st["queue_p99_s"] * 1e3, 2),
        "max_in_flight": st["max_in_flight"],
        "peak_sv_gib": round(st["peak_sv_bytes"] / 2**30, 2),
        "wall_s": round(st["wall_s"], 2),
        "capacity_circuits_per_s": round(cap["throughput"], 2),
        "capacity_without_batch_lane": round(cap_nb["throughput"], 2),
        "capacity_sample": len(sub),
        "batch_lane": {"batches": st["batches"], "batched_requests": st["batched_requests"],
                       "launches": st["batch_launches"]},
        "plain_f64_serial_circuits_per_s": round(pl["throughput"], 2),
        "speedup_vs_plain_f64": round(cap["throughput"] / pl["throughput"], 2),
        "max_norm_error": max(norms),
        "streams": streams,
        "compile_threads": threads,
        "fidelity_vs_f64_min": float(min(fid)),
        "fidelity_vs_f64_median": float(np.median(fid)),
        "fidelity_sample": len(fid),
        "results_identical_in_accuracy_pass": bool(same),
        "below_saturation": {"rate": sus_rate, "requests": len(sus_reqs),
                             "circuits_per_s": round(sus["throughput"], 2),
                             "p50_ms": round(sus["p50_s"] * 1e3, 2),
                             "p99_ms": round(sus["p99_s"] * 1e3, 2),
                             "max_in_flight": sus["max_in_flight"],
                             "without_batch_lane": {
                                 "circuits_per_s": round(sus_nb["throughput"], 2),
                                 "p50_ms": round(sus_nb["p50_s"] * 1e3, 2),
                                 "p99_ms": round(sus_nb["p99_s"] * 1e3, 2)}},
    }


def run_trajectory_launch(cs, n: int, count: int, tuning):
    """The trajectory workload of run_trajectories with a fixed tuning and no
    plain-path comparison: one plan + checkpoints + persistent launch."""
    D, p = 40, 1e-3
    lc = cs.LayeredCircuit.random_layers(n, D, SEED)
    pairs = [q for i in range(n - 1) for q in (i, i + 1)]
    S = cs.sample_trajectories(lc, cs.NoiseSpec(p, 10, D, SEED), count)
    eng = cs.TrajectoryEngine(lc, cs.Precision.F32, pairs)
    r = eng.run(S, tuning)
    sv_bytes = (2 << n) * 4
    alg = r.n_gate_passes * 2 * sv_bytes + r.n_tasks * sv_bytes
    return {"workload": f"{count} trajectories of the {n}-qubit 40-layer circuit, f32",
            "tuning": repr(tuning), "tasks": r.n_tasks, "gate_passes": r.n_gate_passes,
            "slots": r.n_slots, "gpu_s": r.gpu_s, "total_s": r.total_s,
            "circuits_per_s": count / r.total_s, "algorithmic_bytes": alg,
            "kernel_gbs": alg / r.gpu_s / 1e9}


def cpu_info() -> dict:
    """CPU model and the cores this process may run on: logical CPUs in its
    affinity set, and distinct (package, core) pairs among them."""
    try:
        allowed = set(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = set(range(os.cpu_count() or 1))
    model, phys = platform.machine(), set()
    try:
        for block in open("/proc/cpuinfo").read().split("\n\n"):
            f = dict(line.split(":", 1) for line in block.splitlines() if ":" in line)
            f = {k.strip(): v.strip() for k, v in f.items()}
            if "processor" not in f:
                continue
            model = f.get("model name", model)
            if int(f["processor"]) in allowed:
                phys.add((f.get("physical id", "0"), f.get("core id", f["processor"])))
    except OSError:
        pass
    return {"model": model, "logical": len(allowed), "physical": len(phys) or len(allowed)}


def cpu_baseline(gates, n_gpu: int, budget_s: float = 12.0):
    """The oracle's SIMD split-layout kernels (a restatement of the
    reference CPU path) over the first fused gates of the same circuit, on
    the same qubit count when host memory allows (else 28 qubits): once on
    every logical CPU of this process (the reference's default thread count,
    src/cpu/mod.rs:22-28) and once on one thread per physical core (the
    reference docs' advice, docs/tools.md:176-178)."""
    import numpy as np

    import oracle

    info = cpu_info()

    avail = 0
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemAvailable:"):
                avail = int(line.split()[1]) * 1024
    except OSError:
        pass
    n = n_gpu if avail > 3 * (2 << n_gpu) * 8 else 28
    sv = oracle.SimdStatevector(n, f32=False)
    sv.buf[0] = 1.0
    b = 2 * (1 << n) * 2 * 8

    def timed(threads):
        warm = gates[0]
        sv.apply(np.asarray(warm.matrix()), list(warm.qubits()), threads)  # untimed
        done, t0 = 0, time.perf_counter()
        for g in gates[1:]:
            if max(g.qubits()) >= n:
                continue
            sv.apply(np.asarray(g.matrix()), list(g.qubits()), threads)
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        return done, time.perf_counter() - t0

    logical = oracle.num_cpus()
    d_l, t_l = timed(logical)
    physical = max(1, min(info["physical"], logical))
    d_p, t_p = timed(physical) if physical != logical else (d_l, t_l)
    return {
        "value": d_l * b / t_l / 1e9,
        "unit": "GB/s",
        "cores": logical,
        "kind": "port",
        "sample": f"{d_l} fused gates (k<=5) of the same circuit on a {n}-qubit f64 split-layout "
                  f"statevector, {logical} threads (every logical CPU of this process), "
                  f"{t_l:.1f} s",
        "cpu": info["model"],
        "logical_cpus": info["logical"],
        "physical_cores": info["physical"],
        "physical_core_run": {"threads": physical, "gbs": round(d_p * b / t_p / 1e9, 2),
                              "gates": d_p, "seconds": round(t_p, 1)},
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--max-size", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f32", action="store_true")
    ap.add_argument("--workload", choices=["headline", "dense-large", "trajectories"],
                    default="headline",
                    help="dense-large: the 33q f64 / 34q f32 k<=6 matrix-core workload; "
                         "trajectories: the batched-trajectory launch alone with a fixed "
                         "tuning (for profiling); each prints its own JSON line")
    ap.add_argument("--traj-tuning", default="4,6,512,1,2",
                    help="chunk_layers,fusion_max,block,slots_per_cu,waves_per_simd[,mfma_f32] for "
                         "--workload trajectories")
    ap.add_argument("--no-traj", action="store_true")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 33q f64 / 34q f32 dense-fusion object (needs 192 GiB)")
    ap.add_argument("--no-qft", action="store_true", help="skip the 20-qubit QFT object")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the marginal-probability and density-matrix measurements")
    ap.add_argument("--requests", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--traj-qubits", type=int, default=24)
    ap.add_argument("--trajectories", type=int, default=4096)
    ap.add_argument("--copy-probe-qubits", type=int, default=0,
                    help="first copy one f64 statevector of this many qubits device-to-device "
                         "(a known-size transfer that calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (got --gpus {args.gpus}); sharding a "
              "statevector over several GPUs is not implemented.", file=sys.stderr)
        return 2

    import cast_simulator_amd as cs

    if args.copy_probe_qubits > 0:
        from cast_simulator_amd import abi

        nq = args.copy_probe_qubits
        nbytes = (2 << nq) * 8
        src, dst = abi.sv_alloc(nq, abi.F64), abi.sv_alloc(nq, abi.F64)
        abi.memcpy_dtod_async(dst, src, nbytes)
        abi.stream_sync(None)
        abi.sv_free(src)
        abi.sv_free(dst)
        log(f"[bench] copy probe: {nbytes} bytes device-to-device")

    if args.workload == "dense-large":
        print(json.dumps(dense_large_line(run_dense_large(cs))), flush=True)
        return 0
    if args.workload == "trajectories":
        vals = [int(x) for x in args.traj_tuning.split(",")]
        b, k, blk, spc, occ = vals[:5]
        mf32 = vals[5] if len(vals) > 5 else 0
        print(json.dumps(run_trajectory_launch(cs, args.traj_qubits, args.trajectories,
                                               cs.TrajectoryTuning(b, k, blk, spc, 0, occ, mf32))),
              flush=True)
        return 0

    log(f"[bench] f64 {args.qubits}q circuit: warmup {args.warmup}, steps {args.steps}")
    r64 = run_gpu(cs, args.qubits, "f64", args.steps, args.warmup, args.max_size)
    log(f"[bench] f64: {r64['gbs_per_gate']:.1f} GB/s per gate, {r64['n_gates']} gates, "
        f"{r64['ms_per_step']:.1f} ms/step, per-k {r64['per_k_gbs']}")
    r32 = fid = None
    if not args.no_f32:
        r32 = run_gpu(cs, args.qubits, "f32", args.steps, args.warmup, args.max_size)
        log(f"[bench] f32: {r32['gbs_per_gate']:.1f} GB/s per gate, per-k {r32['per_k_gbs']}")
        fid = fidelity_f32_vs_f64(cs, r64, r32, args.qubits)
        log(f"[bench] f32 vs f64 fidelity {fid['fidelity_vs_f64']:.15f}")
    qft = None
    if not args.no_qft:
        log("[bench] 20-qubit QFT (fusion <= 2)")
        qft = run_qft20(cs)
    large = None
    if not args.no_large:
        log("[bench] dense fusion at 33q f64 / 34q f32")
        large = run_dense_large(cs)
    traj = None
    if not args.no_traj:
        log(f"[bench] trajectories: {args.trajectories} x {args.traj_qubits}q")
        traj = run_trajectories(cs, args.traj_qubits, args.trajectories,
                                min(1024, args.trajectories), min(64, args.trajectories))
    serve = None
    if not args.no_serve:
        log(f"[bench] serving: {args.requests} requests at {args.rate}/s")
        serve = run_serving(cs, args.requests, args.rate, min(200, args.requests))
    marg = dm = None
    if not args.no_extra and args.qubits == 30:
        log("[bench] marginals on 30 qubits")
        marg = run_marginals(cs)
        log("[bench] density matrix on 15 qubits (30 virtual)")
        dm = run_density_matrix(cs)
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(r64["gates"], args.qubits)
        log(f"[bench] cpu: {cpu['value']:.2f} GB/s ({cpu['sample']})")
        if qft:
            qft["cpu"] = cpu_qft20(qft, cpu["cores"])
            qft["cpu"]["kind"] = "port"
            qft["cpu_physical_cores"] = cpu_qft20(qft, cpu["physical_core_run"]["threads"])
            log(f"[bench] qft20 cpu: {qft['cpu']['ms_per_pass']} ms/pass, gpu "
                f"{qft['gpu_wall_ms_per_pass']} ms/pass, rel L2 {qft['cpu']['gpu_vs_cpu_rel_l2']:.2e}")

    dom = r64["dominant"]
    # HBM bytes per launch of the dominant kernel, from the committed PMC
    # summary (FETCH_SIZE / WRITE_SIZE runs calibrated by a known-size copy).
    traffic = None
    tpath = ROOT / "profiles" / "gate_f64_30q_traffic.json"
    if args.qubits == 30 and tpath.exists():
        traffic = json.loads(tpath.read_text())["hbm_bytes_per_launch"].get(dom["kernel"])
    line = {
        "metric": METRIC,
        "value": round(r64["gbs_per_gate"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.qubits}-qubit random-layer circuit, depth 20, seed {SEED}, "
                        f"fused to k<={args.max_size} by the hardware-adaptive cost model, "
                        "one resident f64 statevector",
            "qubits": args.qubits,
            "fused_gates": r64["n_gates"],
            "k_histogram": r64["k_hist"],
            "cost_model_profile": r64["profile"],
            "crossover_ai": round(r64["crossover_ai"], 2),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom["kernel"],
            "achieved": round(dom["achieved_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": "profiles/gate_f64_30q_traffic.json" if traffic else None,
            "avg_launch_ms": round(dom["avg_ms"], 4),
            "launches": dom["launches"],
            "stream_ceiling_gbs": round(r64["stream_ceiling_gbs"], 1),
            "value_frac_of_stream_ceiling": round(r64["gbs_per_gate"] / r64["stream_ceiling_gbs"], 4),
        },
        "per_kernel_gbs": {k: round(v, 1) for k, v in r64["per_k_gbs"].items()},
        "autotune": r64["autotune"],
        "gate_updates_per_s": round(r64["gate_updates_per_s"], 2),
        "amplitude_updates_per_s": round(r64["gate_updates_per_s"] * (1 << args.qubits), 0),
        "cpu_baseline": cpu,
    }
    if fid:
        line["fidelity_f32_vs_f64"] = fid["fidelity_vs_f64"]
    if r32:
        d32 = r32["dominant"]
        t32 = None
        tpath32 = ROOT / "profiles" / "gate_f32_30q_traffic.json"
        if args.qubits == 30 and tpath32.exists():
            t32 = json.loads(tpath32.read_text())["hbm_bytes_per_launch"].get(d32["kernel"])
        line["f32"] = {
            "gbs_per_gate": round(r32["gbs_per_gate"], 2),
            "ms_per_step": round(r32["ms_per_step"], 3),
            "roofline": {"bound": "hbm", "kernel": d32["kernel"],
                         "achieved": round(d32["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(d32["achieved_gbs"] / HBM_PEAK_GBS, 4),
                         "traffic": t32,
                         "traffic_source": "profiles/gate_f32_30q_traffic.json" if t32 else None,
                         "avg_launch_ms": round(d32["avg_ms"], 4), "launches": d32["launches"]},
            "per_kernel_gbs": {k: round(v, 1) for k, v in r32["per_k_gbs"].items()},
            "fused_gates": r32["n_gates"],
            "k_histogram": r32["k_hist"],
            **(fid or {}),
        }
    if qft:
        line["qft20"] = {k: v for k, v in qft.items() if not k.startswith("_")}
    if large:
        line["dense_large"] = large
    if traj:
        line["trajectories"] = traj
    if serve:
        line["serving"] = serve
    if marg:
        line["marginals"] = marg
    if dm:
        line["density_matrix"] = dm
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
