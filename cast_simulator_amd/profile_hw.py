"""Roofline profiler for the gfx950 backend (the counterpart of the
reference's ``src/bin/profile_hw.rs``): sweeps random gates of increasing
arithmetic intensity on a resident statevector, fits the piecewise roofline
(memory slope, peak, crossover AI) with the reference's adaptive sweep
(``src/profile.rs:229-344``), and saves one HardwareProfile JSON per
precision.  The hardware-adaptive fusion cost model loads these files.

    python -m cast_simulator_amd.profile_hw [--precision f32|f64|all]
        [-n QUBITS] [--budget SECONDS] [--save-profiles DIR]

Each gate's kernel family is settled by the per-gate autotuner before it is
timed, so the fit describes the kernels a simulation actually runs.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

DEFAULT_DIR = Path(__file__).resolve().parent / "profiles"
# R^2 below which the fit is reported as unreliable (src/profile.rs:69, 323-328).
R2_WARN = 0.90


def auto_qubits(precision: str) -> int:
    """Largest n <= 30 whose statevector fits in 80% of free device memory
    (profile_hw.rs, the --n-qubits default)."""
    from . import abi

    free, _ = abi.free_memory()
    scalar = 8 if precision == "f64" else 4
    n = 30
    while n > 10 and (2 << n) * scalar > 0.8 * free:
        n -= 1
    return n


def r2_warning(label: str, r2: float):
    """The reference's warning for an unreliable fit (src/profile.rs:323-328),
    or None when R^2 >= R2_WARN."""
    if r2 >= R2_WARN:
        return None
    return (f"  {label}: Warning: roofline fit R^2 = {r2:.3f} (< {R2_WARN}); "
            "crossover estimate may be unreliable.")


def profile(precision: str, n: int, budget: float, seed: int = 1):
    import cast_simulator_amd as cs

    spec = cs.HipKernelGenSpec.f64() if precision == "f64" else cs.HipKernelGenSpec.f32()
    return cs.measure_hip(spec, n, budget, seed)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--precision", choices=["f32", "f64", "all"], default="all")
    ap.add_argument("-n", "--n-qubits", type=int, default=0)
    ap.add_argument("--budget", type=float, default=30.0)
    ap.add_argument("--save-profiles", type=Path, default=DEFAULT_DIR)
    args = ap.parse_args(argv)
    precs = ["f32", "f64"] if args.precision == "all" else [args.precision]
    args.save_profiles.mkdir(parents=True, exist_ok=True)
    for p in precs:
        n = args.n_qubits or auto_qubits(p)
        print(f"profiling hip {p} on {n} qubits, budget {args.budget:.0f} s", flush=True)
        prof = profile(p, n, args.budget)
        out = args.save_profiles / f"mi355x_{p}.json"
        prof.save(str(out))
        import cast_simulator_amd as cs

        r2 = cs.fit_roofline(prof.raw)[2] if prof.raw else float("nan")
        print(f"  crossover AI {prof.crossover_ai:.2f}, bandwidth slope "
              f"{prof.bw_slope:.2f} GFLOP/s per AI, peak {prof.peak_gflops_s:.1f} GFLOP/s "
              f"(reference units), {len(prof.raw)} probes, R^2 {r2:.3f} -> {out}", flush=True)
        warn = r2_warning(f"hip {p}", r2)
        if warn:
            print(warn, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
