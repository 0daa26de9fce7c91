"""The roofline profiler on the GPU (src/profile.rs:229-344, 412-447): a
short adaptive sweep of random gates through measure_hip, and profile_hw's
command line end to end.  Checks the sweep's own invariants: probes sorted
by arithmetic intensity over the reference's range, GFLOP/s equal
to 2 AI 2^n / t in the reference's unit (so GFLOP/s / GiB/s is fixed by AI
and the scalar size), bandwidth below the HBM peak on a statevector far
larger than the caches, and the saved profile's crossover equal to the fit
of its own raw sweep."""
from __future__ import annotations

import json

import pytest

import cast_simulator_amd as cs
from cast_simulator_amd import profile_hw

pytestmark = pytest.mark.gpu

HBM_PEAK_GIB = 8000e9 / 2**30


@pytest.mark.parametrize("f32", [False, True], ids=["f64", "f32"])
def test_measure_hip_sweep(f32):
    n = 27  # 1 or 2 GiB: well beyond the 256 MB last-level cache
    spec = cs.HipKernelGenSpec.f32() if f32 else cs.HipKernelGenSpec.f64()
    print(f"[profile] measure_hip n={n} f32={f32}", flush=True)
    prof = cs.measure_hip(spec, n, 4.0, 3)
    raw = list(prof.raw)
    ais = [a for a, _, _ in raw]
    assert len(raw) >= 7  # the base probes at AI 1, 2, 4, ..., 64, then refinements
    assert min(ais) <= 2 and max(ais) >= 32 and ais == sorted(ais)
    s = 4 if f32 else 8
    for ai, gf, gib in raw:
        assert 0 < gib < HBM_PEAK_GIB
        # t = 2 * 2^n * 2 s / (gib * 2^30) and gf = 2 ai 2^n / t / 1e9
        assert gf == pytest.approx(ai * gib * 2**30 / (2 * s) / 1e9, rel=1e-9)
    cx, slope, r2 = cs.fit_roofline(raw)
    assert prof.crossover_ai == pytest.approx(prof.peak_gflops_s / prof.bw_slope, rel=1e-12)
    assert 0 < r2 <= 1 and slope > 0
    assert prof.peak_bw_gib_s == pytest.approx(max(g for _, _, g in raw))


def test_profile_hw_command_writes_profile(tmp_path, capsys):
    rc = profile_hw.main(["--precision", "f64", "-n", "26", "--budget", "3",
                          "--save-profiles", str(tmp_path)])
    assert rc == 0
    out = capsys.readouterr().out
    d = json.loads((tmp_path / "mi355x_f64.json").read_text())
    assert d["config"]["precision"] == "f64" and d["config"]["n_qubits"] == 26
    assert len(d["raw"]) >= 7 and "crossover AI" in out
    prof = cs.HardwareProfile.load(str(tmp_path / "mi355x_f64.json"))
    r2 = cs.fit_roofline(list(prof.raw))[2]
    assert ("Warning: roofline fit" in out) == (r2 < profile_hw.R2_WARN)
