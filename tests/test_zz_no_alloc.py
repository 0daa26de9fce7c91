"""GPU check of cast_hip_observe_noalloc, the allocation-free form of
cast_hip_observe_async: its reduction partials live in the library's static
device array instead of caller scratch.

Both calls run the same kernels over the same grid, so the results must be
bit-identical (the reduction order is fixed).  Two calls in a row, the second
on a smaller statevector with fewer pairs, show that the array's leftover
contents from the first call do not leak into the second.  The norm and
<Z_a Z_b> sums are also checked against the oracle's host sums.
"""
from __future__ import annotations

import numpy as np
import pytest

import cast_simulator_amd as cs
from cast_simulator_amd import abi

pytestmark = pytest.mark.gpu


def host_observe(state, n, pairs):
    """f64 norm and pair sums of an interleaved state on the host."""
    a = state[0::2] + 1j * state[1::2]
    pr = np.abs(a) ** 2
    idx = np.arange(1 << n, dtype=np.uint64)
    out = [pr.sum()]
    for qa, qb in pairs:
        par = ((idx >> np.uint64(qa)) ^ (idx >> np.uint64(qb))) & np.uint64(1)
        out.append(np.where(par == 1, -pr, pr).sum())
    return np.array(out)


def run_both(n, precision, pairs, seed):
    state = np.asarray(cs.random_statevector(n, seed))
    sv = abi.sv_alloc(n, precision)
    out_a = out_b = scratch = 0
    try:
        abi.sv_upload(sv, state, precision)
        w = 1 + len(pairs)
        out_a = abi.malloc(8 * w)
        out_b = abi.malloc(8 * w)
        scratch = abi.malloc(abi.observe_scratch_bytes(len(pairs)))
        abi.observe_async(sv, n, precision, pairs, out_a, scratch)
        abi.observe_noalloc(sv, n, precision, pairs, out_b)
        abi.stream_sync(None)
        a = np.zeros(w)
        b = np.zeros(w)
        abi.memcpy_dtoh(a, out_a)
        abi.memcpy_dtoh(b, out_b)
        stored = np.asarray(abi.sv_download(sv, n, precision))
    finally:
        for p in (out_a, out_b, scratch):
            if p:
                abi.free(p)
        abi.sv_free(sv)
    return a, b, host_observe(stored, n, pairs)


@pytest.mark.parametrize("precision", [abi.F64, abi.F32], ids=["f64", "f32"])
def test_noalloc_observe_matches_scratch_observe_twice(precision):
    big = [(i, i + 1) for i in range(21)] + [(0, 21), (3, 17)]   # 23 pairs on 22 qubits
    small = [(0, 1), (4, 9), (2, 11)]                              # 3 pairs on 12 qubits
    for n, pairs, seed in ((22, big, 5), (12, small, 6)):
        print(f"[noalloc] n={n} precision={precision} pairs={len(pairs)}", flush=True)
        a, b, host = run_both(n, precision, pairs, seed)
        assert np.array_equal(a, b), (a, b)
        np.testing.assert_allclose(b, host, rtol=0, atol=1e-12)


def test_noalloc_observe_rejects_more_pairs_than_the_scratch_holds():
    sv = abi.sv_alloc(8, abi.F64)
    out = abi.malloc(8 * 40)
    try:
        abi.sv_zero(sv, 8, abi.F64)
        with pytest.raises(abi.AbiError, match="at most 32 qubit pairs"):
            abi.observe_noalloc(sv, 8, abi.F64, [(0, 1)] * 33, out)
        abi.observe_noalloc(sv, 8, abi.F64, [(0, 1)] * 32, out)
        abi.stream_sync(None)
        got = np.zeros(33)
        abi.memcpy_dtoh(got, out)
        assert np.array_equal(got, np.ones(33))
    finally:
        abi.free(out)
        abi.sv_free(sv)
