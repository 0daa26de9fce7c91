"""Python view of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the CPU restatement of the
reference's CPU gate path (see cast_oracle.h for the file:line each function
follows).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module, always as the checker or the timed CPU baseline,
never as part of the product path.

Every function here is pinned by tests/test_oracle.py or (marginal,
dm_populations, simulate_dm, ensemble, samples_per_member)
tests/test_ensemble_host.py; none is "parity unpinned".
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

DIR = Path(__file__).resolve().parent
LIB_PATH = DIR / "liboracle.so"

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)


def build() -> Path:
    """Compile the oracle with its Makefile (gcc only)."""
    proc = subprocess.run(["make", "-C", str(DIR), "-B"], capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + proc.stdout + proc.stderr)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_pdep.restype = C.c_uint64
        L.orc_pdep.argtypes = [C.c_uint64, C.c_uint64]
        for nm, p in (("f64", _dp), ("f32", _fp)):
            f = getattr(L, f"orc_apply_gate_{nm}")
            f.restype = None
            f.argtypes = [p, C.c_uint, _dp, _u32p, C.c_uint]
            f = getattr(L, f"orc_norm_squared_{nm}")
            f.restype = C.c_double
            f.argtypes = [p, C.c_size_t]
            f = getattr(L, f"orc_scale_{nm}")
            f.restype = None
            f.argtypes = [p, C.c_size_t, C.c_double]
            f = getattr(L, f"orc_simd_pack_{nm}")
            f.restype = None
            f.argtypes = [p, _dp, C.c_uint, C.c_uint]
            f = getattr(L, f"orc_simd_unpack_{nm}")
            f.restype = None
            f.argtypes = [_dp, p, C.c_uint, C.c_uint]
            f = getattr(L, f"orc_simd_apply_{nm}")
            f.restype = C.c_int
            f.argtypes = [p, C.c_uint, C.c_uint, _dp, _u32p, C.c_uint, C.c_uint]
        L.orc_expect_zz_f64.restype = None
        L.orc_expect_zz_f64.argtypes = [_dp, C.c_uint, _u32p, C.c_size_t, _dp]
        L.orc_marginal_f64.restype = None
        L.orc_marginal_f64.argtypes = [_dp, C.c_uint, _u32p, C.c_uint, _dp]
        L.orc_dm_populations_f64.restype = None
        L.orc_dm_populations_f64.argtypes = [_dp, C.c_uint, _dp]
        L.orc_simd_len.restype = C.c_size_t
        L.orc_simd_len.argtypes = [C.c_uint, C.c_uint]
        L.orc_num_cpus.restype = C.c_uint
        L.orc_num_cpus.argtypes = []
        _lib = L
    return _lib


def _mat(matrix) -> np.ndarray:
    m = np.ascontiguousarray(matrix, dtype=np.complex128)
    return np.ascontiguousarray(m.view(np.float64).reshape(-1))


def _q(qubits) -> C.Array:
    return (C.c_uint32 * len(qubits))(*qubits)


def pdep(src: int, mask: int) -> int:
    return int(lib().orc_pdep(src, mask))


def apply_gate(sv: np.ndarray, matrix, qubits, n: int | None = None) -> np.ndarray:
    """Apply in place to an interleaved float64 or float32 statevector."""
    if n is None:
        n = int(sv.size // 2).bit_length() - 1
    m = _mat(matrix)
    k = len(qubits)
    if sv.dtype == np.float64:
        lib().orc_apply_gate_f64(sv.ctypes.data_as(_dp), n, m.ctypes.data_as(_dp), _q(qubits), k)
    elif sv.dtype == np.float32:
        lib().orc_apply_gate_f32(sv.ctypes.data_as(_fp), n, m.ctypes.data_as(_dp), _q(qubits), k)
    else:
        raise TypeError("statevector must be float64 or float32")
    return sv


def apply_gates(sv: np.ndarray, gates) -> np.ndarray:
    """gates: iterable of (matrix, qubits)."""
    for m, q in gates:
        apply_gate(sv, m, q)
    return sv


def norm_squared(x: np.ndarray) -> float:
    if x.dtype == np.float64:
        return float(lib().orc_norm_squared_f64(x.ctypes.data_as(_dp), x.size))
    return float(lib().orc_norm_squared_f32(x.ctypes.data_as(_fp), x.size))


def scale(x: np.ndarray, factor: float) -> np.ndarray:
    if x.dtype == np.float64:
        lib().orc_scale_f64(x.ctypes.data_as(_dp), x.size, factor)
    else:
        lib().orc_scale_f32(x.ctypes.data_as(_fp), x.size, factor)
    return x


def expect_zz(sv: np.ndarray, pairs) -> np.ndarray:
    x = np.ascontiguousarray(sv, dtype=np.float64)
    n = int(x.size // 2).bit_length() - 1
    flat = [q for p in pairs for q in p]
    out = np.zeros(len(pairs), dtype=np.float64)
    lib().orc_expect_zz_f64(x.ctypes.data_as(_dp), n, (C.c_uint32 * len(flat))(*flat),
                            len(pairs), out.ctypes.data_as(_dp))
    return out


def marginal(sv: np.ndarray, measured) -> np.ndarray:
    """Marginal probabilities over `measured` (bin bit i = qubit measured[i])."""
    x = np.ascontiguousarray(sv, dtype=np.float64)
    n = int(x.size // 2).bit_length() - 1
    out = np.zeros(1 << len(measured), dtype=np.float64)
    lib().orc_marginal_f64(x.ctypes.data_as(_dp), n, _q(list(measured)) if measured else None,
                           len(measured), out.ctypes.data_as(_dp))
    return out


def dm_populations(rho: np.ndarray, n_physical: int) -> np.ndarray:
    """Diagonal of a density matrix held as a 2n-qubit vectorised state."""
    x = np.ascontiguousarray(rho, dtype=np.float64)
    out = np.zeros(1 << n_physical, dtype=np.float64)
    lib().orc_dm_populations_f64(x.ctypes.data_as(_dp), n_physical, out.ctypes.data_as(_dp))
    return out


def simulate_dm(n_physical: int, gates) -> np.ndarray:
    """Density-matrix simulation (src/simulator/mod.rs:523-547 with
    prepare_gates :607-615): rho = |0><0| as a 2n-qubit vector, each gate
    given as its already-lifted (superoperator, qubits) pair."""
    rho = np.zeros(2 << (2 * n_physical))
    rho[0] = 1.0
    for m, q in gates:
        apply_gate(rho, m, q)
    return rho


def ensemble(n: int, gates, max_ensemble: int):
    """Top-M ensemble over noise branches (src/simulator/trajectory.rs:186-265),
    pure Python over the oracle's apply_gate; small cases only.

    gates: list of ("u", matrix, qubits) or ("noise", [(p, kraus), ...], qubits).
    Candidates are ordered by weight, descending, ties in (member, branch)
    order (a stable sort; the reference's unstable sort leaves ties in an
    unspecified order).  Returns [(weight, noise_path, statevector)]."""
    sv = np.zeros(2 << n)
    sv[0] = 1.0
    ens = [(1.0, [], sv)]
    M = max(1, max_ensemble)
    for g in gates:
        if g[0] == "u":
            for _, _, x in ens:
                apply_gate(x, g[1], g[2])
            continue
        branches, q = g[1], g[2]
        cand = [(mi, bi, ens[mi][0] * p) for mi in range(len(ens))
                for bi, (p, _) in enumerate(branches)]
        cand.sort(key=lambda c: -c[2])  # stable
        cand = cand[:M]
        nxt = []
        for mi, bi, w in cand:
            x = ens[mi][2].copy()
            apply_gate(x, branches[bi][1], q)
            nxt.append((w, ens[mi][1] + [bi], x))
        ens = nxt
    return ens


def samples_per_member(weights, n_samples: int):
    """Samples drawn from each ensemble member (trajectory.rs:283-293)."""
    total = sum(weights)
    out, remaining = [], n_samples
    for i, w in enumerate(weights):
        s = remaining if i == len(weights) - 1 else min(int(np.floor(w / total * n_samples)),
                                                         remaining)
        remaining -= min(remaining, s)
        out.append(s)
    return out


class SimdStatevector:
    """Split-layout statevector of the CPU SIMD baseline (512-bit vectors)."""

    def __init__(self, n: int, f32: bool):
        self.n = n
        self.f32 = f32
        self.s = 4 if f32 else 3
        size = int(lib().orc_simd_len(n, self.s))
        self.buf = np.zeros(size, dtype=np.float32 if f32 else np.float64)
        self._p = self.buf.ctypes.data_as(_fp if f32 else _dp)

    def load(self, inter: np.ndarray) -> None:
        x = np.ascontiguousarray(inter, dtype=np.float64)
        f = lib().orc_simd_pack_f32 if self.f32 else lib().orc_simd_pack_f64
        f(self._p, x.ctypes.data_as(_dp), self.n, self.s)

    def interleaved(self) -> np.ndarray:
        out = np.empty(2 << self.n, dtype=np.float64)
        f = lib().orc_simd_unpack_f32 if self.f32 else lib().orc_simd_unpack_f64
        f(out.ctypes.data_as(_dp), self._p, self.n, self.s)
        return out

    def apply(self, matrix, qubits, n_threads: int = 0) -> None:
        m = _mat(matrix)
        f = lib().orc_simd_apply_f32 if self.f32 else lib().orc_simd_apply_f64
        rc = f(self._p, self.n, self.s, m.ctypes.data_as(_dp), _q(qubits), len(qubits), n_threads)
        if rc != 0:
            raise ValueError("orc_simd_apply rejected the gate (need n >= k + s, ascending qubits)")


def num_cpus() -> int:
    return int(lib().orc_num_cpus())
