"""ctypes view of the C ABI in ``include/cast_hip.h``.

This is the binding a foreign host (the reference's Rust crate would use
``extern "C"`` declarations, src/ffi/cuda.rs) sees: plain pointers, sizes,
uint64 device addresses, opaque handles and (return code, err_buf) errors.
Tests use it to drive the kernels directly through the boundary.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

HEADER = Path(__file__).resolve().parent.parent / "include" / "cast_hip.h"
LIB_PATH = Path(__file__).resolve().parent / "lib" / "libcast_hip.so"
ERR_LEN = 1024

F32, F64 = 0, 1
VARIANT_AUTO, VARIANT_DENSE, VARIANT_MONOMIAL, VARIANT_TILE, VARIANT_MFMA = 0, 1, 2, 3, 4


class AbiError(RuntimeError):
    """A cast_hip_* call returned failure; the message is the library's err_buf."""


class Complex64(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class KernelGenSpec(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("ztol", C.c_double),
        ("otol", C.c_double),
        ("variant", C.c_uint32),
        ("flags", C.c_uint32),
    ]


FLAG_TILE_SWIZZLE = 1
FLAG_TILE_GAUSS = 2
FLAG_TILE_XCD = 4
FLAG_MFMA_SPLIT = 8
FLAG_MFMA_F32 = 16


class PlanInfo(C.Structure):
    _fields_ = [
        ("n_qubits", C.c_uint32),
        ("precision", C.c_uint32),
        ("variant", C.c_uint32),
        ("nnz_rows", C.c_uint32),
        ("scalar_nnz", C.c_uint64),
        ("qubits", C.c_uint32 * 8),
    ]


class DeviceInfo(C.Structure):
    _fields_ = [
        ("name", C.c_char * 128),
        ("arch", C.c_char * 64),
        ("n_cu", C.c_uint32),
        ("wavefront_size", C.c_uint32),
        ("lds_per_block", C.c_uint32),
        ("l2_bytes", C.c_uint32),
        ("clock_khz", C.c_uint32),
        ("mem_clock_khz", C.c_uint32),
        ("mem_bus_width", C.c_uint32),
        ("reserved", C.c_uint32),
        ("total_mem", C.c_uint64),
    ]


class TrajGate(C.Structure):
    _fields_ = [
        ("matrix_offset", C.c_uint32),
        ("k", C.c_uint8),
        ("q", C.c_uint8 * 6),
        ("pad", C.c_uint8),
    ]


class TrajTask(C.Structure):
    _fields_ = [
        ("first_gate", C.c_uint32),
        ("n_gates", C.c_uint32),
        ("checkpoint", C.c_uint32),
        ("pad", C.c_uint32),
    ]


class TrajLaunch(C.Structure):
    _fields_ = [
        ("precision", C.c_uint32),
        ("n_qubits", C.c_uint32),
        ("tasks", C.c_uint64),
        ("order", C.c_uint64),
        ("n_tasks", C.c_uint32),
        ("n_gates", C.c_uint32),
        ("gates", C.c_uint64),
        ("matrices", C.c_uint64),
        ("n_matrix_elems", C.c_uint64),
        ("checkpoints", C.c_uint64),
        ("n_checkpoints", C.c_uint32),
        ("n_slots", C.c_uint32),
        ("slots", C.c_uint64),
        ("block", C.c_uint32),
        ("n_pairs", C.c_uint32),
        ("pairs", C.c_uint32 * 64),
        ("out", C.c_uint64),
        ("dump", C.c_uint64),
        ("work", C.c_uint64),
        ("waves_per_simd", C.c_uint32),
        ("flags", C.c_uint32),
    ]


_vp, _sz, _u8, _u32, _u64, _i = C.c_void_p, C.c_size_t, C.c_uint8, C.c_uint32, C.c_uint64, C.c_int
_err = C.c_char_p
_SIGS = {
    "cast_hip_version": (C.c_char_p, []),
    "cast_hip_device_info": (_i, [C.POINTER(DeviceInfo), _err, _sz]),
    "cast_hip_free_memory": (_i, [C.POINTER(_u64), C.POINTER(_u64), _err, _sz]),
    "cast_hip_stream_create": (_vp, [_err, _sz]),
    "cast_hip_stream_create_nonblocking": (_vp, [_err, _sz]),
    "cast_hip_event_query": (_i, [_vp, C.POINTER(C.c_int), _err, _sz]),
    "cast_hip_host_alloc": (_vp, [_sz, _err, _sz]),
    "cast_hip_host_free": (None, [_vp]),
    "cast_hip_memcpy_dtoh_async": (_i, [_vp, _u64, _sz, _vp, _err, _sz]),
    "cast_hip_sv_zero_async": (_i, [_u64, _sz, _u8, _vp, _err, _sz]),
    "cast_hip_observe_scratch_bytes": (_sz, [_u32]),
    "cast_hip_observe_async": (
        _i, [_u64, _u32, _u8, C.POINTER(_u32), _sz, _u64, _u64, _vp, _err, _sz]),
    "cast_hip_observe_noalloc": (
        _i, [_u64, _u32, _u8, C.POINTER(_u32), _sz, _u64, _vp, _err, _sz]),
    "cast_hip_stream_destroy": (None, [_vp]),
    "cast_hip_stream_sync": (_i, [_vp, _err, _sz]),
    "cast_hip_event_create": (_vp, [_err, _sz]),
    "cast_hip_event_record": (_i, [_vp, _vp, _err, _sz]),
    "cast_hip_event_synchronize": (_i, [_vp, _err, _sz]),
    "cast_hip_event_destroy": (None, [_vp]),
    "cast_hip_event_elapsed_ms": (_i, [_vp, _vp, C.POINTER(C.c_float), _err, _sz]),
    "cast_hip_memcpy_dtod_async": (_i, [_u64, _u64, _sz, _vp, _err, _sz]),
    "cast_hip_malloc": (_u64, [_sz, _err, _sz]),
    "cast_hip_free": (None, [_u64]),
    "cast_hip_memcpy_htod_async": (_i, [_u64, _vp, _sz, _vp, _err, _sz]),
    "cast_hip_memcpy_dtoh": (_i, [_vp, _u64, _sz, _vp, _err, _sz]),
    "cast_hip_sv_alloc": (_u64, [_sz, _u8, _err, _sz]),
    "cast_hip_sv_free": (None, [_u64]),
    "cast_hip_sv_zero": (_i, [_u64, _sz, _u8, _err, _sz]),
    "cast_hip_sv_upload": (_i, [_u64, C.POINTER(C.c_double), _sz, _u8, _err, _sz]),
    "cast_hip_sv_download": (_i, [_u64, C.POINTER(C.c_double), _sz, _u8, _err, _sz]),
    "cast_hip_norm_squared": (_i, [_u64, _sz, _u8, C.POINTER(C.c_double), _err, _sz]),
    "cast_hip_scale": (_i, [_u64, _sz, _u8, C.c_double, _err, _sz]),
    "cast_hip_gate_plan_create": (
        _i,
        [C.POINTER(KernelGenSpec), C.POINTER(Complex64), _sz, C.POINTER(_u32), _sz,
         C.POINTER(_vp), _err, _sz],
    ),
    "cast_hip_gate_plan_destroy": (None, [_vp]),
    "cast_hip_gate_plan_info": (_i, [_vp, C.POINTER(PlanInfo), _err, _sz]),
    "cast_hip_gate_plan_family": (_i, [_vp, _u32, C.POINTER(_u32), _err, _sz]),
    "cast_hip_gate_apply": (_i, [_vp, _vp, _u64, _u32, _err, _sz]),
    "cast_hip_gate_apply_batch": (
        _i, [C.POINTER(_vp), C.POINTER(_u64), C.POINTER(_u32), _sz, _vp, _err, _sz]),
    "cast_hip_expect_zz": (
        _i, [_u64, _u32, _u8, C.POINTER(_u32), _sz, C.POINTER(C.c_double), _vp, _err, _sz]),
    "cast_hip_traj_run": (_i, [C.POINTER(TrajLaunch), _vp, _err, _sz]),
    "cast_hip_inner_product": (
        _i, [_u64, _u8, _u64, _u8, _u32, C.POINTER(C.c_double), _vp, _err, _sz]),
    "cast_hip_inner_product_scratch_bytes": (_sz, []),
    "cast_hip_inner_product_async": (
        _i, [_u64, _u8, _u64, _u8, _u32, _u64, _u64, _vp, _err, _sz]),
    "cast_hip_marginal_probabilities": (
        _i, [_u64, _u32, _u8, C.POINTER(_u32), _sz, C.POINTER(C.c_double), _vp, _err, _sz]),
    "cast_hip_dm_populations": (
        _i, [_u64, _u32, _u8, C.POINTER(C.c_double), _vp, _err, _sz]),
}


def declared_symbols(header: Path = HEADER) -> list[str]:
    """Function names declared in a C header of this repository."""
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(cast_[a-z0-9_]+)\s*\(", text)))


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build the gfx950 backend first")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _call(name: str, *args):
    buf = C.create_string_buffer(ERR_LEN)
    rc = getattr(lib(), name)(*args, buf, ERR_LEN)
    if rc != 0:
        raise AbiError(f"{name}: {buf.value.decode(errors='replace')}")


def _ctor(name: str, *args):
    buf = C.create_string_buffer(ERR_LEN)
    h = getattr(lib(), name)(*args, buf, ERR_LEN)
    if not h:
        raise AbiError(f"{name}: {buf.value.decode(errors='replace')}")
    return h


def device_info() -> DeviceInfo:
    d = DeviceInfo()
    _call("cast_hip_device_info", C.byref(d))
    return d


def free_memory() -> tuple[int, int]:
    f, t = _u64(), _u64()
    _call("cast_hip_free_memory", C.byref(f), C.byref(t))
    return f.value, t.value


def sv_alloc(n_qubits: int, precision: int) -> int:
    return _ctor("cast_hip_sv_alloc", 2 << n_qubits, precision)


def sv_free(dptr: int) -> None:
    lib().cast_hip_sv_free(dptr)


def sv_zero(dptr: int, n_qubits: int, precision: int) -> None:
    _call("cast_hip_sv_zero", dptr, 2 << n_qubits, precision)


def sv_upload(dptr: int, host: np.ndarray, precision: int) -> None:
    h = np.ascontiguousarray(host, dtype=np.float64).reshape(-1)
    _call("cast_hip_sv_upload", dptr, h.ctypes.data_as(C.POINTER(C.c_double)), h.size, precision)


def sv_download(dptr: int, n_qubits: int, precision: int) -> np.ndarray:
    h = np.empty(2 << n_qubits, dtype=np.float64)
    _call("cast_hip_sv_download", dptr, h.ctypes.data_as(C.POINTER(C.c_double)), h.size,
          precision)
    return h


def norm_squared(dptr: int, n_qubits: int, precision: int) -> float:
    out = C.c_double()
    _call("cast_hip_norm_squared", dptr, 2 << n_qubits, precision, C.byref(out))
    return out.value


def scale(dptr: int, n_qubits: int, precision: int, factor: float) -> None:
    _call("cast_hip_scale", dptr, 2 << n_qubits, precision, factor)


def plan_create(matrix: np.ndarray, qubits, precision: int, ztol: float | None = None,
                otol: float | None = None, variant: int = VARIANT_AUTO, flags: int = 0) -> int:
    m = np.ascontiguousarray(matrix, dtype=np.complex128).reshape(-1)
    arr = (Complex64 * m.size)(*[(z.real, z.imag) for z in m])
    q = (_u32 * len(qubits))(*qubits)
    tol = 1e-6 if precision == F32 else 1e-12
    spec = KernelGenSpec(precision, tol if ztol is None else ztol, tol if otol is None else otol,
                         variant, flags)
    out = _vp()
    _call("cast_hip_gate_plan_create", C.byref(spec), arr, m.size, q, len(qubits), C.byref(out))
    return out.value


def plan_destroy(plan: int) -> None:
    lib().cast_hip_gate_plan_destroy(plan)


def plan_info(plan: int) -> PlanInfo:
    info = PlanInfo()
    _call("cast_hip_gate_plan_info", plan, C.byref(info))
    return info


def plan_family(plan: int, n_qubits: int) -> int:
    out = _u32()
    _call("cast_hip_gate_plan_family", plan, n_qubits, C.byref(out))
    return out.value


def gate_apply(plan: int, dptr: int, n_qubits: int, stream: int | None = None) -> None:
    _call("cast_hip_gate_apply", plan, stream, dptr, n_qubits)


def gate_apply_batch(plans, svs, n_qubits, stream: int | None = None) -> None:
    n = len(plans)
    _call("cast_hip_gate_apply_batch", (_vp * n)(*plans), (_u64 * n)(*svs),
          (_u32 * n)(*n_qubits), n, stream)


def expect_zz(dptr: int, n_qubits: int, precision: int, pairs, stream: int | None = None):
    flat = [q for p in pairs for q in p]
    out = (C.c_double * len(pairs))()
    _call("cast_hip_expect_zz", dptr, n_qubits, precision, (_u32 * len(flat))(*flat),
          len(pairs), out, stream)
    return np.array(out[:], dtype=np.float64)


def stream_create() -> int:
    return _ctor("cast_hip_stream_create")


def stream_destroy(s: int) -> None:
    lib().cast_hip_stream_destroy(s)


def stream_sync(s: int | None) -> None:
    _call("cast_hip_stream_sync", s)


def event_create() -> int:
    return _ctor("cast_hip_event_create")


def event_record(e: int, s: int | None) -> None:
    _call("cast_hip_event_record", e, s)


def event_synchronize(e: int) -> None:
    _call("cast_hip_event_synchronize", e)


def event_elapsed_ms(a: int, b: int) -> float:
    out = C.c_float()
    _call("cast_hip_event_elapsed_ms", a, b, C.byref(out))
    return out.value


def event_destroy(e: int) -> None:
    lib().cast_hip_event_destroy(e)


def memcpy_dtod_async(dst: int, src: int, n_bytes: int, stream: int | None = None) -> None:
    _call("cast_hip_memcpy_dtod_async", dst, src, n_bytes, stream)


def malloc(n_bytes: int) -> int:
    return _ctor("cast_hip_malloc", n_bytes)


def free(dptr: int) -> None:
    lib().cast_hip_free(dptr)


def memcpy_htod(dst: int, host: np.ndarray, stream: int | None = None) -> None:
    """Copy a C-contiguous host array to device memory and wait."""
    a = np.ascontiguousarray(host)
    _call("cast_hip_memcpy_htod_async", dst, a.ctypes.data, a.nbytes, stream)
    stream_sync(stream)


def memcpy_dtoh(host: np.ndarray, src: int, stream: int | None = None) -> None:
    """Fill a C-contiguous host array from device memory (synchronous)."""
    if not host.flags.c_contiguous:
        raise ValueError("host array must be C-contiguous")
    _call("cast_hip_memcpy_dtoh", host.ctypes.data, src, host.nbytes, stream)


def stream_create_nonblocking() -> int:
    return _ctor("cast_hip_stream_create_nonblocking")


def event_query(e: int) -> bool:
    done = C.c_int(0)
    _call("cast_hip_event_query", e, C.byref(done))
    return bool(done.value)


def sv_zero_async(dptr: int, n_qubits: int, precision: int, stream: int | None) -> None:
    _call("cast_hip_sv_zero_async", dptr, 2 << n_qubits, precision, stream)


def observe_scratch_bytes(n_pairs: int) -> int:
    return int(lib().cast_hip_observe_scratch_bytes(n_pairs))


def observe_async(dptr: int, n_qubits: int, precision: int, pairs, out: int, scratch: int,
                  stream: int | None = None) -> None:
    flat = [q for p in pairs for q in p]
    arr = (_u32 * max(1, len(flat)))(*flat)
    _call("cast_hip_observe_async", dptr, n_qubits, precision, arr, len(pairs), out, scratch,
          stream)


def observe_noalloc(dptr: int, n_qubits: int, precision: int, pairs, out: int,
                    stream: int | None = None) -> None:
    flat = [q for p in pairs for q in p]
    arr = (_u32 * max(1, len(flat)))(*flat)
    _call("cast_hip_observe_noalloc", dptr, n_qubits, precision, arr, len(pairs), out, stream)


def inner_product(a: int, prec_a: int, b: int, prec_b: int, n_qubits: int,
                  stream: int | None = None) -> complex:
    out = (C.c_double * 2)()
    _call("cast_hip_inner_product", a, prec_a, b, prec_b, n_qubits, out, stream)
    return complex(out[0], out[1])


def inner_product_scratch_bytes() -> int:
    return int(lib().cast_hip_inner_product_scratch_bytes())


def inner_product_async(a: int, prec_a: int, b: int, prec_b: int, n_qubits: int, out: int,
                        scratch: int, stream: int | None = None) -> None:
    _call("cast_hip_inner_product_async", a, prec_a, b, prec_b, n_qubits, out, scratch, stream)


def marginal_probabilities(sv: int, n_qubits: int, precision: int, measured,
                           stream: int | None = None) -> np.ndarray:
    out = np.zeros(1 << len(measured), dtype=np.float64)
    q = (_u32 * max(1, len(measured)))(*measured)
    _call("cast_hip_marginal_probabilities", sv, n_qubits, precision, q, len(measured),
          out.ctypes.data_as(C.POINTER(C.c_double)), stream)
    return out


def dm_populations(sv: int, n_physical: int, precision: int,
                   stream: int | None = None) -> np.ndarray:
    out = np.zeros(1 << n_physical, dtype=np.float64)
    _call("cast_hip_dm_populations", sv, n_physical, precision,
          out.ctypes.data_as(C.POINTER(C.c_double)), stream)
    return out


def traj_run(args: TrajLaunch, stream: int | None = None) -> None:
    _call("cast_hip_traj_run", C.byref(args), stream)


def version() -> str:
    return lib().cast_hip_version().decode()
