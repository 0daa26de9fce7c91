"""The C ABI library loads, exports every symbol its headers declare, and
rejects bad arguments with the reference's (return 1, err_buf) convention
before touching a device.  CPU only: no compute call needs a GPU here."""
from __future__ import annotations

import ctypes as C
import subprocess

import numpy as np
import pytest

from cast_simulator_amd import abi


def test_every_declared_symbol_is_exported():
    names = abi.declared_symbols()
    assert len(names) >= 25
    lib = abi.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    # and the ctypes signature table covers the whole header
    assert sorted(abi._SIGS) == names


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in abi.declared_symbols():
        assert n in exported, n  # unmangled extern "C"


def test_version_string():
    assert "gfx950" in abi.version()


# Argument validation of the plan builder mirrors ffi_cuda.cpp:74-94 and
# cpu_gen.cpp:64-67; it happens before any device call.

def _plan_err(**kw):
    with pytest.raises(abi.AbiError) as e:
        abi.plan_create(**kw)
    return str(e.value)


def test_plan_rejects_wrong_matrix_len():
    msg = _plan_err(matrix=np.eye(2), qubits=[0, 1], precision=abi.F64)
    assert "matrix_len must equal (2^n_qubits)^2" in msg


def test_plan_rejects_non_ascending_qubits():
    msg = _plan_err(matrix=np.eye(4), qubits=[1, 0], precision=abi.F64)
    assert "strictly ascending" in msg


def test_plan_rejects_bad_precision():
    msg = _plan_err(matrix=np.eye(2), qubits=[0], precision=7)
    assert "precision" in msg


def test_plan_rejects_oversized_gate():
    msg = _plan_err(matrix=np.eye(128), qubits=list(range(7)), precision=abi.F64)
    assert "exceeds the maximum of 6" in msg


def test_plan_rejects_unknown_flags():
    msg = _plan_err(matrix=np.eye(2), qubits=[0], precision=abi.F64, flags=0x80)
    assert "unknown flags" in msg


def test_plan_rejects_forced_monomial_on_dense_matrix():
    # classification runs on the host, before the device upload
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    with pytest.raises(abi.AbiError, match="MONOMIAL|no HIP device"):
        abi.plan_create(h, [0], abi.F64, variant=abi.VARIANT_MONOMIAL)


def test_null_arguments_are_rejected():
    buf = C.create_string_buffer(abi.ERR_LEN)
    L = abi.lib()
    assert L.cast_hip_gate_plan_create(None, None, 0, None, 0, None, buf, abi.ERR_LEN) == 1
    assert b"must not be null" in buf.value
    assert L.cast_hip_gate_apply(None, None, 0, 4, buf, abi.ERR_LEN) == 1
    assert b"null plan" in buf.value
    assert L.cast_hip_sv_zero(0, 8, 1, buf, abi.ERR_LEN) == 1
    assert b"null device pointer" in buf.value
    assert L.cast_hip_norm_squared(0, 8, 1, None, buf, abi.ERR_LEN) == 1
    assert L.cast_hip_event_elapsed_ms(None, None, None, buf, abi.ERR_LEN) == 1
    assert L.cast_hip_gate_plan_info(None, None, buf, abi.ERR_LEN) == 1
    # batch with zero items is a no-op success
    assert L.cast_hip_gate_apply_batch(None, None, None, 0, None, buf, abi.ERR_LEN) == 0
    assert buf.value == b""  # cleared on success (err_buf.hpp:20-23)


def test_err_buf_is_truncated_and_terminated():
    buf = C.create_string_buffer(8)
    assert abi.lib().cast_hip_sv_zero(0, 8, 1, buf, 8) == 1
    assert len(buf.value) == 7


def test_destroy_functions_accept_null():
    L = abi.lib()
    L.cast_hip_gate_plan_destroy(None)
    L.cast_hip_stream_destroy(None)
    L.cast_hip_event_destroy(None)
    L.cast_hip_sv_free(0)


def test_measurement_entry_points_validate_before_the_device():
    buf = C.create_string_buffer(abi.ERR_LEN)
    L = abi.lib()
    out = (C.c_double * 4)()
    q = (C.c_uint32 * 2)(0, 1)
    assert L.cast_hip_marginal_probabilities(0, 8, 1, q, 2, out, None, buf, abi.ERR_LEN) == 1
    assert b"null statevector" in buf.value
    assert L.cast_hip_marginal_probabilities(64, 8, 3, q, 2, out, None, buf, abi.ERR_LEN) == 1
    assert b"precision" in buf.value
    assert L.cast_hip_marginal_probabilities(64, 41, 1, q, 2, out, None, buf, abi.ERR_LEN) == 1
    assert b"out of range" in buf.value
    assert L.cast_hip_marginal_probabilities(64, 1, 1, q, 2, out, None, buf, abi.ERR_LEN) == 1
    assert b"more measured qubits" in buf.value
    assert L.cast_hip_marginal_probabilities(64, 8, 1, None, 2, out, None, buf, abi.ERR_LEN) == 1
    assert b"null measured" in buf.value
    assert L.cast_hip_dm_populations(64, 21, 1, out, None, buf, abi.ERR_LEN) == 1
    assert b"1..20" in buf.value
    assert L.cast_hip_dm_populations(64, 4, 1, None, None, buf, abi.ERR_LEN) == 1
    assert b"null" in buf.value


def test_misaligned_statevector_pointers_are_rejected():
    buf = C.create_string_buffer(abi.ERR_LEN)
    L = abi.lib()
    assert L.cast_hip_scale(0x1008, 8, 1, 1.0, buf, abi.ERR_LEN) == 1
    assert b"16-byte aligned" in buf.value
    out = (C.c_double * 2)()
    q = (C.c_uint32 * 1)(0)
    assert L.cast_hip_marginal_probabilities(0x1004, 8, 1, q, 1, out, None, buf, abi.ERR_LEN) == 1
    assert b"16-byte aligned" in buf.value
    assert L.cast_hip_inner_product(0x1000, 1, 0x1008, 1, 4, out, None, buf, abi.ERR_LEN) == 1
    assert b"16-byte aligned" in buf.value
