"""cast_simulator_amd — MI355X (gfx950) gate-application backend for CAST.

The compute path is ``lib/libcast_hip.so`` (hand-written HIP kernels behind
the C ABI of ``include/cast_hip.h``); the host layer (gate types, circuit
graph, fusion, cost model, profiler, kernel manager, simulator) is the C++
extension ``_host``.  Both are built in-tree by ``__graft_entry__.build()``.
Importing this package without them raises ImportError: there is no
fallback path.
"""
from __future__ import annotations

from pathlib import Path as _Path

_PKG = _Path(__file__).resolve().parent
_LIB = _PKG / "lib" / "libcast_hip.so"

if not _LIB.exists():
    raise ImportError(
        f"{_LIB} is missing: build the gfx950 backend first "
        "(python -c 'import __graft_entry__ as g; g.build()')"
    )

try:
    from . import _host  # noqa: E402
except ImportError as e:  # pragma: no cover - exercised only on broken builds
    raise ImportError(
        "cast_simulator_amd._host is missing or failed to load; build it with "
        "__graft_entry__.build()"
    ) from e

from ._host import (  # noqa: E402,F401
    CastError,
    CircuitGraph,
    CostModel,
    ExploredBranch,
    FusionConfig,
    FusionDecision,
    FusionLog,
    HardwareAdaptiveCostModel,
    HardwareProfile,
    HipKernelGenSpec,
    HipKernelManager,
    HipStatevector,
    LayeredCircuit,
    NoiseSpec,
    Precision,
    ProfileConfig,
    QasmAngle,
    QasmCircuit,
    QasmGate,
    QuantumCircuit,
    QuantumGate,
    QuantumState,
    Representation,
    Simulator,
    ServeConfig,
    ServeRequest,
    ServeResult,
    Server,
    SizeOnlyCostModel,
    TrajectoryEngine,
    TrajectoryOpts,
    TrajectoryResult,
    TrajectoryRun,
    TrajectorySample,
    TrajectoryTuning,
    adaptive_sweep,
    apply_size_two_fusion,
    batch_sample,
    clean_branch,
    compute_bw_slope,
    expand_matrix,
    fit_from_samples,
    fit_roofline,
    haar_su2,
    measure_hip,
    optimize,
    optimize_with_log,
    parse_qasm,
    qasm_to_quantum_gate,
    poisson_requests,
    random_layer_circuit,
    random_statevector,
    random_unitary_matrix,
    sample_trajectories,
    superoperator_of,
    trajectory_circuit,
)

LIB_PATH = str(_LIB)
HOST_PATH = _host.__file__

__all__ = [n for n in dir() if not n.startswith("_")]
