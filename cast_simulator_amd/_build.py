"""In-tree build of the native parts.

* ``cast_simulator_amd/lib/libcast_hip.so`` — the gfx950 backend (HIP kernels
  + the C ABI of ``include/cast_hip.h``), compiled with hipcc for gfx950.
* ``cast_simulator_amd/_host.<ext>`` — the C++ host layer (gate types,
  circuit graph, fusion, cost model, profiler, kernel manager, simulator) with
  its Python bindings; it calls the backend only through the C ABI.

Objects go to ``_build/`` at the repository root.  Every call recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
BUILD = ROOT / "_build"
LIB_DIR = PKG / "lib"
HIP_LIB = LIB_DIR / "libcast_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 instructions are disabled for all device code; plain float code
# is scheduled as scalar VALU ops instead.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(
            "build step failed:\n  " + " ".join(cmd) + "\n" + proc.stdout + proc.stderr
        )


def _jobs() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def build_hip() -> Path:
    srcs = sorted((PKG / "csrc" / "hip").glob("*.hip"))
    out = BUILD / "hip"
    out.mkdir(parents=True, exist_ok=True)
    objs = [out / (s.stem + ".o") for s in srcs]
    with ThreadPoolExecutor(_jobs()) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", str(so[0]), "-o", str(so[1])]),
                    zip(srcs, objs)))
    LIB_DIR.mkdir(exist_ok=True)
    tmp = HIP_LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs)])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def host_ext_path() -> Path:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return PKG / ("_host" + suffix)


def build_host() -> Path:
    import pybind11

    py_inc = sysconfig.get_paths()["include"]
    # -fcx-limited-range: complex products as (ac - bd, ad + bc), the
    # reference's num-complex formula, instead of a call to the C99 Annex G
    # routine (same values for finite operands, several times faster).
    flags = [
        "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fcx-limited-range",
        "-fvisibility=hidden",
        f"-I{INCLUDE}", f"-I{pybind11.get_include()}", f"-I{py_inc}",
    ]
    srcs = sorted((PKG / "csrc" / "host").glob("*.cpp"))
    out = BUILD / "host"
    out.mkdir(parents=True, exist_ok=True)
    objs = [out / (s.stem + ".o") for s in srcs]
    with ThreadPoolExecutor(_jobs()) as ex:
        list(ex.map(lambda so: _run([CXX, *flags, "-c", str(so[0]), "-o", str(so[1])]),
                    zip(srcs, objs)))
    target = host_ext_path()
    tmp = target.with_name(target.name + ".tmp")
    _run([CXX, "-shared", "-o", str(tmp), *map(str, objs), f"-L{LIB_DIR}", "-lcast_hip",
          "-Wl,-rpath,$ORIGIN/lib"])
    os.replace(tmp, target)
    return target


def build_all() -> None:
    build_hip()
    build_host()
