"""Host sanitizer runs of the C++ host layer (CPU only; no GPU code is
instrumented).

    python -m cast_simulator_amd.sanitize address   # AddressSanitizer + UndefinedBehaviorSanitizer
    python -m cast_simulator_amd.sanitize thread    # ThreadSanitizer

Each builds the host extension ``_host`` with g++ and the sanitizer into
``_build/san-<kind>/cast_simulator_amd`` (a copy of the package's Python
files that loads the in-tree ``lib/libcast_hip.so``), then runs, with the
sanitizer runtime preloaded into Python:

* address: the CPU tests of the host layer (``pytest -m "not gpu"``, all
  files but test_abi.py, which drives the uninstrumented HIP library) and
  the batched trajectory planner on a 4096-trajectory 24-qubit batch (16
  threads);
* thread: the same planner run (the code that shares state between host
  threads).

The run fails (exit status != 0) on the first report.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
FLAGS = {
    "address": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
    "thread": ["-fsanitize=thread"],
}
RUNTIME = {"address": "libasan.so", "thread": "libtsan.so"}

PLAN = """
import cast_simulator_amd as cs
lc = cs.LayeredCircuit.random_layers(24, 40, 20260101)
S = cs.sample_trajectories(lc, cs.NoiseSpec(1e-3, 10, 40, 20260101), 4096)
eng = cs.TrajectoryEngine(lc, cs.Precision.F32, [q for i in range(23) for q in (i, i + 1)])
P = eng.plan(S, cs.TrajectoryTuning(4, 6, 512, 1, 0))
print("plan:", len(P["tasks"]), "tasks,", P["n_dirty_matrices"], "matrices with errors,",
      round(P["plan_s"], 2), "s", flush=True)
"""


def _run(cmd, **kw):
    print("+", " ".join(map(str, cmd)), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build(kind: str) -> Path:
    """Sanitized host extension in a package copy; returns the copy's parent."""
    import pybind11

    cxx = os.environ.get("CXX", "g++")
    dest_root = ROOT / "_build" / f"san-{kind}"
    dest = dest_root / "cast_simulator_amd"
    if dest.exists():
        shutil.rmtree(dest)
    dest.mkdir(parents=True)
    for f in PKG.glob("*.py"):
        shutil.copy2(f, dest / f.name)
    shutil.copytree(PKG / "profiles", dest / "profiles")
    (dest / "lib").symlink_to(PKG / "lib")
    inc = dest_root / "include"  # the package finds the C header beside itself
    if not inc.exists():
        inc.symlink_to(ROOT / "include")
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", *FLAGS[kind], "-std=c++17", "-fPIC",
             "-ffp-contract=off", "-fcx-limited-range", "-fvisibility=hidden",
             f"-I{ROOT / 'include'}", f"-I{pybind11.get_include()}",
             f"-I{sysconfig.get_paths()['include']}"]
    srcs = sorted((PKG / "csrc" / "host").glob("*.cpp"))
    objs = [dest_root / (s.stem + ".o") for s in srcs]
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(lambda so: _run([cxx, *flags, "-c", str(so[0]), "-o", str(so[1])]),
                    zip(srcs, objs)))
    ext = dest / ("_host" + (sysconfig.get_config_var("EXT_SUFFIX") or ".so"))
    _run([cxx, *FLAGS[kind], "-shared", "-o", str(ext), *map(str, objs),
          f"-L{PKG / 'lib'}", "-lcast_hip", f"-Wl,-rpath,{PKG / 'lib'}"])
    return dest_root


def run(kind: str) -> int:
    dest_root = build(kind)
    cxx = os.environ.get("CXX", "g++")
    def lib(name):
        return subprocess.run([cxx, f"-print-file-name={name}"], capture_output=True, text=True,
                              check=True).stdout.strip()

    env = dict(os.environ)
    # The sanitizer runtime first; libstdc++ right after it, so that the
    # runtime's exception interceptors find the real __cxa_throw (Python
    # itself does not load libstdc++).
    env["LD_PRELOAD"] = " ".join([lib(RUNTIME[kind]), lib("libstdc++.so.6")])
    env["PYTHONPATH"] = os.pathsep.join([str(dest_root), str(ROOT)])
    # Python's own allocations are not instrumented: leak checks would report
    # the interpreter, not this library.
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:report_signal_unsafe=0"
    check = "import cast_simulator_amd as cs; assert cs.HOST_PATH.startswith(%r), cs.HOST_PATH" % (
        str(dest_root))
    # Run from the copy's folder: Python puts the working directory first on
    # its path, ahead of the repository (which provides tests and oracle).
    cwd = str(dest_root)
    _run([sys.executable, "-c", check], env=env, cwd=cwd)
    if kind == "address":
        # The host-layer tests; test_abi.py exercises the uninstrumented HIP
        # library and runtime, which are not what this run checks.
        tests = [str(p) for p in sorted((ROOT / "tests").glob("test_*.py"))
                 if p.name != "test_abi.py"]
        _run([sys.executable, "-m", "pytest", *tests, "-m", "not gpu", "-x", "-q",
              "-p", "no:cacheprovider"], env=env, cwd=cwd)
    _run([sys.executable, "-c", PLAN], env=env, cwd=cwd)
    print(f"sanitize {kind}: no reports", flush=True)
    return 0


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1 or argv[0] not in FLAGS:
        print(__doc__)
        return 2
    try:
        return run(argv[0])
    except subprocess.CalledProcessError as e:
        print(f"sanitize {argv[0]}: failed ({e.returncode})", flush=True)
        return 1


if __name__ == "__main__":
    sys.exit(main())
