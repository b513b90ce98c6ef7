"""Build recipe for libautoscoper_hip.so (gfx950) and the CPU oracle.

    python -m autoscoper_amd.build          # build both
    python -m autoscoper_amd.build --lib    # product library only

HIP sources are compiled by hipcc for gfx950 with:
  -ffp-contract=off                        no implicit FMA: expressions keep
                                           the reference's IEEE float order
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division and sqrt
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                           no packed fp32 instructions
Plain C++ host sources are compiled by g++ with -ffp-contract=off.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libautoscoper_hip.so"
FACADE_LIB = PKG / "libautoscoper_xromm.so"
ORACLE_DIR = ROOT / "oracle"

ARCH = os.environ.get("ASC_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "-fvisibility=hidden", "-Wall", "-Wno-unused-function"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    deps = [src] + sorted(CSRC.glob("*.hpp")) + [ROOT / "include" / "autoscoper_hip.h"]
    deps += sorted((ROOT / "include").glob("*.hpp"))
    if obj.exists() and all(obj.stat().st_mtime >= d.stat().st_mtime for d in deps):
        return obj
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXX_FLAGS, f"-I{ROOT / 'include'}", "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


FACADE_SRCS = ("facade_gpu.cpp", "facade.cpp")


def build_lib(force: bool = False) -> Path:
    """libautoscoper_hip.so (the C ABI: kernels, host geometry, optimisers)
    and libautoscoper_xromm.so (the standalone C++ facade on top of it).
    The facade is a library of its own so that libautoscoper can compile
    facade_gpu.cpp against its own Volume/Filter without two definitions of
    the xromm::gpu classes in one process."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    if force:
        for o in BUILD.glob("*.o"):
            o.unlink()
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    core = [o for s, o in zip(srcs, objs) if s.name not in FACADE_SRCS]
    facade = [o for s, o in zip(srcs, objs) if s.name in FACADE_SRCS]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, core)])
    os.replace(tmp, LIB)
    tmp = FACADE_LIB.with_suffix(".so.tmp")
    _run(["g++", "-shared", "-fPIC", "-o", str(tmp), *map(str, facade), f"-L{PKG}", "-lautoscoper_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FACADE_LIB)
    return LIB


FACADE_CHECK = ROOT / "tests" / "cpp" / "facade_check.bin"


def build_facade_check() -> Path:
    """C++ program that drives the xromm::gpu facade (used by a GPU test)."""
    src = ROOT / "tests" / "cpp" / "facade_check.cpp"
    _run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(src), "-o", str(FACADE_CHECK),
          f"-L{PKG}", "-lautoscoper_xromm", "-lautoscoper_hip", f"-Wl,-rpath,{PKG}",
          "-Wl,-rpath,$ORIGIN/../../autoscoper_amd"])
    return FACADE_CHECK


def build_oracle() -> Path:
    """The CPU oracle, and, where the reference sources are present, its
    host C++ for the pose math and optimisers into oracle/_ref."""
    _run(["make", "-s", "-C", str(ORACLE_DIR)])
    _run(["make", "-s", "-C", str(ORACLE_DIR), "ref"])
    return ORACLE_DIR / "liboracle.so"


def main(argv: list[str]) -> int:
    force = "--force" in argv
    if "--oracle" not in argv:
        print("built", build_lib(force=force))
        print("built", build_facade_check())
    if "--lib" not in argv:
        print("built", build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
