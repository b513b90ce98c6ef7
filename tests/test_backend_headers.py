"""The xromm::gpu backend (include/autoscoper_xromm_gpu.hpp,
autoscoper_amd/csrc/facade_gpu.cpp) compiles inside libautoscoper: against
the library's own, unchanged Volume.hpp and Filter.hpp
(AUTOSCOPER_XROMM_HOST_TYPES), which is how it replaces gpu/opencl there.
Also compiles standalone. Syntax-only g++ runs on the CPU; the check against
the library's headers needs the reference sources and skips without them.
"""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path(os.environ.get("ASC_REFERENCE_DIR", "/root/reference"))
SRC = ROOT / "autoscoper_amd" / "csrc" / "facade_gpu.cpp"


def _syntax(extra):
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", f"-I{ROOT / 'include'}", *extra, str(SRC)]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_backend_compiles_standalone():
    r = _syntax([])
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not (REF / "libautoscoper" / "src" / "Filter.hpp").exists(), reason="reference sources absent")
def test_backend_compiles_against_libautoscoper_headers():
    r = _syntax(["-DAUTOSCOPER_XROMM_HOST_TYPES", f"-I{REF / 'libautoscoper' / 'src'}", f"-I{REF / 'math'}"])
    assert r.returncode == 0, r.stderr
