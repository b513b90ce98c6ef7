"""The drop-in boundary against the reference's own code: libautoscoper's
Tracker.cpp and View.cpp (and the headers they include: Tracker.hpp,
View.hpp, Filter.hpp, VolumeDescription.hpp) compile against this backend
with nothing changed but the backend guards, and the backend's
facade_gpu.cpp compiles against those same edited headers, as in the
library build.

The guard edits (the integration recipe for a third backend) are applied
here, to temporary copies outside the repository:

* an OpenCL branch that only includes gpu/opencl headers (an include site)
  gets a HIP branch next to it that includes this backend's header:
  autoscoper_xromm_gpu_types.hpp in Filter.hpp, View.hpp and
  VolumeDescription.hpp (they need Buffer, GLBuffer and Image only, and
  the full header includes them), autoscoper_xromm_gpu.hpp elsewhere;
* every other `#if`/`#elif defined(Autoscoper_RENDERING_USE_OpenCL_BACKEND)`
  also admits Autoscoper_RENDERING_USE_HIP_BACKEND (the OpenCL flavour's
  compute code runs unchanged on this backend), except the two GL-interop
  branches of View.cpp that wrap a GL pixel-buffer object, which stay
  OpenCL-only.

g++ -fsyntax-only on the CPU; skips where the reference sources are absent.
"""
import os
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path(os.environ.get("ASC_REFERENCE_DIR", "/root/reference"))
SRC = REF / "libautoscoper" / "src"
OCL = "defined(Autoscoper_RENDERING_USE_OpenCL_BACKEND)"
HIP = "defined(Autoscoper_RENDERING_USE_HIP_BACKEND)"
FILES = ("Tracker.cpp", "Tracker.hpp", "View.cpp", "View.hpp", "Filter.hpp", "VolumeDescription.hpp")
TYPES_ONLY = ("View.hpp", "Filter.hpp", "VolumeDescription.hpp")

pytestmark = pytest.mark.skipif(not (SRC / "Tracker.cpp").exists(), reason="reference sources absent")


def guard_edits(text: str, header: str):
    """-> (edited text, counts of include sites given a HIP branch, branches
    widened to HIP, GL-interop branches left alone)."""
    lines = text.split("\n")
    out, n_inc, n_or, n_gl = [], 0, 0, 0
    i = 0
    while i < len(lines):
        line = lines[i]
        m = re.match(r"^(\s*)#(\s*)(if|elif)\s+" + re.escape(OCL) + r"\s*$", line)
        if not m:
            out.append(line)
            i += 1
            continue
        j = i + 1
        while not re.match(r"^\s*#\s*(elif|else|endif)\b", lines[j]):
            j += 1
        body = lines[i + 1:j]
        if body and all(re.match(r'^\s*#\s*include\s+"gpu/opencl/', b) for b in body):
            out += [line, *body, f"{m.group(1)}#{m.group(2)}elif {HIP}", f'{m.group(1)}#{m.group(2)}  include "{header}"']
            n_inc += 1
            i = j
            continue
        if any("GLBuffer(pbo" in b for b in body):
            out.append(line)
            n_gl += 1
        else:
            out.append(line.rstrip() + f" || {HIP}")
            n_or += 1
        i += 1
    return "\n".join(out), (n_inc, n_or, n_gl)


@pytest.fixture(scope="module")
def edited(tmp_path_factory):
    d = tmp_path_factory.mktemp("libautoscoper_hip_guards")
    counts = {}
    for f in FILES:
        hdr = "autoscoper_xromm_gpu_types.hpp" if f in TYPES_ONLY else "autoscoper_xromm_gpu.hpp"
        text, counts[f] = guard_edits((SRC / f).read_text(), hdr)
        (d / f).write_text(text)
    return d, counts


def _syntax(d, src):
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-DAutoscoper_RENDERING_USE_HIP_BACKEND",
           "-DAUTOSCOPER_XROMM_HOST_TYPES", f"-I{d}", f"-I{ROOT / 'include'}", f"-I{SRC}", f"-I{REF / 'math'}",
           str(src)]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_guard_edits_are_the_recipe(edited):
    _, counts = edited
    print({f: dict(zip(("include_sites", "widened", "gl_interop_kept"), c)) for f, c in counts.items()})
    # every file has its include site; only View.cpp has GL-interop branches
    for f in FILES:
        assert counts[f][0] == 1, f
    assert counts["View.cpp"][2] == 2 and all(counts[f][2] == 0 for f in FILES if f != "View.cpp")
    assert counts["Tracker.cpp"][1] >= 5 and counts["View.cpp"][1] >= 5


@pytest.mark.parametrize("src", ["Tracker.cpp", "View.cpp"])
def test_reference_source_compiles_on_the_hip_backend(edited, src):
    d, _ = edited
    r = _syntax(d, d / src)
    assert r.returncode == 0, r.stderr[-4000:]


def test_backend_compiles_inside_the_library(edited):
    d, _ = edited
    r = _syntax(d, ROOT / "autoscoper_amd" / "csrc" / "facade_gpu.cpp")
    assert r.returncode == 0, r.stderr[-4000:]
