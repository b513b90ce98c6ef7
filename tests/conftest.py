import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C ABI")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def asc():
    import autoscoper_amd as A

    A.lib()  # raises if libautoscoper_hip.so is missing: no fallback
    return A


@pytest.fixture(scope="session")
def orc():
    import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx(asc):
    if asc.device_count() < 1:
        pytest.fail("GPU test collected but no HIP device is visible")
    c = asc.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def tiny(asc, orc):
    """Tiny biplane scene (32x36x40 volume, 96x80 detectors) with targets
    rendered by the oracle."""
    from autoscoper_amd import synth

    wl = synth.workload("tiny")
    prep = orc.volume_prep(wl.volume, (wl.voxel_mm,) * 3)
    cams = [orc.camera_from_mayacam2(*c.mayacam2()) for c in wl.cameras]
    RW, RH = wl.render
    rads = []
    for i, cam in enumerate(cams):
        _, imv, _, _ = orc.view_geometry(cam, wl.gt_poses[0], prep["inv_scale"], prep["inv_trans"], RW, RH)
        fvp = synth.full_frame_viewport(cam[24:28], RW, RH)
        img, _ = orc.drr(prep["volume"], imv[:12], fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, RW, RH)
        rads.append(synth.target_from_drr(img, 70 + i))
    return dict(wl=wl, prep=prep, cams=cams, rads=rads, render=(RW, RH))
