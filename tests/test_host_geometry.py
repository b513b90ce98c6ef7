"""Host geometry of the product library (pose composition, camera set-up,
inverse model-view, projected bounding box, render size, float pose store)
against the oracle: bit-exact, since these decide integer results.
Runs on the CPU (host functions of the C ABI)."""
import numpy as np
import pytest

from autoscoper_amd import synth


def rand_poses(n, seed, trans=60.0, rot=170.0):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(-trans, trans, (n, 3)), rng.uniform(-rot, rot, (n, 3))], 1)


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def test_frames_bit_exact(asc, orc):
    g = asc.geometry
    for p in rand_poses(200, 1):
        f1, f2 = g.frame_from_xyzypr(p), orc.frame_from_xyzypr(p)
        assert (bits(f1) == bits(f2)).all()
        a1, a2 = g.frame_from_xyz_axis_angle(p), orc.frame_from_xyz_axis_angle(p)
        assert (bits(a1) == bits(a2)).all()
        assert (bits(g.frame_to_xyzypr(f1)) == bits(orc.frame_to_xyzypr(f2))).all()
        assert (bits(g.frame_mul(f1, a1)) == bits(orc.frame_mul(f2, a2))).all()
        assert (bits(g.frame_inverse(f1)) == bits(orc.frame_inverse(f2))).all()


def test_camera_from_mayacam2_bit_exact(asc, orc):
    for spec in synth.biplane(n=3) + synth.biplane(sod_mm=500, size=(512, 384), pixel_mm=0.7, n=2):
        K, R, t, size = spec.mayacam2()
        cam = asc.geometry.camera_from_mayacam2(K, R, t, size)
        ref = orc.camera_from_mayacam2(K, R, t, size)
        assert (bits(cam.to_array()) == bits(ref)).all()


@pytest.mark.parametrize("name", ["tiny", "cfg2"])
def test_view_geometry_bit_exact(asc, orc, name):
    if name == "tiny":
        shape, vox, size, sod, pix = (40, 36, 32), 1.0, (96, 80), 400.0, 0.9
    else:
        shape, vox, size, sod, pix = (320, 256, 256), 0.3, (1024, 1024), 800.0, 0.3
    D, H, W = shape
    # uncropped volume description (VolumeDescription.cpp:255-267)
    desc = asc.AscVolumeDesc()
    inv_scale = np.array([1 / np.float32(vox * W), 1 / np.float32(vox * H), 1 / np.float32(vox * D)], np.float32)
    inv_scale = np.array([np.float32(1.0) / np.float32(vox * n) for n in (W, H, D)], np.float32)
    inv_trans = np.array([-0.0 / W, -((H - (H - 1) - 1) / H), 0.0 / D], np.float32)
    for k in range(3):
        desc.inv_scale[k] = inv_scale[k]
        desc.inv_trans[k] = inv_trans[k]
    base = synth.volume_centre_pose(shape, vox)
    rng = np.random.default_rng(5)
    poses = base + np.concatenate([rng.uniform(-20, 20, (300, 3)), rng.uniform(-40, 40, (300, 3))], 1)
    # a few poses far outside the field of view and a NaN pose
    poses[0, :3] += 5000.0
    poses[1, 3:] = np.nan
    nview = 0
    for spec in synth.biplane(sod_mm=sod, size=size, pixel_mm=pix, sid_mm=sod * 1.375):
        cam = asc.geometry.camera_from_mayacam2(*spec.mayacam2())
        camr = cam.to_array()
        for p in poses:
            i1, imv1, vp1, sz1 = asc.geometry.view_geometry(cam, p, desc, size[0], size[1])
            i2, imv2, vp2, sz2 = orc.view_geometry(camr, p, inv_scale, inv_trans, size[0], size[1])
            assert i1 == i2
            assert (bits(imv1) == bits(imv2[:12])).all() or (np.isnan(imv1).all() and np.isnan(imv2[:12]).all())
            if i1:
                nview += 1
                assert (bits(vp1) == bits(vp2)).all()
                assert sz1 == sz2
    assert nview > 500


def test_pose_store_and_composition_bit_exact(asc, orc):
    pivot = orc.frame_from_xyzypr([3.0, -7.0, 11.0, 0, 0, 0])
    for p in rand_poses(200, 2, rot=80.0):
        s1, s2 = asc.geometry.pose_store_set(p), orc.pose_store_set(p)
        assert (bits(s1) == bits(s2)).all()
        assert (bits(asc.geometry.pose_store_get(s1)) == bits(orc.pose_store_get(s2))).all()
        m = p * 0.01
        assert (bits(asc.geometry.compose_pose(s1, pivot, m)) == bits(orc.compose_pose(s2, pivot, m))).all()
