"""Pins for the CPU oracle: each check is independent of the oracle's own
formulas (closed forms, scipy/numpy references, invariants, published
constants), chosen so that a plausible slip -- a half-texel offset, a
flipped axis, clamp-to-edge instead of the border colour, a transposed
rotation, radius vs radius^2 in the Gaussian, a wrong mean in the NCC --
fails at least one of them. CPU only."""
import math

import numpy as np
import pytest
from scipy import ndimage
from scipy.spatial.transform import Rotation as Rot

from autoscoper_amd import synth


# --------------------------------------------------------------------------
# a5: sampler (OpenCL 1.2 sec. 8.2, normalised coords, CLK_ADDRESS_CLAMP)
# --------------------------------------------------------------------------
def ramp_volume(W=9, H=7, D=5):
    k, j, i = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    return (1000 + 37 * i + 11 * j + 5 * k).astype(np.uint16)


def test_sampler_reproduces_linear_ramp_with_half_texel_offset(orc):
    vol = ramp_volume()
    D, H, W = vol.shape
    rng = np.random.default_rng(0)
    for _ in range(200):
        # interior texel coordinates X in [0, W-1] etc.
        X, Y, Z = rng.uniform(0, W - 1), rng.uniform(0, H - 1), rng.uniform(0, D - 1)
        s, t, r = (X + 0.5) / W, (Y + 0.5) / H, (Z + 0.5) / D
        expect = (1000 + 37 * X + 11 * Y + 5 * Z) / 65535.0
        assert abs(orc.sample3(vol, s, t, r, fp64=True) - expect) <= 1e-12
        assert abs(orc.sample3(vol, s, t, r) - expect) <= 2e-6 * expect


def test_sampler_border_colour_is_zero(orc):
    vol = ramp_volume()
    D, H, W = vol.shape
    # X in [-0.5, 0): weight (X+1) on texel 0, the rest on the zero border
    for X in (-0.5, -0.3, -0.01):
        Y, Z = 3.0, 2.0
        s, t, r = (X + 0.5) / W, (Y + 0.5) / H, (Z + 0.5) / D
        t0 = (1000 + 11 * Y + 5 * Z) / 65535.0
        assert abs(orc.sample3(vol, s, t, r, fp64=True) - (X + 1) * t0) <= 1e-12
    # beyond the last texel centre on the far side
    X = W - 1 + 0.25
    s = (X + 0.5) / W
    tl = (1000 + 37 * (W - 1) + 11 * 3 + 5 * 2) / 65535.0
    assert abs(orc.sample3(vol, s, (3.5) / H, (2.5) / D, fp64=True) - 0.75 * tl) <= 1e-12
    # fully outside -> 0
    assert orc.sample3(vol, -0.5, 0.5, 0.5) == 0.0


# --------------------------------------------------------------------------
# a2/a3/a5 end to end: DRR of a ball against analytic chord lengths
# --------------------------------------------------------------------------
def _ray_world(cam_rec, vp, w, h, x, y):
    """World-space ray of DRR pixel (x, y) from the camera record, computed
    with numpy: film point (u, v, -2) in camera space."""
    Rc = np.asarray(cam_rec[:9]).reshape(3, 3).T  # columns: camera axes in world
    c = np.asarray(cam_rec[9:12])
    u = vp[0] + vp[2] * (x / w)
    v = vp[1] + vp[3] * (y / h)
    d = Rc @ np.array([u, v, -2.0])
    return c, d / np.linalg.norm(d)


def test_drr_of_smooth_ball_matches_line_integrals(orc):
    """Volume value C*(1 - d^2/r^2) inside a ball of radius r (0 outside):
    its line integral at distance d from the centre is
    C*(4/3)*(r^2 - d^2)^(3/2)/r^2. Rays come from the camera model,
    computed independently here with numpy/scipy."""
    vox, n = 0.5, 48
    C_u16 = 60000
    zz, yy, xx = np.meshgrid(*(np.arange(n) + 0.5,) * 3, indexing="ij")
    rad_vox = 15.0
    ctr = np.array([n / 2, n / 2 + 1.5, n / 2 - 2.0])  # (x, y, z) in voxels, off-centre
    q2 = ((xx - ctr[0]) ** 2 + (yy - ctr[1]) ** 2 + (zz - ctr[2]) ** 2) / rad_vox ** 2
    vol = np.rint(C_u16 * np.clip(1 - q2, 0, None)).astype(np.uint16)
    vol[0, 0, 0] = vol[-1, -1, -1] = 1  # keep the box from being cropped
    prep = orc.volume_prep(vol, (vox,) * 3)
    assert list(prep["dims"]) == [n, n, n]
    spec = synth.CameraSpec(centre=(300.0, 40.0, 25.0), size=(64, 64), pixel_mm=1.0, sid_mm=420.0)
    cam = orc.camera_from_mayacam2(*spec.mayacam2())
    pose = np.array([-12.0, -12.0, 12.0, 8.0, -5.0, 11.0])  # volume pose (xyzypr)
    inview, imv, vp, (w, h) = orc.view_geometry(cam, pose, prep["inv_scale"], prep["inv_trans"], 64, 64)
    assert inview and w > 20 and h > 20
    step = 0.25
    img, _ = orc.drr(prep["volume"], imv[:12], vp, step, 1e6, 0.0, w, h, fp64=True)
    # ball centre: voxel-centre coordinates (x, y, z) -> volume mm (x, H - y, -z) -> pose
    H = n
    p_vol = np.array([ctr[0] * vox, (H - ctr[1]) * vox, -ctr[2] * vox])
    Rp = Rot.from_euler("ZYX", pose[3:], degrees=True).as_matrix()
    p_w = Rp @ p_vol + pose[:3]
    r_mm = rad_vox * vox
    errs = []
    for y in range(0, h, 2):
        for x in range(0, w, 2):
            o, d = _ray_world(cam, vp, w, h, x, y)
            q = p_w - o
            b = q @ d
            dist2 = q @ q - b * b
            if dist2 < (0.8 * r_mm) ** 2:
                expect = C_u16 / 65535.0 * (4.0 / 3.0) * (r_mm ** 2 - dist2) ** 1.5 / r_mm ** 2
                errs.append(abs(img[y, x] * 1e6 - expect) / expect)
    assert len(errs) > 50
    assert max(errs) < 0.01, max(errs)
    # a ray that misses the box is exactly 0
    far_vp = np.array([5.0, 5.0, 0.01, 0.01], np.float32)
    img2, n2 = orc.drr(prep["volume"], imv[:12], far_vp, step, 1.0, 0.0, 4, 4)
    assert n2 == 0 and (img2 == 0).all()


def test_sample_count_closed_form(orc):
    """N = #{k >= 0 : far - k > near} = ceil(far - near) for the slab
    distances in step units, computed here in double."""
    vox, n = 1.0, 16
    vol = np.full((n, n, n), 100, np.uint16)
    prep = orc.volume_prep(vol, (vox,) * 3)
    spec = synth.CameraSpec(centre=(120.0, 10.0, 5.0), size=(32, 32), pixel_mm=1.0, sid_mm=170.0)
    cam = orc.camera_from_mayacam2(*spec.mayacam2())
    pose = np.array([-8.0, -8.0, 8.0, 3.0, 2.0, -4.0])
    _, imv, vp, (w, h) = orc.view_geometry(cam, pose, prep["inv_scale"], prep["inv_trans"], 32, 32)
    step = 0.7
    total = 0
    imv64 = imv[:12].astype(np.float64)
    for y in range(h):
        for x in range(w):
            u = vp[0] + vp[2] * (x / w)
            v = vp[1] + vp[3] * (y / h)
            look = step * np.array([u, v, -2.0]) / math.sqrt(u * u + v * v + 4)
            o = imv64[[3, 7, 11]]
            d = imv64.reshape(3, 4)[:, :3] @ look
            with np.errstate(divide="ignore"):
                t0 = (np.array([0.0, 0.0, -1.0]) - o) / d
                t1 = (np.array([1.0, 1.0, 0.0]) - o) / d
            near = np.max(np.minimum(t0, t1))
            far = np.min(np.maximum(t0, t1))
            if far > near:
                near = max(near, 0.0)
                total += math.ceil(far - near)
    _, n_orc = orc.drr(prep["volume"], imv[:12], vp, step, 1.0, 0.0, w, h)
    # float slab distances may move a count by one at exact boundaries
    assert abs(n_orc - total) <= max(2, total // 5000)


# --------------------------------------------------------------------------
# a8-a11: filters against scipy in FP64
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def img():
    rng = np.random.default_rng(4)
    a = ndimage.gaussian_filter(rng.random((37, 53)), 1.5)
    a = (a - a.min()) / (a.max() - a.min())
    a[:4, :6] = 0
    return a.astype(np.float32)


def test_sobel_vs_scipy(orc, img):
    x = img.astype(np.float64)
    kx = np.array([[-1, 0, 1], [-2, 0, 2], [-1, 0, 1]], float)
    ky = np.array([[1, 2, 1], [0, 0, 0], [-1, -2, -1]], float)
    gh = ndimage.correlate(x, kx, mode="nearest")
    gv = ndimage.correlate(x, ky, mode="nearest")
    g = np.hypot(gh, gv)
    for scale, blend in ((1.0, 0.5), (1.7, 0.1), (3.0, 0.43), (0.8, 0.75)):
        ref = scale * g + 2 * blend * x if blend < 0.5 else 2 * (1 - blend) * scale * g + x
        ref = np.clip(ref, 0, 1)
        assert np.abs(orc.sobel(img, scale, blend) - ref).max() <= 2e-6


def test_contrast_vs_scipy(orc, img):
    x = img.astype(np.float64)
    ones = np.ones_like(x)
    s = ndimage.correlate(x, np.ones((3, 3)), mode="constant", cval=0.0)
    cnt = ndimage.correlate(ones, np.ones((3, 3)), mode="constant", cval=0.0)
    a = s / cnt
    for alpha, beta in ((1.0, 1.0), (2.5, 0.9), (1.5, 2.0)):
        with np.errstate(divide="ignore", invalid="ignore"):
            ref = np.where(a > 0.01, a ** (alpha - beta) * x ** beta, 0.0)
        got = orc.contrast(img, alpha, beta, 3)
        assert np.allclose(got, ref, rtol=3e-6, atol=1e-7)


def gauss2d(radius):
    R = int(3 * radius)
    d = np.arange(-R, R + 1)
    w = np.exp(-(d[:, None] ** 2 + d[None, :] ** 2) / (2.0 * radius))  # variance = radius
    return w / w.sum()


def test_gaussian_weights_and_filter_vs_scipy(orc, img):
    x = img.astype(np.float64)
    for radius in (0.5, 1.0, 1.5, 2.0):
        w = gauss2d(radius)
        assert orc.gaussian_weights(radius).shape == w.shape
        assert np.abs(orc.gaussian_weights(radius) - w).max() <= 1e-7
        ref = np.clip(ndimage.correlate(x, w, mode="constant", cval=0.0), 0, 1)
        assert np.abs(orc.gaussian(img, radius) - ref).max() <= 2e-6
    assert np.array_equal(orc.gaussian(img, 0.2), img)  # N == 1 -> copy


def test_sharpen_vs_scipy(orc, img):
    x = img.astype(np.float64)
    for radius, c, thr in ((1.0, 1.0, 0.0), (1.0, 2.5, 0.02), (1.5, 1.7, 0.0)):
        blur = ndimage.correlate(x, gauss2d(radius), mode="constant", cval=0.0)
        ref = np.where(np.abs(x - blur) > thr, np.clip(x + c * (x - blur), 0, 1), x)
        got = orc.sharpen(img, radius, c, thr)
        # pixels within float noise of the threshold may go either way
        close = np.abs(np.abs(x - blur) - thr) > 1e-6
        assert np.abs(got - ref)[close].max() <= 2e-6


# --------------------------------------------------------------------------
# a14: NCC
# --------------------------------------------------------------------------
def test_ncc_is_masked_pearson(orc):
    rng = np.random.default_rng(1)
    for n in (5, 100, 1000, 12345):
        f = rng.random(n)
        g = 0.4 * f + rng.random(n)
        m = (rng.random(n) < 0.7).astype(float)
        m[0] = 1
        v, mom = orc.ncc_exact(f * m, g * m, m)
        sel = m > 0.5
        assert abs(v - np.corrcoef(f[sel], g[sel])[0, 1]) <= 1e-12
        assert mom[6] == sel.sum()
    f = rng.random(50)
    assert abs(orc.ncc_exact(f, 3 * f + 2, np.ones(50))[0] - 1.0) <= 1e-12
    assert abs(orc.ncc_exact(f, -f, np.ones(50))[0] + 1.0) <= 1e-12
    assert orc.ncc_exact(np.full(50, 0.3), f, np.ones(50))[0] == 1e5


def test_tree_sum_models(orc):
    rng = np.random.default_rng(2)
    # power-of-two sizes: both index models sum everything
    for n in (1, 2, 128, 256, 4096, 65536):
        x = rng.random(n).astype(np.float32)
        exact = math.fsum(x.astype(np.float64))
        assert abs(orc.tree_sum(x, 0) - exact) <= 1e-5 * max(exact, 1)
    # the OpenCL index (t + b*B) covers n <= 128 and (16256, 16384] exactly
    x = np.ones(16384, np.float32)
    assert orc.tree_sum(x, 1) == 16384.0
    # ... but not a 256 x 256 image: it is then not a plain sum
    x = np.ones(65536, np.float32)
    assert orc.tree_sum(x, 1) != 65536.0
    # non-power-of-two group sizes drop elements in the LDS tree (n = 100 -> T = 100)
    e = np.zeros(100, np.float32)
    e[24] = 1.0
    assert orc.tree_sum(e, 0) == 0.0


# --------------------------------------------------------------------------
# a2/a16: rigid frames, Euler angles, axis-angle, quaternions
# --------------------------------------------------------------------------
def M(f):
    return np.asarray(f[:9]).reshape(3, 3).T


def test_frames_against_scipy(orc):
    rng = np.random.default_rng(3)
    for _ in range(50):
        p = np.concatenate([rng.uniform(-50, 50, 3), rng.uniform(-80, 80, 3)])
        f = orc.frame_from_xyzypr(p)
        # yaw about z, then pitch about the new y, then roll about the new x
        assert np.allclose(M(f), Rot.from_euler("ZYX", p[3:], degrees=True).as_matrix(), atol=1e-12)
        assert np.allclose(f[9:], p[:3])
        assert np.allclose(orc.frame_to_xyzypr(f), p, atol=1e-9)
        a = orc.frame_from_xyz_axis_angle(p)
        ang = np.linalg.norm(p[3:])
        assert np.allclose(M(a), Rot.from_rotvec(np.deg2rad(ang) * p[3:] / ang).as_matrix(), atol=1e-12)
        fi = orc.frame_inverse(f)
        prod = orc.frame_mul(f, fi)
        assert np.allclose(M(prod), np.eye(3), atol=1e-12) and np.allclose(prod[9:], 0, atol=1e-9)
        # composition acts on points as M_a (M_b p + t_b) + t_a
        q = rng.standard_normal(3)
        ab = orc.frame_mul(f, a)
        assert np.allclose(M(ab) @ q + ab[9:], M(f) @ (M(a) @ q + a[9:]) + f[9:])


def test_quaternion_store(orc):
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = np.concatenate([rng.uniform(-50, 50, 3), rng.uniform(-80, 80, 3)])
        s = orc.pose_store_set(p)
        # Quat(yaw, pitch, roll): the ZYX rotation, scalar first
        q = Rot.from_euler("ZYX", p[3:], degrees=True).as_quat()  # x, y, z, w
        ref = np.array([q[3], q[0], q[1], q[2]])
        assert np.allclose(s[3:], ref, atol=2e-6) or np.allclose(s[3:], -ref, atol=2e-6)
        back = orc.pose_store_get(s)
        assert np.allclose(back[:3], np.float32(p[:3]))
        assert np.allclose(back[3:], p[3:], atol=5e-3)


def test_camera_from_mayacam2(orc):
    spec = synth.CameraSpec(centre=(700.0, 150.0, -40.0), size=(1024, 768), pixel_mm=0.3, sid_mm=1100.0,
                            principal=(500.0, 400.0))
    K, R, t, size = spec.mayacam2()
    rec = orc.camera_from_mayacam2(K, R, t, size)
    # camera centre is -R^T t (pinhole identity)
    assert np.allclose(rec[9:12], -R.T @ t)
    # camera axes in world: x right, y up (= -y of the file's camera), z backwards
    assert np.allclose(M(rec), R.T @ np.diag([1.0, -1.0, -1.0]))
    f = K[0, 0]
    assert np.isclose(rec[26], 2 * size[0] / f) and np.isclose(rec[27], 2 * size[1] / f)
    # principal point: x as is, y flipped to (H-1) - cy
    assert np.isclose(rec[24], -(2 * 500.0 - size[0]) / f)
    assert np.isclose(rec[25], -(2 * ((size[1] - 1) - 400.0) - size[1]) / f)


def test_viewport_contains_projected_box(orc):
    """The bounding box equals min/max of the 8 projected box corners
    (film plane z = -2), clipped to the radiograph."""
    shape, vox = (40, 36, 32), 1.0
    vol = np.full(shape, 7, np.uint16)
    prep = orc.volume_prep(vol, (vox,) * 3)
    spec = synth.CameraSpec(centre=(400.0, 30.0, 10.0), size=(96, 80), pixel_mm=0.9, sid_mm=550.0)
    cam = orc.camera_from_mayacam2(*spec.mayacam2())
    D, H, W = shape
    corners = np.array([[x, y, z] for x in (0, W) for y in (0, H) for z in (0, -D)], float) * vox
    rng = np.random.default_rng(8)
    for _ in range(40):
        pose = synth.volume_centre_pose(shape, vox) + np.concatenate([rng.uniform(-3, 3, 3), rng.uniform(-20, 20, 3)])
        inview, _, vp, (w, h) = orc.view_geometry(cam, pose, prep["inv_scale"], prep["inv_trans"], 96, 80)
        assert inview
        Rp = Rot.from_euler("ZYX", pose[3:], degrees=True).as_matrix()
        world = corners @ Rp.T + pose[:3]
        Rc = M(cam)
        camc = (world - cam[9:12]) @ Rc  # world -> camera coordinates
        film = -2 * camc[:, :2] / camc[:, 2:3]
        lo = np.minimum(film.min(0), 1.0)
        hi = np.maximum(film.max(0), -1.0)
        # radiograph film rectangle from the camera's viewport record
        r0 = cam[24:26] - cam[26:28] / 2
        r1 = cam[24:26] + cam[26:28] / 2
        lo, hi = np.maximum(lo, r0), np.minimum(hi, r1)
        assert np.allclose(vp[:2], lo, atol=1e-9) and np.allclose(vp[2:], hi - lo, atol=1e-9)
        assert (w, h) == (int(vp[2] * 96 / cam[26]), int(vp[3] * 80 / cam[27]))


# --------------------------------------------------------------------------
# a1: volume preparation
# --------------------------------------------------------------------------
def test_volume_prep_brute_force(orc):
    rng = np.random.default_rng(6)
    for flip in [(0, 0, 0), (1, 1, 0), (0, 1, 1)]:
        vol = np.zeros((10, 12, 14), np.uint16)
        vol[2:8, 1:9, 3:12] = rng.integers(0, 50000, (6, 8, 9))
        fl = vol[::-1] if flip[2] else vol
        fl = fl[:, ::-1] if flip[1] else fl
        fl = fl[:, :, ::-1] if flip[0] else fl
        z, y, x = np.nonzero(fl)
        out = orc.volume_prep(vol, (0.5, 0.6, 0.7), flip)
        assert list(out["crop_min"]) == [x.min(), y.min(), z.min()]
        assert list(out["crop_max"]) == [x.max(), y.max(), z.max()]
        crop = fl[z.min():z.max() + 1, y.min():y.max() + 1, x.min():x.max() + 1]
        assert np.array_equal(out["volume"], crop)
        assert np.float32(out["min_value"]) == np.float32(crop.min() / np.float32(65535))
        assert np.float32(out["max_value"]) == np.float32(crop.max() / np.float32(65535))
        # the box [0,1]x[0,1]x[-1,0] maps back to the cropped block's mm extent
        dims = out["dims"]
        assert np.allclose(1 / out["inv_scale"], np.array([0.5, 0.6, 0.7]) * dims, rtol=1e-6)
    with pytest.raises(ValueError):
        orc.volume_prep(np.zeros((3, 3, 3), np.uint16), (1, 1, 1))


# --------------------------------------------------------------------------
# a17/a18: optimisers
# --------------------------------------------------------------------------
def test_glibc_rand_known_values(orc):
    vals, rmax = orc.glibc_rand(1, 5)
    assert rmax == 2147483647
    # srand(1) / rand() under glibc (published reference values)
    assert list(vals) == [1804289383, 846930886, 1681692777, 1714636915, 1957747793]


def test_amoeba_sinc_minimum(orc):
    # Numerical Recipes sample: minimum of sin(R)/R is -0.2172336 (tan R = R)
    P, Y, it = orc.amoeba_sinc(np.array([4.0, 1.0, 5.0, 1.0, 4.0, 2.0]), 1e-8)
    assert abs(Y.min() - (-0.21723362821122166)) <= 1e-9
    r = np.hypot(P[np.argmin(Y), 0], P[np.argmin(Y), 1])
    assert abs(math.tan(r) - r) <= 1e-3 * r
    assert 0 < it <= 1000


def test_pso_finds_bowl_centre(orc):
    centre = np.array([0.5, -1.0, 1.5, -0.7, 0.3, 2.0])
    best, bn, epochs = orc.pso_bowl(centre, 7, -3.0, 3.0, 200, 30)
    assert bn < 1e-3
    assert np.abs(best - centre).max() < 0.05
    assert epochs <= 201


# --------------------------------------------------------------------------
# a6/a12: radiograph resample and background mask on a ramp, closed form
# --------------------------------------------------------------------------
def _ramp_image(IW, IH, kind):
    j, i = np.meshgrid(np.arange(IH), np.arange(IW), indexing="ij")
    if kind == "u16":
        return (1000 + 370 * i + 510 * j).astype(np.uint16), lambda X, Y: (1000 + 370 * X + 510 * Y) / 65535.0
    return (0.1 + 0.3 * i / IW + 0.5 * j / IH).astype(np.float32), lambda X, Y: 0.1 + 0.3 * X / IW + 0.5 * Y / IH


@pytest.mark.parametrize("kind", ["f32", "u16"])
@pytest.mark.parametrize("shift", [0.0, 0.3])
def test_rad_resample_of_ramp_closed_form(orc, kind, shift):
    """RadRenderer.cl:19-29 with the viewport equal to the image plane
    (shifted by `shift` pixels): pixel (x, y) samples texel coordinates
    X = x + shift - 1/2 and Y = IH - y - 1/2 (the half-texel offset of the
    sampler and the t = 1/2 - (v - v0)/v1 flip). A linear ramp is reproduced
    exactly at interior X, Y; within one texel of the image edge it falls
    off linearly to the border colour 0 at the first texel centre outside;
    beyond the image plane the output is 0. The resample returns 1 - value."""
    IW, IH = 64, 48
    img, ramp = _ramp_image(IW, IH, kind)
    u0, v0, u1, v1 = 0.13, -0.07, 0.8, 0.6
    plane = np.array([u0, v0, u1, v1], np.float32)
    vp = np.array([u0 - u1 / 2 + shift * u1 / IW, v0 - v1 / 2, u1, v1], np.float32)
    out = orc.rad(img, plane, vp, IW, IH)
    y, x = np.meshgrid(np.arange(IH), np.arange(IW), indexing="ij")
    X, Y = x + shift - 0.5, IH - y - 0.5
    s = (x + shift) / IW
    val = ramp(np.clip(X, 0, IW - 1), np.clip(Y, 0, IH - 1))
    # border falloff: weight (X + 1) below 0, (IW - X) above IW - 1 (same in Y)
    val = val * np.where(X < 0, X + 1, 1.0) * np.where(X > IW - 1, IW - X, 1.0)
    val = val * np.where(Y < 0, Y + 1, 1.0) * np.where(Y > IH - 1, IH - Y, 1.0)
    expect = np.where(s > 1.0, 0.0, 1.0 - val)
    assert np.abs(out - expect).max() <= 2e-6
    # the first row samples the top texel row of the image (the t flip):
    # a missing flip would read row 0, which is 0.5 below it on this ramp
    assert abs(out[1, IW // 2] - (1 - ramp(IW // 2 + shift - 0.5, IH - 1.5))) <= 2e-6


def test_rad_resample_outside_plane_is_zero(orc):
    IW, IH = 32, 24
    img, _ = _ramp_image(IW, IH, "f32")
    plane = np.array([0.0, 0.0, 1.0, 1.0], np.float32)
    # viewport twice the plane: a quarter-width margin on each side is outside
    vp = np.array([-1.0, -1.0, 2.0, 2.0], np.float32)
    out = orc.rad(img, plane, vp, 2 * IW, 2 * IH)
    u = -1.0 + 2.0 * np.arange(2 * IW) / (2 * IW)
    v = -1.0 + 2.0 * np.arange(2 * IH) / (2 * IH)
    outside = (np.abs(v)[:, None] > 0.5) | (np.abs(u)[None, :] > 0.5)
    assert (out[outside] == 0).all()
    assert (out[~outside] > 0).all()


def test_background_mask_of_ramp_closed_form(orc):
    """BackgroundRenderer.cl: 1 where threshold <= sampled value, else 0,
    with the same sampler rule as the resample."""
    IW, IH = 64, 48
    img, ramp = _ramp_image(IW, IH, "f32")
    u0, v0, u1, v1 = 0.0, 0.0, 1.0, 1.0
    plane = np.array([u0, v0, u1, v1], np.float32)
    vp = np.array([-0.5, -0.5, 1.0, 1.0], np.float32)
    thr = 0.45
    out = orc.background(img, plane, vp, thr, IW, IH)
    y, x = np.meshgrid(np.arange(1, IH), np.arange(1, IW), indexing="ij")
    val = ramp(x - 0.5, IH - y - 0.5)
    m = out[1:, 1:]
    clear = np.abs(val - thr) > 1e-6
    assert np.array_equal(m[clear], (val[clear] >= thr).astype(np.float32))
    assert 0.2 < m.mean() < 0.8


# --------------------------------------------------------------------------
# a19/a20: Compositor and Merger on hand-computed channels
# --------------------------------------------------------------------------
def test_composite_and_merge_hand_values(orc):
    """Compositor.cl:2-24: R = src1, G = src1/2 + src2/2 where src3 >= 0.5
    else 0, B = src2. Merger.cl: min(a + b, 1)."""
    s1 = np.array([0.25, 0.75, 1.0, 0.0, 0.5], np.float32)
    s2 = np.array([0.5, 0.25, 0.0, 1.0, 0.5], np.float32)
    s3 = np.array([0.5, 0.4999, 1.0, 0.0, 0.75], np.float32)
    rgb = orc.composite(s1, s2, s3)
    assert rgb.tolist() == [[0.25, 0.375, 0.5], [0.75, 0.0, 0.25], [1.0, 0.5, 0.0], [0.0, 0.0, 1.0],
                            [0.5, 0.5, 0.5]]
    assert orc.merge(s1, s2).tolist() == [0.75, 1.0, 1.0, 1.0, 1.0]


# --------------------------------------------------------------------------
# a5 + a6 + a15 + a18 closed loop: tracking recovers the hidden pose
# --------------------------------------------------------------------------
def _bead_scene(orc, shift_px=0.0, filters=None):
    """Seven smooth beads (u16, zero background) in a 48^3 volume at 0.5 mm,
    two 160^2 cameras 90 degrees apart. The target of each camera is
    1 - (the all-FP64 full-frame DRR at the hidden pose), rows flipped to
    the radiograph's order, so that with no filters the cost is 1 - NCC of
    a DRR and its own resampled rendering. `shift_px` moves the rendering
    grid of the target by that fraction of a pixel (a planted half-texel
    mistake)."""
    N, vox = 48, 0.5
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    c = (N - 1) / 2
    rng = np.random.default_rng(5)
    vol = np.zeros((N, N, N))
    for _ in range(7):
        p = rng.uniform(-0.3, 0.3, 3) * N + c
        s = rng.uniform(1.5, 2.5) / vox
        vol += rng.uniform(0.5, 1.0) * np.exp(-((i - p[0]) ** 2 + (j - p[1]) ** 2 + (k - p[2]) ** 2) / (2 * s * s))
    vol = np.rint(vol / vol.max() * 30000).astype(np.uint16)
    prep = orc.volume_prep(vol, (vox,) * 3)
    cams = [orc.camera_from_mayacam2(*s.mayacam2())
            for s in synth.biplane(sod_mm=400.0, size=(160, 160), pixel_mm=0.25, sid_mm=550.0)]
    RW = RH = 160
    step, inten = synth.DRR_STEP, 6.0
    gt = synth.volume_centre_pose(vol.shape, vox) + np.array([0.3, -0.2, 0.1, 3.0, -2.0, 4.0])
    rads = []
    for cam in cams:
        _, imv, _, _ = orc.view_geometry(cam, gt, prep["inv_scale"], prep["inv_trans"], RW, RH)
        fvp = synth.full_frame_viewport(cam[24:28], RW, RH).astype(np.float64)
        fvp[0] += shift_px * fvp[2] / RW
        img, _ = orc.drr(prep["volume"], imv[:12], fvp, step, inten, 0.0, RW, RH, fp64=True)
        assert 0.3 < img.max() < 1.0  # not clamped
        rads.append((1.0 - np.flipud(img)).astype(np.float32))
    recs = filters if filters is not None else np.zeros((0, 6), np.float32)
    S = orc.Scene(prep, (step, inten, 0.0), cams, rads, [(recs, recs)] * 2, (RW, RH))
    pivot = orc.frame_from_xyzypr(list(prep["trans_center"]) + [0, 0, 0])
    return S, gt, pivot


def _pose_gap(orc, a, b):
    fa, fb = orc.frame_from_xyzypr(a), orc.frame_from_xyzypr(b)
    Ra, Rb = fa[:9].reshape(3, 3), fb[:9].reshape(3, 3)
    cosang = (np.trace(Ra.T @ Rb) - 1.0) / 2.0
    return float(np.linalg.norm(fa[9:] - fb[9:])), float(np.degrees(np.arccos(np.clip(cosang, -1.0, 1.0))))


START = np.array([1.5, -1.2, 0.8, 2.0, -1.5, 1.0])  # ~2 mm / 2 deg off the hidden pose


def test_tracking_recovers_hidden_pose(orc):
    """Tracker::optimize's downhill simplex (DownhillSimplex.cpp, Tracker.cpp:
    476-570) on the oracle's trackFrame, started 2 mm / 2 deg off, lands on
    the pose the target was rendered at within 0.05 mm / 0.05 deg. This
    holds only if the DRR's pixel grid (RayCaster.cl), the resample's
    sampler rule and t flip (RadRenderer.cl), the mask and NCC composition
    (Tracker.cpp:583-773) and the pose plumbing (minimizationFunc) agree;
    the same target rendered half a pixel off moves the optimum by more
    than the bar."""
    S, gt, pivot = _bead_scene(orc)
    s, st, n = S.optimize_frame(orc.pose_store_set(gt + START), pivot, method=1, max_iter=100,
                                offsets=(1.0,) * 6, seed=7)
    dt, dr = _pose_gap(orc, orc.pose_store_get(s), gt)
    assert dt <= 0.05 and dr <= 0.05, (dt, dr)
    assert st[3] <= S.track_frame(gt)["corr"].sum() + 1e-6
    # planted mistake: the target's grid half a pixel off along u
    S2, _, _ = _bead_scene(orc, shift_px=0.5)
    s2, _, _ = S2.optimize_frame(orc.pose_store_set(gt + START), pivot, method=1, max_iter=100,
                                 offsets=(1.0,) * 6, seed=7)
    dt2, dr2 = _pose_gap(orc, orc.pose_store_get(s2), gt)
    assert dt2 > 0.05 or dr2 > 0.05, (dt2, dr2)


def test_filtered_cost_is_minimal_at_hidden_pose(orc):
    """With the same Sobel chain on both images the cost at the hidden pose
    is below the cost at every pose 0.2 mm or 0.2 deg off along one axis:
    the mask is taken after the DRR filters and both images are masked
    alike (Tracker.cpp:746-752)."""
    from autoscoper_amd import filters as F

    S, gt, _ = _bead_scene(orc, filters=F.to_records([F.sobel(1.7, 0.1)]))
    c0 = S.track_frame(gt)["corr"].sum()
    for ax in range(6):
        for sgn in (-1, 1):
            d = np.zeros(6)
            d[ax] = 0.2 * sgn
            assert S.track_frame(gt + d)["corr"].sum() > c0
