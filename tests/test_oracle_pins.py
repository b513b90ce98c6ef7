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
