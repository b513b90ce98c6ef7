"""Seeded synthetic registration inputs: CT phantoms, biplane cameras, poses
and target radiographs.

Phantom recipe (HU, stored as u16 = clip(HU + 1024, 0, 65535)):
  air -1000; soft-tissue cylinder/ellipsoid ~40 HU; bone with a 1.5 mm
  cortical shell at 1800 HU and a trabecular interior of 400 +/- 150 HU
  (Gaussian-smoothed texture); additive N(0, 20 HU) noise everywhere.
Cameras are built the way a MayaCam 2.0 calibration file describes them
(K, world-to-camera R, t) and converted with the reference's loader
arithmetic (Camera.cpp:229-362). Targets are DRRs at hidden poses,
blurred (sigma 1 px), with intensity-dependent noise and a smooth
background, as float32 detectors plus a u16-quantised copy.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SEED_PHANTOM, SEED_NOISE, SEED_POSES, SEED_START = 1001, 1002, 1003, 1004

# DRR parameters from the reference UI maps (ModelViewTreeWidgetItem.cpp:330-338)
# with SampleDistance 0.62 and RayIntensity 0.49: step = exp(7*0.62-5),
# intensity = exp(15*0.49-5), cutoff 0.
DRR_STEP = float(np.float32(np.exp(7 * 0.62 - 5)))
DRR_INTENSITY = float(np.float32(np.exp(15 * 0.49 - 5)))
DRR_CUTOFF = 0.0


def hu_to_u16(hu: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(hu + 1024.0), 0, 65535).astype(np.uint16)


def phantom(shape_zyx, voxel_mm: float, kind: str = "ellipsoid", seed: int = SEED_PHANTOM,
            noise_seed: int = SEED_NOISE, noise_hu: float = 20.0, segmented: bool = False) -> np.ndarray:
    """Return a (D, H, W) uint16 phantom. segmented: the bone alone, 0
    elsewhere (a segmented bone model, as tracking volumes usually are)."""
    from scipy import ndimage

    D, H, W = shape_zyx
    rng = np.random.default_rng(seed)
    z = (np.arange(D, dtype=np.float32) - (D - 1) / 2) * voxel_mm
    y = (np.arange(H, dtype=np.float32) - (H - 1) / 2) * voxel_mm
    x = (np.arange(W, dtype=np.float32) - (W - 1) / 2) * voxel_mm
    ex, ey, ez = W * voxel_mm / 2, H * voxel_mm / 2, D * voxel_mm / 2
    hu = np.full((D, H, W), -1000.0, np.float32)
    Z, Y, X = z[:, None, None], y[None, :, None], x[None, None, :]
    if kind == "ellipsoid":
        # soft tissue ellipsoid and an off-centre ellipsoidal bone
        st = (X / (0.92 * ex)) ** 2 + (Y / (0.88 * ey)) ** 2 + (Z / (0.95 * ez)) ** 2
        hu = np.where(st <= 1.0, 40.0, hu)
        a = np.array([0.55 * ex, 0.45 * ey, 0.7 * ez], np.float32)
        c = np.array([0.06 * ex, -0.04 * ey, 0.03 * ez], np.float32)
        rho = np.sqrt(((X - c[0]) / a[0]) ** 2 + ((Y - c[1]) / a[1]) ** 2 + ((Z - c[2]) / a[2]) ** 2)
        depth = (1.0 - rho) * float(a.min())  # approximate distance inside the surface (mm)
    else:
        # long bone: capsule along z with a slowly varying radius, in a soft-tissue cylinder
        st = (X / (0.92 * ex)) ** 2 + (Y / (0.92 * ey)) ** 2
        hu = np.where(st <= 1.0, 40.0, hu)
        half = 0.8 * ez
        r0 = 0.38 * min(ex, ey)
        zc = np.clip(Z, -half, half)
        radius = r0 * (1.0 + 0.25 * (zc / half) ** 2)  # flared metaphyses
        # the shaft: an oblique axis and an elliptical cross-section, with a
        # ridge along one side and a condyle off the axis at one end (real
        # long bones are not solids of revolution: without these a rotation
        # about the shaft would hardly change the images)
        cx = 0.05 * ex + 0.1 * zc
        cy = -0.03 * ey - 0.04 * zc
        dist = np.sqrt(((X - cx) / 1.2) ** 2 + ((Y - cy) / 0.85) ** 2 + (Z - zc) ** 2)
        depth = radius - dist
        zr = np.clip(Z, -0.5 * half, 0.6 * half)
        rx = 0.05 * ex + 0.1 * zr + 0.95 * r0 * np.cos(0.5)
        ry = -0.03 * ey - 0.04 * zr + 0.75 * r0 * np.sin(0.5)
        ridge = 0.35 * r0 - np.sqrt((X - rx) ** 2 + (Y - ry) ** 2 + (Z - zr) ** 2)
        tx, ty = 0.05 * ex + 0.1 * half, -0.03 * ey - 0.04 * half
        condyle = 0.6 * r0 - np.sqrt((X - tx + 0.7 * r0) ** 2 + (Y - ty - 0.3 * r0) ** 2 + (Z - half) ** 2)
        depth = np.maximum(depth, np.maximum(ridge, condyle))
    bone = depth >= 0
    tex = rng.standard_normal((D, H, W)).astype(np.float32)
    tex = ndimage.gaussian_filter(tex, sigma=1.5 / voxel_mm * 0.5, mode="nearest")
    tex /= max(float(tex.std()), 1e-6)
    trab = 400.0 + 150.0 * tex
    hu = np.where(bone, np.where(depth < 1.5, 1800.0, trab), hu)
    nrng = np.random.default_rng(noise_seed)
    hu = hu + nrng.normal(0.0, noise_hu, size=hu.shape).astype(np.float32)
    out = hu_to_u16(hu)
    if segmented:
        out[~bone] = 0
    return out


@dataclass
class CameraSpec:
    """Pinhole looking at the world origin from `centre` (mm)."""

    centre: tuple
    up: tuple = (0.0, 0.0, 1.0)
    size: tuple = (1024, 1024)
    pixel_mm: float = 0.3
    sid_mm: float = 1100.0  # source-to-image distance
    principal: tuple | None = None  # pixel coordinates (cx, cy); default near the centre

    def mayacam2(self):
        """-> (K, R, t, size) as a MayaCam 2.0 file lists them (OpenCV convention)."""
        c = np.asarray(self.centre, float)
        back = c / np.linalg.norm(c)
        up = np.asarray(self.up, float)
        up = up - back * np.dot(up, back)
        up /= np.linalg.norm(up)
        right = np.cross(up, back)
        M = np.stack([right, up, back], axis=1)  # camera axes in world (columns)
        R = np.diag([1.0, -1.0, -1.0]) @ M.T  # world -> camera, x right / y down / z forward
        t = -R @ c
        f = self.sid_mm / self.pixel_mm
        w, h = self.size
        cx, cy = self.principal if self.principal is not None else (w / 2 + 3.3, h / 2 - 2.7)
        K = np.array([[f, 0.0, cx], [0.0, f, cy], [0.0, 0.0, 1.0]])
        return K, R, t, np.array([float(w), float(h)])


def biplane(sod_mm: float = 800.0, size=(1024, 1024), pixel_mm: float = 0.3, sid_mm: float = 1100.0,
            n: int = 2):
    """n cameras around the vertical axis (2 -> 90 degrees apart)."""
    specs = []
    for k in range(n):
        ang = np.deg2rad(90.0 * k + 7.0)
        centre = (sod_mm * np.cos(ang), sod_mm * np.sin(ang), 35.0 * (k - 0.5))
        specs.append(CameraSpec(centre=centre, size=size, pixel_mm=pixel_mm, sid_mm=sid_mm))
    return specs


def random_poses(base, n: int, trans_mm: float, rot_deg: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    base = np.asarray(base, float)
    d = np.concatenate([rng.uniform(-trans_mm, trans_mm, (n, 3)), rng.uniform(-rot_deg, rot_deg, (n, 3))], 1)
    return base[None, :] + d


def random_walk(base, n: int, step_mm: float, step_deg: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = [np.asarray(base, float)]
    for _ in range(n - 1):
        d = np.concatenate([rng.normal(0, step_mm / np.sqrt(3), 3), rng.normal(0, step_deg / np.sqrt(3), 3)])
        out.append(out[-1] + d)
    return np.stack(out)


def full_frame_viewport(cam_viewport, width: int, height: int) -> np.ndarray:
    """RayCaster viewport whose pixel (j, i) samples the ray through the
    centre of radiograph texel (j, H-1-i) as RadRenderer addresses it
    (RadRenderer.cl:19-29)."""
    u0, v0, u1, v1 = [float(v) for v in cam_viewport]
    return np.array([u0 - u1 / 2 + 0.5 * u1 / width, v0 - v1 / 2 + 0.5 * v1 / height, u1, v1], np.float32)


def target_from_drr(drr_full: np.ndarray, seed: int, noise: float = 0.002, background: float = 0.05):
    """Full-frame DRR (rows bottom-up) -> (float32 radiograph, u16 copy):
    blurred (sigma 1 px), Poisson-like noise of standard deviation
    noise * sqrt(intensity), a smooth background. The noise level is set
    against the DRRs' contrast at the benchmark's ray intensity (peaks near
    0.09): at 1% the Sobel-filtered radiograph is noise-dominated (about one
    to one per pixel at bone edges) and the masked NCC's minimum leaves the
    hidden pose by a millimetre; at 0.2% the hidden pose is the minimum."""
    from scipy import ndimage

    img = np.flipud(drr_full).astype(np.float64)
    img = ndimage.gaussian_filter(img, 1.0, mode="nearest")
    inten = 1.0 - img
    rng = np.random.default_rng(seed)
    inten = inten + rng.normal(0.0, 1.0, inten.shape) * noise * np.sqrt(np.clip(inten, 0, None))
    h, w = inten.shape
    yy, xx = np.mgrid[0:h, 0:w] / np.array([h, w], float)[:, None, None]
    inten = inten - background * (0.6 * xx + 0.4 * yy * yy)
    inten = np.clip(inten, 0.0, 1.0).astype(np.float32)
    u16 = np.clip(np.rint(inten * 65535.0), 0, 65535).astype(np.uint16)
    return inten, u16


@dataclass
class Workload:
    name: str
    volume: np.ndarray  # (D, H, W) u16
    voxel_mm: float
    cameras: list  # list of CameraSpec
    render: tuple  # render width, height
    gt_poses: np.ndarray  # (frames, 6)
    drr_filters: list = field(default_factory=list)
    rad_filters: list = field(default_factory=list)


def default_filters():
    """DRR: Sobel(scale 1.7, blend 0.1); radiograph: Sobel(scale 3, blend
    0.43) then Contrast(alpha 2.5, beta 0.9) -- the effective values of the
    reference's sample .vie settings."""
    from . import filters

    return [filters.sobel(1.7, 0.1)], [filters.sobel(3.0, 0.43), filters.contrast(2.5, 0.9, 3)]


def volume_centre_pose(shape_zyx, voxel_mm: float) -> np.ndarray:
    """Pose that puts the (uncropped) volume's centre at the world origin:
    the translation is VolumeDescription's transCenter."""
    D, H, W = shape_zyx
    return np.array([-voxel_mm * 0.5 * (W - 1), -voxel_mm * 0.5 * (H - 1), voxel_mm * 0.5 * (D - 1), 0, 0, 0],
                    float)


def workload(name: str, frames: int | None = None) -> Workload:
    """Named workloads: 'small' (128^3 @0.5 mm, 2 x 512^2), 'cfg2'
    (256x256x320 @0.3 mm, 2 x 1024^2), 'cfg2seg' (cfg2's bone segmented:
    0 outside it, so the volume crops to the bone), 'tiny' (test size)."""
    dflt, rflt = default_filters()
    if name == "tiny":
        shape, vox, size, sod, pix, kind, nf = (40, 36, 32), 1.0, (96, 80), 400.0, 0.9, "ellipsoid", 3
    elif name == "small":
        shape, vox, size, sod, pix, kind, nf = (128, 128, 128), 0.5, (512, 512), 800.0, 0.45, "ellipsoid", 10
    elif name in ("cfg2", "cfg2seg"):
        shape, vox, size, sod, pix, kind, nf = (320, 256, 256), 0.3, (1024, 1024), 800.0, 0.3, "longbone", 100
    else:
        raise ValueError(name)
    nf = frames or nf
    vol = phantom(shape, vox, kind, segmented=name.endswith("seg"))
    cams = biplane(sod_mm=sod, size=size, pixel_mm=pix, sid_mm=sod * 1.375)
    base = volume_centre_pose(shape, vox)
    rng = np.random.default_rng(SEED_POSES)
    start = base + np.concatenate([rng.uniform(-2, 2, 3), rng.uniform(-5, 5, 3)])
    gt = random_walk(start, nf, 0.5, 1.0, SEED_POSES + 1)
    return Workload(name, vol, vox, cams, size, gt, dflt, rflt)


@dataclass
class MultiWorkload:
    """Several bones tracked in the same cameras (one volume each)."""
    name: str
    volumes: list      # (D, H, W) u16 each
    voxel_mm: list     # isotropic voxel size of each volume
    cameras: list      # CameraSpec
    render: tuple
    gt_poses: np.ndarray  # (frames, bones, 6)
    drr_filters: list = field(default_factory=list)
    rad_filters: list = field(default_factory=list)


def multi_workload(name: str, frames: int | None = None) -> MultiWorkload:
    """Multi-bone workloads: 'tiny3' (three test-size bones, 2 x 96x80) and
    'cfg3' (96^3 @0.5 mm, 160x160x220 @0.4 mm, 256x256x320 @0.3 mm, 2 x
    1024^2). Bones are stacked along the vertical axis, as at a joint, so
    they project to different detector regions of both cameras."""
    dflt, rflt = default_filters()
    if name == "tiny3":
        specs = [((14, 18, 18), 1.0, "ellipsoid", -17.0), ((20, 22, 20), 0.9, "longbone", 0.0),
                 ((16, 20, 24), 1.0, "ellipsoid", 17.0)]
        size, sod, pix, nf = (96, 80), 400.0, 0.9, 2
    elif name == "cfg3":
        specs = [((96, 96, 96), 0.5, "ellipsoid", -80.0), ((220, 160, 160), 0.4, "longbone", -10.0),
                 ((320, 256, 256), 0.3, "longbone", 65.0)]
        size, sod, pix, nf = (1024, 1024), 800.0, 0.3, 50
    else:
        raise ValueError(name)
    nf = frames or nf
    vols, voxs, gts = [], [], []
    for b, (shape, vox, kind, z) in enumerate(specs):
        vols.append(phantom(shape, vox, kind, seed=SEED_PHANTOM + 10 * b, noise_seed=SEED_NOISE + 10 * b))
        voxs.append(vox)
        base = volume_centre_pose(shape, vox) + np.array([0.0, 0.0, z, 0.0, 0.0, 0.0])
        rng = np.random.default_rng(SEED_POSES + 10 * b)
        start = base + np.concatenate([rng.uniform(-2, 2, 3), rng.uniform(-5, 5, 3)])
        gts.append(random_walk(start, nf, 0.5, 1.0, SEED_POSES + 10 * b + 1))
    cams = biplane(sod_mm=sod, size=size, pixel_mm=pix, sid_mm=sod * 1.375)
    return MultiWorkload(name, vols, voxs, cams, size, np.stack(gts, 1), dflt, rflt)
