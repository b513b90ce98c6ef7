"""ctypes front end of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
and by bench.py's CPU-baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_P = C.POINTER
_d, _f, _i, _u, _vp = C.c_double, C.c_float, C.c_int, C.c_uint, C.c_void_p

_L = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _L
    if _L is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        sig = {
            "orc_volume_prep": ([_vp, _i, _i, _i, _i, _P(_i), _P(_d), _vp, _P(_i), _P(_f), _P(_f), _P(_d), _P(_f),
                                 _P(_i)], _i),
            "orc_drr_f32": ([_vp, _P(_i), _P(_f), _P(_f), _f, _f, _f, _u, _u, _P(_f)], C.c_int64),
            "orc_drr_f64": ([_vp, _P(_i), _P(_f), _P(_f), _f, _f, _f, _u, _u, _P(_d)], C.c_int64),
            "orc_drr_margin_f32": ([_vp, _P(_i), _P(_f), _P(_f), _f, _f, _f, _f, _u, _u, _P(_f), _P(_f)], C.c_int64),
            "orc_rad_f32": ([_vp, _i, _i, _i, _P(_f), _P(_f), _u, _u, _P(_f)], None),
            "orc_background_f32": ([_P(_f), _i, _i, _P(_f), _P(_f), _f, _u, _u, _P(_f)], None),
            "orc_sobel_f32": ([_P(_f), _P(_f), _i, _i, _f, _f], None),
            "orc_contrast_f32": ([_P(_f), _P(_f), _i, _i, _f, _f, _i], None),
            "orc_gaussian_weights": ([_f, _P(_f), _i], _i),
            "orc_gaussian_f32": ([_P(_f), _P(_f), _i, _i, _f], None),
            "orc_sharpen_f32": ([_P(_f), _P(_f), _i, _i, _f, _f, _f], None),
            "orc_filter_chain_f32": ([_P(_f), _i, _P(_f), _P(_f), _i, _i], None),
            "orc_filter_chain_f64": ([_P(_f), _i, _P(_d), _P(_d), _i, _i], None),
            "orc_ncc_exact_f32in": ([_P(_f), _P(_f), _P(_f), C.c_size_t, _P(_d)], _d),
            "orc_ncc_exact_f64in": ([_P(_d), _P(_d), _P(_d), C.c_size_t, _P(_d)], _d),
            "orc_ncc_tree": ([_P(_f), _P(_f), _P(_f), _u, _i], _f),
            "orc_tree_sum": ([_P(_f), _u, _i], _f),
            "orc_frame_from_xyzypr": ([_P(_d), _P(_d)], None),
            "orc_frame_from_xyz_axis_angle": ([_P(_d), _P(_d)], None),
            "orc_frame_to_xyzypr": ([_P(_d), _P(_d)], None),
            "orc_frame_mul": ([_P(_d), _P(_d), _P(_d)], None),
            "orc_frame_inverse": ([_P(_d), _P(_d)], None),
            "orc_frame_linear_extrap": ([_P(_d), _P(_d), _P(_d)], None),
            "orc_camera_from_mayacam2": ([_P(_d), _P(_d), _P(_d), _P(_d), _P(_d)], _i),
            "orc_view_geometry": ([_P(_d), _P(_d), _P(_f), _P(_f), _i, _i, _P(_f), _P(_d), _P(_u)], _i),
            "orc_quatf_roundtrip": ([_P(_d), _P(_d)], None),
            "orc_pose_store_set": ([_P(_f), _P(_d)], None),
            "orc_pose_store_get": ([_P(_f), _P(_d)], None),
            "orc_compose_pose": ([_P(_f), _P(_d), _P(_d), _P(_d)], None),
            "orc_amoeba_sinc": ([_P(_d), _P(_d), _d], _i),
            "orc_pso_bowl": ([_P(_d), _u, _f, _f, _u, _u, _i, _P(_f), _P(_f)], _i),
            "orc_glibc_rand_first": ([_u, _i, _P(_i)], _i),
            "orc_merge_f32": ([_P(_f), _P(_f), _P(_f), C.c_size_t], None),
            "orc_sample3": ([_vp, _P(_i), _d, _d, _d, _i], _d),
            "orc_composite_f32": ([_P(_f), _P(_f), _P(_f), _P(_f), C.c_size_t], None),
            "orc_cost_ripple": ([_P(_d), _vp], _d),
            "orc_cost_bowl": ([_P(_d), _vp], _d),
            "orc_cost_sinc": ([_P(_d), _vp], _d),
            "orc_pso_fn": ([_vp, _vp, _u, _f, _f, _u, _u, _i, _i, _P(_f), _P(_f)], _i),
            "orc_amoeba_fn": ([_vp, _vp, _P(_d), _P(_d), _i, _d], _i),
        }
        scene = [_vp, _P(_i), _P(_f), _P(_f), _i, _P(_d), _P(_vp), _P(_i), _P(_f), _P(_i), _i, _i, _f, _P(_vp),
                 _P(_i)]
        sig["orc_track_frame"] = (scene + [_P(_d), _i, _P(_d), _P(_i), _P(_d), _P(_d)], C.c_int64)
        sig["orc_min_func"] = (scene + [_P(_f), _P(_d), _P(_d), _i], _d)
        sig["orc_optimize_frame"] = (scene + [_P(_f), _P(_d), _i, _u, _d, _d, _u, _P(_d), _u, _i, _i, _i, _P(_d)],
                                     C.c_int64)
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _L = L
    return _L


def _p(a, t):
    return a.ctypes.data_as(_P(t))


def f32(a):
    return np.ascontiguousarray(a, np.float32)


def f64(a):
    return np.ascontiguousarray(a, np.float64)


# ---- a1 --------------------------------------------------------------------
def volume_prep(data: np.ndarray, scale, flip=(0, 0, 0)):
    data = np.ascontiguousarray(data)
    D, H, W = data.shape
    bps = 8 * data.dtype.itemsize
    out = np.zeros(data.size, np.uint16)
    dims = np.zeros(3, np.int32)
    inv_scale = np.zeros(3, np.float32)
    inv_trans = np.zeros(3, np.float32)
    tc = np.zeros(3)
    mm = np.zeros(2, np.float32)
    crop = np.zeros(6, np.int32)
    fl = np.asarray(flip, np.int32)
    sc = f64(scale)
    rc = lib().orc_volume_prep(data.ctypes.data_as(_vp), W, H, D, bps, _p(fl, _i), _p(sc, _d),
                               out.ctypes.data_as(_vp), _p(dims, _i), _p(inv_scale, _f), _p(inv_trans, _f),
                               _p(tc, _d), _p(mm, _f), _p(crop, _i))
    if rc != 0:
        raise ValueError(f"volume_prep failed ({rc})")
    n = int(dims.prod())
    vol = out[:n].reshape(dims[2], dims[1], dims[0]).copy()
    return dict(volume=vol, dims=dims, inv_scale=inv_scale, inv_trans=inv_trans, trans_center=tc,
                min_value=float(mm[0]), max_value=float(mm[1]), crop_min=crop[:3].copy(), crop_max=crop[3:].copy())


# ---- a5/a6/a12 -----------------------------------------------------------------
def drr(vol: np.ndarray, imv, vp, step, intensity, cutoff, w, h, fp64=False):
    vol = np.ascontiguousarray(vol, np.uint16)
    vd = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], np.int32)
    imv, vp = f32(imv), f32(vp)
    if fp64:
        out = np.zeros((h, w))
        n = lib().orc_drr_f64(vol.ctypes.data_as(_vp), _p(vd, _i), _p(imv, _f), _p(vp, _f), step, intensity, cutoff,
                              w, h, _p(out, _d))
    else:
        out = np.zeros((h, w), np.float32)
        n = lib().orc_drr_f32(vol.ctypes.data_as(_vp), _p(vd, _i), _p(imv, _f), _p(vp, _f), step, intensity, cutoff,
                              w, h, _p(out, _f))
    return out, int(n)


def drr_cut_margin(vol: np.ndarray, imv, vp, step, intensity, cutoff, w, h, delta=4e-6):
    """-> (float DRR, per-pixel margin): the margin is the DRR contribution
    of the samples whose value lies within delta * max(val, cutoff) of the
    cutoff, i.e. what a sampler rounding differently could flip."""
    vol = np.ascontiguousarray(vol, np.uint16)
    vd = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], np.int32)
    imv, vp = f32(imv), f32(vp)
    out = np.zeros((h, w), np.float32)
    margin = np.zeros((h, w), np.float32)
    lib().orc_drr_margin_f32(vol.ctypes.data_as(_vp), _p(vd, _i), _p(imv, _f), _p(vp, _f), step, intensity, cutoff,
                             delta, w, h, _p(out, _f), _p(margin, _f))
    return out, margin


def sample3(vol: np.ndarray, s, t, r, fp64=False):
    vol = np.ascontiguousarray(vol, np.uint16)
    vd = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], np.int32)
    return float(lib().orc_sample3(vol.ctypes.data_as(_vp), _p(vd, _i), s, t, r, int(fp64)))


def rad(img: np.ndarray, plane, vp, w, h):
    img = np.ascontiguousarray(img)
    bps = {np.dtype(np.uint8): 8, np.dtype(np.uint16): 16, np.dtype(np.float32): 32}[img.dtype]
    out = np.zeros((h, w), np.float32)
    pl, v = f32(plane), f32(vp)
    lib().orc_rad_f32(img.ctypes.data_as(_vp), bps, img.shape[1], img.shape[0], _p(pl, _f), _p(v, _f), w, h,
                      _p(out, _f))
    return out


def background(img: np.ndarray, plane, vp, threshold, w, h):
    img = f32(img)
    out = np.zeros((h, w), np.float32)
    pl, v = f32(plane), f32(vp)
    lib().orc_background_f32(_p(img, _f), img.shape[1], img.shape[0], _p(pl, _f), _p(v, _f), threshold, w, h,
                             _p(out, _f))
    return out


# ---- a8-a11 ------------------------------------------------------------------------
def sobel(img, scale, blend):
    img = f32(img)
    out = np.zeros_like(img)
    lib().orc_sobel_f32(_p(img, _f), _p(out, _f), img.shape[1], img.shape[0], scale, blend)
    return out


def contrast(img, alpha, beta, size=3):
    img = f32(img)
    out = np.zeros_like(img)
    lib().orc_contrast_f32(_p(img, _f), _p(out, _f), img.shape[1], img.shape[0], alpha, beta, size)
    return out


def gaussian_weights(radius):
    buf = np.zeros(61 * 61, np.float32)
    n = lib().orc_gaussian_weights(radius, _p(buf, _f), buf.size)
    return buf[: n * n].reshape(n, n)


def gaussian(img, radius):
    img = f32(img)
    out = np.zeros_like(img)
    lib().orc_gaussian_f32(_p(img, _f), _p(out, _f), img.shape[1], img.shape[0], radius)
    return out


def sharpen(img, radius, contrast_, threshold):
    img = f32(img)
    out = np.zeros_like(img)
    lib().orc_sharpen_f32(_p(img, _f), _p(out, _f), img.shape[1], img.shape[0], radius, contrast_, threshold)
    return out


def filter_chain(records, img, fp64=False):
    rec = f32(records).reshape(-1, 6)
    if fp64:
        img = f64(img)
        out = np.zeros_like(img)
        lib().orc_filter_chain_f64(_p(rec, _f), rec.shape[0], _p(img, _d), _p(out, _d), img.shape[1], img.shape[0])
    else:
        img = f32(img)
        out = np.zeros_like(img)
        lib().orc_filter_chain_f32(_p(rec, _f), rec.shape[0], _p(img, _f), _p(out, _f), img.shape[1], img.shape[0])
    return out


# ---- a14 ------------------------------------------------------------------------
def ncc_exact(f, g, m):
    mom = np.zeros(7)
    if np.asarray(f).dtype == np.float64:
        f, g, m = f64(f).ravel(), f64(g).ravel(), f64(m).ravel()
        v = lib().orc_ncc_exact_f64in(_p(f, _d), _p(g, _d), _p(m, _d), f.size, _p(mom, _d))
    else:
        f, g, m = f32(f).ravel(), f32(g).ravel(), f32(m).ravel()
        v = lib().orc_ncc_exact_f32in(_p(f, _f), _p(g, _f), _p(m, _f), f.size, _p(mom, _d))
    return v, mom


def ncc_tree(f, g, m, mode):
    f, g, m = f32(f).ravel(), f32(g).ravel(), f32(m).ravel()
    return float(lib().orc_ncc_tree(_p(f, _f), _p(g, _f), _p(m, _f), f.size, mode))


def tree_sum(f, mode):
    f = f32(f).ravel()
    return float(lib().orc_tree_sum(_p(f, _f), f.size, mode))


# ---- geometry ----------------------------------------------------------------------
def _frame_call(name, *arrays, out_n=12):
    args = [f64(a) for a in arrays]
    out = np.zeros(out_n)
    getattr(lib(), name)(*[_p(a, _d) for a in args], _p(out, _d))
    return out


def frame_from_xyzypr(p):
    return _frame_call("orc_frame_from_xyzypr", p)


def frame_from_xyz_axis_angle(p):
    return _frame_call("orc_frame_from_xyz_axis_angle", p)


def frame_to_xyzypr(f):
    return _frame_call("orc_frame_to_xyzypr", f, out_n=6)


def frame_mul(a, b):
    return _frame_call("orc_frame_mul", a, b)


def frame_inverse(a):
    return _frame_call("orc_frame_inverse", a)


def frame_linear_extrap(a, b):
    return _frame_call("orc_frame_linear_extrap", a, b)


def camera_from_mayacam2(K, R, t, size):
    K, R, t, size = f64(K).ravel(), f64(R).ravel(), f64(t).ravel(), f64(size).ravel()
    rec = np.zeros(30)
    if lib().orc_camera_from_mayacam2(_p(K, _d), _p(R, _d), _p(t, _d), _p(size, _d), _p(rec, _d)) != 0:
        raise ValueError("invalid camera")
    return rec


def view_geometry(cam_rec, xyzypr, inv_scale, inv_trans, rw, rh):
    c, p = f64(cam_rec), f64(xyzypr)
    s, t = f32(inv_scale), f32(inv_trans)
    imv = np.zeros(16, np.float32)
    vp = np.zeros(4)
    sz = np.zeros(2, np.uint32)
    r = lib().orc_view_geometry(_p(c, _d), _p(p, _d), _p(s, _f), _p(t, _f), rw, rh, _p(imv, _f), _p(vp, _d),
                                _p(sz, _u))
    return bool(r), imv, vp, (int(sz[0]), int(sz[1]))


def pose_store_set(p):
    s = np.zeros(7, np.float32)
    pp = f64(p)
    lib().orc_pose_store_set(_p(s, _f), _p(pp, _d))
    return s


def pose_store_get(s):
    s = f32(s)
    out = np.zeros(6)
    lib().orc_pose_store_get(_p(s, _f), _p(out, _d))
    return out


def compose_pose(store, pivot, manip):
    s, pv, m = f32(store), f64(pivot), f64(manip)
    out = np.zeros(6)
    lib().orc_compose_pose(_p(s, _f), _p(pv, _d), _p(m, _d), _p(out, _d))
    return out


def amoeba_sinc(P, ftol=1e-10):
    P = f64(P).copy()
    Y = np.zeros(3)
    it = lib().orc_amoeba_sinc(_p(P, _d), _p(Y, _d), ftol)
    return P.reshape(3, 2), Y, it


def pso_bowl(centre, seed, lo, hi, max_epochs, max_stall, nparticles=100):
    c = f64(centre)
    best = np.zeros(6, np.float32)
    bn = np.zeros(1, np.float32)
    ep = lib().orc_pso_bowl(_p(c, _d), seed, lo, hi, max_epochs, max_stall, nparticles, _p(best, _f), _p(bn, _f))
    return best, float(bn[0]), ep


def cost_fn(name):
    """Address of one of the oracle's exported test objectives
    (orc_cost_ripple / orc_cost_bowl / orc_cost_sinc)."""
    return C.cast(getattr(lib(), "orc_cost_" + name), C.c_void_p).value


def pso_fn(name, centre, seed, lo, hi, max_epochs, max_stall, nparticles=100, synchronous=False):
    c = f64(centre).copy()
    best = np.zeros(6, np.float32)
    bn = np.zeros(1, np.float32)
    ep = lib().orc_pso_fn(cost_fn(name), c.ctypes.data, seed, lo, hi, max_epochs, max_stall, nparticles,
                          int(synchronous), _p(best, _f), _p(bn, _f))
    return best, float(bn[0]), ep


def amoeba_fn(name, centre, P, ftol):
    """Simplex from the (ndim+1) x ndim start P; Y is evaluated here."""
    c = f64(centre).copy()
    P = f64(P).copy()
    n = P.shape[1]
    Y = np.array([_cost(name, P[i], c) for i in range(n + 1)])
    it = lib().orc_amoeba_fn(cost_fn(name), c.ctypes.data, _p(P, _d), _p(Y, _d), n, ftol)
    return P, Y, it


def _cost(name, x, centre):
    x = f64(x)
    return getattr(lib(), "orc_cost_" + name)(_p(x, _d), centre.ctypes.data)


# ---------------------------------------------------------------------------
# The reference's own host C++ (CoordFrame, PSO, DownhillSimplex, Quatf),
# compiled from its sources into oracle/_ref/libref_host.so by `make ref`
# where the reference is present. None when it is not built.
# ---------------------------------------------------------------------------
REF_LIB_PATH = HERE / "_ref" / "libref_host.so"
_R = None


def ref_lib():
    global _R
    if _R is None and REF_LIB_PATH.exists():
        R = C.CDLL(str(REF_LIB_PATH))
        sig = {
            "ref_frame_from_xyzypr": ([_P(_d), _P(_d)], None),
            "ref_frame_from_xyz_axis_angle": ([_P(_d), _P(_d)], None),
            "ref_frame_to_xyzypr": ([_P(_d), _P(_d)], None),
            "ref_frame_mul": ([_P(_d), _P(_d), _P(_d)], None),
            "ref_frame_inverse": ([_P(_d), _P(_d)], None),
            "ref_frame_linear_extrap": ([_P(_d), _P(_d), _P(_d)], None),
            "ref_frame_matrix_row_order": ([_P(_d), _P(_d)], None),
            "ref_pose_store_roundtrip": ([_P(_d), _P(_d)], None),
            "ref_compose_pose": ([_P(_d), _P(_d), _P(_d), _P(_d)], None),
            "ref_pso": ([_vp, _vp, _u, _f, _f, _u, _u, _P(_f), _P(_f)], C.c_long),
            "ref_amoeba": ([_vp, _vp, _P(_d), _P(_d), _i, _d], _i),
            "ref_calls": ([], C.c_long),
        }
        for name, (args, res) in sig.items():
            fn = getattr(R, name)
            fn.argtypes = args
            fn.restype = res
        _R = R
    return _R


def glibc_rand(seed, n):
    out = np.zeros(n, np.int32)
    rmax = lib().orc_glibc_rand_first(seed, n, _p(out, _i))
    return out, rmax


def merge(a, b):
    a, b = f32(a), f32(b)
    out = np.zeros_like(a)
    lib().orc_merge_f32(_p(a, _f), _p(b, _f), _p(out, _f), a.size)
    return out


def composite(s1, s2, s3):
    s1, s2, s3 = f32(s1), f32(s2), f32(s3)
    out = np.zeros(s1.shape + (3,), np.float32)
    lib().orc_composite_f32(_p(s1, _f), _p(s2, _f), _p(s3, _f), _p(out, _f), s1.size)
    return out


# ---- scene-level (a15/a16/a17/a18) ------------------------------------------------------
class Scene:
    """Inputs of one tracking scene for orc_track_frame / orc_min_func /
    orc_optimize_frame. cams: list of 30-double camera records;
    rads: list of 2-D images; filters: (drr_records, rad_records) per camera
    as 6-float records."""

    def __init__(self, prep: dict, drr_params, cams, rads, filters, render, bg_threshold=-1.0, bgs=None):
        self.vol = np.ascontiguousarray(prep["volume"], np.uint16)
        self.vdim = np.array([self.vol.shape[2], self.vol.shape[1], self.vol.shape[0]], np.int32)
        self.vdesc = np.concatenate([prep["inv_scale"], prep["inv_trans"]]).astype(np.float32)
        self.drr_params = f32(drr_params)
        self.ncam = len(cams)
        self.cams = f64(np.concatenate([f64(c).ravel() for c in cams]))
        self.rads = [np.ascontiguousarray(r) for r in rads]
        self.rad_ptrs = (_vp * self.ncam)(*[r.ctypes.data for r in self.rads])
        info = []
        for r in self.rads:
            bps = {np.dtype(np.uint8): 8, np.dtype(np.uint16): 16, np.dtype(np.float32): 32}[r.dtype]
            info += [r.shape[1], r.shape[0], bps]
        self.rad_info = np.array(info, np.int32)
        fl = np.zeros((self.ncam, 2, 8, 6), np.float32)
        nf = np.zeros((self.ncam, 2), np.int32)
        for c, (d, r) in enumerate(filters):
            for which, recs in enumerate((d, r)):
                recs = f32(recs).reshape(-1, 6)
                fl[c, which, : recs.shape[0]] = recs
                nf[c, which] = recs.shape[0]
        self.filters, self.nfilters = fl, nf
        self.render = render
        self.bg_threshold = float(bg_threshold)
        self.bgs = [f32(b) for b in bgs] if bgs else None
        if self.bgs:
            self.bg_ptrs = (_vp * self.ncam)(*[b.ctypes.data for b in self.bgs])
            self.bg_info = np.array(sum([[b.shape[1], b.shape[0]] for b in self.bgs], []), np.int32)
        else:
            self.bg_ptrs = None
            self.bg_info = np.zeros(2 * self.ncam, np.int32)

    def _args(self):
        return [self.vol.ctypes.data_as(_vp), _p(self.vdim, _i), _p(self.vdesc, _f), _p(self.drr_params, _f),
                self.ncam, _p(self.cams, _d), self.rad_ptrs, _p(self.rad_info, _i), _p(self.filters, _f),
                _p(self.nfilters, _i), self.render[0], self.render[1], self.bg_threshold, self.bg_ptrs,
                _p(self.bg_info, _i)]

    def track_frame(self, xyzypr, fp64=False):
        p = f64(xyzypr)
        corr = np.zeros(self.ncam)
        info = np.zeros((self.ncam, 8), np.int32)
        vps = np.zeros((self.ncam, 4))
        mom = np.zeros((self.ncam, 7))
        n = lib().orc_track_frame(*self._args(), _p(p, _d), int(fp64), _p(corr, _d), _p(info, _i), _p(vps, _d),
                                  _p(mom, _d))
        return dict(corr=corr, info=info, vp=vps, moments=mom, samples=int(n))

    def min_func(self, store, pivot, manip, fp64=False):
        s, pv, m = f32(store), f64(pivot), f64(manip)
        return float(lib().orc_min_func(*self._args(), _p(s, _f), _p(pv, _d), _p(m, _d), int(fp64)))

    def optimize_frame(self, store, pivot, method=0, max_iter=100, min_limit=-3.0, max_limit=3.0, max_stall=25,
                       offsets=(0.1,) * 6, seed=1, nparticles=100, fp64=False, synchronous=False):
        s = f32(store).copy()
        pv, off = f64(pivot), f64(offsets)
        stats = np.zeros(4)
        n = lib().orc_optimize_frame(*self._args(), _p(s, _f), _p(pv, _d), method, max_iter, min_limit, max_limit,
                                     max_stall, _p(off, _d), seed, nparticles, int(fp64), int(synchronous),
                                     _p(stats, _d))
        return s, stats, int(n)
