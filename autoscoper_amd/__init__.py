"""MI355X (gfx950) backend for Autoscoper's 2D-3D registration cost path.

Python binding of the C ABI declared in include/autoscoper_hip.h
(libautoscoper_hip.so). It mirrors the reference's xromm::gpu operator
surface (libautoscoper/src/gpu/opencl/*.hpp): RayCaster, RadRenderer,
BackgroundRenderer, the four filters, ncc, multiply/drr_background,
merge/composite, plus the batched scene evaluator and the optimizers.

The library must be built (python -m autoscoper_amd.build); importing the
native layer without it raises, there is no CPU fallback.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
from pathlib import Path

import numpy as np

__all__ = [
    "lib", "AscFilter", "AscCamera", "AscVolumeDesc", "AscOptimizeOptions", "AscOptimizeStats",
    "Context", "Volume", "Image", "Scene", "DeviceArray", "filters", "geometry", "io", "Trial", "Server", "Remote",
    "STORAGE_U16", "STORAGE_F16", "STORAGE_BF16", "STORAGE_F32",
    "NCC_MOMENTS", "NCC_TREE_CUDA", "NCC_TREE_OPENCL",
    "FILTER_CONTRAST", "FILTER_SOBEL", "FILTER_MEDIAN", "FILTER_GAUSSIAN", "FILTER_SHARPEN",
    "AscError",
]

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libautoscoper_hip.so"

STORAGE_U16, STORAGE_F16, STORAGE_BF16, STORAGE_F32 = 0, 1, 2, 3
NCC_MOMENTS, NCC_TREE_CUDA, NCC_TREE_OPENCL = 0, 1, 2
FILTER_CONTRAST, FILTER_SOBEL, FILTER_MEDIAN, FILTER_GAUSSIAN, FILTER_SHARPEN = 0, 1, 2, 3, 4


class AscError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"asc error {code}: {msg}")
        self.code = code


class AscFilter(C.Structure):
    _fields_ = [("type", C.c_int), ("enabled", C.c_int), ("p", C.c_float * 4)]


class AscCamera(C.Structure):
    _fields_ = [
        ("rotation", C.c_double * 9), ("translation", C.c_double * 3),
        ("image_plane", C.c_double * 12), ("viewport", C.c_double * 4), ("size", C.c_double * 2),
    ]

    def to_array(self) -> np.ndarray:
        return np.frombuffer(bytes(self), dtype=np.float64).copy()

    @classmethod
    def from_array(cls, a) -> "AscCamera":
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == 30
        return cls.from_buffer_copy(a.tobytes())


class AscVolumeDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_int * 3), ("crop_min", C.c_int * 3), ("crop_max", C.c_int * 3),
        ("inv_scale", C.c_float * 3), ("inv_trans", C.c_float * 3), ("trans_center", C.c_double * 3),
        ("min_value", C.c_float), ("max_value", C.c_float), ("storage", C.c_int), ("reserved", C.c_int),
    ]


class AscOptimizeOptions(C.Structure):
    _fields_ = [
        ("method", C.c_int), ("max_iter", C.c_uint), ("max_stall", C.c_uint),
        ("min_limit", C.c_double), ("max_limit", C.c_double), ("offsets", C.c_double * 6),
        ("seed", C.c_uint), ("particles", C.c_int), ("synchronous", C.c_int), ("volume", C.c_int),
        ("speculative", C.c_int),
    ]


class AscOptimizeStats(C.Structure):
    _fields_ = [
        ("pso_epochs", C.c_int), ("simplex_iterations", C.c_int), ("pso_best_cost", C.c_double),
        ("final_cost", C.c_double), ("initial_cost", C.c_double), ("evaluations", C.c_int64),
    ]


class AscJobDesc(C.Structure):
    _fields_ = [
        ("n_volumes", C.c_int), ("volumes", C.POINTER(C.c_void_p)), ("volume_dims", C.POINTER(C.c_int)),
        ("voxel_mm", C.POINTER(C.c_double)), ("initial_poses", C.POINTER(C.c_double)), ("storage", C.c_int),
        ("drr_step", C.c_float), ("drr_intensity", C.c_float), ("drr_cutoff", C.c_float),
        ("n_cameras", C.c_int), ("cameras", C.POINTER(AscCamera)), ("render_width", C.c_int),
        ("render_height", C.c_int), ("drr_filters", C.POINTER(AscFilter)), ("n_drr_filters", C.c_int),
        ("rad_filters", C.POINTER(AscFilter)), ("n_rad_filters", C.c_int), ("n_frames", C.c_int),
        ("frames", C.POINTER(C.c_void_p)), ("frame_dims", C.POINTER(C.c_int)), ("frame_bps", C.c_int),
        ("options", AscOptimizeOptions), ("guess", C.c_int),
    ]


class AscJobStats(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("total_ms", C.c_double), ("evaluations", C.c_int64), ("frames", C.c_int)]


class AscServeStats(C.Structure):
    _fields_ = [
        ("jobs_done", C.c_int64), ("frames_done", C.c_int64), ("batches", C.c_int64), ("requests", C.c_int64),
        ("evaluations", C.c_int64), ("busy_ms", C.c_double), ("elapsed_ms", C.c_double),
        ("latency_p50_ms", C.c_double), ("latency_p99_ms", C.c_double), ("latency_max_ms", C.c_double),
        ("active_jobs", C.c_int), ("waiting_jobs", C.c_int), ("device_bytes_in_use", C.c_size_t),
    ]


class AscFilterSettings(C.Structure):
    _fields_ = [
        ("has_drr_renderer", C.c_int), ("has_drr_filters", C.c_int), ("has_rad_filters", C.c_int),
        ("n_drr_filters", C.c_int), ("n_rad_filters", C.c_int),
        ("drr_filters", AscFilter * 16), ("rad_filters", AscFilter * 16),
        ("sample_distance", C.c_double), ("ray_intensity", C.c_double), ("cutoff_fraction", C.c_double),
        ("drr_step", C.c_float), ("drr_intensity", C.c_float),
    ]


class AscTuneResult(C.Structure):
    _fields_ = [("unroll", C.c_int), ("wave_patch", C.c_int), ("grid", C.c_int), ("best_ms", C.c_double),
                ("default_ms", C.c_double), ("candidates", C.c_int)]


class AscTrialInfo(C.Structure):
    _fields_ = [
        ("version", C.c_int * 2), ("n_cameras", C.c_int), ("n_volumes", C.c_int), ("n_frames", C.c_int),
        ("render_width", C.c_int), ("render_height", C.c_int), ("offsets", C.c_double * 6),
    ]


_vp = C.c_void_p
_i = C.c_int
_u = C.c_uint
_f = C.c_float
_d = C.c_double
_sz = C.c_size_t
_P = C.POINTER

# name -> argtypes (restype is int unless listed in _RES)
_SIGS = {
    "asc_version": [],
    "asc_device_count": [_P(_i)],
    "asc_context_create": [_i, _P(_vp)],
    "asc_context_destroy": [_vp],
    "asc_context_set_stream": [_vp, _vp],
    "asc_context_set_blocking": [_vp, C.c_int],
    "asc_context_set_ncc_form": [_vp, C.c_int],
    "asc_context_stream": [_vp],
    "asc_context_device": [_vp, _P(_i)],
    "asc_synchronize": [_vp],
    "asc_malloc": [_vp, _sz, _P(_vp)],
    "asc_free": [_vp, _vp],
    "asc_memcpy_h2d": [_vp, _vp, _vp, _sz],
    "asc_memcpy_d2h": [_vp, _vp, _vp, _sz],
    "asc_memcpy_d2d": [_vp, _vp, _vp, _sz],
    "asc_fill_f32": [_vp, _vp, _f, _sz],
    "asc_fill_u8": [_vp, _vp, C.c_ubyte, _sz],
    "asc_volume_create": [_vp, _vp, _i, _i, _i, _i, _P(_i), _P(_d), _i, _P(_vp), _P(AscVolumeDesc)],
    "asc_volume_describe": [_vp, _P(AscVolumeDesc)],
    "asc_volume_set_layout": [_vp, _i],
    "asc_volume_destroy": [_vp],
    "asc_volume_bytes": [_vp, _P(_sz)],
    "asc_image_create": [_vp, _vp, _i, _i, _i, _P(_vp)],
    "asc_image_update": [_vp, _vp],
    "asc_image_destroy": [_vp],
    "asc_frame_from_xyzypr": [_P(_d), _P(_d)],
    "asc_frame_from_xyz_axis_angle": [_P(_d), _P(_d)],
    "asc_frame_to_xyzypr": [_P(_d), _P(_d)],
    "asc_frame_mul": [_P(_d), _P(_d), _P(_d)],
    "asc_frame_inverse": [_P(_d), _P(_d)],
    "asc_camera_from_mayacam2": [_P(_d), _P(_d), _P(_d), _P(_d), _P(AscCamera)],
    "asc_inv_model_view": [_P(_d), _P(_f), _P(_f), _P(_f)],
    "asc_view_geometry": [_P(AscCamera), _P(_d), _P(AscVolumeDesc), _i, _i, _P(_f), _P(_d), _P(_u)],
    "asc_calculate_viewport": [_P(_d), _P(AscCamera), _P(AscVolumeDesc), _P(_d)],
    "asc_drr_render": [_vp, _vp, _P(_f), _P(_f), _f, _f, _f, _vp, _u, _u],
    "asc_rad_render": [_vp, _vp, _P(_f), _P(_f), _vp, _u, _u],
    "asc_background_render": [_vp, _vp, _P(_f), _P(_f), _f, _vp, _u, _u],
    "asc_sobel": [_vp, _vp, _vp, _i, _i, _f, _f],
    "asc_contrast": [_vp, _vp, _vp, _i, _i, _f, _f, _i],
    "asc_gaussian_weights": [_f, _P(_f), _i, _P(_i)],
    "asc_gaussian": [_vp, _vp, _vp, _i, _i, _f],
    "asc_sharpen": [_vp, _vp, _vp, _i, _i, _f, _f, _f],
    "asc_filter_apply": [_vp, _P(AscFilter), _vp, _vp, _i, _i],
    "asc_filter_chain": [_vp, _P(AscFilter), _i, _vp, _vp, _vp, _i, _i],
    "asc_drr_background": [_vp, _vp, _vp, _u, _u],
    "asc_multiply": [_vp, _vp, _vp, _vp, _u, _u],
    "asc_mask_multiply": [_vp, _vp, _vp, _vp, _vp, _u, _u],
    "asc_ncc": [_vp, _vp, _vp, _vp, _u, _P(_d), _P(_d)],
    "asc_ncc_noalloc": [_vp, _vp, _vp, _vp, _u, _vp],
    "asc_ncc_init": [_vp, _u],
    "asc_ncc_deinit": [_vp],
    "asc_merge": [_vp, _vp, _vp, _vp, _u, _u],
    "asc_composite": [_vp, _vp, _vp, _vp, _vp, _vp, _u, _u],
    "asc_scene_create": [_vp, _i, _i, _P(_vp)],
    "asc_scene_destroy": [_vp],
    "asc_scene_add_volume": [_vp, _vp, _f, _f, _f, _P(_d), _P(_i)],
    "asc_scene_volume_pivot": [_vp, _i, _P(_d)],
    "asc_scene_set_drr_params": [_vp, _i, _f, _f, _f],
    "asc_scene_add_camera": [_vp, _P(AscCamera), _vp, _P(_i)],
    "asc_scene_set_rad": [_vp, _i, _vp],
    "asc_scene_set_background": [_vp, _i, _vp, _f],
    "asc_scene_set_filters": [_vp, _i, _i, _P(AscFilter), _i],
    "asc_scene_eval_poses": [_vp, _P(_i), _P(_d), _i, _P(_d), _P(_d)],
    "asc_scene_eval_manips": [_vp, _P(_i), _P(_f), _P(_d), _i, _P(_d), _P(_d)],
    "asc_scene_remove_volume": [_vp, _i],
    "asc_scene_remove_camera": [_vp, _i],
    "asc_scene_add_camera_set": [_vp, _P(_i), _i, _i, _i, _P(_i)],
    "asc_scene_remove_camera_set": [_vp, _i],
    "asc_scene_add_frame": [_vp, _i, _P(_vp), _P(_i)],
    "asc_scene_remove_frame": [_vp, _i],
    "asc_scene_eval_requests": [_vp, _i, _P(_i), _P(_i), _P(_i), _P(_d), _P(_d), _P(_d)],
    "asc_scene_submit_requests": [_vp, _i, _P(_i), _P(_i), _P(_i), _P(_d), _P(_i)],
    "asc_scene_finish_requests": [_vp, _i, _P(_d), _P(_d)],
    "asc_scene_wait_requests": [_vp, _i],
    "asc_scene_use_graph": [_vp, _i],
    "asc_scene_set_debug": [_vp, _i],
    "asc_scene_last_info": [_vp, _P(_i), _i, _P(C.c_int64)],
    "asc_scene_set_timing": [_vp, _i],
    "asc_scene_set_tuning": [_vp, _i, _i],
    "asc_scene_last_timing": [_vp, _P(_f)],
    "asc_scene_last_moments": [_vp, _P(_d), _i],
    "asc_scene_last_images": [_vp, _i, _i, _P(_f), _P(_f), _P(_f)],
    "asc_scene_track_frame_plain": [_vp, _i, _P(_d), _P(_d), _P(_d)],
    "asc_scene_render_view": [_vp, _i, _P(_d), _i, _vp, _sz, _P(_u), _P(_u)],
    "asc_pose_store_set": [_P(_f), _P(_d)],
    "asc_pose_store_get": [_P(_f), _P(_d)],
    "asc_compose_pose": [_P(_f), _P(_d), _P(_d), _P(_d)],
    "asc_optimize_frame": [_vp, _P(_f), _P(AscOptimizeOptions), _P(AscOptimizeStats)],
    "asc_pso": [_vp, _vp, _u, _f, _f, _u, _u, _i, _i, _P(_f), _P(_f), _P(_i)],
    "asc_simplex": [_vp, _vp, _P(_d), _P(_d), _i, _d, _P(_i)],
    "asc_optimize_frame_fn": [_vp, _vp, _P(_f), _P(_d), _P(AscOptimizeOptions), _P(AscOptimizeStats)],
    "asc_frame_linear_extrap": [_P(_d), _P(_d), _P(_d)],
    "asc_server_create": [_vp, _sz, _i, _P(_vp)],
    "asc_server_set_pipeline": [_vp, _i],
    "asc_server_submit": [_vp, _P(AscJobDesc), _P(_i)],
    "asc_server_wait": [_vp, _i, _P(_d), _P(_d), _P(AscJobStats)],
    "asc_server_stats": [_vp, _P(AscServeStats)],
    "asc_server_lock_scene": [_vp, _P(_vp)],
    "asc_server_unlock_scene": [_vp],
    "asc_server_eval_requests": [_vp, _i, _P(_i), _P(_i), _P(_i), _P(_d), _P(_d), _P(_d)],
    "asc_server_stream_begin": [_vp],
    "asc_server_stream_end": [_vp],
    "asc_scene_scratch_bytes": [_vp, _P(C.c_size_t)],
    "asc_remote_create_shared": [_vp, _vp, _P(_vp)],
    "asc_server_destroy": [_vp],
    "asc_scene_autotune": [_vp, _i, _P(_i), _P(_d), _i, _P(AscTuneResult)],
    "asc_remote_create": [_vp, _P(_vp)],
    "asc_remote_listen": [_vp, _i],
    "asc_remote_port": [_vp, _P(_i)],
    "asc_remote_handle": [_vp, _vp, _sz, _vp, _sz, _P(_sz)],
    "asc_remote_destroy": [_vp],
    "asc_remote_set_output_dir": [_vp, C.c_char_p],
    "asc_tiff_info": [C.c_char_p, _P(_i), _P(_i), _P(_i), _P(_i)],
    "asc_tiff_read": [C.c_char_p, _vp, _sz],
    "asc_read_camera": [C.c_char_p, _P(AscCamera)],
    "asc_camera_from_mayacam1": [_P(_d), _P(AscCamera)],
    "asc_camera_from_vtkcam": [_P(_d), _P(_d), _P(_d), _d, _P(_d), _P(AscCamera)],
    "asc_read_filter_settings": [C.c_char_p, _P(AscFilterSettings)],
    "asc_drr_cutoff": [_d, _P(AscVolumeDesc), _P(_f)],
    "asc_trial_open": [C.c_char_p, _P(_vp)],
    "asc_trial_get_info": [_vp, _P(AscTrialInfo)],
    "asc_trial_camera": [_vp, _i, _P(AscCamera)],
    "asc_trial_volume_info": [_vp, _i, _P(_i), _P(_i), _P(_d), _P(_i)],
    "asc_trial_volume_read": [_vp, _i, _vp, _sz],
    "asc_trial_frame_info": [_vp, _i, _i, _P(_i), _P(_i)],
    "asc_trial_frame_read": [_vp, _i, _i, _vp, _sz],
    "asc_trial_frame_path": [_vp, _i, _i, C.c_char_p, _sz],
    "asc_trial_background": [_vp, _i, _P(_f), _sz],
    "asc_trial_close": [_vp],
    "asc_write_tra": [C.c_char_p, _P(_d), _i, _i, _i],
    "asc_write_png": [C.c_char_p, _vp, _i, _i, _i],
    "asc_write_view_png": [C.c_char_p, _P(_f), _i, _i],
    "asc_write_full_drr_pgm": [C.c_char_p, _P(_f), _i, _i],
    "asc_read_tra": [C.c_char_p, _i, _i, _i, _P(_d), _P(_i)],
}
_RES = {"asc_last_error": C.c_char_p, "asc_context_stream": _vp}

_LIB = None


def lib():
    """Load libautoscoper_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        path = Path(os.environ.get("ASC_LIB", LIB_PATH))
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with `python -m autoscoper_amd.build`")
        L = C.CDLL(str(path))
        L.asc_last_error.restype = C.c_char_p
        L.asc_last_error.argtypes = []
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RES.get(name, C.c_int)
        _LIB = L
    return _LIB


def exported_symbols() -> list[str]:
    return ["asc_last_error", *_SIGS.keys()]


def _check(rc: int) -> int:
    if rc < 0:
        raise AscError(rc, lib().asc_last_error().decode())
    return rc


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().asc_device_count(C.byref(n)))
    return n.value


# ---------------------------------------------------------------------------
# geometry (host, double precision)
# ---------------------------------------------------------------------------
class geometry:  # namespace
    @staticmethod
    def frame_from_xyzypr(p) -> np.ndarray:
        p = _f64(p)
        out = np.zeros(12)
        _check(lib().asc_frame_from_xyzypr(_ptr(p, _d), _ptr(out, _d)))
        return out

    @staticmethod
    def frame_from_xyz_axis_angle(p) -> np.ndarray:
        p = _f64(p)
        out = np.zeros(12)
        _check(lib().asc_frame_from_xyz_axis_angle(_ptr(p, _d), _ptr(out, _d)))
        return out

    @staticmethod
    def frame_to_xyzypr(f) -> np.ndarray:
        f = _f64(f)
        out = np.zeros(6)
        _check(lib().asc_frame_to_xyzypr(_ptr(f, _d), _ptr(out, _d)))
        return out

    @staticmethod
    def frame_mul(a, b) -> np.ndarray:
        a, b = _f64(a), _f64(b)
        out = np.zeros(12)
        _check(lib().asc_frame_mul(_ptr(a, _d), _ptr(b, _d), _ptr(out, _d)))
        return out

    @staticmethod
    def frame_inverse(a) -> np.ndarray:
        a = _f64(a)
        out = np.zeros(12)
        _check(lib().asc_frame_inverse(_ptr(a, _d), _ptr(out, _d)))
        return out

    @staticmethod
    def frame_linear_extrap(a, b) -> np.ndarray:
        a, b = _f64(a), _f64(b)
        out = np.zeros(12)
        _check(lib().asc_frame_linear_extrap(_ptr(a, _d), _ptr(b, _d), _ptr(out, _d)))
        return out

    @staticmethod
    def camera_from_mayacam2(K, R, t, size) -> AscCamera:
        K, R, t, size = _f64(K).ravel(), _f64(R).ravel(), _f64(t).ravel(), _f64(size).ravel()
        cam = AscCamera()
        _check(lib().asc_camera_from_mayacam2(_ptr(K, _d), _ptr(R, _d), _ptr(t, _d), _ptr(size, _d),
                                              C.byref(cam)))
        return cam

    @staticmethod
    def inv_model_view(imv16, inv_scale, inv_trans) -> np.ndarray:
        imv16, s, t = _f64(imv16), _f32(inv_scale), _f32(inv_trans)
        out = np.zeros(16, np.float32)
        _check(lib().asc_inv_model_view(_ptr(imv16, _d), _ptr(s, _f), _ptr(t, _f), _ptr(out, _f)))
        return out

    @staticmethod
    def view_geometry(cam: AscCamera, xyzypr, desc: AscVolumeDesc, rw: int, rh: int):
        """-> (in_view, imv[12] float32, viewport[4] float64, (w, h))"""
        p = _f64(xyzypr)
        imv = np.zeros(12, np.float32)
        vp = np.zeros(4)
        sz = np.zeros(2, np.uint32)
        rc = _check(lib().asc_view_geometry(C.byref(cam), _ptr(p, _d), C.byref(desc), rw, rh, _ptr(imv, _f),
                                            _ptr(vp, _d), _ptr(sz, _u)))
        return rc == 1, imv, vp, (int(sz[0]), int(sz[1]))

    @staticmethod
    def pose_store_set(xyzypr) -> np.ndarray:
        p = _f64(xyzypr)
        s = np.zeros(7, np.float32)
        _check(lib().asc_pose_store_set(_ptr(s, _f), _ptr(p, _d)))
        return s

    @staticmethod
    def pose_store_get(store) -> np.ndarray:
        s = _f32(store)
        p = np.zeros(6)
        _check(lib().asc_pose_store_get(_ptr(s, _f), _ptr(p, _d)))
        return p

    @staticmethod
    def compose_pose(store, pivot, manip) -> np.ndarray:
        s, pv, m = _f32(store), _f64(pivot), _f64(manip)
        out = np.zeros(6)
        _check(lib().asc_compose_pose(_ptr(s, _f), _ptr(pv, _d), _ptr(m, _d), _ptr(out, _d)))
        return out


def gaussian_weights(radius: float) -> np.ndarray:
    n = C.c_int(0)
    _check(lib().asc_gaussian_weights(radius, None, 0, C.byref(n)))
    out = np.zeros(n.value * n.value, np.float32)
    _check(lib().asc_gaussian_weights(radius, _ptr(out, _f), out.size, C.byref(n)))
    return out.reshape(n.value, n.value)


ERR_INVALID = -1  # ASC_ERR_INVALID
COST_FN = C.CFUNCTYPE(C.c_int, _P(_d), C.c_int, _P(_d), C.c_void_p)


def _cost_callback(cost, ndim):
    """cost: a Python callable mapping an (n, ndim) float64 array to n costs,
    or the address of a C function of the asc_cost_fn type."""
    if isinstance(cost, int):
        return COST_FN(cost)

    def cb(x, n, out, _user):
        try:
            pts = np.ctypeslib.as_array(x, (n * ndim,)).reshape(n, ndim).copy()
            res = np.asarray(cost(pts), np.float64).reshape(n)
            np.ctypeslib.as_array(out, (n,))[:] = res
            return 0
        except Exception:  # noqa: BLE001 - reported as a status to the optimiser
            return ERR_INVALID

    return COST_FN(cb)


class optimizers:  # namespace: the reference's optimisers on any cost function
    @staticmethod
    def pso(cost, seed, min_limit, max_limit, max_epochs, max_stall, particles=100, synchronous=False, user=None):
        """pso() of PSO.cpp:198-295 (asynchronous order unless synchronous):
        returns (best position float32[6], best cost, epochs)."""
        cb = _cost_callback(cost, 6)
        best = np.zeros(6, np.float32)
        bc = C.c_float()
        ep = C.c_int()
        _check(lib().asc_pso(C.cast(cb, _vp), user, seed, min_limit, max_limit, max_epochs, max_stall, particles,
                             int(bool(synchronous)), _ptr(best, _f), C.byref(bc), C.byref(ep)))
        return best, bc.value, ep.value

    @staticmethod
    def simplex(cost, P, Y, ftol, user=None):
        """AMOEBA of DownhillSimplex.cpp:39-172 on the start simplex P
        ((ndim+1) x ndim) with row costs Y: returns (P, Y, iterations)."""
        P = np.ascontiguousarray(P, np.float64).copy()
        Y = np.ascontiguousarray(Y, np.float64).copy()
        ndim = P.shape[1]
        cb = _cost_callback(cost, ndim)
        it = C.c_int()
        _check(lib().asc_simplex(C.cast(cb, _vp), user, _ptr(P, _d), _ptr(Y, _d), ndim, ftol, C.byref(it)))
        return P, Y, it.value


def optimize_options(method=0, max_iter=100, max_stall=25, min_limit=-3.0, max_limit=3.0, offsets=(0.1,) * 6,
                     seed=1, particles=100, synchronous=False, volume=0, speculative=False) -> "AscOptimizeOptions":
    o = AscOptimizeOptions()
    o.method, o.max_iter, o.max_stall = method, max_iter, max_stall
    o.min_limit, o.max_limit = min_limit, max_limit
    for k in range(6):
        o.offsets[k] = offsets[k]
    o.seed, o.particles, o.synchronous, o.volume = seed, particles, int(synchronous), volume
    o.speculative = int(speculative)
    return o


def optimize_frame_fn(cost, store, pivot, **options):
    """Tracker::optimize for one frame on a Python cost of manipulator
    offsets: cost(manips (n, 6), store) -> n costs, where `store` is the
    current float pose store (updated in place by the driver).
    Returns (final store, AscOptimizeStats)."""
    s = _f32(store).copy()
    pv = _f64(pivot)

    def cb(x, n, out, _user):
        try:
            pts = np.ctypeslib.as_array(x, (n * 6,)).reshape(n, 6).copy()
            np.ctypeslib.as_array(out, (n,))[:] = np.asarray(cost(pts, s.copy()), np.float64).reshape(n)
            return 0
        except Exception:  # noqa: BLE001 - reported to the driver as a status
            return ERR_INVALID

    fn = COST_FN(cb)
    st = AscOptimizeStats()
    o = optimize_options(**options)
    _check(lib().asc_optimize_frame_fn(C.cast(fn, _vp), None, _ptr(s, _f), _ptr(pv, _d), C.byref(o), C.byref(st)))
    return s, st


class Server:
    """Registration server (asc_server_*): continuous batching of tracking
    jobs on one GPU, admitted by a device-memory budget."""

    def __init__(self, ctx: "Context", budget_bytes: int, max_active_jobs: int = 8, pipeline: int = 2):
        h = C.c_void_p()
        _check(lib().asc_server_create(ctx.handle, budget_bytes, max_active_jobs, C.byref(h)))
        self.handle = h.value
        if pipeline != 2:
            self.set_pipeline(pipeline)
        self.ctx = ctx
        self._jobs = {}
        self._scene_keep = []  # volumes and images added to the server's scene by front ends

    def set_pipeline(self, depth: int):
        """Fused launches kept in flight: 2 queues the next while one runs,
        1 waits for each (asc_server_set_pipeline)."""
        _check(lib().asc_server_set_pipeline(self.handle, depth))

    @contextlib.contextmanager
    def scene(self):
        """The server's scene under asc_server_lock_scene, for a front end's
        volumes, cameras, camera sets and frames: a Scene that does not own
        its handle; the batcher starts no launch until the block ends (same
        thread), though one it started before may still run."""
        h = C.c_void_p()
        _check(lib().asc_server_lock_scene(self.handle, C.byref(h)))
        try:
            yield Scene._shared(h, self.ctx, self._scene_keep)
        finally:
            _check(lib().asc_server_unlock_scene(self.handle))

    def eval_requests(self, volumes, sets, frames, poses):
        """n requests (scene volume, camera set, frame or -1, xyzypr), in the
        batcher's next fused launch with every other pending request ->
        (costs[n], per_camera[n, 8])."""
        p = _f64(poses).reshape(-1, 6)
        n = p.shape[0]
        v = np.ascontiguousarray(volumes, np.int32)
        st = np.ascontiguousarray(sets, np.int32)
        fr = np.ascontiguousarray(frames if frames is not None else np.full(n, -1), np.int32)
        costs = np.zeros(n)
        per = np.zeros((n, 8))
        _check(lib().asc_server_eval_requests(self.handle, n, _ptr(v, _i), _ptr(st, _i), _ptr(fr, _i), _ptr(p, _d),
                                              _ptr(costs, _d), _ptr(per, _d)))
        return costs, per

    @contextlib.contextmanager
    def stream(self):
        """asc_server_stream_begin / _end around a stream of requests (an
        optimisation), so that the batcher waits for its next request."""
        _check(lib().asc_server_stream_begin(self.handle))
        try:
            yield
        finally:
            _check(lib().asc_server_stream_end(self.handle))

    def submit(self, volumes, voxel_mm, initial_poses, cameras, frames, render, drr_filters=(), rad_filters=(),
               drr=(0.5, 1.0, 0.0), storage=STORAGE_U16, guess=1, **options) -> int:
        """volumes: list of (D, H, W) u16 arrays; voxel_mm: per volume (float
        or 3-tuple); initial_poses: (n_volumes, 6); cameras: AscCamera list;
        frames: list over frames of per-camera 2-D arrays (u8, u16 or f32);
        render: (width, height); options: optimize_options keywords."""
        vols = [np.ascontiguousarray(v, np.uint16) for v in volumes]
        dims = np.array([[v.shape[2], v.shape[1], v.shape[0]] for v in vols], np.int32)
        vox = np.array([[x] * 3 if np.isscalar(x) else list(x) for x in voxel_mm], np.float64)
        init = np.ascontiguousarray(initial_poses, np.float64).reshape(len(vols), 6)
        cams = (AscCamera * len(cameras))(*cameras)
        first = np.asarray(frames[0][0])
        dt = {np.dtype(np.uint8): (np.uint8, 8), np.dtype(np.uint16): (np.uint16, 16),
              np.dtype(np.float32): (np.float32, 32)}[first.dtype]
        fr = [np.ascontiguousarray(im, dt[0]) for frame in frames for im in frame]
        fdims = np.array([[np.asarray(im).shape[1], np.asarray(im).shape[0]] for im in frames[0]], np.int32)
        dfl = (AscFilter * max(len(drr_filters), 1))(*drr_filters)
        rfl = (AscFilter * max(len(rad_filters), 1))(*rad_filters)
        d = AscJobDesc()
        d.n_volumes = len(vols)
        d.volumes = (C.c_void_p * len(vols))(*[v.ctypes.data for v in vols])
        d.volume_dims = _ptr(dims, _i)
        d.voxel_mm = _ptr(vox, _d)
        d.initial_poses = _ptr(init, _d)
        d.storage = storage
        d.drr_step, d.drr_intensity, d.drr_cutoff = drr
        d.n_cameras = len(cameras)
        d.cameras = cams
        d.render_width, d.render_height = render
        d.drr_filters, d.n_drr_filters = dfl, len(drr_filters)
        d.rad_filters, d.n_rad_filters = rfl, len(rad_filters)
        d.n_frames = len(frames)
        d.frames = (C.c_void_p * len(fr))(*[f.ctypes.data for f in fr])
        d.frame_dims = _ptr(fdims, _i)
        d.frame_bps = dt[1]
        d.options = optimize_options(**options)
        d.guess = guess
        jid = C.c_int()
        _check(lib().asc_server_submit(self.handle, C.byref(d), C.byref(jid)))
        self._jobs[jid.value] = (len(frames), len(vols))
        return jid.value

    def wait(self, job_id: int):
        """-> (poses (frames, volumes, 6), per-frame latency ms, AscJobStats)"""
        nf, nv = self._jobs.pop(job_id)
        poses = np.zeros((nf, nv, 6))
        lat = np.zeros(nf)
        st = AscJobStats()
        _check(lib().asc_server_wait(self.handle, job_id, _ptr(poses, _d), _ptr(lat, _d), C.byref(st)))
        return poses, lat, st

    def stats(self) -> dict:
        st = AscServeStats()
        _check(lib().asc_server_stats(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in AscServeStats._fields_}

    def close(self):
        if self.handle:
            lib().asc_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


# ---------------------------------------------------------------------------
# on-disk inputs and outputs (host only)
# ---------------------------------------------------------------------------
_UINT = {8: np.uint8, 16: np.uint16, 32: np.uint32}


def _path(p) -> bytes:
    return os.fsencode(str(p))


class io:  # namespace: the trial's files (TIFF, MayaCam/VTKCam, .vie, .tra)
    TRA_MATRIX, TRA_ROWS, TRA_COMMAS, TRA_CM, TRA_RADIANS = 1, 2, 4, 8, 16

    @staticmethod
    def tiff_info(path):
        w, h, n, b = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().asc_tiff_info(_path(path), C.byref(w), C.byref(h), C.byref(n), C.byref(b)))
        return w.value, h.value, n.value, b.value

    @staticmethod
    def read_tiff(path) -> np.ndarray:
        """All pages of a TIFF file as (pages, height, width) unsigned ints."""
        w, h, n, b = io.tiff_info(path)
        out = np.empty((n, h, w), _UINT[b])
        _check(lib().asc_tiff_read(_path(path), out.ctypes.data_as(_vp), out.nbytes))
        return out

    @staticmethod
    def read_camera(path) -> AscCamera:
        c = AscCamera()
        _check(lib().asc_read_camera(_path(path), C.byref(c)))
        return c

    @staticmethod
    def camera_from_mayacam1(values) -> AscCamera:
        v = _f64(values).reshape(15)
        c = AscCamera()
        _check(lib().asc_camera_from_mayacam1(_ptr(v, _d), C.byref(c)))
        return c

    @staticmethod
    def camera_from_vtkcam(position, focal_point, view_up, view_angle_deg, size) -> AscCamera:
        p, f, u, s = _f64(position), _f64(focal_point), _f64(view_up), _f64(size)
        c = AscCamera()
        _check(lib().asc_camera_from_vtkcam(_ptr(p, _d), _ptr(f, _d), _ptr(u, _d), float(view_angle_deg),
                                            _ptr(s, _d), C.byref(c)))
        return c

    @staticmethod
    def read_filter_settings(path) -> dict:
        s = AscFilterSettings()
        _check(lib().asc_read_filter_settings(_path(path), C.byref(s)))
        return {"drr_filters": [s.drr_filters[k] for k in range(s.n_drr_filters)],
                "rad_filters": [s.rad_filters[k] for k in range(s.n_rad_filters)],
                "sample_distance": s.sample_distance, "ray_intensity": s.ray_intensity,
                "cutoff_fraction": s.cutoff_fraction, "drr_step": s.drr_step, "drr_intensity": s.drr_intensity,
                "sections": (bool(s.has_drr_renderer), bool(s.has_drr_filters), bool(s.has_rad_filters))}

    @staticmethod
    def drr_cutoff(fraction, desc) -> float:
        out = C.c_float()
        _check(lib().asc_drr_cutoff(float(fraction), C.byref(desc), C.byref(out)))
        return out.value

    @staticmethod
    def write_tra(path, poses, flags=4) -> None:
        """poses (frames, volumes, 6) xyzypr; default flags = comma separated
        x y z yaw pitch roll."""
        p = _f64(poses)
        _check(lib().asc_write_tra(_path(path), _ptr(p, _d), p.shape[0], p.shape[1], int(flags)))

    @staticmethod
    def read_tra(path, n_frames, n_volumes, flags=4) -> np.ndarray:
        out = np.empty((n_frames, n_volumes, 6))
        got = C.c_int()
        _check(lib().asc_read_tra(_path(path), n_frames, n_volumes, int(flags), _ptr(out, _d), C.byref(got)))
        return out

    @staticmethod
    def write_png(path, pixels) -> None:
        """8-bit (H, W) greyscale or (H, W, 3) RGB, top row first."""
        p = np.ascontiguousarray(pixels, np.uint8)
        ch = 1 if p.ndim == 2 else p.shape[2]
        _check(lib().asc_write_png(_path(path), p.ctypes.data_as(_vp), p.shape[1], p.shape[0], ch))

    @staticmethod
    def write_view_png(path, rgb) -> None:
        """A Scene.render_view composite (H, W, 3) float, row 0 at the bottom
        as GL draws it: clamped, rounded to 8 bits, written top row first."""
        p = _f32(rgb)
        _check(lib().asc_write_view_png(_path(path), _ptr(p, _f), p.shape[1], p.shape[0]))

    @staticmethod
    def write_full_drr_pgm(path, image) -> None:
        """save_full_drr's ASCII PGM of an (H, W) float image."""
        p = _f32(image)
        _check(lib().asc_write_full_drr_pgm(_path(path), _ptr(p, _f), p.shape[1], p.shape[0]))


class Trial:
    """A trial configuration (.cfg): cameras, volumes and radiograph frames
    (xromm::Trial)."""

    def __init__(self, cfg_path):
        h = _vp()
        _check(lib().asc_trial_open(_path(cfg_path), C.byref(h)))
        self.h = h
        info = AscTrialInfo()
        _check(lib().asc_trial_get_info(h, C.byref(info)))
        self.version = tuple(info.version)
        self.n_cameras, self.n_volumes, self.n_frames = info.n_cameras, info.n_volumes, info.n_frames
        self.render = (info.render_width, info.render_height)
        self.offsets = np.array(info.offsets[:])
        self.cameras = []
        for i in range(self.n_cameras):
            c = AscCamera()
            _check(lib().asc_trial_camera(h, i, C.byref(c)))
            self.cameras.append(c)

    def volume(self, i) -> dict:
        """{"data": (D, H, W) unsigned ints, "voxel_mm": (3,), "flip": (3,)}"""
        dims, bps = (C.c_int * 3)(), C.c_int()
        vox, flip = (C.c_double * 3)(), (C.c_int * 3)()
        _check(lib().asc_trial_volume_info(self.h, i, dims, C.byref(bps), vox, flip))
        out = np.empty((dims[2], dims[1], dims[0]), _UINT[bps.value])
        _check(lib().asc_trial_volume_read(self.h, i, out.ctypes.data_as(_vp), out.nbytes))
        return {"data": out, "voxel_mm": np.array(vox[:]), "flip": tuple(flip[:])}

    def frame(self, camera, frame) -> np.ndarray:
        dims, bps = (C.c_int * 2)(), C.c_int()
        _check(lib().asc_trial_frame_info(self.h, camera, frame, dims, C.byref(bps)))
        out = np.empty((dims[1], dims[0]), _UINT[bps.value])
        _check(lib().asc_trial_frame_read(self.h, camera, frame, out.ctypes.data_as(_vp), out.nbytes))
        return out

    def frame_path(self, camera, frame) -> str:
        buf = C.create_string_buffer(4096)
        _check(lib().asc_trial_frame_path(self.h, camera, frame, buf, len(buf)))
        return os.fsdecode(buf.value)

    def background(self, camera) -> np.ndarray:
        h, w = self.frame(camera, 0).shape
        out = np.empty((h, w), np.float32)
        _check(lib().asc_trial_background(self.h, camera, _ptr(out, _f), out.size))
        return out

    def close(self):
        if self.h:
            lib().asc_trial_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Remote:
    """A session answering the reference's socket protocol (version 2):
    handle() one message directly, or listen() on 127.0.0.1."""

    def __init__(self, ctx: "Context | None" = None, server: "Server | None" = None):
        """server: answer costs and optimisations through this registration
        server's batcher (launches shared with its other clients and jobs)."""
        h = _vp()
        if server is not None:
            _check(lib().asc_remote_create_shared(ctx.handle, server.handle, C.byref(h)))
        else:
            _check(lib().asc_remote_create(ctx.handle if ctx else None, C.byref(h)))
        self.h = h
        self._ctx = ctx
        self._server = server

    def handle(self, message: bytes) -> bytes:
        n = C.c_size_t()
        buf = C.create_string_buffer(1 << 22)
        rc = lib().asc_remote_handle(self.h, message, len(message), buf, len(buf), C.byref(n))
        _check(rc)
        self.closed = rc == 1  # message 13 asks to close the connection
        return buf.raw[:n.value]

    def listen(self, port: int = 0) -> int:
        _check(lib().asc_remote_listen(self.h, port))
        p = C.c_int()
        _check(lib().asc_remote_port(self.h, C.byref(p)))
        return p.value

    def set_output_dir(self, path):
        """Where message 12 (saveFullDRR) writes image_camNN.pgm."""
        _check(lib().asc_remote_set_output_dir(self.h, _path(path)))

    def close(self):
        if self.h:
            lib().asc_remote_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class filters:  # namespace: filter records
    @staticmethod
    def sobel(scale=1.0, blend=0.5, enabled=True) -> AscFilter:
        return AscFilter(FILTER_SOBEL, int(enabled), (C.c_float * 4)(scale, blend, 0, 0))

    @staticmethod
    def contrast(alpha=1.0, beta=1.0, size=3, enabled=True) -> AscFilter:
        return AscFilter(FILTER_CONTRAST, int(enabled), (C.c_float * 4)(alpha, beta, size, 0))

    @staticmethod
    def gaussian(radius=1.0, enabled=True) -> AscFilter:
        return AscFilter(FILTER_GAUSSIAN, int(enabled), (C.c_float * 4)(radius, 0, 0, 0))

    @staticmethod
    def sharpen(radius=1.0, contrast=1.0, threshold=0.0, enabled=True) -> AscFilter:
        return AscFilter(FILTER_SHARPEN, int(enabled), (C.c_float * 4)(radius, contrast, threshold, 0))

    @staticmethod
    def to_records(fs) -> np.ndarray:
        """6-float records {type, enabled, p0..p3} (the oracle's layout)."""
        out = np.zeros((len(fs), 6), np.float32)
        for k, f in enumerate(fs):
            out[k] = [f.type, f.enabled, *f.p]
        return out


def _filter_array(fs):
    arr = (AscFilter * max(1, len(fs)))(*fs)
    return arr


# ---------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------
class DeviceArray:
    """Device buffer owned by a Context (Buffer in the reference facade)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().asc_malloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().asc_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data_as(_vp), a.nbytes))
        return self

    def download(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().asc_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(_vp), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr and self.ptr.value:
            lib().asc_free(self.ctx.handle, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().asc_context_create(device, C.byref(h)))
        self.handle = h

    def set_stream(self, stream_ptr: int | None):
        _check(lib().asc_context_set_stream(self.handle, C.c_void_p(stream_ptr or 0)))

    @property
    def stream(self) -> int:
        return lib().asc_context_stream(self.handle) or 0

    def set_ncc_form(self, form: int):
        """NCC_MOMENTS (default), or the reference's float tree form
        NCC_TREE_CUDA / NCC_TREE_OPENCL, for ncc() and the C++ facade."""
        _check(lib().asc_context_set_ncc_form(self.handle, int(form)))

    def set_blocking(self, enable: bool):
        """Make every per-operator call wait for its launch before returning
        (the reference's OpenCL Kernel::launch finishes each kernel)."""
        _check(lib().asc_context_set_blocking(self.handle, int(bool(enable))))

    def synchronize(self):
        _check(lib().asc_synchronize(self.handle))

    def alloc(self, nbytes: int) -> DeviceArray:
        return DeviceArray(self, nbytes)

    def to_device(self, a: np.ndarray) -> DeviceArray:
        a = np.ascontiguousarray(a)
        return DeviceArray(self, max(a.nbytes, 4)).upload(a)

    def close(self):
        if self.handle:
            lib().asc_context_destroy(self.handle)
            self.handle = None

    def copy(self, dst: DeviceArray, src: DeviceArray, nbytes: int):
        _check(lib().asc_memcpy_d2d(self.handle, dst.ptr, src.ptr, nbytes))

    # -- per-operator calls on device arrays (raw pointers) --
    def drr_render(self, vol: "Volume", imv, viewport, step, intensity, cutoff, out: DeviceArray, w, h):
        imv, vp = _f32(imv), _f32(viewport)
        _check(lib().asc_drr_render(self.handle, vol.handle, _ptr(imv, _f), _ptr(vp, _f), step, intensity, cutoff,
                                    out.ptr, w, h))

    def rad_render(self, img: "Image", plane, viewport, out: DeviceArray, w, h):
        pl, vp = _f32(plane), _f32(viewport)
        _check(lib().asc_rad_render(self.handle, img.handle, _ptr(pl, _f), _ptr(vp, _f), out.ptr, w, h))

    def background_render(self, img: "Image", plane, viewport, threshold, out: DeviceArray, w, h):
        pl, vp = _f32(plane), _f32(viewport)
        _check(lib().asc_background_render(self.handle, img.handle, _ptr(pl, _f), _ptr(vp, _f), threshold,
                                           out.ptr, w, h))

    def filter_apply(self, f: AscFilter, src: DeviceArray, dst: DeviceArray, w, h):
        _check(lib().asc_filter_apply(self.handle, C.byref(f), src.ptr, dst.ptr, w, h))

    def filter_chain(self, fs, src: DeviceArray, dst: DeviceArray, scratch: DeviceArray | None, w, h):
        arr = _filter_array(fs)
        _check(lib().asc_filter_chain(self.handle, arr, len(fs), src.ptr, dst.ptr,
                                      scratch.ptr if scratch else None, w, h))

    def drr_background(self, src, dst, w, h):
        _check(lib().asc_drr_background(self.handle, src.ptr, dst.ptr, w, h))

    def multiply(self, a, b, dst, w, h):
        _check(lib().asc_multiply(self.handle, a.ptr, b.ptr, dst.ptr, w, h))

    def merge(self, a, b, dst, w, h):
        _check(lib().asc_merge(self.handle, a.ptr, b.ptr, dst.ptr, w, h))

    def composite(self, s1, s2, s3, s4, dst, w, h):
        _check(lib().asc_composite(self.handle, s1.ptr, s2.ptr, s3.ptr, s4.ptr if s4 else None, dst.ptr, w, h))

    def mask_multiply(self, drr, rad, bg, mask, w, h):
        _check(lib().asc_mask_multiply(self.handle, drr.ptr, rad.ptr, bg.ptr if bg else None,
                                       mask.ptr if mask else None, w, h))

    def ncc(self, f: DeviceArray, g: DeviceArray, m: DeviceArray, n: int):
        """-> (ncc, moments[9])"""
        v = C.c_double(0)
        mom = np.zeros(9)
        _check(lib().asc_ncc(self.handle, f.ptr, g.ptr, m.ptr, n, C.byref(v), _ptr(mom, _d)))
        return v.value, mom

    def ncc_noalloc(self, f: DeviceArray, g: DeviceArray, m: DeviceArray, n: int, out10: DeviceArray):
        """asc_ncc_noalloc: moments then NCC into the device buffer out10
        (10 doubles), queued on the context's stream; nothing allocated."""
        _check(lib().asc_ncc_noalloc(self.handle, f.ptr, g.ptr, m.ptr, n, out10.ptr))


class Volume:
    def __init__(self, ctx: Context, data: np.ndarray, scale=(1.0, 1.0, 1.0), flip=(0, 0, 0),
                 storage: int = STORAGE_U16):
        """data: (depth, height, width) uint8/uint16 array, x fastest."""
        data = np.ascontiguousarray(data)
        assert data.ndim == 3 and data.dtype in (np.uint8, np.uint16)
        bps = 8 * data.dtype.itemsize
        d, h, w = data.shape
        fl = np.asarray(flip, np.int32)
        sc = _f64(scale)
        self.desc = AscVolumeDesc()
        hnd = C.c_void_p()
        _check(lib().asc_volume_create(ctx.handle, data.ctypes.data_as(_vp), w, h, d, bps, _ptr(fl, _i),
                                       _ptr(sc, _d), storage, C.byref(hnd), C.byref(self.desc)))
        self.handle = hnd
        self.ctx = ctx
        self.scale = tuple(float(s) for s in scale)

    LINEAR, QUAD, OCTET, BRICK = 0, 1, 2, 3

    def set_layout(self, layout: int):
        _check(lib().asc_volume_set_layout(self.handle, layout))

    @property
    def nbytes(self) -> int:
        n = C.c_size_t(0)
        _check(lib().asc_volume_bytes(self.handle, C.byref(n)))
        return n.value

    def close(self):
        if self.handle:
            lib().asc_volume_destroy(self.handle)
            self.handle = None


class Image:
    def __init__(self, ctx: Context, data: np.ndarray):
        """2-D radiograph/background: uint8, uint16 (UNORM) or float32."""
        data = np.ascontiguousarray(data)
        bps = {np.dtype(np.uint8): 8, np.dtype(np.uint16): 16, np.dtype(np.float32): 32}[data.dtype]
        h, w = data.shape
        hnd = C.c_void_p()
        _check(lib().asc_image_create(ctx.handle, data.ctypes.data_as(_vp), w, h, bps, C.byref(hnd)))
        self.handle = hnd
        self.shape = (h, w)
        self.dtype = data.dtype

    def update(self, data: np.ndarray):
        data = np.ascontiguousarray(data, dtype=self.dtype)
        assert data.shape == self.shape
        _check(lib().asc_image_update(self.handle, data.ctypes.data_as(_vp)))

    def close(self):
        if self.handle:
            lib().asc_image_destroy(self.handle)
            self.handle = None


class Scene:
    """Tracker::load state + batched evaluation (Tracker.cpp:244-335, 583-827)."""

    def __init__(self, ctx: Context, render_width: int, render_height: int):
        h = C.c_void_p()
        _check(lib().asc_scene_create(ctx.handle, render_width, render_height, C.byref(h)))
        self.handle = h
        self.ctx = ctx
        self.ncam = 0
        self.nvol = 0
        self.render = (render_width, render_height)
        self._keep = []
        self._owned = True

    @classmethod
    def _shared(cls, handle, ctx: Context, keep: list) -> "Scene":
        """A scene owned elsewhere (a registration server's): close() leaves it,
        and the objects added through it live in `keep`."""
        sc = cls.__new__(cls)
        sc.handle = handle
        sc.ctx = ctx
        sc.ncam = 0
        sc.nvol = 0
        sc.render = (0, 0)
        sc._keep = keep
        sc._owned = False
        return sc

    def add_volume(self, vol: Volume, step: float, intensity: float, cutoff: float, pivot=None) -> int:
        idx = C.c_int()
        pv = None if pivot is None else _f64(pivot)
        _check(lib().asc_scene_add_volume(self.handle, vol.handle, step, intensity, cutoff,
                                          None if pv is None else _ptr(pv, _d), C.byref(idx)))
        self._keep.append(vol)
        self.nvol += 1
        return idx.value

    def set_drr_params(self, volume: int, step: float, intensity: float, cutoff: float):
        _check(lib().asc_scene_set_drr_params(self.handle, volume, step, intensity, cutoff))

    def pivot(self, volume: int) -> np.ndarray:
        out = np.zeros(12)
        _check(lib().asc_scene_volume_pivot(self.handle, volume, _ptr(out, _d)))
        return out

    def add_camera(self, cam: AscCamera, rad: "Image | None") -> int:
        idx = C.c_int()
        _check(lib().asc_scene_add_camera(self.handle, C.byref(cam), rad.handle if rad else None, C.byref(idx)))
        if rad:
            self._keep.append(rad)
        self.ncam += 1
        return idx.value

    def remove_volume(self, volume: int):
        _check(lib().asc_scene_remove_volume(self.handle, volume))
        self.nvol -= 1

    def remove_camera(self, camera: int):
        _check(lib().asc_scene_remove_camera(self.handle, camera))
        self.ncam -= 1

    def add_camera_set(self, cameras, render_w: int, render_h: int) -> int:
        """One job's views (scene camera indices) and its render size."""
        cams = np.ascontiguousarray(cameras, np.int32)
        out = C.c_int()
        _check(lib().asc_scene_add_camera_set(self.handle, _ptr(cams, _i), len(cams), render_w, render_h,
                                              C.byref(out)))
        return out.value

    def remove_camera_set(self, camera_set: int):
        _check(lib().asc_scene_remove_camera_set(self.handle, camera_set))

    def add_frame(self, camera_set: int, rads) -> int:
        """One radiograph per camera of the set (Image objects, set order)."""
        arr = (C.c_void_p * len(rads))(*[r.handle for r in rads])
        out = C.c_int()
        _check(lib().asc_scene_add_frame(self.handle, camera_set, arr, C.byref(out)))
        self._keep.extend(rads)
        return out.value

    def remove_frame(self, frame: int):
        _check(lib().asc_scene_remove_frame(self.handle, frame))

    def eval_requests(self, volumes, sets, frames, poses):
        """n requests (volume, camera set, frame or -1, xyzypr) in one fused
        launch -> (costs[n], per_camera[n, 8])."""
        p = _f64(poses).reshape(-1, 6)
        n = p.shape[0]
        v = np.ascontiguousarray(volumes, np.int32)
        st = np.ascontiguousarray(sets, np.int32)
        fr = np.ascontiguousarray(frames if frames is not None else np.full(n, -1), np.int32)
        costs = np.zeros(n)
        per = np.zeros((n, 8))
        _check(lib().asc_scene_eval_requests(self.handle, n, _ptr(v, _i), _ptr(st, _i), _ptr(fr, _i), _ptr(p, _d),
                                             _ptr(costs, _d), _ptr(per, _d)))
        return costs, per

    def submit_requests(self, volumes, sets, frames, poses) -> tuple[int, int]:
        """Queue the fused launches of eval_requests' batch behind any batch
        already submitted -> (ticket, n); at most two batches in flight."""
        p = _f64(poses).reshape(-1, 6)
        n = p.shape[0]
        v = np.ascontiguousarray(volumes, np.int32)
        st = np.ascontiguousarray(sets, np.int32)
        fr = np.ascontiguousarray(frames if frames is not None else np.full(n, -1), np.int32)
        t = C.c_int(-1)
        _check(lib().asc_scene_submit_requests(self.handle, n, _ptr(v, _i), _ptr(st, _i), _ptr(fr, _i), _ptr(p, _d),
                                               C.byref(t)))
        return t.value, n

    def finish_requests(self, ticket: tuple[int, int]):
        """Wait for a submitted batch -> (costs[n], per_camera[n, 8])."""
        slot, n = ticket
        costs = np.zeros(n)
        per = np.zeros((n, 8))
        _check(lib().asc_scene_finish_requests(self.handle, slot, _ptr(costs, _d), _ptr(per, _d)))
        return costs, per

    def set_rad(self, camera: int, rad: Image):
        _check(lib().asc_scene_set_rad(self.handle, camera, rad.handle))
        self._keep.append(rad)

    def set_background(self, camera: int, bg: Image | None, threshold: float):
        _check(lib().asc_scene_set_background(self.handle, camera, bg.handle if bg else None, threshold))
        if bg:
            self._keep.append(bg)

    def set_filters(self, camera: int, which: int, fs):
        arr = _filter_array(fs)
        _check(lib().asc_scene_set_filters(self.handle, camera, which, arr, len(fs)))

    def use_graph(self, enable: bool):
        _check(lib().asc_scene_use_graph(self.handle, int(enable)))

    def set_debug(self, enable: bool):
        _check(lib().asc_scene_set_debug(self.handle, int(enable)))

    def eval_poses(self, poses, volumes=None):
        """poses (n, 6) xyzypr -> (costs[n], per_camera[n, ncam])"""
        p = _f64(poses).reshape(-1, 6)
        n = p.shape[0]
        vols = np.zeros(n, np.int32) if volumes is None else np.ascontiguousarray(volumes, np.int32)
        costs = np.zeros(n)
        per = np.zeros((n, self.ncam))
        _check(lib().asc_scene_eval_poses(self.handle, _ptr(vols, _i), _ptr(p, _d), n, _ptr(costs, _d),
                                          _ptr(per, _d)))
        return costs, per

    def eval_manips(self, stores, manips, volumes=None):
        m = _f64(manips).reshape(-1, 6)
        n = m.shape[0]
        s = _f32(stores).reshape(-1, 7)
        if s.shape[0] == 1 and n > 1:
            s = np.ascontiguousarray(np.repeat(s, n, axis=0))
        vols = np.zeros(n, np.int32) if volumes is None else np.ascontiguousarray(volumes, np.int32)
        costs = np.zeros(n)
        per = np.zeros((n, self.ncam))
        _check(lib().asc_scene_eval_manips(self.handle, _ptr(vols, _i), _ptr(s, _f), _ptr(m, _d), n,
                                           _ptr(costs, _d), _ptr(per, _d)))
        return costs, per

    TUNE_UNROLL, TUNE_GRID, TUNE_WAVE_PATCH, TUNE_GATHER_PROBE = 0, 1, 2, 3

    def autotune(self, poses, volumes=None, reps=3) -> dict:
        """Time the launch-shape variants on this batch of poses and keep the
        fastest (asc_scene_autotune)."""
        p = _f64(poses).reshape(-1, 6)
        n = p.shape[0]
        vols = np.zeros(n, np.int32) if volumes is None else np.ascontiguousarray(volumes, np.int32)
        r = AscTuneResult()
        _check(lib().asc_scene_autotune(self.handle, n, _ptr(vols, _i), _ptr(p, _d), int(reps), C.byref(r)))
        return {"unroll": r.unroll, "wave_patch": r.wave_patch, "grid": r.grid, "best_ms": r.best_ms,
                "default_ms": r.default_ms, "candidates": r.candidates}

    def set_tuning(self, key: int, value: int):
        _check(lib().asc_scene_set_tuning(self.handle, key, value))

    def set_timing(self, enable: bool):
        _check(lib().asc_scene_set_timing(self.handle, int(enable)))

    def last_timing(self) -> np.ndarray:
        """ms of {ray march, filters+moments, finalize} of the last eager batch"""
        ms = np.zeros(3, np.float32)
        _check(lib().asc_scene_last_timing(self.handle, _ptr(ms, _f)))
        return ms

    def last_info(self, n_items: int):
        info = np.zeros((n_items, 8), np.int32)
        samples = C.c_int64(0)
        _check(lib().asc_scene_last_info(self.handle, _ptr(info, _i), info.size, C.byref(samples)))
        return info, samples.value

    def last_moments(self, n_items: int) -> np.ndarray:
        m = np.zeros((n_items, 10))
        _check(lib().asc_scene_last_moments(self.handle, _ptr(m, _d), m.size))
        return m

    def last_images(self, pose: int, camera: int, w: int, h: int):
        d = np.zeros((h, w), np.float32)
        r = np.zeros((h, w), np.float32)
        m = np.zeros((h, w), np.float32)
        _check(lib().asc_scene_last_images(self.handle, pose, camera, _ptr(d, _f), _ptr(r, _f), _ptr(m, _f)))
        return d, r, m

    def render_view(self, camera: int, poses) -> np.ndarray:
        """View::render of one camera over its whole image plane with every
        volume at its pose (poses: one xyzypr per volume slot): the
        Compositor's RGB, (H, W, 3) float32, row 0 as rendered."""
        p = _f64(np.atleast_2d(poses))
        w, h = C.c_uint(), C.c_uint()
        _check(lib().asc_scene_render_view(self.handle, camera, _ptr(p, _d), p.shape[0], None, 0, C.byref(w),
                                           C.byref(h)))
        n = 3 * w.value * h.value
        buf = self.ctx.alloc(4 * n)
        try:
            _check(lib().asc_scene_render_view(self.handle, camera, _ptr(p, _d), p.shape[0], buf.ptr, n, C.byref(w),
                                               C.byref(h)))
            return buf.download((h.value, w.value, 3))
        finally:
            buf.free()

    def track_frame_plain(self, pose, volume: int = 0):
        p = _f64(pose)
        per = np.zeros(self.ncam)
        cost = C.c_double(0)
        _check(lib().asc_scene_track_frame_plain(self.handle, volume, _ptr(p, _d), _ptr(per, _d), C.byref(cost)))
        return cost.value, per

    def optimize_frame(self, store, method=0, max_iter=100, max_stall=25, min_limit=-3.0, max_limit=3.0,
                       offsets=(0.1,) * 6, seed=1, particles=100, synchronous=False, volume=0, speculative=False):
        s = _f32(store).copy()
        o = AscOptimizeOptions()
        o.method, o.max_iter, o.max_stall = method, max_iter, max_stall
        o.min_limit, o.max_limit = min_limit, max_limit
        for k in range(6):
            o.offsets[k] = offsets[k]
        o.seed, o.particles, o.synchronous, o.volume = seed, particles, int(synchronous), volume
        o.speculative = int(speculative)
        st = AscOptimizeStats()
        _check(lib().asc_optimize_frame(self.handle, _ptr(s, _f), C.byref(o), C.byref(st)))
        return s, st

    def close(self):
        if self.handle and self._owned:
            lib().asc_scene_destroy(self.handle)
        self.handle = None
