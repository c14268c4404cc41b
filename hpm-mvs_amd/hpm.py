"""ctypes bindings of libhpm_mvs.so (include/hpm_capi.h, hpm_io.h, hpm_scene.h).

`Problem` mirrors the reference's per-problem class ``HPM``
(HPM-MVS/HPM.h:48-112) method for method; numpy arrays cross the boundary
as plain host buffers. The library is the HIP path: if it is not built,
loading fails loudly (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libhpm_mvs.so")

HPM_MAX_SOURCES = 32
HPM_MAX_IMAGES = HPM_MAX_SOURCES + 1


class HpmCamera(C.Structure):
    """`Camera` (HPM-MVS/main.h:44-52)."""

    _fields_ = [
        ("K", C.c_float * 9),
        ("R", C.c_float * 9),
        ("t", C.c_float * 3),
        ("height", C.c_int32),
        ("width", C.c_int32),
        ("depth_min", C.c_float),
        ("depth_max", C.c_float),
    ]


class HpmParams(C.Structure):
    """`PatchMatchParams` (HPM-MVS/HPM.h:29-46) plus settings the reference hard-codes."""

    _fields_ = [
        ("max_iterations", C.c_int32),
        ("patch_size", C.c_int32),
        ("num_images", C.c_int32),
        ("max_image_size", C.c_int32),
        ("radius_increment", C.c_int32),
        ("sigma_spatial", C.c_float),
        ("sigma_color", C.c_float),
        ("top_k", C.c_int32),
        ("baseline", C.c_float),
        ("depth_min", C.c_float),
        ("depth_max", C.c_float),
        ("disparity_min", C.c_float),
        ("disparity_max", C.c_float),
        ("geom_consistency", C.c_int32),
        ("multi_geometry", C.c_int32),
        ("planar_prior", C.c_int32),
        ("knn_k", C.c_int32),
        ("geom_iterations", C.c_int32),
        ("seed_lo", C.c_uint32),
        ("seed_hi", C.c_uint32),
        ("acmm", C.c_int32),
        ("hierarchy", C.c_int32),
    ]


def default_params() -> HpmParams:
    """hpm_params_default (include/hpm_types.h)."""
    p = HpmParams()
    p.max_iterations = 3
    p.patch_size = 11
    p.num_images = 5
    p.max_image_size = 3200
    p.radius_increment = 2
    p.sigma_spatial = 5.0
    p.sigma_color = 3.0
    p.top_k = 4
    p.baseline = 0.54
    p.depth_min = 0.0
    p.depth_max = 1.0
    p.disparity_min = 0.0
    p.disparity_max = 1.0
    p.geom_consistency = 0
    p.multi_geometry = 0
    p.planar_prior = 0
    p.knn_k = 6
    p.geom_iterations = 2
    p.seed_lo = 0x5EED0001
    p.seed_hi = 0x00C0FFEE
    return p


class HpmRunStats(C.Structure):
    _fields_ = [
        ("n_eval", C.c_uint64),
        ("n_eval_propagate", C.c_uint64),
        ("ms_total", C.c_float),
        ("ms_init", C.c_float),
        ("ms_propagate", C.c_float),
        ("ms_finalize_filter", C.c_float),
        ("launches", C.c_int32),
    ]


class HpmSceneConfig(C.Structure):
    """hpm_scene_config (include/hpm_scene.h)."""

    _fields_ = [
        ("seed", C.c_uint32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("fx", C.c_float),
        ("fy", C.c_float),
        ("cx", C.c_float),
        ("cy", C.c_float),
        ("num_views", C.c_int32),
        ("arc_length", C.c_float),
        ("look_depth", C.c_float),
        ("num_planes", C.c_int32),
        ("num_boxes", C.c_int32),
        ("low_texture_fraction", C.c_float),
        ("amp_high", C.c_float),
        ("amp_low", C.c_float),
        ("noise_sigma", C.c_float),
        ("pitch_px", C.c_float),
        ("depth_min", C.c_float),
        ("depth_max", C.c_float),
    ]


class HpmPipelineOptions(C.Structure):
    """hpm_pipeline_options (include/hpm_pipeline.h)."""

    _fields_ = [
        ("params", HpmParams),
        ("planar_prior", C.c_int32),
        ("geom_consistency", C.c_int32),
        ("num_threads", C.c_int32),
        ("verbose", C.c_int32),
        ("scale_bound", C.c_int32),
        ("workers", C.c_int32),
        ("learned_prior", C.c_char_p),
        ("learned_min_accept", C.c_float),
        ("hierarchy_from", C.c_void_p),
        ("stream_base", C.c_int32),
        ("prior_on_device", C.c_int32),
        ("batch", C.c_void_p),
    ]


class HpmAcmmOptions(C.Structure):
    """hpm_acmm_options (include/hpm_acmm.h)."""

    _fields_ = [
        ("params", HpmParams),
        ("scale_bound", C.c_int32),
        ("geom_iterations", C.c_int32),
        ("workers", C.c_int32),
        ("verbose", C.c_int32),
    ]


class HpmPipelineStats(C.Structure):
    _fields_ = [
        ("ms_wall", C.c_double),
        ("ms_phase_a", C.c_double),
        ("ms_phase_b", C.c_double),
        ("ms_prior_host", C.c_double),
        ("ms_prior_device", C.c_double),
        ("ms_device", C.c_double),
        ("n_eval", C.c_uint64),
        ("n_eval_propagate", C.c_uint64),
        ("ms_propagate", C.c_double),
        ("runs", C.c_int32),
        ("propagate_launches", C.c_int32),
        ("learned_views", C.c_int32),
        ("learned_fallbacks", C.c_int32),
        ("learned_accept", C.c_double),
        ("batches", C.c_int32),
    ]


class HpmBatchStats(C.Structure):
    _fields_ = [
        ("ms_device", C.c_double),
        ("ms_host", C.c_double),
        ("n_eval", C.c_uint64),
        ("problems", C.c_int32),
        ("tiles_full", C.c_int32),
        ("tiles_colour", C.c_int32),
        ("grid", C.c_int32),
        ("graph_reused", C.c_int32),
        ("items", C.c_int32),
        ("persistent", C.c_int32),
    ]


_lib: Optional[C.CDLL] = None

_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)
_UP = C.POINTER(C.c_uint32)
_BP = C.POINTER(C.c_uint8)

# name -> (restype, argtypes); every symbol include/*.h declares for the library
SIGNATURES = {
    "hpm_create": (C.c_void_p, [C.c_int]),
    "hpm_destroy": (None, [C.c_void_p]),
    "hpm_last_error": (C.c_char_p, [C.c_void_p]),
    "hpm_set_params": (C.c_int, [C.c_void_p, C.POINTER(HpmParams)]),
    "hpm_get_params": (C.c_int, [C.c_void_p, C.POINTER(HpmParams)]),
    "hpm_set_geom_consistency": (C.c_int, [C.c_void_p, C.c_int]),
    "hpm_set_planar_prior": (C.c_int, [C.c_void_p]),
    "hpm_set_hierarchy_input": (C.c_int, [C.c_void_p, _FP, C.c_int, C.c_int]),
    "hpm_build_planar_prior": (C.c_int, [C.c_void_p, C.c_float, C.c_int, C.POINTER(C.c_double)]),
    "hpm_input_init": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, _IP, C.c_int]),
    "hpm_input_init_memory": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(HpmCamera),
                                        C.POINTER(_FP), C.POINTER(_FP)]),
    "hpm_device_init": (C.c_int, [C.c_void_p, C.c_char_p]),
    "hpm_device_init_state": (C.c_int, [C.c_void_p, _FP, _FP]),
    "hpm_run_patchmatch": (C.c_int, [C.c_void_p]),
    "hpm_get_run_stats": (C.c_int, [C.c_void_p, C.POINTER(HpmRunStats)]),
    "hpm_dims": (C.c_int, [C.c_void_p, _IP, _IP]),
    "hpm_copy_reference_image": (C.c_int, [C.c_void_p, _FP]),
    "hpm_get_camera": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(HpmCamera)]),
    "hpm_get_plane": (C.c_int, [C.c_void_p, C.c_int, _FP]),
    "hpm_get_cost": (C.c_float, [C.c_void_p, C.c_int]),
    "hpm_copy_hypotheses": (C.c_int, [C.c_void_p, _FP]),
    "hpm_copy_costs": (C.c_int, [C.c_void_p, _FP]),
    "hpm_depth_range": (C.c_int, [C.c_void_p, _FP, _FP]),
    "hpm_upload_prior": (C.c_int, [C.c_void_p, _FP, _UP]),
    "hpm_release_prior": (C.c_int, [C.c_void_p]),
    "hpm_release": (C.c_int, [C.c_void_p]),
    "hpm_jbu": (C.c_int, [C.c_void_p, _FP, C.c_int, C.c_int, _FP, C.c_int, C.c_int, _FP, _FP]),
    "hpm_cost_vectors": (C.c_int, [C.c_void_p, _FP, _IP, C.c_int, _FP]),
    "hpm_debug_stage": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "hpm_debug_read_state": (C.c_int, [C.c_void_p, _FP, _FP, _UP]),
    "hpm_debug_write_state": (C.c_int, [C.c_void_p, _FP, _FP, _UP]),
    "hpm_debug_read_prior": (C.c_int, [C.c_void_p, _FP, _UP]),
    "hpm_fusion_run": (C.c_int, [C.c_void_p, C.c_int, _IP, _IP, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_int64), C.c_char_p, C.c_int]),
    "hpm_fusion_run_acmm": (C.c_int, [C.c_void_p, C.c_int, _IP, _IP, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_int64), C.c_char_p, C.c_int]),
    "hpm_fusion_free": (None, [C.c_void_p]),
    "hpm_fusion_run_folder": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int64), C.c_char_p, C.c_int]),
    "hpm_set_stream_index": (C.c_int, [C.c_void_p, C.c_uint32]),
    "hpm_fusion_run_folder_acmm": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int64), C.c_char_p, C.c_int]),
    "hpm_pool_set_state": (C.c_int, [C.c_void_p, C.c_int, _FP, _FP]),
    "hpm_write_ply": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int64]),
    "hpm_device_selftest": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    # hpm_io.h
    "hpm_read_camera": (C.c_int, [C.c_char_p, C.POINTER(HpmCamera)]),
    "hpm_write_camera": (C.c_int, [C.c_char_p, C.POINTER(HpmCamera)]),
    "hpm_read_pair": (C.c_int, [C.c_char_p, _IP, _IP, _IP, _IP, C.c_int]),
    "hpm_read_image_grey": (C.c_int, [C.c_char_p, _FP, _IP, _IP]),
    "hpm_decode_jpeg_grey": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, _IP, _IP]),
    "hpm_read_dmb": (C.c_int, [C.c_char_p, _FP, _IP, _IP, _IP]),
    "hpm_write_dmb": (C.c_int, [C.c_char_p, _FP, C.c_int, C.c_int, C.c_int]),
    "hpm_read_pgm_f32": (C.c_int, [C.c_char_p, _FP, _IP, _IP]),
    "hpm_write_pgm_u8": (C.c_int, [C.c_char_p, _BP, C.c_int, C.c_int]),
    "hpm_resize_bilinear_c": (None, [_FP, C.c_int, C.c_int, C.c_int, _FP, C.c_int, C.c_int]),
    "hpm_resize_bilinear": (None, [_FP, C.c_int, C.c_int, _FP, C.c_int, C.c_int]),
    # hpm_scene.h
    "hpm_scene_default": (None, [C.POINTER(HpmSceneConfig)]),
    "hpm_scene_render": (C.c_int, [C.POINTER(HpmSceneConfig), C.POINTER(HpmCamera), _BP, _FP, _FP, _BP]),
    "hpm_scene_write_folder": (C.c_int, [C.POINTER(HpmSceneConfig), C.c_char_p]),
    # pool (hpm_capi.h)
    "hpm_pool_create": (C.c_void_p, [C.c_int, C.c_int, C.POINTER(HpmCamera), C.POINTER(_FP), C.c_int]),
    "hpm_pool_destroy": (None, [C.c_void_p]),
    "hpm_pool_camera": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(HpmCamera)]),
    "hpm_pool_set_depth": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _FP]),
    "hpm_pool_get_depth": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _FP]),
    "hpm_input_init_pool": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, _IP, C.c_int]),
    # hpm_prior.h
    "hpm_support_points": (C.c_int, [_FP, C.c_int, C.c_int, _IP, C.c_int]),
    "hpm_delaunay": (C.c_int, [_IP, C.c_int, _IP, C.c_int]),
    "hpm_rasterize_triangles": (None, [_IP, _IP, C.c_int, _FP, C.c_int, C.c_int]),
    "hpm_plane_from_points": (None, [C.POINTER(HpmCamera), _IP, _FP, C.c_int, C.c_float, _FP]),
    "hpm_depth_from_plane_factor": (C.c_float, [C.POINTER(HpmCamera), _FP, C.c_int, C.c_int, C.c_float]),
    "hpm_knn_supplement": (C.c_int, [C.POINTER(HpmCamera), _IP, C.c_int, _FP, C.c_int, C.c_int, C.c_float, C.c_int,
                                     C.c_float, C.c_float, _FP, _FP, C.c_int]),
    "hpm_multiscale_count": (C.c_int, [C.c_int, C.c_int]),
    # hpm_pipeline.h
    "hpm_pipeline_default": (None, [C.POINTER(HpmPipelineOptions)]),
    "hpm_pipeline_run": (C.c_int, [C.c_void_p, C.c_int, _IP, _IP, _IP, C.c_int, C.POINTER(HpmPipelineOptions),
                                   C.POINTER(HpmPipelineStats), C.c_char_p, C.c_int]),
    "hpm_pool_get_state": (C.c_int, [C.c_void_p, C.c_int, _FP, _FP]),
    # hpm_unet.h (device pointers, e.g. torch tensors' data_ptr())
    "hpm_unet_mfma_probe": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hpm_unet_conv_fwd": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "hpm_unet_conv_fwd_up": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_void_p]),
    "hpm_unet_conv_wgrad": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                      C.c_int, C.c_int, C.c_void_p]),
    "hpm_unet_avgpool2": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_void_p]),
    "hpm_unet_upsample2": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_void_p]),
    "hpm_unet_to_bf16": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "hpm_lprior_load": (C.c_void_p, [C.c_char_p, C.c_int]),
    "hpm_lprior_destroy": (None, [C.c_void_p]),
    "hpm_lprior_apply": (C.c_int, [C.c_void_p, C.c_void_p, _FP]),
    "hpm_lprior_apply_timed": (C.c_int, [C.c_void_p, C.c_void_p, _FP, _FP]),
    "hpm_lprior_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "hpm_fusion_run_rgb": (C.c_int, [C.c_void_p, C.c_int, _IP, _IP, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                     C.c_void_p, C.POINTER(C.c_int64), C.c_char_p, C.c_int]),
    "hpm_read_image_rgb": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "hpm_decode_jpeg_rgb": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]),
    "hpm_batch_set_schedule": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "hpm_lprior_forward_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "hpm_lprior_noalloc_capacity": (C.c_size_t, []),
    # hpm_batch.h
    "hpm_batch_create": (C.c_void_p, [C.c_int]),
    "hpm_batch_destroy": (None, [C.c_void_p]),
    "hpm_run_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(HpmBatchStats)]),
    "hpm_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "hpm_batch_log": (C.c_char_p, [C.c_void_p]),
    "hpm_pipeline_run_folder": (C.c_int, [C.c_char_p, C.POINTER(HpmPipelineOptions), C.POINTER(HpmPipelineStats),
                                          C.c_char_p, C.c_int]),
    # hpm_acmm.h
    "hpm_acmm_default": (None, [C.c_void_p]),
    "hpm_acmm_run": (C.c_int, [C.c_int, C.c_int, C.POINTER(HpmCamera), C.POINTER(_FP), C.c_int, _IP, _IP, _IP,
                               C.c_int, C.POINTER(HpmAcmmOptions), C.POINTER(C.c_void_p),
                               C.POINTER(HpmPipelineStats), C.c_char_p, C.c_int]),
    "hpm_acmm_run_folder": (C.c_int, [C.c_char_p, C.POINTER(HpmAcmmOptions), C.c_int, C.POINTER(C.c_int64),
                                      C.POINTER(HpmPipelineStats), C.c_char_p, C.c_int]),
}


def lib() -> C.CDLL:
    """Loads libhpm_mvs.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() first")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_FP)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_IP)


def uptr(a: np.ndarray):
    return a.ctypes.data_as(_UP)


def bptr(a: np.ndarray):
    return a.ctypes.data_as(_BP)


class HpmError(RuntimeError):
    pass


def camera_from_arrays(K, R, t, width, height, depth_min, depth_max) -> HpmCamera:
    c = HpmCamera()
    for i, v in enumerate(np.asarray(K, np.float32).reshape(9)):
        c.K[i] = float(v)
    for i, v in enumerate(np.asarray(R, np.float32).reshape(9)):
        c.R[i] = float(v)
    for i, v in enumerate(np.asarray(t, np.float32).reshape(3)):
        c.t[i] = float(v)
    c.width, c.height = int(width), int(height)
    c.depth_min, c.depth_max = float(depth_min), float(depth_max)
    return c


class Problem:
    """One reference-view problem on the GPU: the reference's `class HPM`."""

    def __init__(self, device: int = 0):
        self._h = lib().hpm_create(device)
        if not self._h:
            raise HpmError("hpm_create failed (no HIP device?)")
        self._keep = []

    def close(self):
        if self._h:
            lib().hpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = lib().hpm_last_error(self._h)
            raise HpmError(f"{what}: {msg.decode() if msg else rc}")
        return rc

    # --- configuration (HPM.cpp:468-480) ---
    def set_params(self, p: HpmParams):
        self._check(lib().hpm_set_params(self._h, C.byref(p)), "hpm_set_params")

    def params(self) -> HpmParams:
        p = HpmParams()
        self._check(lib().hpm_get_params(self._h, C.byref(p)), "hpm_get_params")
        return p

    def SetGeomConsistencyParams(self, multi_geometry: bool = False):
        self._check(lib().hpm_set_geom_consistency(self._h, int(multi_geometry)), "SetGeomConsistencyParams")

    def SetPlanarPriorParams(self):
        self._check(lib().hpm_set_planar_prior(self._h), "SetPlanarPriorParams")

    # --- inputs (HPM.cpp:482-700) ---
    def InuputInitialization(self, dense_folder: str, ref_image_id: int, src_image_ids: Sequence[int]):
        src = np.ascontiguousarray(src_image_ids, dtype=np.int32)
        self._check(lib().hpm_input_init(self._h, dense_folder.encode(), int(ref_image_id), iptr(src), len(src)),
                    "InuputInitialization")

    def input_init_memory(self, cams: Sequence[HpmCamera], images: Sequence[np.ndarray],
                          depths: Optional[Sequence[np.ndarray]] = None, ref_image_id: int = 0):
        n = len(cams)
        cam_arr = (HpmCamera * n)(*cams)
        imgs = [np.ascontiguousarray(im, dtype=np.float32) for im in images]
        img_ptrs = (_FP * n)(*[fptr(im) for im in imgs])
        dep_ptrs = None
        deps = []
        if depths is not None:
            deps = [np.ascontiguousarray(d, dtype=np.float32) for d in depths]
            dep_ptrs = (_FP * n)(*[fptr(d) for d in deps])
        self._check(lib().hpm_input_init_memory(self._h, int(ref_image_id), n, cam_arr, img_ptrs, dep_ptrs),
                    "hpm_input_init_memory")

    def CudaSpaceInitialization(self, dense_folder: Optional[str] = None):
        self._check(lib().hpm_device_init(self._h, dense_folder.encode() if dense_folder else None),
                    "CudaSpaceInitialization")

    def device_init_state(self, hyp_world: np.ndarray, cost: np.ndarray):
        hw = np.ascontiguousarray(hyp_world, dtype=np.float32)
        cs = np.ascontiguousarray(cost, dtype=np.float32)
        self._check(lib().hpm_device_init_state(self._h, fptr(hw), fptr(cs)), "hpm_device_init_state")

    # --- run (HPM.cu:1906-1957) ---
    def RunPatchMatch(self):
        self._check(lib().hpm_run_patchmatch(self._h), "RunPatchMatch")

    def run_stats(self) -> HpmRunStats:
        s = HpmRunStats()
        self._check(lib().hpm_get_run_stats(self._h, C.byref(s)), "hpm_get_run_stats")
        return s

    # --- getters (HPM.cpp:724-757) ---
    def dims(self):
        w, h = C.c_int32(), C.c_int32()
        self._check(lib().hpm_dims(self._h, C.byref(w), C.byref(h)), "hpm_dims")
        return w.value, h.value

    def camera(self, image: int) -> HpmCamera:
        """Camera of image `image` of the problem (0 = reference) as the
        device uses it, after the max_image_size downscale."""
        c = HpmCamera()
        self._check(lib().hpm_get_camera(self._h, int(image), C.byref(c)), "hpm_get_camera")
        return c

    def GetReferenceImageWidth(self) -> int:
        return self.dims()[0]

    def GetReferenceImageHeight(self) -> int:
        return self.dims()[1]

    def GetReferenceImage(self) -> np.ndarray:
        w, h = self.dims()
        out = np.empty((h, w), np.float32)
        self._check(lib().hpm_copy_reference_image(self._h, fptr(out)), "GetReferenceImage")
        return out

    def GetPlaneHypothesis(self, index: int) -> np.ndarray:
        out = np.empty(4, np.float32)
        self._check(lib().hpm_get_plane(self._h, int(index), fptr(out)), "GetPlaneHypothesis")
        return out

    def GetCost(self, index: int) -> float:
        return float(lib().hpm_get_cost(self._h, int(index)))

    def hypotheses(self) -> np.ndarray:
        w, h = self.dims()
        out = np.empty((h, w, 4), np.float32)
        self._check(lib().hpm_copy_hypotheses(self._h, fptr(out)), "hpm_copy_hypotheses")
        return out

    def costs(self) -> np.ndarray:
        w, h = self.dims()
        out = np.empty((h, w), np.float32)
        self._check(lib().hpm_copy_costs(self._h, fptr(out)), "hpm_copy_costs")
        return out

    def GetMinDepth(self) -> float:
        a, b = C.c_float(), C.c_float()
        self._check(lib().hpm_depth_range(self._h, C.byref(a), C.byref(b)), "hpm_depth_range")
        return a.value

    def GetMaxDepth(self) -> float:
        a, b = C.c_float(), C.c_float()
        self._check(lib().hpm_depth_range(self._h, C.byref(a), C.byref(b)), "hpm_depth_range")
        return b.value

    # --- prior (HPM.cpp:702-722, 901-921, 1099-1123, 1283-1286) ---
    def upload_prior(self, planes: np.ndarray, mask: np.ndarray):
        pl = np.ascontiguousarray(planes, dtype=np.float32)
        mk = np.ascontiguousarray(mask, dtype=np.uint32)
        self._check(lib().hpm_upload_prior(self._h, fptr(pl), uptr(mk)), "hpm_upload_prior")

    def ReleasePriorCudaMemory(self):
        self._check(lib().hpm_release_prior(self._h), "ReleasePriorCudaMemory")

    def ReleaseProblemCudaMemory(self):
        self._check(lib().hpm_release(self._h), "ReleaseProblemCudaMemory")

    # --- JBU (HPM.cpp:1215-1281) ---
    def JointBilateralUpsampling(self, guide: np.ndarray, src_nd: np.ndarray):
        """guide (h, w) grey; src_nd (sh, sw, 4) = world normal xyz + depth.
        Returns (depth (h, w), normal (h, w, 4)) or None when the scale is 1."""
        g = np.ascontiguousarray(guide, dtype=np.float32)
        s = np.ascontiguousarray(src_nd, dtype=np.float32)
        h, w = g.shape
        sh, sw = s.shape[:2]
        od = np.empty((h, w), np.float32)
        on = np.empty((h, w, 4), np.float32)
        rc = self._check(lib().hpm_jbu(self._h, fptr(g), w, h, fptr(s), sw, sh, fptr(od), fptr(on)),
                         "JointBilateralUpsampling")
        return None if rc == 1 else (od, on)

    # --- checks ---
    def cost_vectors(self, planes: np.ndarray, pix: np.ndarray) -> np.ndarray:
        pl = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
        px = np.ascontiguousarray(pix, dtype=np.int32).reshape(-1, 2)
        nsrc = self.params().num_images - 1
        out = np.empty((len(pl), nsrc), np.float32)
        self._check(lib().hpm_cost_vectors(self._h, fptr(pl), iptr(px), len(pl), fptr(out)), "hpm_cost_vectors")
        return out

    def debug_stage(self, stage: int, it: int = 0, colour: int = 0):
        self._check(lib().hpm_debug_stage(self._h, stage, it, colour), "hpm_debug_stage")

    def device_selftest(self, kind: int, data: np.ndarray) -> np.ndarray:
        if kind == 4:
            a = np.ascontiguousarray(data, dtype=np.uint32).reshape(-1, 6)
            out = np.empty((len(a), 4), np.uint32)
        elif kind == 8:  # quotients of (numerator, denominator) pairs
            a = np.ascontiguousarray(data, dtype=np.float32).reshape(-1, 2)
            out = np.empty(len(a), np.float32)
        elif kind == 9:  # double reciprocals
            a = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
            out = np.empty(len(a), np.float64)
        else:
            a = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
            out = np.empty(len(a), np.float32)
        self._check(lib().hpm_device_selftest(self._h, kind, a.ctypes.data_as(C.c_void_p),
                                              out.ctypes.data_as(C.c_void_p), len(a)), "hpm_device_selftest")
        return out

    def debug_state(self):
        w, h = self.dims()
        hyp = np.empty((h, w, 4), np.float32)
        cost = np.empty((h, w), np.float32)
        sel = np.empty((h, w), np.uint32)
        self._check(lib().hpm_debug_read_state(self._h, fptr(hyp), fptr(cost), uptr(sel)), "hpm_debug_read_state")
        return hyp, cost, sel


# --- I/O (include/hpm_io.h) ---

class Batch:
    """RunPatchMatch of many Problems together (hpm_batch.h): one launch per
    stage over all problems, the schedule replayed from a hipGraph, the
    launch grid autotuned on the first call of a batch shape."""

    def __init__(self, device: int = 0):
        self._b = lib().hpm_batch_create(device)
        if not self._b:
            raise HpmError("hpm_batch_create failed (no HIP device?)")

    def close(self):
        if self._b:
            lib().hpm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, problems: Sequence["Problem"], readback: bool = True) -> HpmBatchStats:
        hs = (C.c_void_p * len(problems))(*[p._h for p in problems])
        st = HpmBatchStats()
        rc = lib().hpm_run_batch(self._b, hs, len(problems), int(readback), C.byref(st))
        if rc < 0:
            raise HpmError(f"hpm_run_batch: {lib().hpm_batch_last_error(self._b).decode()}")
        return st

    def log(self) -> str:
        return lib().hpm_batch_log(self._b).decode()

    def set_schedule(self, grid: int, items: int, persistent: bool = False):
        """hpm_batch_set_schedule: items 0 static ranges over `grid` blocks,
        1/2/4 work-queue items of that many tiles (persistent: one
        cooperative launch, items 1 or 2), -1 autotune."""
        if lib().hpm_batch_set_schedule(self._b, int(grid), int(items), int(persistent)) != 0:
            raise HpmError("hpm_batch_set_schedule: bad schedule")


def read_camera(path: str) -> HpmCamera:
    c = HpmCamera()
    if lib().hpm_read_camera(path.encode(), C.byref(c)) != 0:
        raise HpmError(f"cannot read camera {path}")
    return c


def write_camera(path: str, cam: HpmCamera):
    if lib().hpm_write_camera(path.encode(), C.byref(cam)) != 0:
        raise HpmError(f"cannot write camera {path}")


def read_dmb(path: str) -> np.ndarray:
    w, h, nb = C.c_int32(), C.c_int32(), C.c_int32()
    if lib().hpm_read_dmb(path.encode(), None, C.byref(w), C.byref(h), C.byref(nb)) != 0:
        raise HpmError(f"cannot read {path}")
    out = np.empty((h.value, w.value, nb.value), np.float32)
    if lib().hpm_read_dmb(path.encode(), fptr(out), C.byref(w), C.byref(h), C.byref(nb)) != 0:
        raise HpmError(f"cannot read {path}")
    return out[:, :, 0] if nb.value == 1 else out


def write_dmb(path: str, data: np.ndarray):
    a = np.ascontiguousarray(data, dtype=np.float32)
    nb = 1 if a.ndim == 2 else a.shape[2]
    if lib().hpm_write_dmb(path.encode(), fptr(a), a.shape[1], a.shape[0], nb) != 0:
        raise HpmError(f"cannot write {path}")


def read_pgm(path: str) -> np.ndarray:
    w, h = C.c_int32(), C.c_int32()
    if lib().hpm_read_pgm_f32(path.encode(), None, C.byref(w), C.byref(h)) != 0:
        raise HpmError(f"cannot read {path}")
    out = np.empty((h.value, w.value), np.float32)
    lib().hpm_read_pgm_f32(path.encode(), fptr(out), C.byref(w), C.byref(h))
    return out


def read_image_rgb(path: str) -> np.ndarray:
    """cv::imread(IMREAD_COLOR) as (H, W, 3) RGB bytes (hpm_read_image_rgb)."""
    w, h = C.c_int32(), C.c_int32()
    if lib().hpm_read_image_rgb(path.encode(), None, C.byref(w), C.byref(h)) != 0:
        raise HpmError(f"cannot read {path}")
    out = np.empty((h.value, w.value, 3), np.uint8)
    if lib().hpm_read_image_rgb(path.encode(), bptr(out), C.byref(w), C.byref(h)) != 0:
        raise HpmError(f"cannot read {path}")
    return out


def read_image_grey(path: str) -> np.ndarray:
    """JPEG (luminance, as libjpeg's grayscale decode) or 8-bit PGM, as float."""
    w, h = C.c_int32(), C.c_int32()
    if lib().hpm_read_image_grey(path.encode(), None, C.byref(w), C.byref(h)) != 0:
        raise HpmError(f"cannot read {path}")
    out = np.empty((h.value, w.value), np.float32)
    if lib().hpm_read_image_grey(path.encode(), fptr(out), C.byref(w), C.byref(h)) != 0:
        raise HpmError(f"cannot decode {path}")
    return out


def read_pair(path: str):
    n = C.c_int32()
    if lib().hpm_read_pair(path.encode(), C.byref(n), None, None, None, 0) != 0:
        raise HpmError(f"cannot read {path}")
    refs = np.zeros(n.value, np.int32)
    ns = np.zeros(n.value, np.int32)
    srcs = np.zeros((n.value, HPM_MAX_SOURCES), np.int32)
    if lib().hpm_read_pair(path.encode(), C.byref(n), iptr(refs), iptr(ns), iptr(srcs), HPM_MAX_SOURCES) != 0:
        raise HpmError(f"cannot read {path}")
    return [(int(refs[i]), [int(v) for v in srcs[i, : ns[i]]]) for i in range(n.value)]


def resize_bilinear(src: np.ndarray, dw: int, dh: int) -> np.ndarray:
    a = np.ascontiguousarray(src, dtype=np.float32)
    ch = 1 if a.ndim == 2 else a.shape[2]
    out = np.empty((dh, dw) + (() if a.ndim == 2 else (ch,)), np.float32)
    lib().hpm_resize_bilinear_c(fptr(a), a.shape[1], a.shape[0], ch, fptr(out), dw, dh)
    return out


# --- synthetic scenes (include/hpm_scene.h) ---

def scene_config(**kw) -> HpmSceneConfig:
    cfg = HpmSceneConfig()
    lib().hpm_scene_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def workload_config(name: str) -> HpmSceneConfig:
    """Scene settings of the workloads:
    'c1': 1 ref + 4 src, 640x480, 4 planes (fx=fy=500);
    'c2': 11 views, 1600x1066, 20 planes + 50 boxes, 30% low texture;
    'c5': 21 views, 3840x2160 (the field of view of c2), 20 planes + 50
    boxes, 50% low texture."""
    if name == "c1":
        return scene_config(seed=1)
    if name == "c2":
        return scene_config(seed=2, width=1600, height=1066, fx=1250.0, fy=1250.0, cx=800.0, cy=533.0,
                            num_views=11, arc_length=0.5, num_planes=20, num_boxes=50,
                            low_texture_fraction=0.3)
    if name == "c5":
        return scene_config(seed=5, width=3840, height=2160, fx=3000.0, fy=3000.0, cx=1920.0, cy=1080.0,
                            num_views=21, arc_length=0.5, num_planes=20, num_boxes=50,
                            low_texture_fraction=0.5)
    raise ValueError(name)


def render_scene(cfg: HpmSceneConfig, want_gt: bool = True):
    """Returns (cams list, images (n, H, W) uint8, gt_depth (n, H, W) | None,
    gt_normal (n, H, W, 3) | None, low_texture (H, W) uint8 of the middle view)."""
    n, H, W = cfg.num_views, cfg.height, cfg.width
    cams = (HpmCamera * n)()
    imgs = np.empty((n, H, W), np.uint8)
    gd = np.empty((n, H, W), np.float32) if want_gt else None
    gn = np.empty((n, H, W, 3), np.float32) if want_gt else None
    low = np.empty((H, W), np.uint8)
    rc = lib().hpm_scene_render(C.byref(cfg), cams, bptr(imgs), fptr(gd) if want_gt else None,
                                fptr(gn) if want_gt else None, bptr(low))
    if rc != 0:
        raise HpmError("hpm_scene_render failed")
    return list(cams), imgs, gd, gn, low


def write_scene_folder(cfg: HpmSceneConfig, folder: str):
    if lib().hpm_scene_write_folder(C.byref(cfg), folder.encode()) != 0:
        raise HpmError(f"hpm_scene_write_folder failed for {folder}")


# --- device pool and pass schedule (hpm_capi.h, hpm_pipeline.h) ---

class Pool:
    """All views of a scene resident on the device."""

    def __init__(self, cams: Sequence[HpmCamera], images: Sequence[np.ndarray], device: int = 0,
                 max_image_size: int = 3200, _handle=None):
        if _handle is not None:  # a pool the library made (hpm_acmm_run)
            self._p = _handle
            self.n = len(cams)
            return
        n = len(cams)
        self._imgs = [np.ascontiguousarray(im, dtype=np.float32) for im in images]
        ptrs = (_FP * n)(*[fptr(im) for im in self._imgs])
        self._p = lib().hpm_pool_create(device, n, (HpmCamera * n)(*cams), ptrs, max_image_size)
        if not self._p:
            raise HpmError("hpm_pool_create failed")
        self.n = n

    def close(self):
        if self._p:
            lib().hpm_pool_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def camera(self, v: int) -> HpmCamera:
        c = HpmCamera()
        if lib().hpm_pool_camera(self._p, v, C.byref(c)) != 0:
            raise HpmError("hpm_pool_camera")
        return c

    def state(self, v: int):
        c = self.camera(v)
        hyp = np.empty((c.height, c.width, 4), np.float32)
        cost = np.empty((c.height, c.width), np.float32)
        if lib().hpm_pool_get_state(self._p, v, fptr(hyp), fptr(cost)) != 0:
            raise HpmError(f"no state for view {v}")
        return hyp, cost

    def depth(self, v: int, which: int) -> np.ndarray:
        c = self.camera(v)
        out = np.empty((c.height, c.width), np.float32)
        if lib().hpm_pool_get_depth(self._p, v, which, fptr(out)) != 0:
            raise HpmError("hpm_pool_get_depth")
        return out

    def set_depth(self, v: int, which: int, depth: np.ndarray):
        d = np.ascontiguousarray(depth, dtype=np.float32)
        if lib().hpm_pool_set_depth(self._p, v, which, fptr(d)) != 0:
            raise HpmError("hpm_pool_set_depth")


def pipeline_options(**kw) -> HpmPipelineOptions:
    o = HpmPipelineOptions()
    lib().hpm_pipeline_default(C.byref(o))
    params = kw.pop("params", None)
    if params is not None:
        o.params = params
    for k, v in kw.items():
        if hasattr(o.params, k) and not hasattr(o, k):
            setattr(o.params, k, v)
        else:
            setattr(o, k, v)
    return o


def run_pipeline(pool: Pool, problems, opt: Optional[HpmPipelineOptions] = None) -> HpmPipelineStats:
    """problems: list of (ref_view, [source views])."""
    n = len(problems)
    refs = np.array([p[0] for p in problems], np.int32)
    ns = np.array([len(p[1]) for p in problems], np.int32)
    srcs = np.zeros((n, HPM_MAX_SOURCES), np.int32)
    for i, (_, s) in enumerate(problems):
        srcs[i, : len(s)] = s
    st = HpmPipelineStats()
    err = C.create_string_buffer(512)
    o = opt if opt is not None else pipeline_options()
    rc = lib().hpm_pipeline_run(pool._p, n, iptr(refs), iptr(ns), iptr(srcs), HPM_MAX_SOURCES, C.byref(o),
                                C.byref(st), err, 512)
    if rc != 0:
        raise HpmError(f"hpm_pipeline_run: {err.value.decode()}")
    return st


# --- fusion (hpm_fusion.h) ---

POINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("nx", "<f4"), ("ny", "<f4"), ("nz", "<f4"),
                        ("r", "u1"), ("g", "u1"), ("b", "u1"), ("pad", "u1")])


def _problem_arrays(problems):
    n = len(problems)
    ns = np.array([len(p[1]) for p in problems], np.int32)
    srcs = np.zeros((n, HPM_MAX_SOURCES), np.int32)
    for i, (_, s) in enumerate(problems):
        srcs[i, : len(s)] = s
    return ns, srcs


def fusion(pool: Pool, problems, geom_consistency: bool = True, acmm: bool = False) -> np.ndarray:
    """RunFusion over the pool's maps (acmm: ACMM+NESP's rule,
    hpm_fusion_run_acmm); problems[i] = (i, [source views]). Returns the
    points as a POINT_DTYPE array in the reference's order."""
    for i, (r, _) in enumerate(problems):
        if r != i:
            raise HpmError("fusion: problem i must have reference view i")
    ns, srcs = _problem_arrays(problems)
    pts = C.c_void_p()
    npts = C.c_int64()
    err = C.create_string_buffer(512)
    fn = lib().hpm_fusion_run_acmm if acmm else lib().hpm_fusion_run
    rc = fn(pool._p, len(problems), iptr(ns), iptr(srcs), HPM_MAX_SOURCES, int(geom_consistency), C.byref(pts),
            C.byref(npts), err, 512)
    if rc != 0:
        raise HpmError(f"hpm_fusion_run: {err.value.decode()}")
    try:
        n = npts.value
        out = np.empty(n, POINT_DTYPE)
        if n:
            C.memmove(out.ctypes.data, pts.value, n * POINT_DTYPE.itemsize)
    finally:
        lib().hpm_fusion_free(pts)
    return out


def write_ply(path: str, points: np.ndarray):
    p = np.ascontiguousarray(points, dtype=POINT_DTYPE)
    if lib().hpm_write_ply(path.encode(), C.c_void_p(p.ctypes.data), len(p)) != 0:
        raise HpmError(f"hpm_write_ply failed for {path}")


def acmm_options(**kw) -> HpmAcmmOptions:
    """hpm_acmm_default, then the given fields (params fields by name)."""
    o = HpmAcmmOptions()
    lib().hpm_acmm_default(C.byref(o))
    for k, v in kw.items():
        if hasattr(o.params, k) and not hasattr(o, k):
            setattr(o.params, k, v)
        else:
            setattr(o, k, v)
    return o


def run_acmm(cams, images, problems, opt: Optional[HpmAcmmOptions] = None, device: int = 0):
    """The ACMM+NESP coarse-to-fine schedule (include/hpm_acmm.h) on full-size
    views; problems: list of (ref_view, [source views]). Returns (Pool of the
    last level, HpmPipelineStats)."""
    n = len(cams)
    imgs = [np.ascontiguousarray(im, dtype=np.float32) for im in images]
    ptrs = (_FP * n)(*[fptr(im) for im in imgs])
    refs = np.array([p[0] for p in problems], np.int32)
    ns, srcs = _problem_arrays(problems)
    st = HpmPipelineStats()
    err = C.create_string_buffer(512)
    out = C.c_void_p()
    o = opt if opt is not None else acmm_options()
    rc = lib().hpm_acmm_run(device, n, (HpmCamera * n)(*cams), ptrs, len(problems), iptr(refs), iptr(ns),
                            iptr(srcs), HPM_MAX_SOURCES, C.byref(o), C.byref(out), C.byref(st), err, 512)
    if rc != 0:
        raise HpmError(f"hpm_acmm_run: {err.value.decode()}")
    return Pool(cams, [], _handle=out.value), st


def run_acmm_folder(folder: str, opt: Optional[HpmAcmmOptions] = None, fuse: bool = True):
    """hpm_acmm_run_folder: returns (HpmPipelineStats, fused point count)."""
    st = HpmPipelineStats()
    err = C.create_string_buffer(512)
    npts = C.c_int64(0)
    o = opt if opt is not None else acmm_options()
    if lib().hpm_acmm_run_folder(folder.encode(), C.byref(o), int(fuse), C.byref(npts), C.byref(st), err, 512) != 0:
        raise HpmError(f"hpm_acmm_run_folder: {err.value.decode()}")
    return st, npts.value


def run_pipeline_folder(folder: str, opt: Optional[HpmPipelineOptions] = None) -> HpmPipelineStats:
    st = HpmPipelineStats()
    err = C.create_string_buffer(512)
    o = opt if opt is not None else pipeline_options()
    if lib().hpm_pipeline_run_folder(folder.encode(), C.byref(o), C.byref(st), err, 512) != 0:
        raise HpmError(f"hpm_pipeline_run_folder: {err.value.decode()}")
    return st


# --- host prior construction (hpm_prior.h) ---

def support_points(cost: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(cost, dtype=np.float32)
    h, w = c.shape
    out = np.empty(((w // 5 + 1) * (h // 5 + 1), 2), np.int32)
    n = lib().hpm_support_points(fptr(c), w, h, iptr(out), len(out))
    return out[:n].copy()


def delaunay(points: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(points, dtype=np.int32).reshape(-1, 2)
    cap = max(2 * len(p), 1)  # a triangulation of n points has fewer than 2n triangles
    out = np.empty((cap, 3), np.int32)
    n = lib().hpm_delaunay(iptr(p), len(p), iptr(out), cap)
    if n < 0:
        raise HpmError("hpm_delaunay")
    return out[:n]


def rasterize_triangles(points: np.ndarray, tris: np.ndarray, w: int, h: int) -> np.ndarray:
    p = np.ascontiguousarray(points, dtype=np.int32)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    mask = np.zeros((h, w), np.float32)
    lib().hpm_rasterize_triangles(iptr(p), iptr(t), len(t), fptr(mask), w, h)
    return mask


def plane_from_points(ref: HpmCamera, pts3: np.ndarray, depth: np.ndarray, factor: float) -> np.ndarray:
    p = np.ascontiguousarray(pts3, dtype=np.int32).reshape(6)
    d = np.ascontiguousarray(depth, dtype=np.float32)
    out = np.empty(4, np.float32)
    lib().hpm_plane_from_points(C.byref(ref), iptr(p), fptr(d), d.shape[1], factor, fptr(out))
    return out


def knn_supplement(ref: HpmCamera, points: np.ndarray, depth: np.ndarray, factor: float, k: int, dmin: float,
                   dmax: float, mask: np.ndarray, threads: int = 0):
    p = np.ascontiguousarray(points, dtype=np.int32)
    d = np.ascontiguousarray(depth, dtype=np.float32)
    h, w = d.shape
    m = np.ascontiguousarray(mask, dtype=np.float32).copy()
    planes = np.zeros((h, w, 4), np.float32)
    rc = lib().hpm_knn_supplement(C.byref(ref), iptr(p), len(p), fptr(d), w, h, factor, k, dmin, dmax, fptr(m),
                                  fptr(planes), threads)
    if rc != 0:
        raise HpmError("hpm_knn_supplement")
    return m, planes
