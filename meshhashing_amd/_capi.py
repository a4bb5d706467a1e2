"""ctypes binding of include/meshhashing.h.

Loads meshhashing_amd/libmeshhashing_amd.so and raises if it is missing:
there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libmeshhashing_amd.so"

MH_OK = 0
MH_ERR_ARG = -1
MH_ERR_HIP = -2
MH_ERR_CAPACITY = -3
MH_ERR_STATE = -4
MH_ERR_EMPTY = -5

STATUS_BITS = {
    "BLOCK_HEAP_EXHAUSTED": 1 << 0,
    "INSERT_DROPPED": 1 << 1,
    "DDA_CAP": 1 << 2,
    "VERTEX_HEAP_EXHAUSTED": 1 << 3,
    "TRIANGLE_HEAP_EXHAUSTED": 1 << 4,
    "COORD_RANGE": 1 << 5,
    "OVERFLOW_QUEUE_FULL": 1 << 6,
    "STENCIL_RANGE": 1 << 7,
}


class HashParams(C.Structure):
    _fields_ = [("bucket_count", C.c_uint32), ("bucket_size", C.c_uint32),
                ("entry_count", C.c_uint32), ("linked_list_size", C.c_uint32),
                ("value_capacity", C.c_uint32)]


class MeshParams(C.Structure):
    _fields_ = [("max_vertex_count", C.c_uint32), ("max_triangle_count", C.c_uint32)]


class VolumeParams(C.Structure):
    _fields_ = [("voxel_size", C.c_float), ("truncation_distance_scale", C.c_float),
                ("truncation_distance", C.c_float), ("sdf_upper_bound", C.c_float),
                ("weight_sample", C.c_uint32), ("weight_upper_bound", C.c_uint32)]


class SensorParams(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("width", C.c_uint32), ("height", C.c_uint32),
                ("min_depth_range", C.c_float), ("max_depth_range", C.c_float),
                ("range_factor", C.c_float)]


class RayCasterParams(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("width", C.c_uint32), ("height", C.c_uint32),
                ("min_raycast_depth", C.c_float), ("max_raycast_depth", C.c_float),
                ("raycast_step", C.c_float), ("sample_sdf_threshold", C.c_float),
                ("sdf_threshold", C.c_float), ("enable_gradients", C.c_int32)]


class McParams(C.Structure):
    _fields_ = [("sdf_threshold", C.c_float), ("rho_min", C.c_float),
                ("weight_min", C.c_float), ("snap_distance", C.c_double),
                ("enable_sdf_gradient", C.c_int32)]


class Status(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("dropped_inserts", C.c_uint32),
                ("dda_cap_hits", C.c_uint32), ("list_inserts", C.c_uint32),
                ("block_count", C.c_uint32), ("candidate_count", C.c_uint32),
                ("vertex_count", C.c_uint32), ("triangle_count", C.c_uint32),
                ("voxel_updates", C.c_uint64), ("blocks_integrated", C.c_uint64),
                ("mc_vertex_writes", C.c_uint64), ("mc_triangle_writes", C.c_uint64),
                ("mc_blocks", C.c_uint64), ("ray_samples", C.c_uint64), ("ray_blocks", C.c_uint64)]


class RaycastGrad(C.Structure):
    _fields_ = [("loss", C.c_double), ("grad", C.c_double * 6), ("hessian", C.c_double * 36),
                ("pixels", C.c_uint32), ("outliers", C.c_uint32), ("observed", C.c_uint32)]


class PointToPsdf(C.Structure):
    _fields_ = [("A", C.c_double * 36), ("b", C.c_double * 6), ("err", C.c_double),
                ("count", C.c_uint32)]


MH_P2S_REFERENCE = 0
MH_P2S_CORRECTED = 1


class TuneResult(C.Structure):
    _fields_ = [("tiling", C.c_int32), ("workgroup_size", C.c_int32),
                ("blocks_per_workgroup", C.c_int32), ("voxels_per_lane", C.c_int32),
                ("workgroups_per_cu", C.c_int32), ("candidates", C.c_int32),
                ("best_ms", C.c_float), ("default_ms", C.c_float), ("early_loads", C.c_int32),
                ("ray_shape", C.c_int32)]


class TuneMcResult(C.Structure):
    _fields_ = [("done", C.c_int32), ("shape", C.c_int32), ("workgroups_per_cu", C.c_int32),
                ("static_quarters", C.c_int32), ("candidates", C.c_int32), ("best_ms", C.c_float),
                ("default_ms", C.c_float), ("blocks_per_ms", C.c_float)]


class TuneShapeResult(C.Structure):
    _fields_ = [("shape", C.c_int32), ("tile_width", C.c_int32), ("tile_height", C.c_int32),
                ("candidates", C.c_int32), ("best_ms", C.c_float), ("default_ms", C.c_float)]


MH_NUM_TILINGS = 18
MH_NUM_TUNE_GRIDS = 5
MH_NUM_RAY_SHAPES = 4
MH_NUM_MC_SHAPES = 9
MH_FIELD_SDF, MH_FIELD_WEIGHT, MH_FIELD_COLOR, MH_FIELD_BETA_A, MH_FIELD_BETA_B = range(5)
MH_DIAG_KERNEL_TIMES = 1
MH_DIAG_RAY_BLOCKS = 2


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double)]


MH_RUN_MESH = 1 << 0
MH_RUN_RAYCAST = 1 << 1
MH_RUN_GRAPH = 1 << 2
MH_RUN_TUNE = 1 << 3


class SynthSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("kind", C.c_int32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("frames", C.c_uint32),
                ("hole_fraction", C.c_float), ("outlier_fraction", C.c_float),
                ("noise", C.c_int32)]


VP = C.c_void_p
P_I32 = C.POINTER(C.c_int32)
P_U32 = C.POINTER(C.c_uint32)
P_U16 = C.POINTER(C.c_uint16)
P_U8 = C.POINTER(C.c_uint8)
P_F32 = C.POINTER(C.c_float)
P_F64 = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "mh_volume_create": (C.c_int, [C.POINTER(HashParams), C.POINTER(VolumeParams),
                                   C.POINTER(MeshParams), C.POINTER(SensorParams), C.c_int,
                                   C.POINTER(VP)]),
    "mh_volume_destroy": (C.c_int, [VP]),
    "mh_volume_reset": (C.c_int, [VP]),
    "mh_volume_set_stream": (C.c_int, [VP, VP]),
    "mh_synchronize": (C.c_int, [VP]),
    "mh_get_status": (C.c_int, [VP, C.POINTER(Status)]),
    "mh_mc_params_default": (None, [C.POINTER(McParams)]),
    "mh_last_error": (C.c_char_p, []),
    "mh_frame_upload_depth_u16": (C.c_int, [VP, P_U16, P_F32]),
    "mh_frame_depth_u16_device": (C.c_int, [VP, VP, P_F32]),
    "mh_frame_set_pose": (C.c_int, [VP, P_F32]),
    "mh_frame_upload_color_bgra8": (C.c_int, [VP, P_U8]),
    "mh_frame_color_bgra8_device": (C.c_int, [VP, VP]),
    "mh_frame_depth_device_ptr": (C.c_int, [VP, C.POINTER(VP)]),
    "mh_alloc_blocks": (C.c_int, [VP]),
    "mh_collect_visible": (C.c_int, [VP, P_U32]),
    "mh_collect_all": (C.c_int, [VP, P_U32]),
    "mh_integrate_tsdf": (C.c_int, [VP]),
    "mh_collect_low_surfel_blocks": (C.c_int, [VP]),
    "mh_starve_blocks": (C.c_int, [VP]),
    "mh_collect_garbage_blocks": (C.c_int, [VP]),
    "mh_recycle_garbage_blocks": (C.c_int, [VP, P_U32]),
    "mh_predict_inlier": (C.c_int, [VP]),
    "mh_integrate_psdf": (C.c_int, [VP]),
    "mh_inlier_ratio_device_ptr": (C.c_int, [VP, C.POINTER(VP)]),
    "mh_marching_cubes": (C.c_int, [VP, C.POINTER(McParams)]),
    "mh_raycast": (C.c_int, [VP, C.POINTER(RayCasterParams), P_F32, VP, VP, VP]),
    "mh_primal_dual_init": (C.c_int, [VP]),
    "mh_primal_dual_iterate": (C.c_int, [VP, C.c_float, C.c_float, C.c_float, P_F64]),
    "mh_dump_blocks": (C.c_int, [VP, C.c_uint32, P_I32, P_F32, P_F32, P_U32, P_F32, P_F32,
                                 P_U32]),
    "mh_load_blocks": (C.c_int, [VP, C.c_uint32, P_I32, P_F32, P_F32, P_U32, P_F32, P_F32]),
    "mh_dump_hash": (C.c_int, [VP, C.c_uint32, P_I32, P_U32]),
    "mh_dump_candidates": (C.c_int, [VP, C.c_uint32, P_I32, P_U32]),
    "mh_dump_candidate_flags": (C.c_int, [VP, C.c_uint32, P_I32, P_U32]),
    "mh_dump_mesh": (C.c_int, [VP, C.c_uint32, P_I32, P_I32, C.c_uint32, P_I32, P_F32, P_I32,
                               P_U32, P_U32]),
    "mh_dump_cube_indices": (C.c_int, [VP, C.c_uint32, P_U16, P_U32]),
    "mh_dump_primal_dual": (C.c_int, [VP, C.c_uint32, P_F32, P_U8, P_U32]),
    "mh_frame_get_pose": (C.c_int, [VP, P_F32, P_F32]),
    "mh_compress_mesh": (C.c_int, [VP, P_U32, P_U32]),
    "mh_compact_mesh_get": (C.c_int, [VP, C.c_uint32, P_F32, P_F32, P_F32, C.c_uint32, P_I32]),
    "mh_save_ply": (C.c_int, [VP, C.c_char_p]),
    "mh_point_to_psdf": (C.c_int, [VP, P_F32, C.c_int, C.POINTER(PointToPsdf)]),
    "mh_localize": (C.c_int, [VP, C.c_uint32, C.c_float, C.c_int, P_F32, P_F64, P_U32, P_U32]),
    "mh_volume_set_diagnostics": (C.c_int, [VP, C.c_uint32]),
    "mh_volume_field_device_ptr": (C.c_int, [VP, C.c_int, C.POINTER(VP), C.POINTER(C.c_uint64)]),
    "mh_kernel_times": (C.c_int, [VP, C.c_uint32, C.POINTER(KernelTime), C.POINTER(C.c_uint32),
                                  C.c_int]),
    "mh_raycast_backward": (C.c_int, [VP, C.POINTER(RayCasterParams), P_F32, VP, VP,
                                      C.POINTER(RaycastGrad)]),
    "mh_raycast_backward_noalloc": (C.c_int, [VP, C.POINTER(RayCasterParams), P_F32, VP, VP,
                                              C.POINTER(RaycastGrad)]),
    "mh_refine_pose": (C.c_int, [VP, C.POINTER(RayCasterParams), P_F32, VP, C.c_uint32, P_F32,
                                 P_F64]),
    "mh_dump_voxel_field": (C.c_int, [VP, VP, C.c_uint32, P_F32, P_U32]),
    "mh_tune_integrate": (C.c_int, [VP, C.c_uint32, C.POINTER(TuneResult)]),
    "mh_set_integrate_tiling": (C.c_int, [VP, C.c_int, C.c_int]),
    "mh_tune_raycast": (C.c_int, [VP, C.POINTER(RayCasterParams), P_F32, C.c_uint32,
                                  C.POINTER(TuneShapeResult)]),
    "mh_set_raycast_shape": (C.c_int, [VP, C.c_int]),
    "mh_tune_marching_cubes": (C.c_int, [VP, C.c_uint32]),
    "mh_marching_cubes_tuning": (C.c_int, [VP, C.POINTER(TuneMcResult)]),
    "mh_set_marching_cubes_shape": (C.c_int, [VP, C.c_int]),
    "mh_batch_set_marching_cubes_shape": (C.c_int, [VP, C.c_int]),
    "mh_batch_tune_marching_cubes": (C.c_int, [VP, C.c_uint32]),
    "mh_batch_marching_cubes_tuning": (C.c_int, [VP, C.POINTER(TuneMcResult)]),
    "mh_batch_create": (C.c_int, [C.c_uint32, C.POINTER(HashParams), C.POINTER(VolumeParams),
                                  C.POINTER(MeshParams), C.POINTER(SensorParams),
                                  C.POINTER(VP)]),
    "mh_batch_destroy": (C.c_int, [VP]),
    "mh_batch_set_stream": (C.c_int, [VP, VP]),
    "mh_batch_volume": (C.c_int, [VP, C.c_uint32, C.POINTER(VP)]),
    "mh_batch_frame_u16": (C.c_int, [VP, VP, C.c_int, P_F32]),
    "mh_batch_set_raycast": (C.c_int, [VP, C.POINTER(RayCasterParams), VP, VP, VP]),
    "mh_batch_set_integrate_tiling": (C.c_int, [VP, C.c_int, C.c_int]),
    "mh_batch_run": (C.c_int, [VP, C.POINTER(McParams), C.c_uint32, C.POINTER(TuneResult)]),
    "mh_batch_synchronize": (C.c_int, [VP]),
    "mh_synth_pose": (C.c_int, [C.POINTER(SynthSpec), C.c_uint32, P_F32]),
    "mh_synth_depth": (C.c_int, [C.POINTER(SynthSpec), C.c_uint32, P_U16]),
    "mh_synth_depth_exact": (C.c_int, [C.POINTER(SynthSpec), C.c_uint32, P_F32, P_F64]),
    "mh_synth_color": (C.c_int, [C.POINTER(SynthSpec), C.c_uint32, P_U8]),
}

_lib = None


def load() -> C.CDLL:
    """Load the library (no GPU call is made here)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m meshhashing_amd.build`; "
                "there is no CPU fallback")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class MHError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> int:
    if rc != MH_OK:
        msg = load().mh_last_error()
        raise MHError(f"{what} failed with {rc}: {msg.decode() if msg else ''}")
    return rc


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the C ABI must be contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def f32x16(m) -> C.Array:
    a = np.ascontiguousarray(np.asarray(m, dtype=np.float32).reshape(16))
    return (C.c_float * 16)(*a.tolist())
