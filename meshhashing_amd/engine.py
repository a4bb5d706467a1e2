"""Python view of the C ABI: a Volume handle plus the reference engine's
per-frame driver (MainEngine::Mapping / Meshing / Recycle,
engine/main_engine.cc:100-192).

Every method calls libmeshhashing_amd.so; nothing here computes on the CPU
except array bookkeeping for dumps.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A
from .configs import Config


def synth_pose(cfg: Config, frame: int) -> np.ndarray:
    lib = A.load()
    out = (C.c_float * 16)()
    A.check(lib.mh_synth_pose(C.byref(cfg.synth_spec()), frame, out), "mh_synth_pose")
    return np.array(out[:], dtype=np.float32).reshape(4, 4)


def synth_depth(cfg: Config, frame: int) -> np.ndarray:
    lib = A.load()
    d = np.zeros((cfg.height, cfg.width), dtype=np.uint16)
    A.check(lib.mh_synth_depth(C.byref(cfg.synth_spec()), frame, A.ptr(d, C.c_uint16)),
            "mh_synth_depth")
    return d


def synth_color(cfg: Config, frame: int) -> np.ndarray:
    """(H, W, 4) uint8 BGRA colour of a frame (0 where there is no surface)."""
    lib = A.load()
    c = np.zeros((cfg.height, cfg.width, 4), dtype=np.uint8)
    A.check(lib.mh_synth_color(C.byref(cfg.synth_spec()), frame, A.ptr(c, C.c_uint8)),
            "mh_synth_color")
    return c


def synth_depth_exact(cfg: Config, wTc: np.ndarray) -> np.ndarray:
    lib = A.load()
    d = np.zeros((cfg.height, cfg.width), dtype=np.float64)
    A.check(lib.mh_synth_depth_exact(C.byref(cfg.synth_spec()), 0, A.f32x16(wTc),
                                     A.ptr(d, C.c_double)), "mh_synth_depth_exact")
    return d


def mc_tune_calls(reps: int = 1) -> int:
    """Meshing calls an online marching-cubes tuning of reps rounds takes."""
    return max(reps, 1) * 2 * (A.MH_NUM_MC_SHAPES - 1) + 1


class Volume:
    """Hash table + block pool + mesh + frame buffers on the current GPU."""

    owned = True

    @classmethod
    def wrap(cls, cfg: Config, handle: C.c_void_p) -> "Volume":
        """A view of a volume owned by someone else (a batch's scene)."""
        v = cls.__new__(cls)
        v.cfg, v.lib, v.h, v.owned = cfg, A.load(), handle, False
        return v

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.lib = A.load()
        self.h = C.c_void_p()
        A.check(self.lib.mh_volume_create(C.byref(cfg.hash_params()), C.byref(cfg.volume_params()),
                                          C.byref(cfg.mesh_params()), C.byref(cfg.sensor_params()),
                                          1 if cfg.psdf else 0, C.byref(self.h)),
                "mh_volume_create")

    def close(self):
        if self.h and self.owned:
            self.lib.mh_volume_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- frame -------------------------------------------------------------
    def upload_depth(self, depth_u16: np.ndarray, wTc: np.ndarray):
        d = np.ascontiguousarray(depth_u16, dtype=np.uint16)
        A.check(self.lib.mh_frame_upload_depth_u16(self.h, A.ptr(d, C.c_uint16), A.f32x16(wTc)),
                "mh_frame_upload_depth_u16")

    def upload_color(self, bgra: np.ndarray):
        c = np.ascontiguousarray(bgra, dtype=np.uint8)
        A.check(self.lib.mh_frame_upload_color_bgra8(self.h, A.ptr(c, C.c_uint8)),
                "mh_frame_upload_color_bgra8")

    def set_depth_device(self, dev_ptr: int, wTc: np.ndarray):
        A.check(self.lib.mh_frame_depth_u16_device(self.h, C.c_void_p(dev_ptr), A.f32x16(wTc)),
                "mh_frame_depth_u16_device")

    def field_ptr(self, field: int = A.MH_FIELD_SDF) -> tuple[int, int]:
        """(device pointer, number of values) of a voxel field."""
        p, n = C.c_void_p(), C.c_uint64(0)
        A.check(self.lib.mh_volume_field_device_ptr(self.h, field, C.byref(p), C.byref(n)),
                "mh_volume_field_device_ptr")
        return p.value, n.value

    def set_diagnostics(self, kernel_times: bool = False, ray_blocks: bool = False):
        flags = (A.MH_DIAG_KERNEL_TIMES if kernel_times else 0) | (A.MH_DIAG_RAY_BLOCKS if ray_blocks else 0)
        A.check(self.lib.mh_volume_set_diagnostics(self.h, flags), "mh_volume_set_diagnostics")

    def kernel_times(self, reset: bool = True) -> dict:
        """{kernel: (launches, total ms)} since the last reset (synchronises)."""
        n = C.c_uint32(0)
        A.check(self.lib.mh_kernel_times(self.h, 0, None, C.byref(n), 0), "mh_kernel_times")
        buf = (A.KernelTime * max(1, n.value))()
        A.check(self.lib.mh_kernel_times(self.h, n.value, buf, C.byref(n), 1 if reset else 0),
                "mh_kernel_times")
        return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms) for i in range(n.value)}

    def set_stream(self, stream_ptr: int):
        A.check(self.lib.mh_volume_set_stream(self.h, C.c_void_p(stream_ptr)), "set_stream")

    def reset(self):
        A.check(self.lib.mh_volume_reset(self.h), "mh_volume_reset")

    # -- stages --------------------------------------------------------------
    def alloc(self):
        A.check(self.lib.mh_alloc_blocks(self.h), "mh_alloc_blocks")

    def collect_visible(self, sync: bool = False) -> int | None:
        n = C.c_uint32(0)
        A.check(self.lib.mh_collect_visible(self.h, C.byref(n) if sync else None),
                "mh_collect_visible")
        return n.value if sync else None

    def collect_all(self, sync: bool = False) -> int | None:
        n = C.c_uint32(0)
        A.check(self.lib.mh_collect_all(self.h, C.byref(n) if sync else None), "mh_collect_all")
        return n.value if sync else None

    def integrate_tsdf(self):
        A.check(self.lib.mh_integrate_tsdf(self.h), "mh_integrate_tsdf")

    def collect_low_surfel_blocks(self):
        A.check(self.lib.mh_collect_low_surfel_blocks(self.h), "mh_collect_low_surfel_blocks")

    def starve_blocks(self):
        A.check(self.lib.mh_starve_blocks(self.h), "mh_starve_blocks")

    def collect_garbage_blocks(self):
        A.check(self.lib.mh_collect_garbage_blocks(self.h), "mh_collect_garbage_blocks")

    def recycle_garbage_blocks(self, sync: bool = True) -> int | None:
        """Blocks freed (sync=True waits for the stream)."""
        n = C.c_uint32(0)
        A.check(self.lib.mh_recycle_garbage_blocks(self.h, C.byref(n) if sync else None),
                "mh_recycle_garbage_blocks")
        return n.value if sync else None

    def tune_integrate(self, reps: int = 3) -> A.TuneResult:
        r = A.TuneResult()
        A.check(self.lib.mh_tune_integrate(self.h, reps, C.byref(r)), "mh_tune_integrate")
        return r

    def set_integrate_tiling(self, tiling: int, workgroups_per_cu: int = 8):
        A.check(self.lib.mh_set_integrate_tiling(self.h, tiling, workgroups_per_cu),
                "mh_set_integrate_tiling")

    def tune_raycast(self, cTw: np.ndarray, reps: int = 3,
                     params: A.RayCasterParams | None = None) -> A.TuneShapeResult:
        """Times the raycast tile shapes at cTw and keeps the fastest (mh_tune_raycast)."""
        p = params if params is not None else self.cfg.raycaster_params()
        r = A.TuneShapeResult()
        A.check(self.lib.mh_tune_raycast(self.h, C.byref(p), A.f32x16(cTw), reps, C.byref(r)),
                "mh_tune_raycast")
        return r

    def set_raycast_shape(self, shape: int):
        A.check(self.lib.mh_set_raycast_shape(self.h, shape), "mh_set_raycast_shape")

    def tune_marching_cubes(self, reps: int = 1):
        """The next mc_tune_calls(reps) marching_cubes() calls alternate launch
        shape 0 with each other shape in turn; the call after them keeps the
        shape that beat the shape-0 trials around it by 3 % in every trial,
        the best such, or shape 0."""
        A.check(self.lib.mh_tune_marching_cubes(self.h, reps), "mh_tune_marching_cubes")

    def marching_cubes_tuning(self) -> A.TuneMcResult:
        r = A.TuneMcResult()
        A.check(self.lib.mh_marching_cubes_tuning(self.h, C.byref(r)), "mh_marching_cubes_tuning")
        return r

    def set_marching_cubes_shape(self, shape: int):
        A.check(self.lib.mh_set_marching_cubes_shape(self.h, shape), "mh_set_marching_cubes_shape")

    def pose(self) -> tuple[np.ndarray, np.ndarray]:
        """(wTc, cTw) of the current frame, FP32, as the kernels use them."""
        w = np.zeros(16, np.float32)
        c = np.zeros(16, np.float32)
        A.check(self.lib.mh_frame_get_pose(self.h, A.ptr(w, C.c_float), A.ptr(c, C.c_float)),
                "mh_frame_get_pose")
        return w.reshape(4, 4), c.reshape(4, 4)

    def predict_inlier(self):
        A.check(self.lib.mh_predict_inlier(self.h), "mh_predict_inlier")

    def integrate_psdf(self):
        A.check(self.lib.mh_integrate_psdf(self.h), "mh_integrate_psdf")

    def marching_cubes(self, params: A.McParams | None = None):
        p = params if params is not None else self.cfg.mc_params()
        A.check(self.lib.mh_marching_cubes(self.h, C.byref(p)), "mh_marching_cubes")

    def raycast(self, cTw: np.ndarray, depth_ptr: int, vertex_ptr: int = 0, normal_ptr: int = 0,
                params: A.RayCasterParams | None = None):
        p = params if params is not None else self.cfg.raycaster_params()
        A.check(self.lib.mh_raycast(self.h, C.byref(p), A.f32x16(cTw), C.c_void_p(depth_ptr or None),
                                    C.c_void_p(vertex_ptr or None), C.c_void_p(normal_ptr or None)),
                "mh_raycast")

    def raycast_backward(self, cTw: np.ndarray, obs_ptr: int, grad_sdf_ptr: int = 0,
                         params: A.RayCasterParams | None = None, noalloc: bool = False) -> dict:
        """Adjoint of the raycast depth (mh_raycast_backward): loss, dL/dxi,
        Gauss-Newton J^T J and pixel count; dL/dsdf added to grad_sdf_ptr.
        noalloc: the allocation-free entry point (static scratch)."""
        p = params if params is not None else self.cfg.raycaster_params()
        g = A.RaycastGrad()
        fn = self.lib.mh_raycast_backward_noalloc if noalloc else self.lib.mh_raycast_backward
        A.check(fn(self.h, C.byref(p), A.f32x16(cTw), C.c_void_p(obs_ptr),
                   C.c_void_p(grad_sdf_ptr or None), C.byref(g)),
                "mh_raycast_backward_noalloc" if noalloc else "mh_raycast_backward")
        return {"loss": g.loss, "grad": np.array(g.grad[:]),
                "hessian": np.array(g.hessian[:]).reshape(6, 6), "pixels": g.pixels,
                "outliers": g.outliers, "observed": g.observed}

    def refine_pose(self, wTc: np.ndarray, obs_ptr: int, iterations: int = 10,
                    params: A.RayCasterParams | None = None):
        """Gauss-Newton pose refinement (mh_refine_pose): (wTc, cost trace)."""
        p = params if params is not None else self.cfg.raycaster_params()
        out = (C.c_float * 16)()
        trace = np.zeros(iterations + 1, np.float64)
        A.check(self.lib.mh_refine_pose(self.h, C.byref(p), A.f32x16(wTc), C.c_void_p(obs_ptr),
                                        iterations, out, A.ptr(trace, C.c_double)),
                "mh_refine_pose")
        return np.array(out[:], np.float32).reshape(4, 4), trace

    def point_to_psdf(self, wTc: np.ndarray, corrected: bool = False) -> dict:
        """PointToSurface on the current frame (mh_point_to_psdf): A (6x6),
        b (6), err, count; the tangent is (translation, rotation)."""
        r = A.PointToPsdf()
        A.check(self.lib.mh_point_to_psdf(self.h, A.f32x16(wTc),
                                          A.MH_P2S_CORRECTED if corrected else A.MH_P2S_REFERENCE,
                                          C.byref(r)), "mh_point_to_psdf")
        return {"A": np.array(r.A[:]).reshape(6, 6), "b": np.array(r.b[:]), "err": r.err,
                "count": r.count}

    def localize(self, wTc: np.ndarray, iterations: int = 20, lam: float = 1e5,
                 corrected: bool = False):
        """MainEngine::Localizing (mh_localize): the refined wTc (also the
        frame's pose from now on), per-step err and count, and the number
        of steps taken (fewer than asked if a step used no pixel)."""
        T = (C.c_float * 16)(*np.asarray(wTc, np.float32).reshape(16).tolist())
        err = np.zeros(iterations, np.float64)
        cnt = np.zeros(iterations, np.uint32)
        done = C.c_uint32(0)
        A.check(self.lib.mh_localize(self.h, iterations, lam,
                                     A.MH_P2S_CORRECTED if corrected else A.MH_P2S_REFERENCE, T,
                                     A.ptr(err, C.c_double), A.ptr(cnt, C.c_uint32), C.byref(done)),
                "mh_localize")
        return {"wTc": np.array(T[:], np.float32).reshape(4, 4), "err": err, "count": cnt,
                "steps": done.value}

    def compress_mesh(self) -> dict:
        """CompressMesh over the current worklist: dense vertices (pos,
        normal, colour) and triangles (vertex indices)."""
        nv, nt = C.c_uint32(0), C.c_uint32(0)
        A.check(self.lib.mh_compress_mesh(self.h, C.byref(nv), C.byref(nt)), "mh_compress_mesh")
        pos = np.zeros((nv.value, 3), np.float32)
        nrm = np.zeros((nv.value, 3), np.float32)
        col = np.zeros((nv.value, 3), np.float32)
        tri = np.zeros((nt.value, 3), np.int32)
        A.check(self.lib.mh_compact_mesh_get(self.h, nv.value, A.ptr(pos, C.c_float),
                                             A.ptr(nrm, C.c_float), A.ptr(col, C.c_float), nt.value,
                                             A.ptr(tri, C.c_int32)), "mh_compact_mesh_get")
        return {"pos": pos, "normal": nrm, "color": col, "tri": tri}

    def save_ply(self, path: str):
        A.check(self.lib.mh_save_ply(self.h, str(path).encode()), "mh_save_ply")

    def dump_voxel_field(self, field_ptr: int) -> np.ndarray:
        n = C.c_uint32(0)
        A.check(self.lib.mh_dump_voxel_field(self.h, C.c_void_p(field_ptr), 0, None, C.byref(n)),
                "count")
        out = np.zeros((n.value, 512), np.float32)
        A.check(self.lib.mh_dump_voxel_field(self.h, C.c_void_p(field_ptr), n.value,
                                             A.ptr(out, C.c_float), C.byref(n)), "mh_dump_voxel_field")
        return out

    def pd_init(self):
        A.check(self.lib.mh_primal_dual_init(self.h), "mh_primal_dual_init")

    def pd_iterate(self, lam: float, sigma: float, tau: float, energy: bool = False):
        e = (C.c_double * 3)()
        A.check(self.lib.mh_primal_dual_iterate(self.h, lam, sigma, tau, e if energy else None),
                "mh_primal_dual_iterate")
        return tuple(e) if energy else None

    def synchronize(self):
        A.check(self.lib.mh_synchronize(self.h), "mh_synchronize")

    def status(self) -> dict:
        s = A.Status()
        A.check(self.lib.mh_get_status(self.h, C.byref(s)), "mh_get_status")
        d = {k: getattr(s, k) for k, _ in A.Status._fields_}
        d["flag_names"] = [n for n, b in A.STATUS_BITS.items() if s.flags & b]
        return d

    def inlier_ptr(self) -> int:
        p = C.c_void_p()
        A.check(self.lib.mh_inlier_ratio_device_ptr(self.h, C.byref(p)), "inlier ptr")
        return p.value

    def depth_ptr(self) -> int:
        p = C.c_void_p()
        A.check(self.lib.mh_frame_depth_device_ptr(self.h, C.byref(p)), "depth ptr")
        return p.value

    # -- dumps -------------------------------------------------------------------
    def dump_blocks(self) -> dict:
        n = C.c_uint32(0)
        A.check(self.lib.mh_dump_blocks(self.h, 0, None, None, None, None, None, None,
                                        C.byref(n)), "count")
        nb = n.value
        pos = np.zeros((nb, 3), np.int32)
        sdf = np.zeros((nb, 512), np.float32)
        w = np.zeros((nb, 512), np.float32)
        col = np.zeros((nb, 512), np.uint32)
        a = np.zeros((nb, 512), np.float32) if self.cfg.psdf else None
        b = np.zeros((nb, 512), np.float32) if self.cfg.psdf else None
        A.check(self.lib.mh_dump_blocks(self.h, nb, A.ptr(pos, C.c_int32), A.ptr(sdf, C.c_float),
                                        A.ptr(w, C.c_float), A.ptr(col, C.c_uint32),
                                        A.ptr(a, C.c_float), A.ptr(b, C.c_float), C.byref(n)),
                "mh_dump_blocks")
        out = {"pos": pos, "sdf": sdf, "weight": w, "color": col}
        if self.cfg.psdf:
            out["a"], out["b"] = a, b
        return out

    def load_blocks(self, pos, sdf, weight, color=None, a=None, b=None):
        pos = np.ascontiguousarray(pos, np.int32)
        sdf = np.ascontiguousarray(sdf, np.float32)
        weight = np.ascontiguousarray(weight, np.float32)
        color = None if color is None else np.ascontiguousarray(color, np.uint32)
        a = None if a is None else np.ascontiguousarray(a, np.float32)
        b = None if b is None else np.ascontiguousarray(b, np.float32)
        A.check(self.lib.mh_load_blocks(self.h, len(pos), A.ptr(pos, C.c_int32),
                                        A.ptr(sdf, C.c_float), A.ptr(weight, C.c_float),
                                        A.ptr(color, C.c_uint32), A.ptr(a, C.c_float),
                                        A.ptr(b, C.c_float)), "mh_load_blocks")

    def dump_hash(self) -> np.ndarray:
        n = C.c_uint32(0)
        A.check(self.lib.mh_dump_hash(self.h, 0, None, C.byref(n)), "count")
        rec = np.zeros((n.value, 6), np.int32)
        A.check(self.lib.mh_dump_hash(self.h, n.value, A.ptr(rec, C.c_int32), C.byref(n)),
                "mh_dump_hash")
        return rec

    def dump_candidates(self) -> np.ndarray:
        n = C.c_uint32(0)
        A.check(self.lib.mh_dump_candidates(self.h, 0, None, C.byref(n)), "count")
        pos = np.zeros((n.value, 3), np.int32)
        A.check(self.lib.mh_dump_candidates(self.h, n.value, A.ptr(pos, C.c_int32), C.byref(n)),
                "mh_dump_candidates")
        return pos

    def dump_candidate_flags(self) -> np.ndarray:
        n = C.c_uint32(0)
        A.check(self.lib.mh_dump_candidate_flags(self.h, 0, None, C.byref(n)), "count")
        rec = np.zeros((n.value, 4), np.int32)
        A.check(self.lib.mh_dump_candidate_flags(self.h, n.value, A.ptr(rec, C.c_int32), C.byref(n)),
                "mh_dump_candidate_flags")
        return rec

    def dump_cube_indices(self) -> np.ndarray:
        n = C.c_uint32(0)
        A.check(self.lib.mh_dump_cube_indices(self.h, 0, None, C.byref(n)), "count")
        cube = np.zeros((n.value, 512), np.uint16)
        A.check(self.lib.mh_dump_cube_indices(self.h, n.value, A.ptr(cube, C.c_uint16),
                                              C.byref(n)), "mh_dump_cube_indices")
        return cube

    def dump_mesh(self) -> dict:
        nt, nv = C.c_uint32(0), C.c_uint32(0)
        A.check(self.lib.mh_dump_mesh(self.h, 0, None, None, 0, None, None, None, C.byref(nt),
                                      C.byref(nv)), "count")
        tri = np.zeros((nt.value, 12), np.int32)
        tcube = np.zeros((nt.value, 4), np.int32)
        vkey = np.zeros((nv.value, 4), np.int32)
        vval = np.zeros((nv.value, 10), np.float32)
        vref = np.zeros(nv.value, np.int32)
        A.check(self.lib.mh_dump_mesh(self.h, nt.value, A.ptr(tri, C.c_int32),
                                      A.ptr(tcube, C.c_int32), nv.value, A.ptr(vkey, C.c_int32),
                                      A.ptr(vval, C.c_float), A.ptr(vref, C.c_int32),
                                      C.byref(nt), C.byref(nv)), "mh_dump_mesh")
        return {"tri": tri, "tri_cube": tcube, "vkey": vkey, "vval": vval, "vref": vref}

    def dump_primal_dual(self) -> dict:
        n = C.c_uint32(0)
        A.check(self.lib.mh_dump_primal_dual(self.h, 0, None, None, C.byref(n)), "count")
        pdv = np.zeros((n.value, 6, 512), np.float32)
        mask = np.zeros((n.value, 512), np.uint8)
        A.check(self.lib.mh_dump_primal_dual(self.h, n.value, A.ptr(pdv, C.c_float),
                                             A.ptr(mask, C.c_uint8), C.byref(n)),
                "mh_dump_primal_dual")
        return {"sdf0": pdv[:, 0], "sdf_bar": pdv[:, 1], "w_pd": pdv[:, 2], "p": pdv[:, 3:6],
                "mask": mask}


def refine_pose_pyramid(levels, wTc: np.ndarray, obs_ptr: int) -> np.ndarray:
    """Coarse-to-fine pose refinement: mh_refine_pose on each level in turn,
    each starting from the previous level's result.  levels: (Volume,
    RayCasterParams or None, iterations) from coarsest to finest, volumes of
    the same scene at growing resolution; their wider truncation bands keep a
    pose that is centimetres off inside the coarse SDFs' valid range."""
    T = np.asarray(wTc, np.float32)
    for vol, params, iterations in levels:
        T, _ = vol.refine_pose(T, obs_ptr, iterations, params=params)
    return T


class MainEngine:
    """Per-frame driver with the reference's stage order and recycling
    schedule (engine/main_engine.cc:100-192).

    Mapping: AllocBlockArray -> CollectBlocksInFrustum -> UpdateBlocksSimple
    (or PredictOutlierRatio + UpdateBlocksBayesian), after which the frame
    counter advances.  Meshing: MarchingCubes on the same worklist, then
    CollectLowSurfelBlocks, and RecycleGarbageBlockArray when the counter is
    a multiple of 10.  Recycle (plain TSDF only): when counter % 15 == 14,
    StarveOccupiedBlockArray -> CollectGarbageBlockArray ->
    RecycleGarbageBlockArray.  recycle=False leaves blocks allocated for
    good (Meshing is then MarchingCubes alone).
    """

    def __init__(self, cfg: Config, recycle: bool = True):
        self.cfg = cfg
        self.vol = Volume(cfg)
        self.recycle_blocks = recycle
        self.integrated_frame_count = 0

    def mapping(self):
        v = self.vol
        v.alloc()
        v.collect_visible()
        if self.cfg.psdf:
            v.predict_inlier()
            v.integrate_psdf()
        else:
            v.integrate_tsdf()
        self.integrated_frame_count += 1

    def meshing(self):
        v = self.vol
        v.marching_cubes()
        if not self.recycle_blocks:
            return
        v.collect_low_surfel_blocks()
        if self.integrated_frame_count % 10 == 0:
            v.recycle_garbage_blocks(sync=False)

    def recycle(self):
        if not self.recycle_blocks or self.cfg.psdf or self.integrated_frame_count % 15 != 14:
            return
        v = self.vol
        v.starve_blocks()
        v.collect_garbage_blocks()
        v.recycle_garbage_blocks(sync=False)

    def frame(self, depth_u16: np.ndarray, wTc: np.ndarray, mesh: bool = True):
        self.vol.upload_depth(depth_u16, wTc)
        self.mapping()
        if mesh:
            self.meshing()
        self.recycle()


class Batch:
    """Independent scenes of one image size, every stage one launch over all
    scenes (mh_batch_*).  volumes[i] views scene i for status and dumps."""

    def __init__(self, cfgs: list[Config]):
        self.cfgs = list(cfgs)
        self.lib = A.load()
        n = len(self.cfgs)
        arr = lambda T, xs: (T * n)(*xs)
        self._mc = arr(A.McParams, [c.mc_params() for c in self.cfgs])
        self.h = C.c_void_p()
        A.check(self.lib.mh_batch_create(
            n, arr(A.HashParams, [c.hash_params() for c in self.cfgs]),
            arr(A.VolumeParams, [c.volume_params() for c in self.cfgs]),
            arr(A.MeshParams, [c.mesh_params() for c in self.cfgs]),
            arr(A.SensorParams, [c.sensor_params() for c in self.cfgs]), C.byref(self.h)),
            "mh_batch_create")
        self.volumes = []
        for i, c in enumerate(self.cfgs):
            vh = C.c_void_p()
            A.check(self.lib.mh_batch_volume(self.h, i, C.byref(vh)), "mh_batch_volume")
            self.volumes.append(Volume.wrap(c, vh))

    def close(self):
        if self.h:
            self.lib.mh_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr: int):
        A.check(self.lib.mh_batch_set_stream(self.h, C.c_void_p(stream_ptr)), "mh_batch_set_stream")

    def frame(self, depth, wTc: np.ndarray):
        """depth: (n, H, W) uint16 host array, or an int device pointer to
        the same layout (kept valid until the run completes)."""
        w = np.ascontiguousarray(np.asarray(wTc, np.float32).reshape(-1, 16))
        if isinstance(depth, int):
            rc = self.lib.mh_batch_frame_u16(self.h, C.c_void_p(depth), 1, A.ptr(w, C.c_float))
        else:
            d = np.ascontiguousarray(depth, dtype=np.uint16)
            rc = self.lib.mh_batch_frame_u16(self.h, d.ctypes.data_as(C.c_void_p), 0,
                                             A.ptr(w, C.c_float))
        A.check(rc, "mh_batch_frame_u16")

    def set_raycast(self, depth_ptr: int = 0, vertex_ptr: int = 0, normal_ptr: int = 0):
        n = len(self.cfgs)
        rp = (A.RayCasterParams * n)(*[c.raycaster_params() for c in self.cfgs])
        A.check(self.lib.mh_batch_set_raycast(self.h, rp, C.c_void_p(depth_ptr or None),
                                              C.c_void_p(vertex_ptr or None),
                                              C.c_void_p(normal_ptr or None)),
                "mh_batch_set_raycast")

    def set_integrate_tiling(self, tiling: int, workgroups_per_cu: int = 8):
        A.check(self.lib.mh_batch_set_integrate_tiling(self.h, tiling, workgroups_per_cu),
                "mh_batch_set_integrate_tiling")

    def set_marching_cubes_shape(self, shape: int):
        A.check(self.lib.mh_batch_set_marching_cubes_shape(self.h, shape),
                "mh_batch_set_marching_cubes_shape")

    def tune_marching_cubes(self, reps: int = 1):
        """The next mc_tune_calls(reps) graph replays with meshing alternate
        launch shape 0 with each other shape, each replay timed whole; the
        replay after them keeps the chosen shape (the rule of
        Volume.tune_marching_cubes)."""
        A.check(self.lib.mh_batch_tune_marching_cubes(self.h, reps), "mh_batch_tune_marching_cubes")

    def marching_cubes_tuning(self) -> A.TuneMcResult:
        r = A.TuneMcResult()
        A.check(self.lib.mh_batch_marching_cubes_tuning(self.h, C.byref(r)),
                "mh_batch_marching_cubes_tuning")
        return r

    def run(self, mesh: bool = True, raycast: bool = False, graph: bool = False,
            tune: bool = False) -> A.TuneResult | None:
        flags = ((A.MH_RUN_MESH if mesh else 0) | (A.MH_RUN_RAYCAST if raycast else 0) |
                 (A.MH_RUN_GRAPH if graph else 0) | (A.MH_RUN_TUNE if tune else 0))
        r = A.TuneResult()
        A.check(self.lib.mh_batch_run(self.h, self._mc, flags, C.byref(r) if tune else None),
                "mh_batch_run")
        return r if tune else None

    def synchronize(self):
        A.check(self.lib.mh_batch_synchronize(self.h), "mh_batch_synchronize")
