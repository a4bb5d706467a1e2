"""ctypes wrapper of the FP64 oracle (oracle/_build/libmh_oracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline -- never by the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libmh_oracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp = C.c_void_p
        P = C.POINTER
        sig = {
            "orc_create": (vp, [vp, vp, vp, C.c_int]),
            "orc_destroy": (None, [vp]),
            "orc_hash_bucket": (C.c_uint32, [C.c_int32, C.c_int32, C.c_int32, C.c_uint32]),
            "orc_invert4f": (None, [P(C.c_float), P(C.c_float)]),
            "orc_invert4d": (None, [P(C.c_double), P(C.c_double)]),
            "orc_set_frame": (C.c_int, [vp, P(C.c_uint16), P(C.c_float)]),
            "orc_alloc": (C.c_int, [vp]),
            "orc_collect": (C.c_uint32, [vp, C.c_int]),
            "orc_integrate_tsdf": (C.c_int, [vp]),
            "orc_predict_inlier": (C.c_int, [vp]),
            "orc_integrate_psdf": (C.c_int, [vp]),
            "orc_marching_cubes": (C.c_int, [vp, vp]),
            "orc_raycast": (C.c_int, [vp, vp, P(C.c_float), P(C.c_float), P(C.c_float),
                                      P(C.c_float)]),
            "orc_pd_init": (C.c_int, [vp]),
            "orc_pd_iterate": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, P(C.c_double)]),
            "orc_block_count": (C.c_uint32, [vp]),
            "orc_dump_blocks": (C.c_int, [vp, P(C.c_int32), P(C.c_double), P(C.c_double),
                                          P(C.c_uint32), P(C.c_double), P(C.c_double),
                                          P(C.c_uint16)]),
            "orc_dump_pd": (C.c_int, [vp, P(C.c_double), P(C.c_uint8)]),
            "orc_load_blocks": (C.c_int, [vp, C.c_uint32, P(C.c_int32), P(C.c_float),
                                          P(C.c_float), P(C.c_uint32), P(C.c_float),
                                          P(C.c_float)]),
            "orc_dump_hash": (C.c_uint32, [vp, P(C.c_int32)]),
            "orc_dump_candidates": (C.c_uint32, [vp, P(C.c_int32)]),
            "orc_tri_count": (C.c_uint32, [vp]),
            "orc_vtx_count": (C.c_uint32, [vp]),
            "orc_dump_mesh": (C.c_int, [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32),
                                        P(C.c_double), P(C.c_int32)]),
            "orc_inlier": (None, [vp, P(C.c_double)]),
            "orc_depth": (None, [vp, P(C.c_float)]),
            "orc_dropped": (C.c_uint32, [vp]),
            "orc_spatial_gradient": (C.c_int, [vp, P(C.c_float), P(C.c_double)]),
            "orc_voxel_updates": (C.c_uint64, [vp]),
            "orc_list_inserts": (C.c_uint32, [vp]),
            "orc_mc_ntri": (C.c_int, [C.c_int]),
            "orc_mc_tri_edge": (C.c_int, [C.c_int, C.c_int]),
            "orc_mc_edge_mask": (C.c_uint32, [C.c_int]),
            "orc_mc_corner": (C.c_int, [C.c_int, C.c_int]),
            "orc_mc_edge_a": (C.c_int, [C.c_int]),
            "orc_mc_edge_b": (C.c_int, [C.c_int]),
            "orc_raycast_adjoint": (C.c_int, [vp, vp, C.c_int, P(C.c_float), P(C.c_double),
                                              P(C.c_float), C.c_double, P(C.c_double),
                                              P(C.c_double), P(C.c_int64), P(C.c_double),
                                              P(C.c_double)]),
            "orc_perturb_sdf": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_double]),
            "orc_set_color": (C.c_int, [vp, P(C.c_uint8)]),
            "orc_point_to_psdf": (C.c_int, [vp, P(C.c_float), C.c_int, P(C.c_double)]),
            "orc_collect_low_surfel": (C.c_int, [vp]),
            "orc_starve": (C.c_int, [vp]),
            "orc_collect_garbage": (C.c_int, [vp]),
            "orc_recycle_garbage": (C.c_uint32, [vp]),
            "orc_dump_cand_flags": (C.c_uint32, [vp, P(C.c_int32)]),
            "orc_val_to_rgb": (None, [C.c_double, C.c_double, C.c_double, P(C.c_double)]),
            "orc_set_threads": (None, [C.c_int]),
            "orc_load_vertices": (C.c_int, [vp, C.c_uint32, P(C.c_int32), P(C.c_double)]),
            "orc_get_threads": (C.c_int, []),
            "orc_set_vertex": (C.c_int, [vp, P(C.c_int32), P(C.c_float), P(C.c_double),
                                         C.c_double]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, ct):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


def _m16(m):
    a = np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))
    return (C.c_float * 16)(*a.tolist())


class Oracle:
    """CPU volume with the same stage functions as the HIP library."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        self._hp = cfg.hash_params()
        self._vp = cfg.volume_params()
        self._sp = cfg.sensor_params()
        self.h = self.L.orc_create(C.byref(self._hp), C.byref(self._vp), C.byref(self._sp),
                                   1 if cfg.psdf else 0)

    def __del__(self):
        try:
            if self.h:
                self.L.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_frame(self, depth_u16, wTc):
        d = np.ascontiguousarray(depth_u16, np.uint16)
        self.L.orc_set_frame(self.h, _p(d, C.c_uint16), _m16(wTc))

    def set_color(self, bgra):
        c = np.ascontiguousarray(bgra, np.uint8)
        self.L.orc_set_color(self.h, _p(c, C.c_uint8))

    def alloc(self):
        self.L.orc_alloc(self.h)

    def collect_visible(self) -> int:
        return self.L.orc_collect(self.h, 1)

    def collect_all(self) -> int:
        return self.L.orc_collect(self.h, 0)

    def integrate_tsdf(self):
        self.L.orc_integrate_tsdf(self.h)

    def predict_inlier(self):
        self.L.orc_predict_inlier(self.h)

    def integrate_psdf(self):
        self.L.orc_integrate_psdf(self.h)

    def marching_cubes(self, params=None):
        p = params if params is not None else self.cfg.mc_params()
        self.L.orc_marching_cubes(self.h, C.byref(p))

    def raycast(self, cTw, params=None):
        p = params if params is not None else self.cfg.raycaster_params()
        n = p.width * p.height
        d = np.zeros(n, np.float32)
        vtx = np.zeros(3 * n, np.float32)
        nrm = np.zeros(3 * n, np.float32)
        self.L.orc_raycast(self.h, C.byref(p), _m16(cTw), _p(d, C.c_float), _p(vtx, C.c_float),
                           _p(nrm, C.c_float))
        return (d.reshape(p.height, p.width), vtx.reshape(p.height, p.width, 3),
                nrm.reshape(p.height, p.width, 3))

    def set_vertex(self, key, pos, normal, radius):
        """Place a vertex record in voxel key[:3]'s slot key[3] (test hook)."""
        k = np.ascontiguousarray(key, np.int32)
        p = np.ascontiguousarray(pos, np.float32)
        n = np.ascontiguousarray(normal, np.float64)
        assert self.L.orc_set_vertex(self.h, _p(k, C.c_int32), _p(p, C.c_float),
                                     _p(n, C.c_double), float(radius)) == 0

    def load_vertices(self, vkey, vval):
        """Replace the payloads of the given vertices (dump_mesh layout),
        keeping their reference counts."""
        k = np.ascontiguousarray(vkey, np.int32)
        v = np.ascontiguousarray(vval, np.float64)
        assert self.L.orc_load_vertices(self.h, len(k), _p(k, C.c_int32), _p(v, C.c_double)) == 0

    def pd_init(self):
        self.L.orc_pd_init(self.h)

    def pd_iterate(self, lam, sigma, tau):
        e = np.zeros(3, np.float64)
        self.L.orc_pd_iterate(self.h, lam, sigma, tau, _p(e, C.c_double))
        return e

    def frame(self, depth_u16, wTc, mesh=True):
        self.set_frame(depth_u16, wTc)
        self.alloc()
        self.collect_visible()
        if self.cfg.psdf:
            self.predict_inlier()
            self.integrate_psdf()
        else:
            self.integrate_tsdf()
        if mesh:
            self.marching_cubes()

    # -- localisation (localizing/point_to_psdf.cu) ---------------------------
    def point_to_psdf(self, wTc, corrected: bool = False) -> dict:
        """Normal equations of the current frame's depth against the volume
        at camera->world pose wTc: A (6x6), b (6), err, count."""
        out = np.zeros(44, np.float64)
        self.L.orc_point_to_psdf(self.h, _m16(wTc), 1 if corrected else 0, _p(out, C.c_double))
        return {"A": out[:36].reshape(6, 6).copy(), "b": out[36:42].copy(), "err": out[42],
                "count": int(out[43])}

    # -- recycling (mapping/recycle.cu) ---------------------------------------
    def collect_low_surfel_blocks(self):
        self.L.orc_collect_low_surfel(self.h)

    def starve_blocks(self):
        self.L.orc_starve(self.h)

    def collect_garbage_blocks(self):
        self.L.orc_collect_garbage(self.h)

    def recycle_garbage_blocks(self) -> int:
        return int(self.L.orc_recycle_garbage(self.h))

    def dump_cand_flags(self) -> np.ndarray:
        n = self.L.orc_dump_cand_flags(self.h, None)
        rec = np.zeros((n, 4), np.int32)
        self.L.orc_dump_cand_flags(self.h, _p(rec, C.c_int32))
        return rec

    # -- dumps ---------------------------------------------------------------
    def dump_blocks(self) -> dict:
        n = self.L.orc_block_count(self.h)
        pos = np.zeros((n, 3), np.int32)
        sdf = np.zeros((n, 512), np.float64)
        w = np.zeros((n, 512), np.float64)
        col = np.zeros((n, 512), np.uint32)
        a = np.zeros((n, 512), np.float64)
        b = np.zeros((n, 512), np.float64)
        cube = np.zeros((n, 512), np.uint16)
        self.L.orc_dump_blocks(self.h, _p(pos, C.c_int32), _p(sdf, C.c_double),
                               _p(w, C.c_double), _p(col, C.c_uint32), _p(a, C.c_double),
                               _p(b, C.c_double), _p(cube, C.c_uint16))
        return {"pos": pos, "sdf": sdf, "weight": w, "color": col, "a": a, "b": b, "cube": cube}

    def dump_pd(self) -> dict:
        n = self.L.orc_block_count(self.h)
        pdv = np.zeros((n, 6, 512), np.float64)
        mask = np.zeros((n, 512), np.uint8)
        self.L.orc_dump_pd(self.h, _p(pdv, C.c_double), _p(mask, C.c_uint8))
        return {"sdf0": pdv[:, 0], "sdf_bar": pdv[:, 1], "w_pd": pdv[:, 2], "p": pdv[:, 3:6],
                "mask": mask}

    def load_blocks(self, pos, sdf, weight, color=None, a=None, b=None):
        pos = np.ascontiguousarray(pos, np.int32)
        f = lambda x: None if x is None else np.ascontiguousarray(x, np.float32)
        col = None if color is None else np.ascontiguousarray(color, np.uint32)
        sdf, weight, a, b = f(sdf), f(weight), f(a), f(b)
        rc = self.L.orc_load_blocks(self.h, len(pos), _p(pos, C.c_int32), _p(sdf, C.c_float),
                                    _p(weight, C.c_float), _p(col, C.c_uint32),
                                    _p(a, C.c_float), _p(b, C.c_float))
        assert rc == 0

    def dump_hash(self) -> np.ndarray:
        n = self.L.orc_dump_hash(self.h, None)
        rec = np.zeros((n, 6), np.int32)
        self.L.orc_dump_hash(self.h, _p(rec, C.c_int32))
        return rec

    def dump_candidates(self) -> np.ndarray:
        n = self.L.orc_dump_candidates(self.h, None)
        pos = np.zeros((n, 3), np.int32)
        self.L.orc_dump_candidates(self.h, _p(pos, C.c_int32))
        return pos

    def dump_mesh(self) -> dict:
        nt = self.L.orc_tri_count(self.h)
        nv = self.L.orc_vtx_count(self.h)
        tri = np.zeros((nt, 12), np.int32)
        tcube = np.zeros((nt, 4), np.int32)
        vkey = np.zeros((nv, 4), np.int32)
        vval = np.zeros((nv, 10), np.float64)
        vref = np.zeros(nv, np.int32)
        self.L.orc_dump_mesh(self.h, _p(tri, C.c_int32), _p(tcube, C.c_int32),
                             _p(vkey, C.c_int32), _p(vval, C.c_double), _p(vref, C.c_int32))
        return {"tri": tri, "tri_cube": tcube, "vkey": vkey, "vval": vval, "vref": vref}

    def inlier(self) -> np.ndarray:
        out = np.zeros(self.cfg.width * self.cfg.height, np.float64)
        self.L.orc_inlier(self.h, _p(out, C.c_double))
        return out.reshape(self.cfg.height, self.cfg.width)

    def depth(self) -> np.ndarray:
        out = np.zeros(self.cfg.width * self.cfg.height, np.float32)
        self.L.orc_depth(self.h, _p(out, C.c_float))
        return out.reshape(self.cfg.height, self.cfg.width)

    def dropped(self) -> int:
        return self.L.orc_dropped(self.h)

    def spatial_gradient(self, pos):
        p = (C.c_float * 3)(*[float(x) for x in pos])
        g = (C.c_double * 3)()
        ok = self.L.orc_spatial_gradient(self.h, p, g)
        return np.array(g[:]), bool(ok)

    def voxel_updates(self) -> int:
        return self.L.orc_voxel_updates(self.h)

    def raycast_adjoint(self, fp64: bool, pose, obs=None, params=None, jac=False, grad_sdf=False,
                        max_residual=float("inf")):
        """Raycast adjoint.  fp64=False: FP32 march from pose = cTw (float32
        4x4, as the GPU); fp64=True: FP64 march from pose = wTc (float64).
        Returns depth, path (int64, -1: no hit), optional per-pixel tangents
        (H, W, 11), optional dL/dsdf (blocks, 512) and the summary: loss,
        grad[6], hessian (6x6), pixels."""
        p = params if params is not None else self.cfg.raycaster_params()
        n = p.width * p.height
        depth = np.zeros(n, np.float64)
        path = np.zeros(n, np.int64)
        J = np.zeros((n, 11), np.float64) if jac else None
        nb = self.L.orc_block_count(self.h)
        G = np.zeros((nb, 512), np.float64) if grad_sdf else None
        summ = np.zeros(30, np.float64)
        o = None if obs is None else np.ascontiguousarray(obs, np.float32).reshape(-1)
        if fp64:
            w = np.ascontiguousarray(np.asarray(pose, np.float64).reshape(16))
            c32 = None
        else:
            w = None
            c32 = np.ascontiguousarray(np.asarray(pose, np.float32).reshape(16))
        self.L.orc_raycast_adjoint(self.h, C.byref(p), 1 if fp64 else 0, _p(c32, C.c_float),
                                   _p(w, C.c_double), _p(o, C.c_float), max_residual,
                                   _p(depth, C.c_double),
                                   _p(J, C.c_double), _p(path, C.c_int64), _p(G, C.c_double),
                                   _p(summ, C.c_double))
        H = np.zeros((6, 6))
        t = 7
        for i in range(6):
            for j in range(i, 6):
                H[i, j] = H[j, i] = summ[t]
                t += 1
        out = {"depth": depth.reshape(p.height, p.width), "path": path.reshape(p.height, p.width),
               "loss": summ[0], "grad": summ[1:7].copy(), "hessian": H, "pixels": int(summ[28]),
               "outliers": int(summ[29])}
        if jac:
            out["jac"] = J.reshape(p.height, p.width, 11)
        if grad_sdf:
            out["grad_sdf"] = G
        return out

    def perturb_sdf(self, block: int, voxel: int, delta: float):
        assert self.L.orc_perturb_sdf(self.h, block, voxel, delta) == 0


def set_threads(n: int) -> None:
    """Threads of the oracle's block- and pixel-parallel stages (results do
    not depend on it)."""
    lib().orc_set_threads(int(n))


def val_to_rgb(v: float, vmin: float, vmax: float) -> np.ndarray:
    out = np.zeros(3, np.float64)
    lib().orc_val_to_rgb(float(v), float(vmin), float(vmax), _p(out, C.c_double))
    return out


def hash_bucket(x: int, y: int, z: int, n: int) -> int:
    return lib().orc_hash_bucket(x, y, z, n)


def invert4f(m) -> np.ndarray:
    out = (C.c_float * 16)()
    lib().orc_invert4f(_m16(m), out)
    return np.array(out[:], np.float32).reshape(4, 4)
