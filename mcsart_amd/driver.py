"""McSART drivers: the ``entry_AUMISART`` gateway mirror and device-resident jobs.

``entry_aumisart(img0, proj, geo, it)`` follows ``src/entry_AUMISART.cu``: it parses GEO_PARA /
ITER_PARA dictionaries with the reference's field names, aliases and defaults (:18-160), picks the
static or motion-compensated overload by the presence of ``alpha_x`` (:193-207) and returns
``(img, nrm, alpha_x, residual_proj)`` (:172-190).  ``Job`` wraps the library's device-resident
job API for callers that keep a scan on the GPU (benchmarks, batching, serving).
"""
from __future__ import annotations

import ctypes
from typing import Mapping, Optional, Sequence

import numpy as np

from ._lib import McsartError, check, load
from .geometry import Geometry


class MexBadInput(ValueError):
    """Raised where the reference calls mexErrMsgIdAndTxt("MATLAB:badInput", ...)."""


def _get(d: Mapping, *names):
    for n in names:
        if n in d and d[n] is not None:
            return d[n]
    return None


_SCALAR_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p,
                              ctypes.POINTER(ctypes.c_double))


class CGeoParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nx", "ny", "nz", "na", "nb")] + \
               [(n, ctypes.c_float) for n in ("dx", "da", "db", "ai", "bi", "SO", "SD", "angle")] + \
               [("ind", ctypes.c_int), ("defaulted", ctypes.c_uint)]


class CIterParams(ctypes.Structure):
    _fields_ = [("n_iter", ctypes.c_int), ("n_views", ctypes.c_int), ("outIter", ctypes.c_int),
                ("lam", ctypes.c_float), ("motion", ctypes.c_int), ("defaulted", ctypes.c_uint)]


GATEWAYS = {"aumisart": 0, "backprojection": 1, "projection": 2, "diff": 3}


def _lookup(d: Mapping):
    """A mcsart_scalar_fn over a mapping: present if the key exists and is not None; the value
    is the first element, as MATLAB's mxGetScalar."""
    def fn(_user, name, out):
        v = d.get(name.decode())
        if v is None:
            return 0
        if out:
            a = np.asarray(v, dtype=np.float64).ravel()
            out[0] = float(a[0]) if a.size else 0.0
        return 1
    return _SCALAR_FN(fn)


def _raise_bad_input(rc, what):
    msg = load().mcsart_last_error().decode()
    if msg.startswith("MATLAB:badInput: "):
        raise MexBadInput(msg[len("MATLAB:badInput: "):])
    check(rc, what)


def parse_geo_params(geo: Mapping, gateway: str = "aumisart") -> CGeoParams:
    """GEO_PARA of one of the gateways, parsed by the library (mcsart_parse_geo)."""
    p = CGeoParams()
    cb = _lookup(geo)
    rc = load().mcsart_parse_geo(GATEWAYS[gateway], ctypes.cast(cb, ctypes.c_void_p), None,
                                 ctypes.byref(p))
    if rc:
        _raise_bad_input(rc, "mcsart_parse_geo")
    return p


def parse_geo(geo: Mapping, gateway: str = "aumisart") -> Geometry:
    """GEO_PARA -> Geometry (src/entry_AUMISART.cu:18-108; aliases na|nv, nb|nu, SO|SI,
    SD|DI+SO; defaults da = db = dx = 1, ai = bi = 0)."""
    p = parse_geo_params(geo, gateway)
    return Geometry(p.nx, p.ny, p.nz, p.na, p.nb, float(p.da), float(p.db), float(p.ai),
                    float(p.bi), float(p.SO), float(p.SD))


def parse_iter(it: Mapping):
    """ITER_PARA (src/entry_AUMISART.cu:111-160), scalars and presence checks by the library
    (mcsart_parse_iter); the arrays are taken from the mapping."""
    p = CIterParams()
    cb = _lookup(it)
    rc = load().mcsart_parse_iter(ctypes.cast(cb, ctypes.c_void_p), None, ctypes.byref(p))
    if rc:
        _raise_bad_input(rc, "mcsart_parse_iter")
    out = {k: np.ascontiguousarray(it[k], dtype=np.float32)
           for k in ("volumes", "flows", "err_weights", "angles")}
    op = np.ascontiguousarray(it["op_iter"], dtype=np.int32).ravel()
    if op.size < p.n_iter:
        raise MexBadInput("op_iter has fewer than n_iter entries")
    return dict(n_iter=p.n_iter, op_iter=op[:p.n_iter], n_views=p.n_views, lam=float(p.lam),
                outIter=p.outIter, motion=bool(p.motion), **out)


def entry_aumisart(img0: np.ndarray, proj: np.ndarray, geo: Mapping, it: Mapping):
    """``[img, nrm, alpha_x, res] = entry_AUMISART(img0, proj, geo, iter)``."""
    g = parse_geo(geo)
    p = parse_iter(it)
    V = p["n_views"]
    lib = load()
    img0 = np.ascontiguousarray(img0, dtype=np.float32).reshape(g.vol_shape)
    proj = np.ascontiguousarray(proj, dtype=np.float32).reshape((V,) + g.det_shape)
    out_img = np.empty(g.vol_shape, np.float32)
    out_norm = np.empty(p["n_iter"] * V, np.float32)
    out_ax = np.empty(g.vol_shape, np.float32)
    out_proj = np.empty((V,) + g.det_shape, np.float32)
    args = [out_img.ctypes.data, out_proj.ctypes.data, out_norm.ctypes.data, out_ax.ctypes.data,
            img0.ctypes.data, proj.ctypes.data, g.nx, g.ny, g.nz, g.na, g.nb, p["outIter"], V,
            p["n_iter"], p["op_iter"].ctypes.data, g.da, g.db, g.ai, g.bi, g.SO, g.SD,
            float(parse_geo_params(geo).dx), p["lam"], p["volumes"].ctypes.data,
            p["flows"].ctypes.data, p["err_weights"].ctypes.data, p["angles"].ctypes.data]
    if not p["motion"]:
        check(lib.mcsart_aumisart(*args), "mcsart_aumisart")
    else:
        fs = [np.ascontiguousarray(it[k], dtype=np.float32).reshape(g.vol_shape)
              for k in ("alpha_x", "alpha_y", "alpha_z", "beta_x", "beta_y", "beta_z")]
        check(lib.mcsart_aumisart_motion(*args, *[f.ctypes.data for f in fs]),
              "mcsart_aumisart_motion")
    return out_img, out_norm, out_ax, out_proj


# ---- device-resident jobs -------------------------------------------------------------------
class CGeom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nx", "ny", "nz", "na", "nb")] + \
               [(n, ctypes.c_float) for n in ("da", "db", "ai", "bi", "SO", "SD")]

    @staticmethod
    def of(g: Geometry) -> "CGeom":
        return CGeom(g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD)


class COpts(ctypes.Structure):
    _fields_ = [("motion", ctypes.c_int), ("nonneg", ctypes.c_int),
                ("keep_residual", ctypes.c_int), ("lam", ctypes.c_float)]


class CTuning(ctypes.Structure):
    """Launch shapes of the footprint kernels (0 = automatic)."""
    _fields_ = [("proj_split", ctypes.c_int), ("back_zslab", ctypes.c_int),
                ("back_stage", ctypes.c_int), ("proj_rows", ctypes.c_int),
                ("back_zslab_motion", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


TIMING_KINDS = {"projector": 0, "backprojector": 1, "warp": 2, "motion": 3}


class Context:
    """A stream on one device (mcsart_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = load()
        p = ctypes.c_void_p()
        check(self._lib.mcsart_ctx_create(ctypes.byref(p), device), "mcsart_ctx_create")
        self.ptr = p.value

    def synchronize(self):
        check(self._lib.mcsart_ctx_synchronize(self.ptr), "mcsart_ctx_synchronize")

    def enable_timing(self, on: bool = True):
        check(self._lib.mcsart_ctx_enable_timing(self.ptr, int(on)), "mcsart_ctx_enable_timing")

    def reset_timing(self):
        check(self._lib.mcsart_ctx_reset_timing(self.ptr), "mcsart_ctx_reset_timing")

    def kernel_time(self, kind: str):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        check(self._lib.mcsart_ctx_kernel_time(self.ptr, TIMING_KINDS[kind], ctypes.byref(ms),
                                               ctypes.byref(n)), "mcsart_ctx_kernel_time")
        return ms.value, n.value

    def close(self):
        if self.ptr:
            self._lib.mcsart_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_bytes(g: Geometry, n_views: int, n_iter: int, keep_residual: bool = False) -> int:
    cg = CGeom.of(g)
    return int(load().mcsart_job_bytes(ctypes.byref(cg), n_views, n_iter, int(keep_residual)))


def view_voxels(g: Geometry, angles) -> np.ndarray:
    """(n_views, 5) int64 per view: sizes of the culled sets (voxels in its projector's read
    set; the same with the volume's edge shell; voxels where its ratio backprojection can be
    nonzero; the union of the latter with the next view's read set and shell) and the layouts its
    rays read (bit 0 natural, bit 1 transposed) (mcsart_view_voxels; host arithmetic, no device
    needed)."""
    ang = np.ascontiguousarray(angles, np.float32)
    out = np.zeros((ang.size, 5), np.int64)
    cg = CGeom.of(g)
    check(load().mcsart_view_voxels(ctypes.byref(cg), int(ang.size), ang.ctypes.data,
                                    out.ctypes.data), "mcsart_view_voxels")
    return out


class Job:
    """One scan resident on the device (mcsart_job)."""

    def __init__(self, ctx: Context, g: Geometry, angles, volumes, flows, op_iter,
                 lam: float = 1.0, motion: bool = True, nonneg: bool = False,
                 keep_residual: bool = False):
        self._lib = load()
        self.ctx = ctx
        self.g = g
        self.angles = np.ascontiguousarray(angles, np.float32)
        self.volumes = np.ascontiguousarray(volumes, np.float32)
        self.flows = np.ascontiguousarray(flows, np.float32)
        self.op_iter = np.ascontiguousarray(op_iter, np.int32)
        self.V = int(self.angles.size)
        self.n_iter = int(self.op_iter.size)
        self.keep_residual = keep_residual
        cg = CGeom.of(g)
        opts = COpts(int(motion), int(nonneg), int(keep_residual), float(lam))
        p = ctypes.c_void_p()
        check(self._lib.mcsart_job_create(ctx.ptr, ctypes.byref(cg), self.V,
                                          self.angles.ctypes.data, self.volumes.ctypes.data,
                                          self.flows.ctypes.data, ctypes.byref(opts),
                                          self.n_iter, self.op_iter.ctypes.data, ctypes.byref(p)),
              "mcsart_job_create")
        self.ptr = p.value

    def set_projections(self, proj: np.ndarray):
        proj = np.ascontiguousarray(proj, np.float32)
        assert proj.size == self.V * self.g.ndet
        check(self._lib.mcsart_job_set_projections(self.ptr, proj.ctypes.data),
              "mcsart_job_set_projections")

    def set_projections_device(self, d_proj):
        check(self._lib.mcsart_job_set_projections_device(self.ptr, int(getattr(d_proj, "ptr", d_proj))),
              "mcsart_job_set_projections_device")

    def set_image(self, img: Optional[np.ndarray]):
        if img is None:
            check(self._lib.mcsart_job_set_image(self.ptr, None), "mcsart_job_set_image")
        else:
            img = np.ascontiguousarray(img, np.float32)
            assert img.size == self.g.nvox
            check(self._lib.mcsart_job_set_image(self.ptr, img.ctypes.data), "mcsart_job_set_image")

    def set_fields(self, fields: Optional[Sequence[np.ndarray]]):
        arr = (ctypes.c_void_p * 6)()
        keep = []
        for c in range(6):
            if fields is None or fields[c] is None:
                arr[c] = None
            else:
                f = np.ascontiguousarray(fields[c], np.float32)
                assert f.size == self.g.nvox
                keep.append(f)
                arr[c] = f.ctypes.data
        check(self._lib.mcsart_job_set_fields(self.ptr, arr), "mcsart_job_set_fields")

    def reset(self):
        check(self._lib.mcsart_job_reset(self.ptr), "mcsart_job_reset")

    def set_graphs(self, enable: bool):
        check(self._lib.mcsart_job_set_graphs(self.ptr, int(enable)), "mcsart_job_set_graphs")

    def set_culling(self, mode: int):
        """0: every voxel every view; 1: skip work no later kernel sees (default); 2: as 1, with
        the skipped writes made as NaN (checks that nothing reads them)."""
        check(self._lib.mcsart_job_set_culling(self.ptr, int(mode)), "mcsart_job_set_culling")

    def set_tuning(self, proj_split: int = 0, back_zslab: int = 0, back_stage: int = 0,
                   proj_rows: int = 0, back_zslab_motion: int = 0):
        t = CTuning(proj_split, back_zslab, back_stage, proj_rows, back_zslab_motion)
        check(self._lib.mcsart_job_set_tuning(self.ptr, ctypes.byref(t)), "mcsart_job_set_tuning")

    def tuning(self) -> dict:
        t = CTuning()
        check(self._lib.mcsart_job_get_tuning(self.ptr, ctypes.byref(t)), "mcsart_job_get_tuning")
        return t.as_dict()

    def autotune(self) -> dict:
        """Time candidate launch shapes on a few views, keep the fastest, return it."""
        t = CTuning()
        check(self._lib.mcsart_job_autotune(self.ptr, ctypes.byref(t)), "mcsart_job_autotune")
        return t.as_dict()

    KERNEL_KINDS = {"projector": 0, "backprojector": 1, "warp": 2, "warp_add": 3, "motion": 4,
                    "projector_motion": 5}

    def kernel_time(self, kind: str, reps: int = 3) -> float:
        """Average microseconds per launch of one of KERNEL_KINDS over all views of the scan,
        launched back to back between two HIP events on the job's stream (scratch outputs)."""
        us = ctypes.c_double()
        check(self._lib.mcsart_job_kernel_time(self.ptr, self.KERNEL_KINDS[kind], int(reps),
                                               ctypes.byref(us)),
              "mcsart_job_kernel_time")
        return us.value

    def view_layouts(self) -> np.ndarray:
        """Per view: bit 0 = rays read the natural layout, bit 1 = the transposed layout."""
        out = np.zeros(self.V, np.int32)
        check(self._lib.mcsart_job_view_layouts(self.ptr, out.ctypes.data),
              "mcsart_job_view_layouts")
        return out

    def run(self):
        check(self._lib.mcsart_job_run(self.ptr), "mcsart_job_run")

    def run_sweep(self, it: int):
        check(self._lib.mcsart_job_run_sweep(self.ptr, it), "mcsart_job_run_sweep")

    def image(self) -> np.ndarray:
        out = np.empty(self.g.vol_shape, np.float32)
        check(self._lib.mcsart_job_get_image(self.ptr, out.ctypes.data), "mcsart_job_get_image")
        return out

    def norms(self) -> np.ndarray:
        out = np.empty(self.n_iter * self.V, np.float32)
        check(self._lib.mcsart_job_get_norms(self.ptr, out.ctypes.data), "mcsart_job_get_norms")
        return out

    def residual(self) -> np.ndarray:
        out = np.empty((self.V,) + self.g.det_shape, np.float32)
        check(self._lib.mcsart_job_get_residual(self.ptr, out.ctypes.data), "mcsart_job_get_residual")
        return out

    def fields(self):
        outs = [np.empty(self.g.vol_shape, np.float32) for _ in range(6)]
        arr = (ctypes.c_void_p * 6)(*[o.ctypes.data for o in outs])
        check(self._lib.mcsart_job_get_fields(self.ptr, arr), "mcsart_job_get_fields")
        return outs

    def device_image_ptr(self) -> int:
        p = ctypes.c_void_p()
        check(self._lib.mcsart_job_device_image(self.ptr, ctypes.byref(p)), "mcsart_job_device_image")
        return p.value

    def close(self):
        if self.ptr:
            self._lib.mcsart_job_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["entry_aumisart", "parse_geo", "parse_iter", "Context", "Job", "job_bytes",
           "MexBadInput", "McsartError"]
