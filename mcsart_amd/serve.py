"""Python side of the reconstruction server (mcsart_server in libmcsart_hip).

A ``Server`` owns a scheduler thread in the library and ``slots`` streams on one device.
``submit`` queues one scan (geometry, views, schedule, measured projections, optional initial
image and 5DCT fields); the server admits it when a slot is free and its device footprint fits
the HBM budget, runs it as a device-resident job concurrently with the other admitted jobs, and
writes the results into host arrays owned by the returned ``Ticket``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from ._lib import check, load
from .driver import COpts, CGeom, CTuning
from .geometry import Geometry

_FP = ctypes.POINTER(ctypes.c_float)


class CServerOpts(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("autotune", ctypes.c_int), ("proj_split", ctypes.c_int),
                ("back_zslab", ctypes.c_int), ("back_stage", ctypes.c_int),
                ("proj_rows", ctypes.c_int), ("back_zslab_motion", ctypes.c_int)]


class CRequest(ctypes.Structure):
    _fields_ = [("geom", CGeom), ("n_views", ctypes.c_int),
                ("angles", _FP), ("volumes", _FP), ("flows", _FP),
                ("opts", COpts), ("n_iter", ctypes.c_int), ("op_iter", ctypes.POINTER(ctypes.c_int)),
                ("h_proj", _FP), ("h_img", _FP), ("h_fields", _FP * 6),
                ("out_img", _FP), ("out_norm", _FP), ("out_fields", _FP * 6)]


class CRequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("status", ctypes.c_int), ("slot", ctypes.c_int),
                ("bytes", ctypes.c_size_t), ("queue_ms", ctypes.c_double),
                ("run_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("tuning", CTuning), ("admit_at_ms", ctypes.c_double),
                ("done_at_ms", ctypes.c_double)]


STATES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}


def _fp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_FP)


@dataclass
class Ticket:
    """A submitted scan: keeps its host inputs alive and receives its results."""
    id: int
    image: np.ndarray
    norms: np.ndarray
    fields: Optional[list]
    _keep: list = field(default_factory=list, repr=False)


class Server:
    def __init__(self, slots: int = 4, hbm_budget: int = 0, autotune: bool = False,
                 tuning: Sequence[int] = (0, 0, 0, 0, 0), device: int = 0):
        self._lib = load()
        o = CServerOpts(int(slots), int(hbm_budget), int(autotune), *[int(t) for t in tuning])
        p = ctypes.c_void_p()
        check(self._lib.mcsart_server_create(int(device), ctypes.byref(o), ctypes.byref(p)),
              "mcsart_server_create")
        self.ptr = p.value

    def submit(self, g: Geometry, angles, volumes, flows, op_iter, proj: np.ndarray,
               image: Optional[np.ndarray] = None, fields: Optional[Sequence[np.ndarray]] = None,
               lam: float = 1.0, motion: bool = True, nonneg: bool = False,
               want_fields: bool = False) -> Ticket:
        f32 = lambda x: np.ascontiguousarray(x, dtype=np.float32)
        angles, volumes, flows = f32(angles), f32(volumes), f32(flows)
        op = np.ascontiguousarray(op_iter, dtype=np.int32)
        V = int(angles.size)
        proj = f32(proj)
        if proj.size != V * g.ndet:
            raise ValueError("projections must hold n_views * na * nb floats")
        img = None if image is None else f32(image)
        flds = [None] * 6 if fields is None else [None if f is None else f32(f) for f in fields]
        out_img = np.empty(g.vol_shape, np.float32)
        out_norm = np.empty(len(op) * V, np.float32)   # reference nrm(n_iter*V, 1) order
        out_f = [np.empty(g.vol_shape, np.float32) for _ in range(6)] if want_fields else None
        r = CRequest()
        r.geom = CGeom.of(g)
        r.n_views = V
        r.angles, r.volumes, r.flows = _fp(angles), _fp(volumes), _fp(flows)
        r.opts = COpts(int(motion), int(nonneg), 0, float(lam))
        r.n_iter = int(op.size)
        r.op_iter = op.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        r.h_proj, r.h_img = _fp(proj), _fp(img)
        for c in range(6):
            r.h_fields[c] = _fp(flds[c])
            r.out_fields[c] = _fp(out_f[c]) if out_f else None
        r.out_img, r.out_norm = _fp(out_img), _fp(out_norm)
        rid = ctypes.c_longlong()
        check(self._lib.mcsart_server_submit(self.ptr, ctypes.byref(r), ctypes.byref(rid)),
              "mcsart_server_submit")
        keep = [angles, volumes, flows, op, proj, img, flds, r]
        return Ticket(rid.value, out_img, out_norm, out_f, keep)

    def wait(self, t: Ticket, timeout_s: float = -1.0) -> Ticket:
        check(self._lib.mcsart_server_wait(self.ptr, t.id, float(timeout_s)), "mcsart_server_wait")
        return t

    def cancel(self, t: Ticket):
        """Cancel a queued or running request (mcsart_server_cancel): a running one stops after
        the sweeps already on the device; its results are not copied."""
        check(self._lib.mcsart_server_cancel(self.ptr, t.id), "mcsart_server_cancel")

    def drain(self):
        check(self._lib.mcsart_server_drain(self.ptr), "mcsart_server_drain")

    def info(self, t: Ticket) -> dict:
        i = CRequestInfo()
        check(self._lib.mcsart_server_info(self.ptr, t.id, ctypes.byref(i)), "mcsart_server_info")
        return {"state": STATES.get(i.state, i.state), "status": i.status, "slot": i.slot,
                "bytes": i.bytes, "queue_ms": i.queue_ms, "run_ms": i.run_ms,
                "total_ms": i.total_ms, "tuning": i.tuning.as_dict(),
                "admit_at_ms": i.admit_at_ms, "done_at_ms": i.done_at_ms}

    def release(self, t: Ticket):
        check(self._lib.mcsart_server_release(self.ptr, t.id), "mcsart_server_release")
        t._keep.clear()

    def memory(self):
        u, b = ctypes.c_size_t(), ctypes.c_size_t()
        check(self._lib.mcsart_server_memory(self.ptr, ctypes.byref(u), ctypes.byref(b)),
              "mcsart_server_memory")
        return u.value, b.value

    def close(self):
        if getattr(self, "ptr", None):
            check(self._lib.mcsart_server_destroy(self.ptr), "mcsart_server_destroy")
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
