"""Host-side mirror of the reference operator surface (``include/cu_*.h``).

Each ``host_*`` function has the reference's name and argument order and calls the matching
``mcsart_*`` entry point of libmcsart_hip with device pointers.  ``DeviceArray`` is a small owner
of device memory allocated through the library itself (no other GPU runtime is involved).

Example (the reference's ``host_projection(d_proj, d_img, angle, SO, SD, da, na, ai, db, nb, bi,
nx, ny, nz)``, ``include/cu_projection.h:7``)::

    d_img = DeviceArray.from_host(vol)            # (nz, ny, nx) float32
    d_proj = DeviceArray.empty((nb, na))
    host_projection(d_proj, d_img, angle, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi,
                    g.nx, g.ny, g.nz)
    p = d_proj.to_host()
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from ._lib import check, load


class DeviceArray:
    """float32 device buffer owned by this object (freed on close / garbage collection)."""

    def __init__(self, shape, ptr: int, owner: bool = True):
        self.shape = tuple(int(s) for s in shape)
        self.ptr = ptr
        self._owner = owner

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    @property
    def nbytes(self) -> int:
        return 4 * self.size

    @staticmethod
    def empty(shape) -> "DeviceArray":
        lib = load()
        p = ctypes.c_void_p()
        n = 4 * int(np.prod(shape))
        check(lib.mcsart_malloc(ctypes.byref(p), max(n, 4)), "mcsart_malloc")
        return DeviceArray(shape, p.value)

    @staticmethod
    def zeros(shape) -> "DeviceArray":
        a = DeviceArray.empty(shape)
        check(load().mcsart_memset(a.ptr, 0, a.nbytes), "mcsart_memset")
        return a

    @staticmethod
    def from_host(x: np.ndarray) -> "DeviceArray":
        x = np.ascontiguousarray(x, dtype=np.float32)
        a = DeviceArray.empty(x.shape)
        check(load().mcsart_memcpy_htod(a.ptr, x.ctypes.data, x.nbytes), "mcsart_memcpy_htod")
        return a

    def upload(self, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.size == self.size
        check(load().mcsart_memcpy_htod(self.ptr, x.ctypes.data, x.nbytes), "mcsart_memcpy_htod")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float32)
        check(load().mcsart_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes), "mcsart_memcpy_dtoh")
        return out

    def close(self) -> None:
        if self._owner and self.ptr:
            load().mcsart_free(self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _p(x) -> int:
    return x.ptr if isinstance(x, DeviceArray) else int(x)


def device_count() -> int:
    n = ctypes.c_int(0)
    check(load().mcsart_device_count(ctypes.byref(n)), "mcsart_device_count")
    return n.value


def host_projection(d_proj, d_img, angle, SO, SD, da, na, ai, db, nb, bi, nx, ny, nz):
    """``include/cu_projection.h:7``"""
    check(load().mcsart_projection(_p(d_proj), _p(d_img), angle, SO, SD, da, na, ai, db, nb, bi,
                                   nx, ny, nz), "mcsart_projection")


def host_backprojection(d_img, d_proj, angle, SO, SD, da, na, ai, db, nb, bi, nx, ny, nz):
    """``include/cu_backprojection.h:7``"""
    check(load().mcsart_backprojection(_p(d_img), _p(d_proj), angle, SO, SD, da, na, ai, db, nb,
                                       bi, nx, ny, nz), "mcsart_backprojection")


def host_deform(d_img1, d_img, nx, ny, nz, volume, flow, ax, ay, az, bx, by, bz):
    """``include/cu_deform.h:8``"""
    check(load().mcsart_deform(_p(d_img1), _p(d_img), nx, ny, nz, volume, flow, _p(ax), _p(ay),
                               _p(az), _p(bx), _p(by), _p(bz)), "mcsart_deform")


def host_deform2(d_img1, d_img, nx, ny, nz, volume, flow, ax, ay, az, bx, by, bz):
    """``include/cu_deform.h:9``"""
    check(load().mcsart_deform2(_p(d_img1), _p(d_img), nx, ny, nz, volume, flow, _p(ax), _p(ay),
                                _p(az), _p(bx), _p(by), _p(bz)), "mcsart_deform2")


def host_update_udvf(ax, ay, az, bx, by, bz, img, img0, volume_diff, flow_diff, nx, ny, nz, it=0):
    """``include/cu_update_udvf.h:8``"""
    check(load().mcsart_update_udvf(_p(ax), _p(ay), _p(az), _p(bx), _p(by), _p(bz), _p(img),
                                    _p(img0), volume_diff, flow_diff, nx, ny, nz, it),
          "mcsart_update_udvf")


def host_add(img1, img, nx, ny, nz, weight):
    """``include/cu_add.h:4``"""
    check(load().mcsart_add(_p(img1), _p(img), nx, ny, nz, weight), "mcsart_add")


def host_add2(img1, img, nx, ny, nz, img0, weight, ind):
    """``include/cu_add.h:8``"""
    check(load().mcsart_add2(_p(img1), _p(img), nx, ny, nz, _p(img0), weight, ind), "mcsart_add2")


def host_initial(img, nx, ny, nz, value):
    """``include/cu_initial.h:5``"""
    check(load().mcsart_initial(_p(img), nx, ny, nz, value), "mcsart_initial")


def host_initial2(img, nx, ny, nz, img0, volume, flow):
    """``include/cu_initial.h:8``"""
    check(load().mcsart_initial2(_p(img), nx, ny, nz, _p(img0), volume, flow), "mcsart_initial2")


def host_division(img1, img, nx, ny, nz):
    """``include/cu_division.h:5``"""
    check(load().mcsart_division(_p(img1), _p(img), nx, ny, nz), "mcsart_division")


def host_diff(img1, img, nx, ny, nz, ind):
    """Backward difference of ``src/mex_diff.cu:67-112``."""
    check(load().mcsart_diff(_p(img1), _p(img), nx, ny, nz, ind), "mcsart_diff")


def diff_forward(img1, img, nx, ny, nz, ind):
    """Volume form of ``diff_x/diff_y/diff_z`` (``include/cu_diff.h:4-6``)."""
    check(load().mcsart_diff_forward(_p(img1), _p(img), nx, ny, nz, ind), "mcsart_diff_forward")


def sart_residual(d_rp, d_norm, d_vol, motion_pairs, d_meas, angle, SO, SD, da, na, ai, db, nb,
                  bi, nx, ny, nz):
    """Fused residual projector (src/AUMISART.cu:194-209 / :246-263): d_rp gets (r, n) pairs.
    motion_pairs False: d_vol is x, n = P(1); True: d_vol holds (g, q) pairs, n = P(q)."""
    check(load().mcsart_sart_residual(_p(d_rp), _p(d_norm), _p(d_vol), int(bool(motion_pairs)),
                                      _p(d_meas), angle, SO, SD, da, na, ai, db, nb, bi, nx, ny,
                                      nz), "mcsart_sart_residual")


def sart_backprojection(d_u, d_rp, angle, SO, SD, da, na, ai, db, nb, bi, nx, ny, nz):
    """u = B(r)/B(n), NaN/Inf -> 0 (src/AUMISART.cu:212-218 / :265-270)."""
    check(load().mcsart_sart_backprojection(_p(d_u), _p(d_rp), angle, SO, SD, da, na, ai, db, nb,
                                            bi, nx, ny, nz), "mcsart_sart_backprojection")


def motion_prep(d_pairs, d_img, ax, ay, az, bx, by, bz, vd, fd, nx, ny, nz):
    """(g, q) pairs: g = x - img0, q = (D1+D2+D3)x (-vd-fd)  (src/AUMISART.cu:252-258, :267)."""
    check(load().mcsart_motion_prep(_p(d_pairs), _p(d_img), _p(ax), _p(ay), _p(az), _p(bx),
                                    _p(by), _p(bz), vd, fd, nx, ny, nz), "mcsart_motion_prep")


def motion_update(ax, ay, az, bx, by, bz, d_u, d_img, wv, wf, nx, ny, nz):
    """fields += u^T (wv|wf) D_c x  (src/AUMISART.cu:272-277)."""
    check(load().mcsart_motion_update(_p(ax), _p(ay), _p(az), _p(bx), _p(by), _p(bz), _p(d_u),
                                      _p(d_img), wv, wf, nx, ny, nz), "mcsart_motion_update")


# ---- convenience wrappers on numpy arrays --------------------------------------------------
def _sized(x: np.ndarray, n: int, what: str) -> np.ndarray:
    """x as float32, refused unless it holds exactly n elements (the kernels read n of them)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if x.size != n:
        raise ValueError(f"{what} holds {x.size} elements, {n} expected")
    return x


def project(img: np.ndarray, g, angle) -> np.ndarray:
    img = _sized(img, g.nvox, "volume")
    with DeviceArray.from_host(img) as d_img, DeviceArray.empty(g.det_shape) as d_p:
        host_projection(d_p, d_img, angle, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi,
                        g.nx, g.ny, g.nz)
        return d_p.to_host()


def backproject(proj: np.ndarray, g, angle) -> np.ndarray:
    proj = _sized(proj, g.ndet, "projection")
    with DeviceArray.from_host(proj) as d_p, DeviceArray.empty(g.vol_shape) as d_img:
        host_backprojection(d_img, d_p, angle, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi,
                            g.nx, g.ny, g.nz)
        return d_img.to_host()


def deform(img: np.ndarray, volume, flow, fields: Sequence[np.ndarray]) -> np.ndarray:
    nz, ny, nx = img.shape
    if len(fields) != 6:
        raise ValueError("six fields (ax, ay, az, bx, by, bz) expected")
    d = [DeviceArray.from_host(_sized(f, img.size, "field")) for f in fields]
    with DeviceArray.from_host(img) as d_img, DeviceArray.empty(img.shape) as d_out:
        host_deform(d_out, d_img, nx, ny, nz, volume, flow, *d)
        out = d_out.to_host()
    for a in d:
        a.close()
    return out


def host_multiple(img1, nx, ny, nz, img0, weight, ind):
    """``host_multiple`` (``src/cu_multiple.cu:2-9``): img1 *= weight * D_ind(img0)."""
    check(load().mcsart_multiple(_p(img1), nx, ny, nz, _p(img0), float(weight), int(ind)),
          "mcsart_multiple")


def invert_dvf(mx: np.ndarray, my: np.ndarray, mz: np.ndarray, niter: int):
    """Fixed-point DVF inverse on the device (mcsart_invert_dvf); returns (mx2, my2, mz2)."""
    nz, ny, nx = mx.shape
    ins = [DeviceArray.from_host(_sized(a, mx.size, "DVF component")) for a in (mx, my, mz)]
    outs = [DeviceArray.empty(mx.shape) for _ in range(3)]
    try:
        check(load().mcsart_invert_dvf(*[o.ptr for o in outs], *[i.ptr for i in ins], nx, ny, nz,
                                       int(niter)), "mcsart_invert_dvf")
        return [o.to_host() for o in outs]
    finally:
        for a in ins + outs:
            a.close()


def generate_proj(img: np.ndarray, fields: Sequence[np.ndarray], volumes, flows, angles, g,
                  niter_inverse: int = 0) -> np.ndarray:
    """Forward simulator on the device (mcsart_generate_proj); returns (V, nb, na)."""
    vol, flo, ang = (np.ascontiguousarray(x, dtype=np.float32).ravel()
                     for x in (volumes, flows, angles))
    V = int(ang.size)
    d_f = [DeviceArray.from_host(_sized(f, g.nvox, "field")) for f in fields]
    d_img = DeviceArray.from_host(_sized(img, g.nvox, "volume"))
    d_p = DeviceArray.empty((V, g.nb, g.na))
    try:
        check(load().mcsart_generate_proj(
            d_p.ptr, d_img.ptr, *[f.ptr for f in d_f], vol.ctypes.data, flo.ctypes.data,
            ang.ctypes.data, V, int(niter_inverse), g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb,
            g.bi, g.nx, g.ny, g.nz), "mcsart_generate_proj")
        return d_p.to_host()
    finally:
        for a in d_f + [d_img, d_p]:
            a.close()
