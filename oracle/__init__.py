"""CPU oracle for the McSART hot path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``liboracle_mcsart.so`` (built from ``mcsart_oracle.cpp`` by the Makefile in
this directory).  Only the test-suite, the smoke check and the CPU-baseline leg of ``bench.py``
import this module, and only as the checker; the product library never calls it.

Array conventions match the reference's MATLAB column-major layout
(``src/cu_projection.cu:20,87``): a volume is a C-ordered numpy array of shape ``(nz, ny, nx)``
(so ``vol[iz, iy, ix] == img[ix + iy*nx + iz*nx*ny]``), one projection is ``(nb, na)`` and a
projection stack is ``(V, nb, na)``.

``dtype=np.float64`` selects the FP64 parity instance; ``np.float32`` selects the instance that
follows the reference's own single-precision arithmetic.  Geometry objects need the attributes
``nx ny nz na nb da db ai bi SO SD`` with ``ai``/``bi`` the reference's *internal* detector
offsets (``src/entry_AUMISART.cu:76-94``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_mcsart.so")
_lib = None


def build() -> str:
    """Compile the oracle library in place (g++ + OpenMP) and return its path."""
    subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I = ctypes.c_int
_LL = ctypes.c_longlong


def _declare(lib):
    for sfx, R in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
        geo = [R, R, R, R, _I, R, R, _I, R, _I, _I, _I]
        getattr(lib, f"oracle_projection_{sfx}").argtypes = [_P, _P] + geo
        getattr(lib, f"oracle_backprojection_{sfx}").argtypes = [_P, _P] + geo
        for nm in ("deform", "deform2"):
            getattr(lib, f"oracle_{nm}_{sfx}").argtypes = [_P, _P, _I, _I, _I, R, R] + [_P] * 6
        getattr(lib, f"oracle_update_udvf_{sfx}").argtypes = [_P] * 8 + [R, R, _I, _I, _I]
        getattr(lib, f"oracle_add_{sfx}").argtypes = [_P, _P, _LL, R]
        getattr(lib, f"oracle_add2_{sfx}").argtypes = [_P, _P, _I, _I, _I, _P, R, _I]
        getattr(lib, f"oracle_initial2_{sfx}").argtypes = [_P, _I, _I, _I, _P, R, R]
        getattr(lib, f"oracle_division_{sfx}").argtypes = [_P, _P, _LL]
        getattr(lib, f"oracle_diff_forward_{sfx}").argtypes = [_P, _P, _I, _I, _I, _I]
        getattr(lib, f"oracle_diff_backward_{sfx}").argtypes = [_P, _P, _I, _I, _I, _I]
        f = getattr(lib, f"oracle_norm2_{sfx}")
        f.argtypes = [_P, _LL]
        f.restype = R
        getattr(lib, f"oracle_multiple_{sfx}").argtypes = [_P, _I, _I, _I, _P, R, _I]
        getattr(lib, f"oracle_invert_dvf_{sfx}").argtypes = [_P] * 6 + [_I] * 4
        getattr(lib, f"oracle_generate_proj_{sfx}").argtypes = (
            [_P] * 11 + [_I, _I, R, R, R, _I, R, R, _I, R, _I, _I, _I])
        getattr(lib, f"oracle_motion_sweep_{sfx}").argtypes = (
            [_P] * 9 + [_I] * 7 + [_P] * 3 + [R] * 6)
        getattr(lib, f"oracle_aumisart_{sfx}").argtypes = (
            [_P] * 6 + [_I] * 7 + [_P] + [R] * 7 + [_P] * 3 + [_P] * 6 + [_I, _I])
    lib.oracle_num_threads.restype = _I
    lib.oracle_set_num_threads.argtypes = [_I]


def _sfx(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _arr(x, dtype):
    return np.ascontiguousarray(x, dtype=dtype)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def num_threads() -> int:
    return _load().oracle_num_threads()


def set_num_threads(n: int) -> None:
    _load().oracle_set_num_threads(int(n))


def _geo_args(g, angle):
    return (float(angle), float(g.SO), float(g.SD), float(g.da), int(g.na), float(g.ai),
            float(g.db), int(g.nb), float(g.bi), int(g.nx), int(g.ny), int(g.nz))


def projection(img, g, angle, dtype=np.float64):
    """One view of the forward projector (``src/cu_projection.cu:15-145``)."""
    img = _arr(img, dtype).reshape(g.nz, g.ny, g.nx)
    out = np.empty((g.nb, g.na), dtype=dtype)
    getattr(_load(), f"oracle_projection_{_sfx(dtype)}")(_p(out), _p(img), *_geo_args(g, angle))
    return out


def backprojection(proj, g, angle, dtype=np.float64):
    """One view of the backprojector (``src/cu_backprojection.cu:49-170``)."""
    proj = _arr(proj, dtype).reshape(g.nb, g.na)
    out = np.empty((g.nz, g.ny, g.nx), dtype=dtype)
    getattr(_load(), f"oracle_backprojection_{_sfx(dtype)}")(_p(out), _p(proj), *_geo_args(g, angle))
    return out


def _fields(fields, shape, dtype):
    fs = [_arr(f, dtype).reshape(shape) for f in fields]
    assert len(fs) == 6
    return fs


def deform(img, volume, flow, fields, dtype=np.float64):
    """5DCT warp (``src/cu_deform.cu:66-90``); fields = (ax, ay, az, bx, by, bz)."""
    img = _arr(img, dtype)
    nz, ny, nx = img.shape
    fs = _fields(fields, img.shape, dtype)
    out = np.empty_like(img)
    getattr(_load(), f"oracle_deform_{_sfx(dtype)}")(
        _p(out), _p(img), nx, ny, nz, float(volume), float(flow), *[_p(f) for f in fs])
    return out


def deform2(img, volume, flow, fields, dtype=np.float64):
    """First-order Taylor warp (``src/cu_deform.cu:92-113``)."""
    img = _arr(img, dtype)
    nz, ny, nx = img.shape
    fs = _fields(fields, img.shape, dtype)
    out = np.empty_like(img)
    getattr(_load(), f"oracle_deform2_{_sfx(dtype)}")(
        _p(out), _p(img), nx, ny, nz, float(volume), float(flow), *[_p(f) for f in fs])
    return out


def update_udvf(fields, img, img0, volume_diff, flow_diff, dtype=np.float64):
    """Kaczmarz step on the six fields (``src/cu_update_udvf.cu:11-62``); returns new fields."""
    img = _arr(img, dtype)
    img0 = _arr(img0, dtype).reshape(img.shape)
    nz, ny, nx = img.shape
    fs = [f.copy() for f in _fields(fields, img.shape, dtype)]
    getattr(_load(), f"oracle_update_udvf_{_sfx(dtype)}")(
        *[_p(f) for f in fs], _p(img), _p(img0), float(volume_diff), float(flow_diff), nx, ny, nz)
    return fs


def add(a, b, w, dtype=np.float64):
    """``a + w*b`` (``src/cu_add.cu:9-18``)."""
    a = _arr(a, dtype).copy()
    b = _arr(b, dtype).reshape(a.shape)
    getattr(_load(), f"oracle_add_{_sfx(dtype)}")(_p(a), _p(b), a.size, float(w))
    return a


def add2(a, b, c, w, ind, dtype=np.float64):
    """``a[ix,iy,iz] + b[iy,ix,iz]*w*D_ind(c)`` (``src/cu_add.cu:28-58``)."""
    a = _arr(a, dtype).copy()
    nz, ny, nx = a.shape
    b = _arr(b, dtype).reshape(a.shape)
    c = _arr(c, dtype).reshape(a.shape)
    getattr(_load(), f"oracle_add2_{_sfx(dtype)}")(_p(a), _p(b), nx, ny, nz, _p(c), float(w), int(ind))
    return a


def multiple(a, b, w, ind, dtype=np.float64):
    """``a * (w * D_ind(b))``, D_ind the forward difference (``src/cu_multiple.cu:11-43``)."""
    a = _arr(a, dtype).copy()
    nz, ny, nx = a.shape
    b = _arr(b, dtype).reshape(a.shape)
    getattr(_load(), f"oracle_multiple_{_sfx(dtype)}")(_p(a), nx, ny, nz, _p(b), float(w), int(ind))
    return a


def invert_dvf(mx, my, mz, niter, dtype=np.float64):
    """Fixed-point DVF inverse (``waste/kernel_invertDVF.cu:1-18``); returns (mx2, my2, mz2)."""
    mx = _arr(mx, dtype)
    nz, ny, nx = mx.shape
    my = _arr(my, dtype).reshape(mx.shape)
    mz = _arr(mz, dtype).reshape(mx.shape)
    out = [np.empty_like(mx) for _ in range(3)]
    getattr(_load(), f"oracle_invert_dvf_{_sfx(dtype)}")(
        *[_p(o) for o in out], _p(mx), _p(my), _p(mz), nx, ny, nz, int(niter))
    return out


def generate_proj(img, fields, volumes, flows, angles, g, niter, dtype=np.float64):
    """Forward simulator: per view, forwardDVF -> invertDVF (niter > 0) -> warp -> projection
    (the kernel order of ``include/generateProj.h:21-24``); returns (V, nb, na)."""
    img = _arr(img, dtype).reshape(g.nz, g.ny, g.nx)
    fs = _fields(fields, img.shape, dtype)
    vol, flo, ang = (_arr(x, dtype).ravel() for x in (volumes, flows, angles))
    V = ang.size
    out = np.empty((V, g.nb, g.na), dtype=dtype)
    getattr(_load(), f"oracle_generate_proj_{_sfx(dtype)}")(
        _p(out), _p(img), *[_p(f) for f in fs], _p(vol), _p(flo), _p(ang), V, int(niter),
        float(g.SO), float(g.SD), float(g.da), int(g.na), float(g.ai), float(g.db), int(g.nb),
        float(g.bi), int(g.nx), int(g.ny), int(g.nz))
    return out


def initial2(img0, volume, flow, dtype=np.float64):
    """``(D1+D2+D3)(img0)*(volume+flow)`` (``src/cu_initial.cu:27-49``)."""
    img0 = _arr(img0, dtype)
    nz, ny, nx = img0.shape
    out = np.empty_like(img0)
    getattr(_load(), f"oracle_initial2_{_sfx(dtype)}")(_p(out), nx, ny, nz, _p(img0), float(volume), float(flow))
    return out


def division(a, b, dtype=np.float64):
    """``a/b`` with NaN/Inf mapped to 0 (``src/cu_division.cu:11-25``)."""
    a = _arr(a, dtype).copy()
    b = _arr(b, dtype).reshape(a.shape)
    getattr(_load(), f"oracle_division_{_sfx(dtype)}")(_p(a), _p(b), a.size)
    return a


def diff_forward(img, ind, dtype=np.float64):
    """Forward difference, zero at the last index (``src/cu_diff.cu:2-33``)."""
    img = _arr(img, dtype)
    nz, ny, nx = img.shape
    out = np.empty_like(img)
    getattr(_load(), f"oracle_diff_forward_{_sfx(dtype)}")(_p(out), _p(img), nx, ny, nz, int(ind))
    return out


def diff_backward(img, ind, dtype=np.float64):
    """Backward difference, zero at both ends (``src/mex_diff.cu:75-112``)."""
    img = _arr(img, dtype)
    nz, ny, nx = img.shape
    out = np.empty_like(img)
    getattr(_load(), f"oracle_diff_backward_{_sfx(dtype)}")(_p(out), _p(img), nx, ny, nz, int(ind))
    return out


def norm2(x, dtype=np.float64):
    x = _arr(x, dtype)
    return float(getattr(_load(), f"oracle_norm2_{_sfx(dtype)}")(_p(x), x.size))


def aumisart(h_img, h_proj, g, angles, volumes, flows, op_iter, lam=1.0, fields=None,
             nonneg=False, dtype=np.float64):
    """Both host_AUMISART overloads (``src/AUMISART.cu:3-141`` when ``fields`` is None, else
    ``:143-313``).  Returns ``(img, norms, alpha_x, residual_proj)`` shaped like the MEX outputs
    of ``src/entry_AUMISART.cu:172-190``."""
    V = len(angles)
    op = np.ascontiguousarray(op_iter, dtype=np.int32)
    n_iter = op.size
    h_img = _arr(h_img, dtype).reshape(g.nz, g.ny, g.nx)
    h_proj = _arr(h_proj, dtype).reshape(V, g.nb, g.na)
    ang = _arr(angles, dtype)
    vol = _arr(volumes, dtype)
    flo = _arr(flows, dtype)
    motion = fields is not None
    if motion:
        fs = _fields(fields, h_img.shape, dtype)
    else:
        fs = [np.zeros(1, dtype=dtype)] * 6
    out_img = np.empty_like(h_img)
    out_proj = np.empty_like(h_proj)
    out_norm = np.empty(n_iter * V, dtype=dtype)
    out_ax = np.empty_like(h_img)
    getattr(_load(), f"oracle_aumisart_{_sfx(dtype)}")(
        _p(out_img), _p(out_proj), _p(out_norm), _p(out_ax), _p(h_img), _p(h_proj),
        g.nx, g.ny, g.nz, g.na, g.nb, V, n_iter, op.ctypes.data_as(ctypes.c_void_p),
        float(g.da), float(g.db), float(g.ai), float(g.bi), float(g.SO), float(g.SD), float(lam),
        _p(vol), _p(flo), _p(ang), *[_p(f) for f in fs], int(motion), int(bool(nonneg)))
    return out_img, out_norm, out_ax, out_proj


def motion_sweep(img, proj, g, angles, volumes, flows, fields, stride=1, dtype=np.float64):
    """One motion (udvf) sweep (``src/AUMISART.cu:242-288``; ``stride=10`` is overload 1's form,
    ``:72-118``) started from ``fields`` = (ax, ay, az, bx, by, bz) and image ``img``.  Returns
    ``(fields, norms)``: the six updated fields and the n_views residual ratios (entry 0 is 0)."""
    V = len(angles)
    img = _arr(img, dtype).reshape(g.nz, g.ny, g.nx)
    proj = _arr(proj, dtype).reshape(V, g.nb, g.na)
    fs = [f.copy() for f in _fields(fields, img.shape, dtype)]
    ang, vol, flo = (_arr(x, dtype).ravel() for x in (angles, volumes, flows))
    out_norm = np.empty(V, dtype=dtype)
    getattr(_load(), f"oracle_motion_sweep_{_sfx(dtype)}")(
        *[_p(f) for f in fs], _p(out_norm), _p(img), _p(proj), g.nx, g.ny, g.nz, g.na, g.nb, V,
        int(stride), _p(ang), _p(vol), _p(flo), float(g.da), float(g.db), float(g.ai),
        float(g.bi), float(g.SO), float(g.SD))
    return fs, out_norm
