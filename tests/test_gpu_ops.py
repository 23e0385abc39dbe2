"""GPU parity of each operator against the FP64 oracle (rel-L2 <= 1e-5 unless stated).

All calls go through the C ABI of libmcsart_hip (device pointers, reference argument order).
"""
import numpy as np
import pytest

import oracle
from mcsart_amd import configs
from mcsart_amd.geometry import Geometry
from mcsart_amd.ops import (DeviceArray, backproject, deform, diff_forward, host_add, host_add2,
                            host_deform, host_deform2, host_diff, host_division, host_initial,
                            host_initial2, host_update_udvf, project)
from mcsart_amd.synth import ellipsoid_phantom
from _util import random_fields, rel_l2, smooth_random

pytestmark = pytest.mark.gpu
TOL = 1e-5   # per-operator tolerance against the FP64 oracle


def _phantom(spec):
    g = spec.geometry
    return ellipsoid_phantom(g.nx, g.ny, g.nz, spec.voxel, spec.seed)


@pytest.mark.parametrize("spec_fn,views", [
    (configs.c1_static64, "all"),
    (configs.c2_mcsart128, [0, 1, 24, 25, 26, 50, 75, 99, 124, 175]),
    (configs.c3_clinical, [0, 82, 200]),
])
def test_projection_parity(gpu_lib, spec_fn, views):
    spec = spec_fn()
    g = spec.geometry
    ph = _phantom(spec)
    vs = range(spec.n_views) if views == "all" else views
    worst = 0.0
    for v in vs:
        ang = spec.angles[v]
        got = project(ph, g, ang)
        ref = oracle.projection(ph, g, ang)
        e = rel_l2(got, ref)
        worst = max(worst, e)
        assert e <= TOL, f"{spec.name} view {v}: rel-L2 {e:.3e}"
    print(f"{spec.name}: worst projection rel-L2 {worst:.3e}")


def test_projection_ones_closed_form(gpu_lib):
    """A ones-volume gives the number of planes the ray's normalised footprint covers
    (src/cu_projection.cu:83,86); central rays of the 64^3 geometry stay inside every plane."""
    spec = configs.c1_static64()
    g = spec.geometry
    ones = np.ones(g.vol_shape, np.float32)
    for ang in (0.0, 0.3, np.pi / 2):
        p = project(ones, g, np.float32(ang))
        c = p[g.nb // 2 - 2:g.nb // 2 + 2, g.na // 2 - 2:g.na // 2 + 2]
        assert np.allclose(c, g.nx, rtol=1e-5), c


@pytest.mark.parametrize("spec_fn,views", [
    (configs.c1_static64, [0, 7, 15, 22, 30, 45]),
    (configs.c2_mcsart128, [0, 25, 60, 133]),
    (configs.c3_clinical, [0, 82]),
])
def test_backprojection_parity(gpu_lib, spec_fn, views):
    spec = spec_fn()
    g = spec.geometry
    for v in views:
        ang = spec.angles[v]
        proj = smooth_random(g.det_shape, spec.seed + v)
        got = backproject(proj, g, ang)
        ref = oracle.backprojection(proj, g, ang)
        e = rel_l2(got, ref)
        assert e <= TOL, f"{spec.name} view {v}: rel-L2 {e:.3e}"


# Irregular geometries (voxel units): a detector much finer than the voxel shadow (footprints
# wider than the fixed windows: the projector's and backprojector's generic loops), a truncated
# detector offset from the axis (rays and footprints clipped at the detector edges), and sizes
# that are multiples of nothing.
IRREGULAR = {
    "fine_detector": Geometry.from_user(24, 24, 16, 150, 70, da=0.2, db=0.25, SO=60.0, SD=100.0),
    "truncated_offset": Geometry.from_user(40, 40, 24, 30, 14, da=1.5, db=1.5, ai=9.0, bi=-4.5,
                                           SO=70.0, SD=120.0),
    "odd_sizes": Geometry.from_user(37, 37, 19, 53, 31, da=1.3, db=1.1, ai=0.35, bi=-0.2,
                                    SO=55.0, SD=97.0),
    # nx != ny with a fan narrower than the volume (78 % of its width at the isocentre): views
    # whose rays run along x band the backprojector's tiles by column over unequal tile counts
    "non_square": Geometry.from_user(48, 32, 20, 40, 24, da=1.0, db=1.0, SO=100.0, SD=160.0),
}


@pytest.mark.parametrize("name", sorted(IRREGULAR))
def test_irregular_geometry_parity(gpu_lib, name):
    g = IRREGULAR[name]
    vol = smooth_random(g.vol_shape, 5)
    for k, ang in enumerate((0.0, 0.4, np.pi / 4, 1.3, np.pi / 2, 2.6, 4.0)):
        a = np.float32(ang)
        e = rel_l2(project(vol, g, a), oracle.projection(vol, g, a))
        assert e <= TOL, f"{name} projection angle {ang}: rel-L2 {e:.3e}"
        proj = smooth_random(g.det_shape, 40 + k)
        e = rel_l2(backproject(proj, g, a), oracle.backprojection(proj, g, a))
        assert e <= TOL, f"{name} backprojection angle {ang}: rel-L2 {e:.3e}"


def test_backprojection_constant(gpu_lib):
    """A constant projection c backprojects to c times the clipped footprint area, which for a
    voxel near the isocentre is the full (unnormalised) footprint; it must equal the oracle."""
    spec = configs.c1_static64()
    g = spec.geometry
    proj = np.full(g.det_shape, 2.0, np.float32)
    got = backproject(proj, g, np.float32(0.0))
    ref = oracle.backprojection(proj, g, np.float32(0.0))
    assert rel_l2(got, ref) <= TOL
    assert np.all(got >= 0)


# (nz, ny, nx): partial tiles, full tiles, odd sizes with a one-voxel edge tile, and a 1 x 1
# column (the warps' general 32-bit index form: the fast form needs nx >= 2)
@pytest.mark.parametrize("shape", [(16, 24, 24), (32, 64, 64), (19, 33, 33), (7, 1, 1)])
def test_deform_parity(gpu_lib, shape):
    img = smooth_random(shape, 11)
    fields = random_fields(shape, 23)
    for vol, flow in ((0.7, -0.3), (-1.3, 0.45), (0.0, 0.0)):
        got = deform(img, vol, flow, fields)
        ref = oracle.deform(img, vol, flow, fields)
        assert rel_l2(got, ref) <= TOL, (vol, flow)
    # zero DVF is an exact identity
    got = deform(img, 0.0, 0.0, fields)
    assert np.array_equal(got, img)


def test_deform_in_place(gpu_lib):
    shape = (12, 20, 20)
    nz, ny, nx = shape
    img = smooth_random(shape, 5)
    fields = random_fields(shape, 7)
    d = [DeviceArray.from_host(f) for f in fields]
    x = DeviceArray.from_host(img)
    host_deform(x, x, nx, ny, nz, 0.9, 0.2, *d)
    ref = oracle.deform(img, 0.9, 0.2, fields)
    assert rel_l2(x.to_host(), ref) <= TOL


def test_deform2_parity(gpu_lib):
    shape = (10, 16, 16)
    nz, ny, nx = shape
    img = smooth_random(shape, 3)
    fields = random_fields(shape, 4)
    d = [DeviceArray.from_host(f) for f in fields]
    x = DeviceArray.from_host(img)
    out = DeviceArray.empty(shape)
    host_deform2(out, x, nx, ny, nz, 0.5, -0.25, *d)
    ref = oracle.deform2(img, 0.5, -0.25, fields)
    assert rel_l2(out.to_host(), ref) <= TOL


def test_elementwise_parity(gpu_lib):
    shape = (9, 14, 14)
    nz, ny, nx = shape
    a = smooth_random(shape, 1)
    b = smooth_random(shape, 2)
    c = smooth_random(shape, 3)
    # add
    da, db = DeviceArray.from_host(a), DeviceArray.from_host(b)
    host_add(da, db, nx, ny, nz, -0.75)
    assert rel_l2(da.to_host(), oracle.add(a, b, -0.75)) <= 1e-6
    # add2, all three directions (transposed read of b)
    for ind in (1, 2, 3):
        da, db, dc = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(c)
        host_add2(da, db, nx, ny, nz, dc, 1.7, ind)
        assert rel_l2(da.to_host(), oracle.add2(a, b, c, 1.7, ind)) <= 1e-6, ind
    # initial / initial2
    d = DeviceArray.empty(shape)
    host_initial(d, nx, ny, nz, 3.25)
    assert np.all(d.to_host() == np.float32(3.25))
    dc = DeviceArray.from_host(c)
    host_initial2(d, nx, ny, nz, dc, -0.4, 0.15)
    assert rel_l2(d.to_host(), oracle.initial2(c, -0.4, 0.15)) <= 1e-6
    # division with NaN/Inf mapping
    num = a.copy()
    den = b.copy()
    den[0, 0, :5] = 0.0
    num[0, 0, :2] = 0.0
    dn, dd = DeviceArray.from_host(num), DeviceArray.from_host(den)
    host_division(dn, dd, nx, ny, nz)
    got = dn.to_host()
    ref = oracle.division(num, den)
    assert np.all(got[0, 0, :5] == 0.0)
    assert rel_l2(got, ref) <= 1e-6
    # diffs
    for ind in (1, 2, 3):
        dx = DeviceArray.from_host(c)
        out = DeviceArray.empty(shape)
        host_diff(out, dx, nx, ny, nz, ind)
        assert np.allclose(out.to_host(), oracle.diff_backward(c, ind), atol=1e-6), ind
        diff_forward(out, dx, nx, ny, nz, ind)
        assert np.allclose(out.to_host(), oracle.diff_forward(c, ind), atol=1e-6), ind


def test_update_udvf_parity(gpu_lib):
    shape = (8, 12, 12)
    nz, ny, nx = shape
    img = smooth_random(shape, 8)
    img0 = smooth_random(shape, 9)
    img[2, 3, :] = img[2, 3, 0]   # flat rows: zero gradient (w == 0 early return) at some voxels
    fields = random_fields(shape, 10)
    d = [DeviceArray.from_host(f) for f in fields]
    host_update_udvf(*d, DeviceArray.from_host(img), DeviceArray.from_host(img0), 0.3, -0.8,
                     nx, ny, nz, 0)
    ref = oracle.update_udvf(fields, img, img0, 0.3, -0.8)
    for k in range(6):
        assert rel_l2(d[k].to_host(), ref[k]) <= 1e-5, k


def test_invalid_arguments(gpu_lib):
    from mcsart_amd._lib import McsartError
    x = DeviceArray.zeros((4, 6, 8))
    y = DeviceArray.zeros((4, 6, 8))
    with pytest.raises(McsartError):
        host_add2(x, y, 8, 6, 4, y, 1.0, 1)          # nx != ny
    with pytest.raises(McsartError):
        host_diff(x, y, 8, 6, 4, 7)                  # bad ind
    with pytest.raises(McsartError):
        host_add(x, y, 0, 6, 4, 1.0)                 # empty volume


def test_async_operators_on_two_contexts(gpu_lib):
    """The _async operators enqueue on a context's stream and give exactly the synchronous
    operators' results; two contexts run the same chain side by side (scratch is per context:
    the in-place warp and the transposed copy of the projector each use it)."""
    from mcsart_amd.driver import Context
    spec = configs.c1_static64()
    g = spec.geometry
    vol = smooth_random(g.vol_shape, 3)
    fields = random_fields(g.vol_shape, 9)
    ang = np.float32(0.3)
    geo = (ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, g.nx, g.ny, g.nz)
    dF = [DeviceArray.from_host(f) for f in fields]
    # synchronous chain: warp in place, project, backproject
    dx = DeviceArray.from_host(vol)
    host_deform(dx, dx, g.nx, g.ny, g.nz, 0.7, 0.2, *dF)
    dp = DeviceArray.empty(g.det_shape)
    gpu_lib.mcsart_projection(dp.ptr, dx.ptr, *geo)
    db = DeviceArray.empty(g.vol_shape)
    gpu_lib.mcsart_backprojection(db.ptr, dp.ptr, *geo)
    want = (dx.to_host(), dp.to_host(), db.to_host())
    ctxs = [Context(0), Context(0)]
    bufs = []
    for c in ctxs:   # enqueue both chains before waiting on either
        x = DeviceArray.from_host(vol)
        p = DeviceArray.empty(g.det_shape)
        b = DeviceArray.empty(g.vol_shape)
        assert gpu_lib.mcsart_deform_async(c.ptr, x.ptr, x.ptr, g.nx, g.ny, g.nz, 0.7, 0.2,
                                           *[f.ptr for f in dF]) == 0
        assert gpu_lib.mcsart_projection_async(c.ptr, p.ptr, x.ptr, *geo) == 0
        assert gpu_lib.mcsart_backprojection_async(c.ptr, b.ptr, p.ptr, *geo) == 0
        bufs.append((x, p, b))
    for c, (x, p, b) in zip(ctxs, bufs):
        c.synchronize()
        for got, ref in zip((x.to_host(), p.to_host(), b.to_host()), want):
            assert np.array_equal(got, ref)
        c.close()
    assert gpu_lib.mcsart_projection_async(None, dp.ptr, dx.ptr, *geo) != 0


def test_multiple_parity(gpu_lib):
    from mcsart_amd.ops import host_multiple
    shape = (12, 20, 20)
    a, b = smooth_random(shape, 31), smooth_random(shape, 32)
    nz, ny, nx = shape
    for ind in (1, 2, 3):
        with DeviceArray.from_host(a) as da, DeviceArray.from_host(b) as db:
            host_multiple(da, nx, ny, nz, db, 0.75, ind)
            got = da.to_host()
        assert rel_l2(got, oracle.multiple(a, b, 0.75, ind)) <= TOL
    with DeviceArray.from_host(a) as da, DeviceArray.from_host(b) as db:
        assert gpu_lib.mcsart_multiple(da.ptr, nx, ny, nz, db.ptr, 1.0, 4) != 0   # ind 1..3


def test_invert_dvf_parity(gpu_lib):
    from mcsart_amd.ops import invert_dvf
    shape = (10, 24, 24)
    m = [2.0 * smooth_random(shape, 40 + k) for k in range(3)]
    for niter in (1, 5):
        got = invert_dvf(*m, niter)
        ref = oracle.invert_dvf(*m, niter)
        for g_, r_ in zip(got, ref):
            assert rel_l2(g_, r_) <= TOL, niter
    c = np.full(shape, 0.75, np.float32)
    assert np.array_equal(invert_dvf(c, c, c, 3)[0], -c)


def test_generate_proj_parity(gpu_lib):
    from mcsart_amd.ops import generate_proj
    g = Geometry.from_user(32, 32, 24, 48, 40, da=1.1, db=1.1, SO=120.0, SD=200.0)
    img = smooth_random(g.vol_shape, 50)
    fields = random_fields(g.vol_shape, 51)
    vol = np.array([0.4, -0.8, 1.2, 0.0], np.float32)
    flo = np.array([0.1, -0.2, 0.3, 0.0], np.float32)
    ang = np.array([0.0, 0.7, 2.0, 4.4], np.float32)
    for niter in (0, 3):
        got = generate_proj(img, fields, vol, flo, ang, g, niter)
        ref = oracle.generate_proj(img, fields, vol, flo, ang, g, niter)
        for i in range(len(ang)):
            assert rel_l2(got[i], ref[i]) <= TOL, (niter, i)
