"""Culling changes no result (mcsart_job_set_culling).

With culling on, a view's warp writes only the voxels its projector can read, the motion step
prepares g only there and updates the accumulators only where u can be nonzero, and the warp-add
leaves x alone where W(u) is 0 for that reason.  Whole schedules must then give the same image,
per-view residual ratios and fields bit for bit as with culling off.  Mode 2 writes a NaN into
every cell culling skips: if any kernel read one, the norms and images would carry it, so equality
with mode 0 also shows that nothing reads a skipped cell.  The CPU side (every cell the projector
kernels read lies in the read set) is tests/test_cull.py.
"""
import numpy as np
import pytest

from mcsart_amd import configs
from mcsart_amd.configs import make_scan_inputs, scaled
from mcsart_amd.driver import Context, Job
from mcsart_amd.geometry import Geometry, uniform_angles
from mcsart_amd.simulate import measure
from mcsart_amd.synth import Voxel, breathing_trace, ellipsoid_phantom, motion_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _run(ctx, g, ang, vol, flo, op, proj, fields, mode, motion=True, nonneg=True):
    job = Job(ctx, g, ang, vol, flo, op, lam=1.0, motion=motion, nonneg=nonneg)
    job.set_culling(mode)
    job.set_projections(proj)
    job.set_image(None)
    job.set_fields(fields)
    job.run()
    out = job.image(), job.norms(), job.fields()
    job.close()
    return out


def _same(a, b):
    img_a, nrm_a, f_a = a
    img_b, nrm_b, f_b = b
    assert np.all(np.isfinite(img_a)) and np.all(np.isfinite(nrm_a))
    assert np.array_equal(img_a, img_b)
    assert np.array_equal(nrm_a, nrm_b)
    for x, y in zip(f_a, f_b):
        assert np.array_equal(x, y)


def _check_modes(ctx, g, ang, vol, flo, op, proj, fields, **kw):
    print(f"culling modes 0/1/2: {g.nx}x{g.ny}x{g.nz}, {g.na}x{g.nb}, {len(ang)} views, op {op}",
          flush=True)
    ref = _run(ctx, g, ang, vol, flo, op, proj, fields, 0, **kw)
    _same(_run(ctx, g, ang, vol, flo, op, proj, fields, 1, **kw), ref)
    _same(_run(ctx, g, ang, vol, flo, op, proj, fields, 2, **kw), ref)


def test_culling_changes_nothing_256x256x128(ctx):
    """The 256x256x128 scan's geometry (its detector's fan misses a third of the volume), 48
    views over the full circle, warped image sweeps and motion sweeps."""
    spec = scaled(configs.c3_clinical(), n_views=48, op_iter=[1, 0, 1, 0])
    data = make_scan_inputs(spec)
    proj = measure(spec, data, spec.seed + 7)
    _check_modes(ctx, spec.geometry, spec.angles, data.alpha, data.beta, spec.op_iter, proj,
                 data.fields)


def test_culling_changes_nothing_once_the_fields_overflow(ctx):
    """The whole 256x256x128 scan (656 views) over [1, 0, 1, 0, 1]: its motion sweeps take the
    fields to FP32 overflow (the sweep divides by B(P q), which has near-zero values), so g holds
    Inf and NaN in places.  The motion step then prepares g in the projector's edge shell too (read
    there with zero weights, where a stale Inf would make NaN), and the runs still agree bit for
    bit, NaNs included."""
    spec = scaled(configs.c3_clinical(), op_iter=[1, 0, 1, 0, 1])
    data = make_scan_inputs(spec)
    proj = measure(spec, data, spec.seed + 404)
    g = spec.geometry
    print("culling modes 0/1/2: 256x256x128, 656 views, op [1, 0, 1, 0, 1]", flush=True)
    runs = {mode: _run(ctx, g, spec.angles, data.alpha, data.beta, spec.op_iter, proj,
                       data.fields, mode) for mode in (0, 1, 2)}
    img0, nrm0, f0 = runs[0]
    assert not all(np.isfinite(f).all() for f in f0), "the fields did not overflow"
    for mode in (1, 2):
        img, nrm, f = runs[mode]
        assert np.array_equal(img, img0)
        assert np.array_equal(nrm, nrm0, equal_nan=True)
        for x, y in zip(f, f0):
            assert np.array_equal(x, y, equal_nan=True)


def test_culling_changes_nothing_offset_detector(ctx):
    """A truncated, offset detector finer than the voxels (the read set then misses most of the
    volume on one side), odd detector counts, 45-degree views among the angles."""
    g = Geometry.from_user(48, 48, 40, 161, 51, da=0.4, db=0.6, ai=23.5, bi=-7.25, SO=120.0,
                           SD=200.0)
    V = 32
    ang = uniform_angles(V)
    vox = Voxel(2.0, 2.0)
    ph = ellipsoid_phantom(g.nx, g.ny, g.nz, vox, 11)
    vol, flo = breathing_trace(V, 12, dt_s=0.1)
    fields = list(motion_fields(g.nx, g.ny, g.nz, vox, 13)) + \
        list(motion_fields(g.nx, g.ny, g.nz, vox, 14, scale=0.3))
    import oracle
    proj = np.stack([oracle.projection(oracle.deform(ph, vol[i], flo[i], fields), g, ang[i])
                     for i in range(V)]).astype(np.float32)
    _check_modes(ctx, g, ang, vol, flo, [1, 0, 1], proj, fields)
    _check_modes(ctx, g, ang, vol, flo, [0, 1], proj, fields, nonneg=False)


@pytest.mark.parametrize("shape", ["odd sizes", "coarse detector"])
def test_culling_changes_nothing_other_kernels(ctx, shape):
    """Odd volume sizes (planes half a voxel off the integer grid), and a detector coarser than
    the voxels, whose footprints outgrow the fixed windows (the wide projector kernel reads
    exactly each footprint's cells)."""
    if shape == "odd sizes":
        g = Geometry.from_user(37, 37, 21, 61, 29, da=1.1, db=0.9, ai=0.3, bi=0.7, SO=90.0,
                               SD=150.0)
    else:
        g = Geometry.from_user(40, 40, 32, 20, 16, da=4.0, db=4.0, SO=100.0, SD=160.0)
    V = 24
    ang = uniform_angles(V)
    vox = Voxel(2.0, 2.0)
    ph = ellipsoid_phantom(g.nx, g.ny, g.nz, vox, 31)
    vol, flo = breathing_trace(V, 32, dt_s=0.1)
    fields = list(motion_fields(g.nx, g.ny, g.nz, vox, 33)) + \
        list(motion_fields(g.nx, g.ny, g.nz, vox, 34, scale=0.3))
    import oracle
    proj = np.stack([oracle.projection(oracle.deform(ph, vol[i], flo[i], fields), g, ang[i])
                     for i in range(V)]).astype(np.float32)
    _check_modes(ctx, g, ang, vol, flo, [1, 0, 1], proj, fields)


def test_culling_changes_nothing_static_overload(ctx):
    """Overload 1 (no warps; its motion sweep steps views by 10, so each motion step prepares a
    view ten ahead)."""
    g = Geometry.from_user(40, 40, 24, 72, 40, da=0.8, db=0.8, ai=4.0, SO=100.0, SD=170.0)
    V = 36
    ang = uniform_angles(V)
    ph = ellipsoid_phantom(g.nx, g.ny, g.nz, Voxel(2.0, 2.0), 21)
    import oracle
    proj = np.stack([oracle.projection(ph, g, ang[i]) for i in range(V)]).astype(np.float32)
    vol, flo = breathing_trace(V, 22, dt_s=0.1)
    fields = list(motion_fields(g.nx, g.ny, g.nz, Voxel(2.0, 2.0), 23)) + \
        list(motion_fields(g.nx, g.ny, g.nz, Voxel(2.0, 2.0), 24, scale=0.3))
    _check_modes(ctx, g, ang, vol, flo, [1, 0, 1], proj, fields, motion=False)


def test_culling_shortens_the_warp_at_256x256x128(ctx):
    """The forward warp skips the voxels outside the read set (about a third here): its isolated
    time must drop (loose bound: timing differs between boxes)."""
    spec = scaled(configs.c3_clinical(), n_views=16, op_iter=[1])
    data = make_scan_inputs(spec)
    g = spec.geometry
    job = Job(ctx, g, spec.angles, data.alpha, data.beta, [1], motion=True, nonneg=True)
    job.set_image(data.phantom)
    job.set_fields(data.fields)
    t = {}
    for mode in (0, 1):
        job.set_culling(mode)
        t[mode] = job.kernel_time("warp", reps=3)
    job.close()
    print(f"warp at 256x256x128: {t[0]:.1f} us uncut, {t[1]:.1f} us culled", flush=True)
    assert t[1] < 0.9 * t[0], t
