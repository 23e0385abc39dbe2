"""Launch shapes of the footprint kernels: every shape gives the same reconstruction.

The projector's split along each ray changes only the order of its partial sums, the
backprojector's z-slab only where its axial bound lines are anchored, and LDS staging only where
its detector values are read from, so all shapes agree to float rounding (staged or not: bit for
bit), and all of them match the FP64 oracle to the reconstruction bar.
"""
import numpy as np
import pytest

import oracle
from mcsart_amd import configs
from mcsart_amd._lib import McsartError
from mcsart_amd.driver import Context, Job
from mcsart_amd.synth import ellipsoid_phantom
from _util import rel_l2

pytestmark = pytest.mark.gpu
# (proj_split, back_zslab, back_stage, proj_rows)
# back_stage: 0 automatic (staged), 1 staged in LDS, 2 detector read from global memory
SHAPES = [(1, 1, 2, 1), (1, 1, 1, 1), (8, 16, 2, 2), (8, 16, 1, 2), (2, 4, 1, 4), (4, 2, 2, 4),
          (16, 64, 1, 4), (16, 8, 2, 1), (0, 0, 0, 1), (0, 0, 0, 2)]
AUTO = {"proj_split": 0, "back_zslab": 0, "back_stage": 0, "proj_rows": 0, "back_zslab_motion": 0}


@pytest.fixture(scope="module")
def scan():
    spec = configs.c1_static64()
    g = spec.geometry
    ph = ellipsoid_phantom(g.nx, g.ny, g.nz, spec.voxel, spec.seed)
    ang = spec.angles
    proj = np.stack([oracle.projection(ph, g, a) for a in ang]).astype(np.float32)
    ctx = Context(0)
    yield g, ang, proj, ctx
    ctx.close()


def _job(scan, op_iter=(1, 1)):
    g, ang, proj, ctx = scan
    zeros = np.zeros(len(ang), np.float32)
    job = Job(ctx, g, ang, zeros, zeros, np.array(op_iter, np.int32), lam=1.0, motion=False,
              nonneg=True)
    job.set_projections(proj)
    job.set_image(None)
    job.set_fields(None)
    return job


def test_shapes_agree(scan):
    g, ang, proj, ctx = scan
    job = _job(scan)
    assert job.tuning() == AUTO
    job.run()
    base = job.image()
    imgs = {}
    for sh in SHAPES:
        job.reset()
        job.set_tuning(*sh)
        assert tuple(job.tuning().values())[:4] == sh
        for _ in range(2):   # second run replays the graphs captured with this shape
            job.reset()
            job.run()
            imgs[sh] = job.image()
            e = rel_l2(imgs[sh], base)
            assert e <= 1e-5, f"shape {sh}: rel-L2 {e:.3e} against the automatic shape"
    job.close()
    for sp, zs, rw in {(s[0], s[1], s[3]) for s in SHAPES}:
        if (sp, zs, 1, rw) in imgs and (sp, zs, 2, rw) in imgs:
            assert np.array_equal(imgs[(sp, zs, 1, rw)], imgs[(sp, zs, 2, rw)]), (sp, zs, rw)
    zeros = np.zeros(len(ang), np.float32)
    ref, _, _, _ = oracle.aumisart(np.zeros(g.vol_shape), proj, g, ang, zeros, zeros,
                                   np.array([1, 1], np.int32), lam=1.0, nonneg=True)
    assert rel_l2(base, ref) <= 1e-3


def test_autotune_keeps_state(scan):
    g, ang, proj, ctx = scan
    job = _job(scan, op_iter=(1,))
    job.run()   # non-trivial image and norms
    img0, nrm0 = job.image(), job.norms()
    chosen = job.autotune()
    assert chosen["proj_split"] in (1, 2, 4, 8, 16) and chosen["proj_rows"] in (1, 2, 4)
    assert 1 <= chosen["back_zslab"] <= 16 and chosen["back_stage"] in (1, 2)
    assert 1 <= chosen["back_zslab_motion"] <= 16
    assert job.tuning() == chosen
    assert np.array_equal(job.image(), img0)
    assert np.array_equal(job.norms(), nrm0)
    job.reset()
    job.run()
    assert rel_l2(job.image(), img0) <= 1e-5
    job.close()


def test_bad_shapes_rejected(scan):
    job = _job(scan, op_iter=(1,))
    for bad in [(3, 0, 0, 0), (32, 0, 0, 0), (0, -1, 0, 0), (0, 65, 0, 0), (0, 0, 3, 0),
                (0, 0, 0, 3), (0, 0, 0, 8), (0, 0, 0, 0, -1), (0, 0, 0, 0, 65)]:
        with pytest.raises(McsartError):
            job.set_tuning(*bad)
    assert job.tuning() == AUTO
    job.close()


def test_motion_backprojector_slabs_agree():
    """The motion sweep's backprojector (with the fused update and next prep) under every z-slab:
    after one update the fields agree with the automatic shape to rounding (the slab only moves
    where the axial bound lines are anchored).  The update divides by B(P q) with
    q = (D1 + D2 + D3) x (-dv - df), so x is a positive ramp (q constant inside the volume, the
    ratio well conditioned); voxels whose footprint only grazes the detector, where the ratio
    switches between 0/0 -> 0 and r/n, are allowed to differ (at most 1 % of them)."""
    spec = configs.c1_static64()
    g = spec.geometry
    V = 2
    ang = np.array([0.3, 0.9], np.float32)
    ctx = Context(0)
    rng = np.random.default_rng(3)
    proj = rng.random((V,) + g.det_shape, dtype=np.float32)
    iz, iy, ix = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in g.vol_shape],
                             indexing="ij")
    x0 = (0.01 * (ix + iy + iz) + 0.1).astype(np.float32)
    fields = [0.2 * rng.standard_normal(g.vol_shape).astype(np.float32) for _ in range(6)]
    out = {}
    for zs in (0, 1, 2, 8, 16, 64):
        job = Job(ctx, g, ang, np.array([0.0, 0.4], np.float32),
                  np.array([0.0, 0.1], np.float32), np.array([0], np.int32), lam=1.0,
                  motion=True, nonneg=True)
        job.set_projections(proj)
        job.set_image(x0)
        job.set_fields(fields)
        job.set_tuning(back_zslab_motion=zs)
        job.run()
        out[zs] = job.fields(), job.norms()
        job.close()
    ctx.close()
    for zs, (f, nrm) in out.items():
        assert np.array_equal(nrm, out[0][1]), zs   # the projector's ratios do not depend on it
        for a, b in zip(f, out[0][0]):
            off = np.abs(a - b) > 1e-5 * np.abs(b).max()
            assert off.mean() <= 0.01, (zs, off.mean())
            assert rel_l2(a[~off], b[~off]) <= 1e-5, zs


def test_kernel_time_all_kinds_leave_job_untouched():
    """Isolated timing of every sweep kernel writes only scratch: a motion scan run after the
    timings gives bit for bit the image, norms and fields of the same scan run before them."""
    spec = configs.c1_static64()
    g = spec.geometry
    V = 12
    ang = np.linspace(0, 2 * np.pi, V, endpoint=False).astype(np.float32)
    vol = np.linspace(-1.0, 1.0, V).astype(np.float32)
    ctx = Context(0)
    job = Job(ctx, g, ang, vol, 0.2 * vol, np.array([1, 0], np.int32), lam=1.0, motion=True,
              nonneg=True)
    rng = np.random.default_rng(7)
    job.set_projections(rng.random((V,) + g.det_shape, dtype=np.float32))
    job.set_image(None)
    job.set_fields([0.3 * rng.standard_normal(g.vol_shape).astype(np.float32)
                    for _ in range(6)])
    job.run()
    before = job.image(), job.norms(), job.fields()
    layouts = job.view_layouts()
    assert layouts.shape == (V,) and ((layouts >= 1) & (layouts <= 3)).all()
    for kind in Job.KERNEL_KINDS:
        us = job.kernel_time(kind, reps=2)
        assert 0.0 < us < 1e5, (kind, us)
    assert np.array_equal(job.image(), before[0])
    assert np.array_equal(job.norms(), before[1])
    for a, b in zip(job.fields(), before[2]):
        assert np.array_equal(a, b)
    job.reset()
    job.run()
    assert np.array_equal(job.image(), before[0])
    assert np.array_equal(job.norms(), before[1])
    with pytest.raises(McsartError):
        job.kernel_time("projector", reps=0)
    job.close()
    ctx.close()


@pytest.mark.parametrize("which", ["clinical256_48views", "mcsart128_overflow"])
def test_staged_equals_global_reads(which):
    """The backprojector's LDS path reads each row at its wavefront's widest footprint and carries
    a voxel's last row sum into the next voxel; the global-memory path reads every row at the
    view's widest footprint.  The slots the LDS path leaves out carry weight 0 on a cell already
    counted, so whole schedules agree bit for bit: at the 256x256x128 geometry (rows of 4, 5 and
    6 cells) and on a 128^3 schedule whose motion sweeps overflow the fields (NaN and Inf reach
    the residuals the backprojector reads)."""
    from mcsart_amd.configs import make_scan_inputs, scaled
    from mcsart_amd.simulate import measure
    if which == "clinical256_48views":
        spec, salt = scaled(configs.c3_clinical(), n_views=48, op_iter=[1, 0, 1, 0]), 7
    else:   # with these projections 20 % of the per-view norms are non-finite
        spec, salt = scaled(configs.c2_mcsart128(), op_iter=[1, 0, 1, 0]), 11
    data = make_scan_inputs(spec)
    proj = measure(spec, data, spec.seed + salt)
    print(f"staged vs global backprojector reads: {spec.name}, {spec.n_views} views",
          flush=True)
    ctx = Context(0)
    out = []
    for stage in (1, 2):
        job = Job(ctx, spec.geometry, spec.angles, data.alpha, data.beta, spec.op_iter, lam=1.0,
                  motion=True, nonneg=True)
        job.set_tuning(back_zslab=8, back_stage=stage, back_zslab_motion=8)
        job.set_projections(proj)
        job.set_image(None)
        job.set_fields(data.fields)
        job.run()
        out.append((job.image(), job.norms(), job.fields()))
        job.close()
    ctx.close()
    (img1, nrm1, f1), (img2, nrm2, f2) = out
    assert np.all(np.isfinite(img1))
    assert np.array_equal(img1, img2)
    assert np.array_equal(nrm1, nrm2, equal_nan=True)
    for a, b in zip(f1, f2):
        assert np.array_equal(a, b, equal_nan=True)
    if which == "mcsart128_overflow":   # the case this test is for
        assert not np.all(np.isfinite(nrm1))


@pytest.mark.parametrize("which", ["clinical256_48views", "mcsart128"])
def test_projector_rows_agree_bit_for_bit(which):
    """A ray's projection is the same arithmetic whether its lane walks one, two or four rays and
    whether those rays share their z-rows (rows_share: at 256x256x128 two and four rays share,
    at 128^3 only two), so with the split fixed whole schedules agree bit for bit."""
    from mcsart_amd.configs import make_scan_inputs, scaled
    from mcsart_amd.simulate import measure
    if which == "clinical256_48views":
        spec = scaled(configs.c3_clinical(), n_views=48, op_iter=[1, 0, 1])
    else:
        spec = scaled(configs.c2_mcsart128(), n_views=60, op_iter=[1, 0, 1])
    data = make_scan_inputs(spec)
    proj = measure(spec, data, spec.seed + 9)
    print(f"projector rows 1/2/4: {spec.name}, {spec.n_views} views", flush=True)
    ctx = Context(0)
    out = {}
    for rows in (1, 2, 4):
        job = Job(ctx, spec.geometry, spec.angles, data.alpha, data.beta, spec.op_iter, lam=1.0,
                  motion=True, nonneg=True)
        job.set_tuning(proj_split=8, proj_rows=rows)
        job.set_projections(proj)
        job.set_image(None)
        job.set_fields(data.fields)
        job.run()
        out[rows] = (job.image(), job.norms(), job.fields())
        job.close()
    ctx.close()
    for rows in (2, 4):
        assert np.array_equal(out[rows][0], out[1][0]), rows
        assert np.array_equal(out[rows][1], out[1][1], equal_nan=True), rows
        for a, b in zip(out[rows][2], out[1][2]):
            assert np.array_equal(a, b, equal_nan=True), rows
