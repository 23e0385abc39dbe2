"""Edge cases of the path: degenerate sizes, empty and single-view schedules, exact branch ties,
and the largest served size (512x512x256 with a 1024x768 detector): the operators against the
FP64 oracle on a few views (seconds each on the host's cores), and size-independent properties
of projection, backprojection and whole sweeps.
"""
import numpy as np
import pytest

import oracle
from mcsart_amd import configs
from mcsart_amd.driver import Context, Job
from mcsart_amd.geometry import Geometry, uniform_angles
from mcsart_amd.ops import DeviceArray, backproject, deform, project
from mcsart_amd.synth import Voxel, ellipsoid_phantom
from _util import rel_l2, smooth_random, smooth_random_large

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.mark.parametrize("name,g", [
    ("single_slice", Geometry.from_user(24, 24, 1, 36, 8, da=1.0, db=1.0, SO=80.0, SD=130.0)),
    ("two_slices", Geometry.from_user(16, 16, 2, 30, 6, da=1.2, db=0.8, SO=60.0, SD=100.0)),
    ("one_cell_detector", Geometry.from_user(12, 12, 6, 1, 1, da=4.0, db=4.0, SO=50.0, SD=90.0)),
    ("one_row_detector", Geometry.from_user(12, 12, 6, 40, 1, da=0.7, db=2.0, SO=50.0, SD=90.0)),
    ("one_column_detector", Geometry.from_user(12, 12, 6, 1, 20, da=2.0, db=0.7, SO=50.0,
                                               SD=90.0)),
    ("tiny_volume", Geometry.from_user(2, 2, 2, 8, 8, da=0.5, db=0.5, SO=20.0, SD=35.0)),
])
def test_degenerate_sizes(gpu_lib, name, g):
    vol = smooth_random(g.vol_shape, 3) + 0.5
    for ang in (0.0, 0.6, np.pi / 2, 2.2):
        a = np.float32(ang)
        p_ref = oracle.projection(vol, g, a)
        assert rel_l2(project(vol, g, a), p_ref) <= TOL, (name, ang)
        proj = smooth_random(g.det_shape, 7) + 0.5 if g.ndet > 1 else np.ones(g.det_shape)
        b_ref = oracle.backprojection(proj, g, a)
        got = backproject(proj, g, a)
        if np.linalg.norm(b_ref) == 0:
            assert np.array_equal(got, np.zeros_like(got)), (name, ang)
        else:
            assert rel_l2(got, b_ref) <= TOL, (name, ang)


def test_exact_branch_ties(gpu_lib):
    """Angles at exact multiples of 45 degrees (|x21| == |y21| for the central ray) and just
    beside them: the x/y branch choice must be the reference's (strict |x21| > |y21|)."""
    spec = configs.c1_static64()
    g = spec.geometry
    vol = smooth_random(g.vol_shape, 11)
    for k in range(8):
        for eps in (0.0, 1e-6, -1e-6):
            a = np.float32(k * np.pi / 4 + eps)
            assert rel_l2(project(vol, g, a), oracle.projection(vol, g, a)) <= TOL, (k, eps)


def _job(ctx, g, V, op_iter, motion):
    ang = uniform_angles(V)
    z = np.zeros(V, np.float32)
    job = Job(ctx, g, ang, z, z, np.array(op_iter, np.int32), lam=1.0, motion=motion)
    job.set_projections(np.ones((V,) + g.det_shape, np.float32))
    return job, ang


def test_projection_of_a_far_input(gpu_lib):
    """An input more than 4 GiB away from the context's scratch, where its transposed copy goes:
    the rows projector addresses both layouts from one base with 32-bit offsets, so the entry
    point stages a copy of the input beside the transposed one.  Results are bit-identical to the
    near case (usually the allocator places the input above the 5 GiB spacer, far from the
    scratch that the first call allocated)."""
    g = Geometry.from_user(32, 32, 16, 48, 24, da=1.0, db=1.0, SO=100.0, SD=160.0)
    vol = smooth_random(g.vol_shape, 21)
    angles = [np.float32(a) for a in (0.0, 0.8, 1.6, 2.4)]
    near = [project(vol, g, a) for a in angles]
    with DeviceArray.empty(((5 << 30) // 4,)):
        far = [project(vol, g, a) for a in angles]
    for a, n, f in zip(angles, near, far):
        assert np.array_equal(f, n), a
        assert rel_l2(f, oracle.projection(vol, g, a)) <= TOL, a


def test_empty_and_single_view_schedules(gpu_lib):
    g = Geometry.from_user(16, 16, 8, 24, 16, SO=60.0, SD=100.0)
    ctx = Context(0)
    x0 = smooth_random(g.vol_shape, 5)
    # no sweeps at all: the image stays the initial one
    job, _ = _job(ctx, g, 5, [], True)
    job.set_image(x0)
    job.set_fields(None)
    job.run()
    assert np.array_equal(job.image(), x0)
    job.close()
    # one view: image sweeps update from it; motion sweeps (views 1..V-1) do nothing
    for motion in (True, False):
        job, ang = _job(ctx, g, 1, [1, 0, 1], motion)
        job.set_image(x0)
        job.set_fields(None)
        job.run()
        img, nrm = job.image(), job.norms()
        job.close()
        ref_img, ref_nrm, _, _ = oracle.aumisart(x0.astype(np.float64),
                                                 np.ones((1,) + g.det_shape), g, ang,
                                                 np.zeros(1), np.zeros(1), [1, 0, 1])
        assert rel_l2(img, ref_img) <= 1e-5
        np.testing.assert_allclose(nrm, ref_nrm, rtol=1e-4, atol=1e-7)
        assert nrm[1] == 0.0   # the motion sweep of a one-view scan touches no view
    ctx.close()


LARGE = configs.c5_large().geometry   # 512 x 512 x 256, detector 1024 x 768


def _rot90(img):
    """The object rotated by +90 degrees about z: out[:, a, b] = img[:, n-1-b, a] (even n)."""
    return np.ascontiguousarray(img[:, ::-1, :].transpose(0, 2, 1))


def test_large_operators_parity(gpu_lib):
    """At the largest served size, against the FP64 oracle: projection of a sharp-edged phantom
    on a y-branch and an x-branch view, backprojection of that projection, and the warp with
    smooth 5DCT fields; rel-L2 <= 1e-5 each."""
    g = LARGE
    ph = ellipsoid_phantom(g.nx, g.ny, g.nz, Voxel(1.0, 1.0), 20261204)
    p = None
    for ang in (0.3, 1.3):   # |cos| > |sin|, then |sin| > |cos|: both ray branches
        a = np.float32(ang)
        print(f"large: projection at {ang}", flush=True)
        p = oracle.projection(ph, g, a)
        e = rel_l2(project(ph, g, a), p)
        assert e <= TOL, f"projection at {ang}: rel-L2 {e:.3e}"
    print("large: backprojection at 1.3", flush=True)
    e = rel_l2(backproject(p, g, np.float32(1.3)), oracle.backprojection(p, g, np.float32(1.3)))
    assert e <= TOL, f"backprojection: rel-L2 {e:.3e}"
    base = [smooth_random_large(g.vol_shape, 60 + k, scale=0.8) for k in range(3)]
    fields = base + [(-0.4 * f).astype(np.float32) for f in base[::-1]]
    print("large: warp (0.8, -0.4)", flush=True)
    e = rel_l2(deform(ph, 0.8, -0.4, fields), oracle.deform(ph, 0.8, -0.4, fields))
    assert e <= TOL, f"warp: rel-L2 {e:.3e}"


def test_large_projection_properties(gpu_lib):
    """At the largest served size: linearity, P(ones) = number of planes on central rays, and
    the 90-degree rotation symmetry (x and y branches mirror each other)."""
    g = LARGE
    x = smooth_random_large(g.vol_shape, 21)
    y = smooth_random_large(g.vol_shape, 22)
    a = np.float32(0.4)
    px, py = project(x, g, a), project(y, g, a)
    pxy = project((2.0 * x - 0.5 * y).astype(np.float32), g, a)
    assert rel_l2(pxy, 2.0 * px - 0.5 * py) <= 1e-5
    ones = np.ones(g.vol_shape, np.float32)
    p1 = project(ones, g, np.float32(0.3))
    c = p1[g.nb // 2 - 2:g.nb // 2 + 2, g.na // 2 - 2:g.na // 2 + 2]
    assert np.allclose(c, g.nx, rtol=1e-5), c
    # the object rotated by 90 degrees, seen from phi + 90 degrees, projects as the original
    phi = 0.25
    p_rot = project(_rot90(x), g, np.float32(phi + np.pi / 2))
    assert rel_l2(p_rot, project(x, g, np.float32(phi))) <= 1e-4


def test_large_backprojection_properties(gpu_lib):
    g = LARGE
    a = np.float32(1.1)
    p = smooth_random(g.det_shape, 31)
    q = smooth_random(g.det_shape, 32)   # detector-sized: the regular generator is quick
    bp, bq = backproject(p, g, a), backproject(q, g, a)
    bpq = backproject((3.0 * p + q).astype(np.float32), g, a)
    assert rel_l2(bpq, 3.0 * bp + bq) <= 1e-5
    assert np.isfinite(bp).all() and (backproject(np.abs(p), g, a) >= 0).all()


def test_large_job_sweeps_deterministic(gpu_lib):
    """Two views of an image and a motion sweep at the largest size: finite and bit-identical
    across two runs (graph replays included)."""
    g = LARGE
    V = 3
    ctx = Context(0)
    ang = uniform_angles(V)
    vol = np.array([0.0, 0.5, 1.0], np.float32)
    job = Job(ctx, g, ang, vol, 0.1 * vol, np.array([1, 0], np.int32), lam=1.0, motion=True,
              nonneg=True)
    rng = np.random.default_rng(1)
    job.set_projections(rng.random((V,) + g.det_shape, dtype=np.float32))
    job.set_image(None)
    fields = [smooth_random_large(g.vol_shape, 40 + k, scale=0.5) for k in range(6)]
    job.set_fields(fields)
    job.run()
    img1, n1 = job.image(), job.norms()
    job.reset()
    job.run()
    img2, n2 = job.image(), job.norms()
    job.close()
    ctx.close()
    assert np.isfinite(img1).all() and np.isfinite(n1).all()
    assert np.array_equal(img1, img2) and np.array_equal(n1, n2)


def test_warp_extreme_and_nonfinite_displacements(gpu_lib):
    """Regression for the warp's float->int conversions: displacements of +-1e20 and +-Inf sample
    the edge voxel on their side exactly, NaN the lower edge (the displacement is clamped to
    +-(n+1) first, which changes no clamp-to-edge sample), as the FP64 oracle now defines; a mix
    of an in-range fractional displacement and an extreme one matches the oracle to 1e-5."""
    import oracle
    from mcsart_amd.ops import deform
    from _util import smooth_random
    nz, ny, nx = 9, 20, 20
    img = smooth_random((nz, ny, nx), 40)
    zero = np.zeros((nz, ny, nx), np.float32)
    for comp, axis in ((1, 2), (0, 1), (2, 0)):
        n = img.shape[axis]
        for val, edge in ((1e20, n - 1), (-1e20, 0), (np.inf, n - 1), (-np.inf, 0), (np.nan, 0)):
            f = [zero] * 6
            f[comp] = np.full((nz, ny, nx), val, np.float32)
            print(f"warp: field {comp} = {val}", flush=True)
            got = deform(img, 1.0, 0.0, f)
            ref = np.broadcast_to(np.take(img, [edge], axis=axis), img.shape)
            assert np.array_equal(got, ref), (comp, val)
            assert np.array_equal(got, oracle.deform(img, 1.0, 0.0, f).astype(np.float32))
    f = [zero] * 6
    f[1] = np.full((nz, ny, nx), 0.25, np.float32)
    f[2] = np.full((nz, ny, nx), -3e30, np.float32)
    f[3] = np.broadcast_to(np.where(np.arange(nx) % 3 == 0, np.inf, 0.7),
                           (nz, ny, nx)).astype(np.float32)
    got = deform(img, 1.0, 0.5, f)
    ref = oracle.deform(img, 1.0, 0.5, f)
    assert np.isfinite(got).all()
    assert rel_l2(got, ref) <= 1e-5
