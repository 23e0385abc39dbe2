"""Pins for the CPU oracle (no GPU).

The reference ships no tests, fixtures or golden vectors, so each oracle function is pinned by
what the mathematics fixes: closed forms derived from the scan geometry, symmetries of the
discretisation, textbook routines (scipy trilinear interpolation, numpy differences) and
fixed points of the SART loop.  Each check is chosen so that a plausible transcription mistake
(a swapped axis, a wrong sign, an off-by-half offset, a dropped transpose) fails it.
"""
import numpy as np
import pytest
from scipy import ndimage

import oracle
from mcsart_amd.geometry import Geometry, uniform_angles
from _closed_form import _T, ramp_motion_scan
from _util import random_fields, rel_l2, smooth_random


def small_geo(n=16, nz=12, na=24, nb=20, da=1.0, db=1.0, SO=60.0, SD=90.0):
    return Geometry.from_user(n, n, nz, na, nb, da=da, db=db, SO=SO, SD=SD)


# ---------------------------------------------------------------- projector
def test_projection_zero_and_linearity():
    g = small_geo()
    a = smooth_random(g.vol_shape, 1).astype(np.float64)
    b = smooth_random(g.vol_shape, 2).astype(np.float64)
    ang = np.float32(0.7)
    assert np.all(oracle.projection(np.zeros(g.vol_shape), g, ang) == 0)
    lhs = oracle.projection(2.0 * a - 3.0 * b, g, ang)
    rhs = 2.0 * oracle.projection(a, g, ang) - 3.0 * oracle.projection(b, g, ang)
    assert rel_l2(lhs, rhs) < 1e-12


@pytest.mark.parametrize("ang", [0.0, np.pi / 2, 0.3, 2.0])
def test_projection_ones_counts_planes(ang):
    """Normalised footprints: a ray whose footprint stays inside every plane sums to the number
    of planes it steps through (nx for the x-branch, ny for the y-branch)."""
    g = small_geo()
    p = oracle.projection(np.ones(g.vol_shape), g, np.float32(ang))
    c = p[g.nb // 2 - 1:g.nb // 2 + 1, g.na // 2 - 1:g.na // 2 + 1]
    assert np.allclose(c, g.nx, rtol=1e-12)
    assert np.all(p <= g.nx + 1e-9) and np.all(p >= 0)


def test_projection_single_voxel_phi0():
    """At phi = 0 the footprint of detector cell [y20 -+ da/2] x [z2 -+ db/2] in x-plane ix is the
    cell scaled by xc/SD (similar triangles), xc = ix + 0.5 - nx/2 + SO, offset by n/2 into
    index space; a unit voxel contributes the product of its normalised overlaps."""
    g = small_geo()
    ix0, iy0, iz0 = 10, 9, 7
    img = np.zeros(g.vol_shape)
    img[iz0, iy0, ix0] = 1.0
    p = oracle.projection(img, g, np.float32(0.0))
    xc = ix0 + 0.5 - g.nx / 2 + g.SO
    expected = np.zeros((g.nb, g.na))
    for ib in range(g.nb):
        for ia in range(g.na):
            y20 = (ia + g.ai) * g.da
            z2 = (ib + g.bi) * g.db
            ylo, yhi = sorted(((y20 - g.da / 2) * xc / g.SD + g.ny // 2,
                               (y20 + g.da / 2) * xc / g.SD + g.ny // 2))
            zlo, zhi = sorted(((z2 - g.db / 2) * xc / g.SD + g.nz // 2,
                               (z2 + g.db / 2) * xc / g.SD + g.nz // 2))
            wy = max(0.0, min(iy0 + 1, yhi) - max(iy0, ylo)) / (yhi - ylo)
            wz = max(0.0, min(iz0 + 1, zhi) - max(iz0, zlo)) / (zhi - zlo)
            expected[ib, ia] = wy * wz
    assert np.abs(p - expected).max() < 1e-12
    assert p.max() > 0.1


def _rot90_volume(img):
    """The object rotated by +90 deg about z: img'(x, y) = img(y, -x) on the centred grid, i.e.
    img'[:, iy', ix'] = img[:, n-1-ix', iy'] for even n."""
    n = img.shape[1]
    out = np.empty_like(img)
    for ixp in range(n):
        for iyp in range(n):
            out[:, iyp, ixp] = img[:, n - 1 - ixp, iyp]
    return out


@pytest.mark.parametrize("ang", [0.2, 0.9, 1.3])
def test_projection_rotation_symmetry(ang):
    """Rotating source, detector and object together by 90 deg leaves the projection unchanged;
    the x- and y-branches of the projector must mirror each other exactly."""
    g = small_geo()
    img = smooth_random(g.vol_shape, 3).astype(np.float64)
    p0 = oracle.projection(img, g, np.float64(ang))
    p1 = oracle.projection(_rot90_volume(img), g, np.float64(ang) + np.pi / 2)
    assert rel_l2(p1, p0) < 1e-10


def test_projection_z_flip_symmetry():
    g = small_geo()
    img = smooth_random(g.vol_shape, 4)
    ang = np.float32(1.1)
    p0 = oracle.projection(img, g, ang)
    p1 = oracle.projection(img[::-1].copy(), g, ang)
    assert rel_l2(p1[::-1], p0) < 1e-12


def test_projection_f32_instance_close_to_f64():
    g = small_geo()
    img = smooth_random(g.vol_shape, 5)
    for ang in uniform_angles(7):
        assert rel_l2(oracle.projection(img, g, ang, dtype=np.float32),
                      oracle.projection(img, g, ang)) < 1e-5


# ---------------------------------------------------------------- backprojector
def test_backprojection_linearity():
    g = small_geo()
    a = smooth_random(g.det_shape, 6).astype(np.float64)
    b = smooth_random(g.det_shape, 7).astype(np.float64)
    ang = np.float32(0.4)
    lhs = oracle.backprojection(a + 0.5 * b, g, ang)
    rhs = oracle.backprojection(a, g, ang) + 0.5 * oracle.backprojection(b, g, ang)
    assert rel_l2(lhs, rhs) < 1e-12


def test_backprojection_isocentre_voxel_closed_form():
    """phi = 0, voxel with centre (0.5, 0.5, 0.5): its six points are the rotated-box corners
    (x, y) in {0, 1}^2 at z = 0.5 and (0.5, 0.5, {0, 1}); mapping a = SD y/(x + SO),
    b = SD z/(x + SO) gives a in [0, SD/SO] and b in [0, SD/(SO + 0.5)], so a unit projection
    backprojects to the unnormalised footprint area (SD/SO/da) * (SD/(SO+0.5)/db)."""
    g = small_geo(da=0.7, db=0.9)
    ones = np.ones(g.det_shape)
    v = oracle.backprojection(ones, g, np.float64(0.0))
    got = v[g.nz // 2, g.ny // 2, g.nx // 2]
    want = (g.SD / g.SO / g.da) * (g.SD / (g.SO + 0.5) / g.db)
    assert abs(got - want) < 1e-12 * want


def test_backprojection_rotation_symmetry():
    g = small_geo()
    p = smooth_random(g.det_shape, 8)
    ang = 0.6
    b0 = oracle.backprojection(p, g, np.float64(ang))
    b1 = oracle.backprojection(p, g, np.float64(ang) + np.pi / 2)
    assert rel_l2(b1, _rot90_volume(b0)) < 1e-10


def test_backprojection_not_adjoint_of_projection():
    """P is normalised per ray and B is not, so <Px, y> != <x, By> (documented property)."""
    g = small_geo()
    x = smooth_random(g.vol_shape, 9)
    y = smooth_random(g.det_shape, 10)
    ang = np.float32(0.3)
    lhs = float(np.sum(oracle.projection(x, g, ang) * y))
    rhs = float(np.sum(x * oracle.backprojection(y, g, ang)))
    assert abs(lhs - rhs) > 1e-3 * abs(lhs)


# ---------------------------------------------------------------- warp
def _scipy_warp(img, fields, vol, flow):
    """Textbook trilinear (scipy, order=1, clamp-to-edge) on the index grid.  The reference
    displaces dim 1 (x, fastest) by m_y and dim 2 (y) by m_x (src/cu_deform.cu:86-89)."""
    ax, ay, az, bx, by, bz = [f.astype(np.float64) for f in fields]
    mx, my, mz = ax * vol + bx * flow, ay * vol + by * flow, az * vol + bz * flow
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in img.shape], indexing="ij")
    return ndimage.map_coordinates(img.astype(np.float64), [z + mz, y + mx, x + my], order=1,
                                   mode="nearest")


def test_deform_matches_scipy_trilinear():
    shape = (10, 14, 14)
    img = smooth_random(shape, 11)
    fields = random_fields(shape, 12, amp=3.0)
    for vol, flow in ((0.8, -0.2), (-1.5, 0.6)):
        got = oracle.deform(img, vol, flow, fields)
        assert rel_l2(got, _scipy_warp(img, fields, vol, flow)) < 1e-12


def test_deform_identity_and_integer_shift():
    shape = (6, 8, 8)
    img = smooth_random(shape, 13)
    zero = [np.zeros(shape, np.float32)] * 6
    assert np.array_equal(oracle.deform(img, 1.0, 1.0, zero), img.astype(np.float64))
    # integer shift along dim 2 (numpy axis 1) by +2 via alpha_x = 1, volume = 2
    f = [np.ones(shape, np.float32), *[np.zeros(shape, np.float32)] * 5]
    out = oracle.deform(img, 2.0, 0.0, f)
    ref = img[:, np.minimum(np.arange(8) + 2, 7), :]
    assert np.array_equal(out, ref.astype(np.float64))


# ---------------------------------------------------------------- deform2
def test_deform2_on_ramp_with_swapped_checks():
    """On img = a*i1 + b*i2 + c*i3 the forward differences are the slopes, except where the
    reference's swapped boundary checks (src/cu_deform.cu:100-107) let a difference wrap into the
    next row / slice, and zero past the end of the volume."""
    n, nz = 6, 4
    a, b, c = 1.0, 10.0, 100.0
    i3, i2, i1 = np.meshgrid(np.arange(nz), np.arange(n), np.arange(n), indexing="ij")
    img = (a * i1 + b * i2 + c * i3).astype(np.float32)
    rng = [smooth_random((nz, n, n), s) for s in (14, 15, 16, 17, 18, 19)]
    vol, flow = 0.5, -0.25
    out = oracle.deform2(img, vol, flow, rng)
    ax, ay, az, bx, by, bz = [f.astype(np.float64) for f in rng]
    mx, my, mz = ax * vol + bx * flow, ay * vol + by * flow, az * vol + bz * flow
    flat = img.ravel().astype(np.float64)
    N = flat.size
    exp = np.empty(N)
    for idx in range(N):
        iy, ix, iz = idx % n, (idx // n) % n, idx // (n * n)   # id = iy + ix*ny + iz*nx*ny
        def nb(j):
            return flat[j] if j < N else flat[idx]
        dx = 0.0 if ix == n - 1 else nb(idx + 1) - flat[idx]
        dy = 0.0 if iy == n - 1 else nb(idx + n) - flat[idx]
        dz = 0.0 if iz == nz - 1 else nb(idx + n * n) - flat[idx]
        exp[idx] = flat[idx] + dy * mx.ravel()[idx] + dx * my.ravel()[idx] + dz * mz.ravel()[idx]
    assert np.abs(out.ravel() - exp).max() < 1e-12
    # the wrap really happens: at iy = n-1 (x = last column) dx is not the slope a
    assert not np.allclose(out, img + a * my + b * mx + c * mz)


# ---------------------------------------------------------------- differences, add2, initial2
def test_differences_match_numpy():
    img = smooth_random((5, 7, 9), 20).astype(np.float64)
    for ind, axis in ((1, 2), (2, 1), (3, 0)):
        f = np.zeros_like(img)
        sl = [slice(None)] * 3
        sl[axis] = slice(0, -1)
        f[tuple(sl)] = np.diff(img, axis=axis)
        assert np.array_equal(oracle.diff_forward(img, ind), f)
        bwd = np.zeros_like(img)
        sl[axis] = slice(1, -1)
        sl2 = [slice(None)] * 3
        sl2[axis] = slice(0, -2)
        sl3 = [slice(None)] * 3
        sl3[axis] = slice(1, -1)
        bwd[tuple(sl3)] = img[tuple(sl3)] - img[tuple(sl2)]
        assert np.array_equal(oracle.diff_backward(img, ind), bwd)


def test_add2_transposed_read():
    shape = (4, 6, 6)
    a, b, c = (smooth_random(shape, s).astype(np.float64) for s in (21, 22, 23))
    for ind in (1, 2, 3):
        d = oracle.diff_forward(c, ind)
        want = a + np.swapaxes(b, 1, 2) * 0.7 * d
        assert np.abs(oracle.add2(a, b, c, 0.7, ind) - want).max() < 1e-14


def test_initial2_and_division():
    x = smooth_random((4, 6, 6), 24).astype(np.float64)
    s = sum(oracle.diff_forward(x, k) for k in (1, 2, 3))
    assert np.abs(oracle.initial2(x, 0.3, -0.1) - s * 0.2).max() < 1e-14
    num = np.array([1.0, 0.0, -2.0, 5.0, 3.0]).reshape(1, 1, 5)
    den = np.array([2.0, 0.0, 0.0, 1e-320, np.inf]).reshape(1, 1, 5)
    out = oracle.division(num, den).ravel()
    assert out.tolist() == [0.5, 0.0, 0.0, 0.0, 0.0]   # NaN, -Inf, Inf -> 0


def test_add_and_norm():
    a, b = smooth_random((3, 4, 5), 25), smooth_random((3, 4, 5), 26)
    assert np.abs(oracle.add(a, b, -2.0) - (a - 2.0 * b.astype(np.float64))).max() < 1e-12
    assert abs(oracle.norm2(a) - np.linalg.norm(a.astype(np.float64))) < 1e-12


# ---------------------------------------------------------------- update_udvf
def test_update_udvf_residual_contraction():
    """With c = (D x) vd, (D x) fd and the linearised residual b = dfI - sum c*field (fields read
    with x/y swapped, src/cu_update_udvf.cu:45-51), the step of :54-60 changes b by the factor
    1 - (2 cax cay + caz^2 + 2 cbx cby + cbz^2) / (6 w); a step without the swap would give 5/6."""
    shape = (5, 8, 8)
    x = smooth_random(shape, 27).astype(np.float64)
    x0 = smooth_random(shape, 28).astype(np.float64)
    F = [f.astype(np.float64) for f in random_fields(shape, 29)]
    vd, fd = 0.4, -0.9

    def lin_res(fs):
        d1, d2, d3 = (oracle.diff_forward(x, k) for k in (1, 2, 3))
        cax, cay, caz, cbx, cby, cbz = d1 * vd, d2 * vd, d3 * vd, d1 * fd, d2 * fd, d3 * fd
        ax, ay, az, bx, by, bz = fs
        b = (x0 - x) - (cax * ay + cay * ax + caz * az + cbx * by + cby * bx + cbz * bz)
        w = cax ** 2 + cay ** 2 + caz ** 2 + cbx ** 2 + cby ** 2 + cbz ** 2
        factor = 1 - (2 * cax * cay + caz ** 2 + 2 * cbx * cby + cbz ** 2) / (6 * np.where(w > 0, w, 1))
        return b, w, factor

    b0, w, factor = lin_res(F)
    F1 = oracle.update_udvf(F, x, x0, vd, fd)
    b1, _, _ = lin_res(F1)
    m = w > 1e-12
    assert np.allclose(b1[m], b0[m] * factor[m], rtol=1e-9, atol=1e-12)
    # w == 0 voxels are left untouched
    flat = x.copy()
    flat[:] = 1.0
    F2 = oracle.update_udvf(F, flat, x0, vd, fd)
    assert all(np.array_equal(a, b) for a, b in zip(F2, F))


# ---------------------------------------------------------------- driver
def _consistent_scan(motion):
    g = small_geo(n=16, nz=12, na=24, nb=20)
    V = 6
    ang = uniform_angles(V)
    x = smooth_random(g.vol_shape, 30, scale=0.05).astype(np.float64)
    if motion:
        F = random_fields(g.vol_shape, 31, amp=1.5)
        alpha = np.float32([0.0, 0.3, 0.5, 0.2, -0.1, -0.4])
        beta = np.float32([0.1, 0.2, -0.3, 0.0, 0.4, 0.2])
        # the image sweep models view i as W(x; v_i - v_0): data consistent with x as reference
        a64, b64 = alpha.astype(np.float64), beta.astype(np.float64)
        proj = np.stack([oracle.projection(oracle.deform(x, a64[i] - a64[0], b64[i] - b64[0], F),
                                           g, ang[i]) for i in range(V)])
    else:
        F, alpha, beta = None, np.zeros(V, np.float32), np.zeros(V, np.float32)
        proj = np.stack([oracle.projection(x, g, a) for a in ang])
    return g, ang, x, F, alpha, beta, proj


@pytest.mark.parametrize("motion", [False, True])
def test_sart_fixed_point(motion):
    """Starting at the image that produced consistent data, the residual is zero in every view
    and the image does not move (pins the sign of the update and the warp direction/pairing)."""
    g, ang, x, F, alpha, beta, proj = _consistent_scan(motion)
    img, nrm, _, res = oracle.aumisart(x, proj, g, ang, alpha, beta, [1, 1], fields=F)
    assert np.abs(img - x).max() < 1e-10
    assert np.abs(nrm).max() < 1e-10 and np.abs(res).max() < 1e-10


def test_sart_converges_and_schedule():
    g, ang, x, F, alpha, beta, proj = _consistent_scan(False)
    V = len(ang)
    img, nrm, ax, res = oracle.aumisart(np.zeros(g.vol_shape), proj, g, ang, alpha, beta,
                                        [1, 1, 0, 1], lam=1.0)
    r = nrm.reshape(4, V)
    assert r[1].mean() < r[0].mean() and r[3].mean() < r[1].mean()
    # overload 1's motion sweep visits views 1, 11, ... only and leaves the image alone
    assert np.count_nonzero(r[2]) <= len(range(1, V, 10))
    img2 = oracle.aumisart(np.zeros(g.vol_shape), proj, g, ang, alpha, beta, [1, 1, 1])[0]
    assert np.array_equal(img, img2)
    # nonneg clamp
    img3 = oracle.aumisart(np.zeros(g.vol_shape), proj, g, ang, alpha, beta, [1], nonneg=True)[0]
    assert img3.min() >= 0


def test_zero_fields_overloads_agree():
    g, ang, x, _, alpha, beta, proj = _consistent_scan(False)
    zero = [np.zeros(g.vol_shape, np.float32)] * 6
    a = oracle.aumisart(np.zeros(g.vol_shape), proj, g, ang, alpha, beta, [1, 1])
    b = oracle.aumisart(np.zeros(g.vol_shape), proj, g, ang, alpha, beta, [1, 1], fields=zero)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---- host_multiple, DVF inverse, forward simulator ---------------------------------------------

def test_multiple_closed_form():
    """img1 * (w * forward difference of img0 along ind, zero at the last index)."""
    rng = np.random.default_rng(3)
    a, b = rng.random((5, 6, 7)), rng.random((5, 6, 7))
    for ind, axis in ((1, 2), (2, 1), (3, 0)):
        d = np.diff(b, axis=axis, append=np.take(b, [-1], axis=axis))   # zero at the end
        np.testing.assert_allclose(oracle.multiple(a, b, 0.7, ind), a * (0.7 * d), rtol=0,
                                   atol=1e-15)


def test_invert_dvf_zero_and_constant():
    z = np.zeros((4, 5, 6))
    for out in oracle.invert_dvf(z, z, z, 4):
        assert np.array_equal(out, z)
    c = np.ones((4, 5, 6))
    mx2, my2, mz2 = oracle.invert_dvf(0.3 * c, -1.25 * c, 2.0 * c, 3)
    assert np.array_equal(mx2, -0.3 * c) and np.array_equal(my2, 1.25 * c)
    assert np.array_equal(mz2, -2.0 * c)


def test_invert_dvf_linear_field_geometric_series():
    """m_x = a*(ix - c0): x_k = -a*(ix + x_{k-1} - c0) from x_0 = 0, so
    x_k = x* (1 - (-a)^k) with x* = -a (ix - c0) / (1 + a), exactly where no sample is clamped
    (trilinear interpolation reproduces linear functions)."""
    nz, ny, nx = 3, 4, 24
    a, c0, k = 0.2, 11.0, 5
    ix = np.arange(nx, dtype=np.float64)
    mx = np.broadcast_to(a * (ix - c0), (nz, ny, nx)).copy()
    z = np.zeros_like(mx)
    got = oracle.invert_dvf(mx, z, z, k)[0][1, 1]
    xs = -a * (ix - c0) / (1 + a) * (1 - (-a) ** k)
    interior = slice(4, 20)
    np.testing.assert_allclose(got[interior], xs[interior], rtol=0, atol=1e-12)


def test_invert_dvf_gauss_seidel_order():
    """y is sampled at the x just computed (not the previous sweep's): with m_x = c and
    m_y = b*ix, one sweep gives x = -c and y = -b*(ix - c)."""
    nz, ny, nx = 2, 3, 16
    b, c = 0.25, 0.5
    ix = np.arange(nx, dtype=np.float64)
    mx = np.full((nz, ny, nx), c)
    my = np.broadcast_to(b * ix, (nz, ny, nx)).copy()
    x, y, _ = oracle.invert_dvf(mx, my, np.zeros_like(mx), 1)
    np.testing.assert_array_equal(x, -mx)
    np.testing.assert_allclose(y[0, 1, 2:-2], -b * (ix[2:-2] - c), rtol=0, atol=1e-14)


def test_generate_proj_composition():
    """Without inversion the simulator is the projection of the warped image per view; with zero
    fields it is the plain projection."""
    g = Geometry.from_user(16, 16, 12, 24, 20, da=1.2, db=1.2, SO=60.0, SD=100.0)
    rng = np.random.default_rng(5)
    img = rng.random(g.vol_shape)
    fields = [0.3 * rng.standard_normal(g.vol_shape) for _ in range(6)]
    vol, flo = np.array([0.5, -1.0, 1.5]), np.array([0.1, 0.2, -0.3])
    ang = np.array([0.0, 1.1, 2.5])
    p = oracle.generate_proj(img, fields, vol, flo, ang, g, 0)
    for i in range(3):
        w = oracle.deform(img, vol[i], flo[i], fields)
        np.testing.assert_allclose(p[i], oracle.projection(w, g, ang[i]), rtol=1e-12, atol=1e-12)
    zf = [np.zeros(g.vol_shape)] * 6
    p0 = oracle.generate_proj(img, zf, vol, flo, ang, g, 4)
    for i in range(3):
        np.testing.assert_allclose(p0[i], oracle.projection(img, g, ang[i]), rtol=1e-12,
                                   atol=1e-12)


# ---- extreme displacements ---------------------------------------------------------------------

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_deform_extreme_and_nonfinite_displacements(dtype):
    """A displacement far outside the volume samples the edge voxel on its side (clamp-to-edge
    addressing), whatever its size: +-1e20 and +-Inf give the edge values exactly; NaN goes to the
    lower edge, as in the HIP warp.  Before the clamp the float->int conversion of such a
    displacement was undefined (1e20 gave index-0 values, +Inf gave NaN)."""
    nz, ny, nx = 5, 6, 6
    img = smooth_random((nz, ny, nx), 40).astype(dtype)
    zero = np.zeros((nz, ny, nx), dtype)
    # (field index, numpy axis it displaces): ay -> dim 1 (x, axis 2), ax -> dim 2 (y, axis 1),
    # az -> dim 3 (z, axis 0)  (src/cu_deform.cu:86-89)
    for comp, axis in ((1, 2), (0, 1), (2, 0)):
        n = img.shape[axis]
        for val, edge in ((1e20, n - 1), (-1e20, 0), (np.inf, n - 1), (-np.inf, 0), (np.nan, 0)):
            f = [zero] * 6
            f[comp] = np.full((nz, ny, nx), val, dtype)
            out = oracle.deform(img, 1.0, 0.0, f, dtype=dtype)
            ref = np.broadcast_to(np.take(img, [edge], axis=axis), img.shape)
            assert np.array_equal(out, ref.astype(dtype)), (comp, val)
    # mixed: an in-range fractional displacement on one axis, an extreme one on another
    f = [zero] * 6
    f[1] = np.full((nz, ny, nx), 0.25, dtype)
    f[2] = np.full((nz, ny, nx), -3e30, dtype)
    out = oracle.deform(img, 1.0, 0.0, f, dtype=dtype)
    z0 = np.broadcast_to(img[:1], img.shape).astype(np.float64)
    x = np.arange(nx)
    lo, hi = x, np.minimum(x + 1, nx - 1)
    ref = 0.75 * z0[:, :, lo] + 0.25 * z0[:, :, hi]
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-6 if dtype == np.float32 else 1e-14)


# ---- motion (udvf) sweep: closed form -----------------------------------------------------------

@pytest.mark.parametrize("stride", [1, 10])
def test_motion_sweep_closed_form(stride):
    """The oracle's motion sweep equals the ramp closed form (fields and residual ratios), with a
    detector narrow enough that B(P q)'s support is not xy-symmetric (a dropped transpose shows),
    fields with ax != ay and ay != T(ay), and vd != fd."""
    g = Geometry.from_user(16, 16, 12, 6, 20, da=1.0, db=1.0, SO=60.0, SD=90.0)
    V = 5 if stride == 1 else 13
    x, F0, proj, ang, vol, flo, F, ratios = ramp_motion_scan(g, V, stride)
    got, nrm = oracle.motion_sweep(x, proj, g, ang, vol, flo, F0, stride=stride)
    for c in range(6):
        np.testing.assert_allclose(got[c], F[c], rtol=0, atol=1e-12)
    np.testing.assert_allclose(nrm, ratios, rtol=1e-10, atol=0)
    upd = list(range(1, V, stride))
    assert np.all(nrm[upd] > 0) and np.count_nonzero(nrm) == len(upd)
    # every field moved (the increments are not trivially zero) and the support is one-sided
    assert all(np.abs(F[c] - F0[c]).max() > 1e-4 for c in range(6))
    inc = np.abs(F[1] - F0[1]) > 0
    assert inc.any() and not np.array_equal(inc, _T(inc))
    # the closed form discriminates the plausible mistakes
    for err in ("swap_xy", "vd_sign", "no_transpose"):
        Fw = ramp_motion_scan(g, V, stride, model_errors=(err,))[6]
        assert max(np.abs(Fw[c] - got[c]).max() for c in range(6)) > 1e-5, err


def test_motion_sweep_equals_driver_alpha_x():
    """The standalone sweep is the driver's motion sweep: aumisart with op_iter [0] (overload 2)
    gives the same alpha_x and ratios."""
    g = Geometry.from_user(16, 16, 12, 24, 20, da=1.0, db=1.0, SO=60.0, SD=90.0)
    V = 4
    x, F0, proj, ang, vol, flo, _, _ = ramp_motion_scan(g, V, 1)
    got, nrm = oracle.motion_sweep(x, proj, g, ang, vol, flo, F0)
    _, nrm2, ax2, _ = oracle.aumisart(x, proj, g, ang, vol, flo, [0], fields=F0)
    assert np.array_equal(got[0], ax2) and np.array_equal(nrm, nrm2)


def test_oracle_under_address_and_ub_sanitizers(tmp_path):
    """Every oracle function, including the warp on +-1e20 / +-Inf / NaN displacements, runs clean
    under AddressSanitizer + UndefinedBehaviorSanitizer (float-cast-overflow included; every report
    aborts): built by `make sanitize` from oracle/sanitize_check.cpp."""
    import os
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    subprocess.run(["make", "-s", "-C", here, "sanitize", f"OUT={tmp_path}"], check=True,
                   timeout=300)
    r = subprocess.run([str(tmp_path / "oracle_sanitize_check")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
    assert r.stdout.strip().endswith("SANITIZE_OK")
