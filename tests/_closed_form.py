"""Closed-form scans for the motion (udvf) sweep, shared by the oracle pins and the GPU tests."""
import numpy as np

import oracle
from mcsart_amd.geometry import uniform_angles
from _util import random_fields


def _T(v):
    """xy transpose as add2 reads its second operand: T(v)[iz, iy, ix] = v[iz, ix, iy]
    (src/cu_add.cu:58)."""
    return np.swapaxes(v, 1, 2)


def ramp_motion_scan(g, V, stride, slopes=(0.02, 0.035, 0.05), b0=0.5, cs=None, seed=50,
                     volumes=None, flows=None, angles=None, model_errors=(), shell=False):
    """A motion-sweep scan whose result has a closed form.

    The image is a ramp x = b0 + a1 ix + a2 iy + a3 iz, so its forward differences are the
    constants D1 = a1, D2 = a2, D3 = a3 (zero on each axis' last plane).  Then, for view i with
    vd = v_i - v_{i-1}, fd = f_i - f_{i-1} (src/AUMISART.cu:249-250):
      img0 = vd (T(ay) D1 + T(ax) D2 + T(az) D3) + fd (T(by) D1 + T(bx) D2 + T(bz) D3)  (:252-258)
      q    = (D1 + D2 + D3) (-vd - fd)                                                   (:267)
    and measured data p_i = P_i(x - img0 + c_i q) make the residual r = p_i - P x + P img0
    (:246-260) equal to c_i P q, so u = B r / B(P q) = c_i wherever B(P q) != 0 and 0 elsewhere
    (0/0 -> 0, :270), by the linearity of P and B.  The update (:272-277) then adds
    T(u)(-vd) D_c to (ay, ax, az) and T(u)(-fd) D_c to (by, bx, bz), and the view's residual ratio
    is ||c_i P q|| / ||p_i|| (:263).

    ``model_errors`` plants one mistake in this model (not in the oracle) to show the check
    discriminates: "swap_xy" pairs ax with D1 and ay with D2, "vd_sign" flips vd in q,
    "no_transpose" drops T in the update.  Returns (x, fields0, proj, angles, volumes, flows,
    expected fields, expected ratios), and with ``shell`` also a boolean volume: the field voxels
    whose u (read at the transposed voxel) lies within one voxel of the edge of B(P q)'s support
    in some updated view.  There the ratio jumps between c_i and 0 (0/0 -> 0) when a footprint
    bound moves by rounding across a detector or field-of-view edge, so FP32 results may differ
    from FP64 by O(c_i) at those voxels whatever the arithmetic."""
    nz, ny, nx = g.vol_shape
    a1, a2, a3 = slopes
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = b0 + a1 * ix + a2 * iy + a3 * iz
    D1 = np.where(ix < nx - 1, a1, 0.0)
    D2 = np.where(iy < ny - 1, a2, 0.0)
    D3 = np.where(iz < nz - 1, a3, 0.0)
    F0 = [f.astype(np.float64) for f in random_fields(g.vol_shape, seed, amp=1.5)]
    ang = uniform_angles(V) if angles is None else np.asarray(angles, np.float32)
    k = np.arange(V)
    # default trace: vd = -0.06 - 0.02 sin(k), fd = 0.03 (-1)^k, so -vd - fd >= 0.01
    vol = np.asarray(volumes, np.float64) if volumes is not None else \
        np.concatenate([[0.0], np.cumsum(-0.06 - 0.02 * np.sin(k[1:]))])
    flo = np.asarray(flows, np.float64) if flows is not None else \
        np.concatenate([[0.0], np.cumsum(0.03 * (-1.0) ** k[1:])])
    if cs is None:
        cs = 0.8 * np.cos(0.7 * np.arange(V)) + 0.3
    ax, ay, az, bx, by, bz = [f.copy() for f in F0]
    proj = np.zeros((V, g.nb, g.na))
    ratios = np.zeros(V)
    T = (lambda v: v) if "no_transpose" in model_errors else _T
    edge = np.zeros(g.vol_shape, bool)
    pair1, pair2 = ((ax, ay), (bx, by)) if "swap_xy" in model_errors else ((ay, ax), (by, bx))
    for i in range(1, V, stride):
        vd, fd = vol[i] - vol[i - 1], flo[i] - flo[i - 1]
        assert abs(-vd - fd) > 1e-3
        img0 = (vd * (_T(ay) * D1 + _T(ax) * D2 + _T(az) * D3)
                + fd * (_T(by) * D1 + _T(bx) * D2 + _T(bz) * D3))
        qs = (vd - fd) if "vd_sign" in model_errors else (-vd - fd)
        q = (D1 + D2 + D3) * (-vd - fd)
        p = oracle.projection(x - img0 + cs[i] * q, g, ang[i])
        proj[i] = p
        Pq = oracle.projection((D1 + D2 + D3) * qs, g, ang[i])
        S = oracle.backprojection(Pq, g, ang[i]) != 0
        if shell:
            from scipy import ndimage
            edge |= ndimage.binary_dilation(S) & \
                ~ndimage.binary_erosion(S, border_value=1)
        u = cs[i] * (-vd - fd) / qs * S
        ratios[i] = np.linalg.norm(cs[i] * oracle.projection(q, g, ang[i])) / np.linalg.norm(p)
        Tu = T(u)
        c1, c2 = pair1
        c1 += Tu * (-vd) * D1
        c2 += Tu * (-vd) * D2
        az += Tu * (-vd) * D3
        d1, d2 = pair2
        d1 += Tu * (-fd) * D1
        d2 += Tu * (-fd) * D2
        bz += Tu * (-fd) * D3
    out = (x, F0, proj, ang, vol, flo, [ax, ay, az, bx, by, bz], ratios)
    return out + (_T(edge),) if shell else out


def well_conditioned_trace(V):
    """A trace and per-view constants for ramp_motion_scan that keep FP32 rounding small: same-sign
    differences (vd in [-0.4, -0.2], fd in [0.05, 0.15], so the summed field updates do not cancel)
    and c_i in [75, 225], so the residual c P q is of the size of P x instead of a small
    difference of large projections.  Returns (volumes, flows, cs)."""
    k = np.arange(V)
    vol = np.concatenate([[0.0], np.cumsum(-0.3 - 0.1 * np.sin(k[1:]))])
    flo = np.concatenate([[0.0], np.cumsum(0.1 + 0.05 * np.sin(1.7 * k[1:]))])
    cs = 150.0 * (1.0 + 0.5 * np.cos(0.7 * k))
    return vol, flo, cs
