"""GPU parity of the motion (udvf) sweep (src/AUMISART.cu:236-288) at depth.

The GPU runs the sweep in accumulator form: the six field updates of every view are carried as
two transposed accumulators, and each view's update and the next view's prep run in the epilogue
of the ratio backprojector (first view: ``first``; middle views; last view: ``last``).  From the
second view on this groups the same sums differently from the reference's add2 sequence, so these
tests run many views through it:

* a well-conditioned sweep with a closed form (ramp image: u = c exactly), 23 motion views at the
  128^3 scan's geometry and 9 at the 256x256x128 scan's, against the FP64 oracle: residual ratios to 1e-4, field increments to
  rel-L2 1e-4 away from the edge of B(P q)'s support (where u jumps between c and 0 under
  rounding, 0/0 -> 0) and within 3x the reference's own FP32 error everywhere;
* teacher forcing at depth on the real 128^3 scan: the FP64 oracle runs the sweep to views 49, 99
  and 149, and from each of those states a five-view job (updates of the views 50-53, 100-103,
  150-153: first, middle and last steps) is held to 3x the error of the reference's own FP32
  arithmetic from the same state (the sweep divides by B(P q), which has near-zero values, so on
  this scan the FP64 fields themselves grow to ~1e6 by view 100);
* a full schedule with a motion sweep at 64^3 against the same bar.
"""
import numpy as np
import pytest

import oracle
from mcsart_amd import configs
from mcsart_amd.configs import make_scan_inputs, scaled
from mcsart_amd.driver import Context, Job
from mcsart_amd.geometry import Geometry, uniform_angles
from mcsart_amd.simulate import measure
from mcsart_amd.synth import Voxel, breathing_trace, ellipsoid_phantom, motion_fields
from _closed_form import ramp_motion_scan, well_conditioned_trace
from _util import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _sweep(ctx, g, ang, vol, flo, proj, x, fields):
    """One motion sweep (op_iter [0], overload 2) as a device-resident job."""
    job = Job(ctx, g, ang, vol, flo, [0], lam=1.0, motion=True, nonneg=False)
    job.set_projections(proj)
    job.set_image(x)
    job.set_fields(fields)
    job.run()
    out = job.fields(), job.norms()
    job.close()
    return out


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def test_motion_sweep_closed_form_c2_geometry(ctx):
    """23 motion views on the 128^3 scan's geometry (128^3 at 2 mm, 256x192 detector at 1.5 mm):
    ramp image, smooth random fields, the well-conditioned trace (-vd - fd in [0.1, 0.35],
    c_i in [75, 225]), data built so that u = c_i exactly."""
    _closed_form_sweep(ctx, configs.c2_mcsart128().geometry, 24)


def test_motion_sweep_closed_form_c3_geometry(ctx):
    """The same construction on the 256x256x128 scan's geometry (1.5 x 1.5 x 2 mm voxels,
    512x384 detector at 0.78 mm), 9 motion views: the headline workload's shapes (window widths,
    x- and y-branch views, both prep layouts) through the accumulator form."""
    _closed_form_sweep(ctx, configs.c3_clinical().geometry, 10)


def _closed_form_sweep(ctx, g, V):
    vol0, flo0, cs = well_conditioned_trace(V)
    x, F0, proj, ang, vol, flo, F_cf, ratios_cf, shell = ramp_motion_scan(
        g, V, 1, volumes=vol0, flows=flo0, cs=cs, shell=True)
    off = ~shell
    x32, p32, F32 = _f32(x), _f32(proj), [_f32(f) for f in F0]
    vol32, flo32 = _f32(vol), _f32(flo)
    print("closed-form sweep: oracle FP64 ...", flush=True)
    F64, n64 = oracle.motion_sweep(x32, p32, g, ang, vol32, flo32, F32)
    print("closed-form sweep: GPU ...", flush=True)
    Fg, ng = _sweep(ctx, g, ang, vol32, flo32, p32, x32, F32)
    Fr, nr = oracle.motion_sweep(x32, p32, g, ang, vol32, flo32, F32, dtype=np.float32)
    e_ratio = float(np.max(np.abs(ng[1:] / n64[1:] - 1)))
    e_inc = [rel_l2(Fg[c].astype(np.float64) - F32[c], F64[c] - F32[c]) for c in range(6)]
    e_ref = [rel_l2(Fr[c].astype(np.float64) - F32[c], F64[c] - F32[c]) for c in range(6)]
    e_off = [rel_l2((Fg[c].astype(np.float64) - F32[c])[off], (F64[c] - F32[c])[off])
             for c in range(6)]
    e_cf = [rel_l2((F64[c] - F32[c])[off], (F_cf[c] - F0[c])[off]) for c in range(6)]
    print(f"ratios max rel err {e_ratio:.2e} (reference FP32 "
          f"{float(np.max(np.abs(nr[1:] / n64[1:] - 1))):.2e}); field increments rel-L2 vs FP64: "
          f"{max(e_off):.2e} off the support-edge shell ({100 * shell.mean():.0f}% of voxels in "
          f"it), {max(e_inc):.2e} everywhere (reference FP32 {max(e_ref):.2e}); FP64 (float32 "
          f"inputs) vs closed form off the shell {max(e_cf):.2e}")
    assert np.all(ng[1:] > 0)
    assert e_ratio <= 1e-4
    # off the shell the sweep is well conditioned: the per-result parity bar
    assert max(e_off) <= 1e-4
    # in the shell u jumps between c_i and 0 with rounding (0/0 -> 0): the reference's own FP32
    # envelope
    assert max(e_inc) <= 3 * max(e_ref)
    assert max(e_cf) <= 1e-4   # the float32 rounding of the inputs moves the closed form little


@pytest.fixture(scope="module")
def c2_deep(ctx):
    """The 128^3 scan (noise-free data from the GPU forward simulator), the image after one
    warped image sweep on the GPU, and the FP64 oracle's motion-sweep state after views 49, 99
    and 149 of a sweep from that image with the true fields."""
    spec = scaled(configs.c2_mcsart128())
    spec.noise_photons = None
    g = spec.geometry
    d = make_scan_inputs(spec)
    proj = measure(spec, d, spec.seed + 5)
    job = Job(ctx, g, spec.angles, d.alpha, d.beta, [1], lam=1.0, motion=True, nonneg=True)
    job.set_projections(proj)
    job.set_image(None)
    job.set_fields(d.fields)
    job.run()
    x = job.image()
    job.close()
    states = {}
    F = [f.astype(np.float64) for f in d.fields]
    lo = 0
    for k in (50, 100, 150):
        print(f"FP64 oracle motion sweep to view {k - 1} ...", flush=True)
        F, _ = oracle.motion_sweep(x, proj[lo:k], g, spec.angles[lo:k], d.alpha[lo:k],
                                   d.beta[lo:k], F)
        states[k] = [_f32(f) for f in F]
        lo = k - 1
    return spec, g, d, proj, x, states


@pytest.mark.parametrize("k", [50, 100, 150])
def test_motion_sweep_teacher_forced_at_depth(ctx, c2_deep, k):
    """From the FP64 state after view k-1, the updates of views k..k+3 (the accumulator form's
    first, middle and last steps): field increments no further from FP64 than 3x the reference's
    own FP32 arithmetic (oracle FP32 instance) from the same state, residual ratios likewise."""
    spec, g, d, proj, x, states = c2_deep
    sl = slice(k - 1, k + 4)
    ang, vol, flo, p = spec.angles[sl], d.alpha[sl], d.beta[sl], proj[sl]
    F0 = states[k]
    print(f"teacher-forced views {k}..{k + 3}: max|field| {max(np.abs(f).max() for f in F0):.3g}",
          flush=True)
    Fg, ng = _sweep(ctx, g, ang, vol, flo, p, x, F0)
    F64, n64 = oracle.motion_sweep(x, p, g, ang, vol, flo, F0)
    F32, n32 = oracle.motion_sweep(x, p, g, ang, vol, flo, F0, dtype=np.float32)
    inc64 = [F64[c] - F0[c] for c in range(6)]
    e_gpu = max(rel_l2(Fg[c].astype(np.float64) - F0[c], inc64[c]) for c in range(6))
    e_32 = max(rel_l2(F32[c].astype(np.float64) - F0[c], inc64[c]) for c in range(6))
    r_gpu = float(np.max(np.abs(ng[1:] / n64[1:] - 1)))
    r_32 = float(np.max(np.abs(n32[1:] / n64[1:] - 1)))
    print(f"  field increments vs FP64: GPU {e_gpu:.2e}, reference FP32 {e_32:.2e}; "
          f"ratios: GPU {r_gpu:.2e}, reference FP32 {r_32:.2e}")
    assert all(np.isfinite(f).all() for f in Fg)
    assert e_gpu <= max(3 * e_32, 1e-5)
    assert r_gpu <= max(3 * r_32, 1e-5)


def test_full_schedule_with_motion_sweep_64(ctx):
    """(image, motion, image) at 64^3 with the 5DCT model: the image and alpha_x after the
    schedule are no further from FP64 than 3x the reference's own FP32 arithmetic (or 1e-3 for
    the image), the accuracy bar for schedules that contain motion sweeps."""
    n = 64
    vox = Voxel(4.0, 4.0)
    g = Geometry.from_mm(n, n, n, 4.0, 4.0, 96, 96, 4.0, 4.0)
    V = 60
    ang = uniform_angles(V)
    seed = 20261203 + 2000 + 77
    ph = ellipsoid_phantom(n, n, n, vox, seed)
    alpha, beta = breathing_trace(V, seed + 1, period_s=4.0, dt_s=0.1)
    fields = list(motion_fields(n, n, n, vox, seed + 2)) + \
        list(motion_fields(n, n, n, vox, seed + 3, scale=0.3))
    proj = np.stack([oracle.projection(oracle.deform(ph, alpha[i], beta[i], fields), g, ang[i])
                     for i in range(V)]).astype(np.float32)
    op = [1, 0, 1]
    job = Job(ctx, g, ang, alpha, beta, op, lam=1.0, motion=True, nonneg=True)
    job.set_projections(proj)
    job.set_image(None)
    job.set_fields(fields)
    job.run()
    img, ax = job.image(), job.fields()[0]
    job.close()
    r64 = oracle.aumisart(np.zeros(g.vol_shape), proj, g, ang, alpha, beta, op, fields=fields,
                          nonneg=True)
    r32 = oracle.aumisart(np.zeros(g.vol_shape), proj, g, ang, alpha, beta, op, fields=fields,
                          nonneg=True, dtype=np.float32)
    e_img, e_img32 = rel_l2(img, r64[0]), rel_l2(r32[0], r64[0])
    e_ax, e_ax32 = rel_l2(ax, r64[2]), rel_l2(r32[2], r64[2])
    print(f"[1,0,1] at 64^3: image GPU {e_img:.2e} / reference FP32 {e_img32:.2e}; "
          f"alpha_x GPU {e_ax:.2e} / reference FP32 {e_ax32:.2e}")
    assert np.isfinite(img).all() and np.isfinite(ax).all()
    assert e_img <= max(3 * e_img32, 1e-3)
    assert e_ax <= max(3 * e_ax32, 1e-3)
