"""The batched workload's scan sizes on the GPU: 320x320x160 and 192x192x128 volumes with the
512x384 detector (the largest shapes of the 16-scan uneven batch), against the FP64 oracle, and a
reduced batch of the real sizes through the server, each request the same alone and in the mix.

Operators: projection and backprojection on three views each (0, 45 deg and an oblique view),
the warp with the 5DCT blob fields; rel-L2 <= 1e-5 per result.
"""
import numpy as np
import pytest

import oracle
from mcsart_amd.configs import ScanSpec, make_scan_inputs
from mcsart_amd.geometry import Geometry
from mcsart_amd.ops import backproject, deform, project
from mcsart_amd.serve import Server
from mcsart_amd.simulate import measure
from mcsart_amd.synth import Voxel, ellipsoid_phantom, motion_fields
from _util import rel_l2, smooth_random_large

pytestmark = pytest.mark.gpu
TOL = 1e-5

# (nx, ny, nz, voxel xy mm, voxel z mm); detector 512x384 with the pitch that makes the
# transaxial detector field of view at the isocentre equal the volume's (magnification 1.5)
C4_LARGE = {
    "320x320x160": (320, 320, 160, 1.2, 1.6),
    "192x192x128": (192, 192, 128, 1.5, 2.0),
}


def _geometry(nx, ny, nz, vx, vz, na=512, nb=384):
    pitch = nx * vx * 1.5 / na
    return Geometry.from_mm(nx, ny, nz, vx, vz, na, nb, pitch, pitch)


@pytest.mark.parametrize("name", sorted(C4_LARGE))
def test_c4_operators_parity(gpu_lib, name):
    nx, ny, nz, vx, vz = C4_LARGE[name]
    g = _geometry(nx, ny, nz, vx, vz)
    vox = Voxel(vx, vz)
    ph = ellipsoid_phantom(nx, ny, nz, vox, 20261203 + 4000 + nx)
    for ang in (0.0, np.pi / 4, 2.2):
        a = np.float32(ang)
        print(f"{name}: projection at {ang:.3f}", flush=True)
        e = rel_l2(project(ph, g, a), oracle.projection(ph, g, a))
        assert e <= TOL, f"{name} projection angle {ang}: rel-L2 {e:.3e}"
        print(f"{name}: backprojection at {ang:.3f}", flush=True)
        proj = smooth_random_large(g.det_shape + (1,), nx + int(10 * ang))[..., 0]
        e = rel_l2(backproject(proj, g, a), oracle.backprojection(proj, g, a))
        assert e <= TOL, f"{name} backprojection angle {ang}: rel-L2 {e:.3e}"
    fields = list(motion_fields(nx, ny, nz, vox, 20261203 + 4100 + nx)) + \
        list(motion_fields(nx, ny, nz, vox, 20261203 + 4200 + nx, scale=0.3))
    for vol, flow in ((0.8, -0.4), (-1.1, 0.9)):
        print(f"{name}: warp ({vol}, {flow})", flush=True)
        e = rel_l2(deform(ph, vol, flow, fields), oracle.deform(ph, vol, flow, fields))
        assert e <= TOL, f"{name} warp ({vol}, {flow}): rel-L2 {e:.3e}"


# a reduced batch: each of the batch's four volume shapes with a detector of the batch, short
# schedules (one image sweep and one motion sweep), few views
C4_REDUCED = [
    ScanSpec("b320", 320, 320, 160, 1.2, 1.6, 512, 384, 1.125, 12, True, [1, 0], seed=20265001),
    ScanSpec("b192", 192, 192, 128, 1.5, 2.0, 384, 288, 1.125, 16, True, [1, 0], seed=20265002),
    ScanSpec("b128", 128, 128, 128, 2.0, 2.0, 256, 192, 1.5, 20, True, [1, 0], seed=20265003),
    ScanSpec("b256", 256, 256, 128, 1.5, 2.0, 512, 384, 1.125, 10, True, [1, 0], seed=20265004),
]


def test_c4_reduced_batch_alone_equals_mix(gpu_lib):
    scans = []
    for s in C4_REDUCED:
        print(f"reduced batch: inputs of {s.name}", flush=True)
        d = make_scan_inputs(s)
        scans.append((s, d, measure(s, d, s.seed + 5)))

    def submit(srv, sc):
        s, d, p = sc
        return srv.submit(s.geometry, s.angles, d.alpha, d.beta, s.op_iter, p, fields=d.fields,
                          lam=s.lam, motion=s.motion, nonneg=s.nonneg, want_fields=True)

    alone = []
    with Server(slots=1) as srv:
        for sc in scans:
            print(f"reduced batch: {sc[0].name} alone", flush=True)
            t = srv.wait(submit(srv, sc))
            alone.append((t.image.copy(), t.norms.copy(), [f.copy() for f in t.fields]))
    print("reduced batch: all four in a mix of three slots", flush=True)
    with Server(slots=3) as srv:
        tickets = [submit(srv, sc) for sc in scans]
        srv.drain()
        for sc, t, (img, nrm, flds) in zip(scans, tickets, alone):
            assert srv.info(t)["state"] == "done"
            assert np.isfinite(t.image).all()
            assert np.array_equal(t.image, img), f"{sc[0].name}: image differs in the mix"
            assert np.array_equal(t.norms, nrm), f"{sc[0].name}: norms differ in the mix"
            for a, b in zip(t.fields, flds):
                assert np.array_equal(a, b), f"{sc[0].name}: fields differ in the mix"
