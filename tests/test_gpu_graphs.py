"""Graph-captured sweeps replay exactly the eager launches.

Each replay is separated by eager work of the program itself and a synchronisation, and every
replay is checked bit for bit against the eager result (captured graphs must stay correct on
replays after the first)."""
import numpy as np
import pytest

import oracle
from mcsart_amd import configs
from mcsart_amd.driver import Context, Job
from mcsart_amd.ops import project
from mcsart_amd.synth import ellipsoid_phantom
from _util import random_fields

pytestmark = pytest.mark.gpu


def _scan(motion):
    spec = configs.c1_static64()
    g = spec.geometry
    V = 24
    ang = spec.angles[::60 // V][:V]
    ph = ellipsoid_phantom(g.nx, g.ny, g.nz, spec.voxel, spec.seed)
    proj = np.stack([oracle.projection(ph, g, a) for a in ang]).astype(np.float32)
    if motion:
        F = random_fields(g.vol_shape, 5, amp=1.0)
        alpha = np.linspace(0, 0.6, V).astype(np.float32)
        beta = np.linspace(0.2, -0.2, V).astype(np.float32)
        op = [1, 0, 1]
    else:
        F, alpha, beta = None, np.zeros(V, np.float32), np.zeros(V, np.float32)
        op = [1, 1, 0]
    return g, ang, proj, F, alpha, beta, op


@pytest.mark.parametrize("motion", [False, True])
def test_graph_replays_match_eager(motion):
    g, ang, proj, F, alpha, beta, op = _scan(motion)
    ctx = Context(0)

    def make(graphs):
        job = Job(ctx, g, ang, alpha, beta, op, lam=1.0, motion=motion, nonneg=True)
        job.set_graphs(graphs)
        job.set_projections(proj)
        job.set_image(None)
        job.set_fields(F)
        return job

    eager = make(False)
    eager.run()
    ref = (eager.image(), eager.norms(), eager.fields())
    eager.close()

    job = make(True)
    for rep in range(4):
        if rep == 2:
            ctx.enable_timing(True)   # re-captured with timing events
        job.reset()
        job.run()
        ctx.synchronize()
        img, nrm, fl = job.image(), job.norms(), job.fields()
        assert np.array_equal(img, ref[0]), rep
        assert np.array_equal(nrm, ref[1]), rep
        assert all(np.array_equal(a, b) for a, b in zip(fl, ref[2])), rep
        # eager work of the program between replays
        _ = project(np.ones(g.vol_shape, np.float32), g, ang[rep])
    ms, n = ctx.kernel_time("projector")
    assert n == 2 * sum(len(ang) if o == 1 else len(range(1, len(ang), 1 if motion else 10))
                        for o in op) and ms > 0
    job.close()
    ctx.close()
