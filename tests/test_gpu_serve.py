"""The reconstruction server: continuous batching of uneven scans on one GPU.

A request's results are the same alone and in any mix (bit for bit with fixed launch shapes:
the same kernels run in the same order on the request's own stream), the HBM budget bounds what
runs at once, and requests that cannot fit are refused at submit.
"""
import numpy as np
import pytest

from mcsart_amd._lib import McsartError
from mcsart_amd.configs import ScanSpec, make_scan_inputs, scaled
from mcsart_amd.driver import Context, Job, job_bytes
from mcsart_amd.serve import Server
from mcsart_amd.simulate import measure
from _util import rel_l2

pytestmark = pytest.mark.gpu

SPECS = [
    ScanSpec("s32", 32, 32, 32, 4.0, 4.0, 48, 40, 4.0, 24, True, [1, 0, 1], seed=71),
    ScanSpec("s40", 40, 40, 24, 3.0, 4.0, 64, 32, 3.0, 30, True, [1, 0, 1, 0], seed=72),
    ScanSpec("s24", 24, 24, 32, 4.0, 3.0, 40, 48, 4.5, 20, False, [1, 1, 0], seed=73),
    ScanSpec("s48", 48, 48, 16, 2.5, 5.0, 72, 24, 2.5, 16, True, [0, 1, 1], seed=74),
]


@pytest.fixture(scope="module")
def scans(gpu_lib):
    out = []
    for s in SPECS:
        d = make_scan_inputs(s)
        out.append((s, d, measure(s, d, s.seed + 5)))
    return out


def _submit(srv, scan, want_fields=False):
    s, d, p = scan
    return srv.submit(s.geometry, s.angles, d.alpha, d.beta, s.op_iter, p, fields=d.fields,
                      lam=s.lam, motion=s.motion, nonneg=s.nonneg, want_fields=want_fields)


def _alone(scans):
    res = []
    with Server(slots=1) as srv:
        for sc in scans:
            t = srv.wait(_submit(srv, sc, want_fields=True))
            res.append((t.image.copy(), t.norms.copy(), [f.copy() for f in t.fields]))
    return res


def test_alone_equals_plain_job(scans):
    """One request through the server equals the same scan run as a plain job."""
    alone = _alone(scans[:2])
    ctx = Context(0)
    for (s, d, p), (img, nrm, flds) in zip(scans[:2], alone):
        job = Job(ctx, s.geometry, s.angles, d.alpha, d.beta, s.op_iter, lam=s.lam,
                  motion=s.motion, nonneg=s.nonneg)
        job.set_projections(p)
        job.set_image(None)
        job.set_fields(d.fields)
        job.run()
        assert np.array_equal(job.image(), img)
        assert np.array_equal(job.norms(), nrm)
        for a, b in zip(job.fields(), flds):
            assert np.array_equal(a, b)
        job.close()
    ctx.close()


def test_mix_equals_alone(scans):
    alone = _alone(scans)
    order = [0, 1, 2, 3, 2, 0, 3, 1]   # duplicates: the same scan twice in flight
    with Server(slots=3) as srv:
        tickets = [_submit(srv, scans[k], want_fields=True) for k in order]
        srv.drain()
        infos = [srv.info(t) for t in tickets]
        for k, t, inf in zip(order, tickets, infos):
            assert inf["state"] == "done", inf
            img, nrm, flds = alone[k]
            assert np.array_equal(t.image, img), f"scan {SPECS[k].name} differs in the mix"
            assert np.array_equal(t.norms, nrm)
            for a, b in zip(t.fields, flds):
                assert np.array_equal(a, b)
    # the three slots really ran jobs side by side
    iv = [(i["admit_at_ms"], i["done_at_ms"]) for i in infos]
    assert max(sum(1 for a, b in iv if a <= t0 < b) for t0, _ in iv) >= 2


def test_budget_bounds_concurrency(scans):
    s, d, p = scans[0]
    nb = job_bytes(s.geometry, s.n_views, len(s.op_iter))
    with Server(slots=4, hbm_budget=int(2.5 * nb)) as srv:
        tickets = [_submit(srv, scans[0]) for _ in range(5)]
        srv.drain()
        infos = [srv.info(t) for t in tickets]
        used, budget = srv.memory()
        assert used == 0 and budget == int(2.5 * nb)
        ref = tickets[0].image
        for t in tickets[1:]:
            assert np.array_equal(t.image, ref)
    iv = [(i["admit_at_ms"], i["done_at_ms"]) for i in infos]   # one clock for all requests
    assert all(0.0 < a <= b for a, b in iv), iv
    overlap = max(sum(1 for a, b in iv if a <= t0 < b) for t0, _ in iv)
    assert overlap <= 2, f"{overlap} jobs ran at once under a budget for 2"
    assert all(i["bytes"] == nb for i in infos)


def test_oversize_and_invalid_rejected(scans):
    s, d, p = scans[0]
    nb = job_bytes(s.geometry, s.n_views, len(s.op_iter))
    with Server(slots=2, hbm_budget=nb - 1) as srv:
        with pytest.raises(McsartError):
            _submit(srv, scans[0])
    with Server(slots=1) as srv:
        with pytest.raises(McsartError):   # nx != ny
            bad = ScanSpec("bad", 32, 30, 16, 4.0, 4.0, 48, 40, 4.0, 8, False, [1], seed=1)
            srv.submit(bad.geometry, bad.angles, np.zeros(8), np.zeros(8), [1],
                       np.zeros((8, 40, 48), np.float32))
        t = srv.wait(_submit(srv, scans[2]))   # the server keeps serving
        assert srv.info(t)["state"] == "done"


def test_autotuned_requests_match(scans):
    """Autotuned launch shapes change only rounding: on image-sweep schedules (well conditioned)
    the results match the automatic shapes to 1e-5.  (Motion sweeps amplify any rounding
    difference, as the reference's own FP32 does against FP64; there, only fixed shapes give
    reproducible results, which is why the server's default is the geometry-only choice.)"""
    img_only = [(scaled(s, op_iter=[1, 1]), d, p) for s, d, p in scans[:2]]
    alone = _alone(img_only)
    with Server(slots=2, autotune=True) as srv:
        tickets = [_submit(srv, sc) for sc in img_only]
        srv.drain()
        for t, (img, nrm, _) in zip(tickets, alone):
            inf = srv.info(t)
            assert inf["tuning"]["proj_split"] in (1, 2, 4, 8, 16)
            assert rel_l2(t.image, img) <= 1e-5


@pytest.fixture(scope="module")
def long_scan(gpu_lib):
    """A 128^3 scan with 40 warped image sweeps (a few hundred ms on the device; image sweeps
    only, so its result stays finite over the whole schedule)."""
    s = ScanSpec("long128", 128, 128, 128, 2.0, 2.0, 256, 192, 1.5, 200, True, [1] * 40,
                 seed=75)
    d = make_scan_inputs(s)
    return s, d, measure(s, d, s.seed + 5)


def test_short_requests_cycle_past_a_long_one(scans, long_scan):
    """Continuous batching: with 2 slots, short requests submitted behind a long one keep cycling
    through the free slot while the long one runs (no completion waits for the whole device), and
    every result equals its alone run."""
    shorts = [scans[0], scans[2], scans[3], scans[0], scans[2]]
    alone = _alone([long_scan] + shorts)
    with Server(slots=2) as srv:
        tl = _submit(srv, long_scan, want_fields=True)
        ts = [_submit(srv, sc, want_fields=True) for sc in shorts]
        srv.drain()
        il = srv.info(tl)
        infos = [srv.info(t) for t in ts]
        for t, (img, nrm, flds) in zip([tl] + ts, alone):
            assert np.array_equal(t.image, img) and np.array_equal(t.norms, nrm)
    print(f"long: admitted {il['admit_at_ms']:.1f} ms, done {il['done_at_ms']:.1f} ms; shorts "
          f"admitted at {[round(i['admit_at_ms'], 1) for i in infos]}, done at "
          f"{[round(i['done_at_ms'], 1) for i in infos]}")
    assert all(i["state"] == "done" for i in infos + [il])
    # the last short request could only start after the earlier ones had finished on the second
    # slot, and still started (and finished) while the long one was running
    assert infos[-1]["admit_at_ms"] > infos[0]["done_at_ms"]
    assert infos[-1]["done_at_ms"] < il["done_at_ms"]


def test_cancel_running_and_queued(scans, long_scan):
    """A running request cancelled mid-schedule stops after the sweeps already on the device and
    frees its slot for the next request; a queued one leaves the queue without running.  Waiting
    on either reports the cancellation (status and message on the waiting thread)."""
    import time
    with Server(slots=1) as srv:
        t_full = _submit(srv, long_scan)
        srv.wait(t_full)
        full_ms = srv.info(t_full)["run_ms"]
        tl = _submit(srv, long_scan)
        tq = _submit(srv, scans[3])   # queued behind it
        tc = _submit(srv, scans[0])   # cancelled while queued
        t0 = time.perf_counter()
        while srv.info(tl)["state"] != "running" and time.perf_counter() - t0 < 30:
            time.sleep(0.001)
        time.sleep(0.02)
        srv.cancel(tc)
        srv.cancel(tl)
        with pytest.raises(McsartError) as e:
            srv.wait(tl)
        assert e.value.status == -6 and "cancelled" in str(e.value)
        with pytest.raises(McsartError):
            srv.wait(tc)
        srv.wait(tq)
        il, iq, ic = srv.info(tl), srv.info(tq), srv.info(tc)
        with pytest.raises(McsartError):   # an ended request cannot be cancelled
            srv.cancel(tl)
    print(f"full run {full_ms:.1f} ms; cancelled run {il['run_ms']:.1f} ms")
    assert il["state"] == "cancelled" and ic["state"] == "cancelled" and iq["state"] == "done"
    assert ic["slot"] == -1                     # never admitted
    assert il["run_ms"] < 0.5 * full_ms         # stopped early
    # the slot went straight to the next request
    assert il["done_at_ms"] - 1.0 <= iq["admit_at_ms"] < il["done_at_ms"] + 50.0
    img, nrm, _ = _alone([scans[3]])[0]
    assert np.array_equal(tq.image, img) and np.array_equal(tq.norms, nrm)
