"""Batched and served McSART on one GPU: scans/hour and p50/p99 latency against plain
one-job-at-a-time processing of the same requests.

    python -m mcsart_amd.serve_bench batch [--jobs 16] [--outer 10] [--slots 1,2,4,8]
    python -m mcsart_amd.serve_bench serve [--rates 0.5,1,2,4] [--requests 20] [--outer 2]
    python -m mcsart_amd.serve_bench ceiling [--outer 2]

``batch``: the 16-scan uneven mix (volumes 128^3 .. 320x320x160, 200-900 views, detectors
256x192 .. 512x384), all submitted at t = 0; one run per slot count (slots = 1 is plain
one-at-a-time processing through the same code).  ``serve``: Poisson arrivals at each rate drawn
from the same mix plus 10% large scans (512x512x256, 900 views, 1024x768 detector), each arrival
submitted at its arrival time; plain (1 slot) and batched runs see the same stream.
``ceiling``: the most any packing of the batch's jobs (concurrent streams or one ragged launch)
could gain over one at a time, from each kernel's measured cost per unit of work by job size
(run_ceiling).  Inputs are
synthetic (seeded phantoms, traces, fields; projections made on the GPU with Poisson noise).
Prints one JSON object per run and a summary line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import configs
from .configs import make_scan_inputs, scaled
from .serve import Server
from .simulate import measure
from .synth import SplitMix64, seed_for


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class Scans:
    """Inputs per distinct scan spec, made once and shared by every request that uses it."""

    def __init__(self):
        self.cache = {}

    def get(self, spec):
        key = (spec.name, spec.n_views, tuple(spec.op_iter))
        if key not in self.cache:
            t = time.perf_counter()
            log(f"[serve_bench] making inputs {spec.name} {spec.nx}x{spec.ny}x{spec.nz} "
                f"V={spec.n_views} ...")
            d = make_scan_inputs(spec)
            log(f"[serve_bench] {spec.name}: phantom and fields in {time.perf_counter() - t:.1f} s")
            p = measure(spec, d, spec.seed + 404, log=log)
            self.cache[key] = (spec, d, p)
            log(f"[serve_bench] inputs {spec.name} {spec.nx}x{spec.ny}x{spec.nz} det "
                f"{spec.na}x{spec.nb} V={spec.n_views}: {time.perf_counter() - t:.1f} s")
        return self.cache[key]


def submit(srv, scan):
    s, d, p = scan
    return srv.submit(s.geometry, s.angles, d.alpha, d.beta, s.op_iter, p, fields=d.fields,
                      lam=s.lam, motion=s.motion, nonneg=s.nonneg)


def stats(infos, t_first, t_last_done, n):
    tot = np.array([i["total_ms"] for i in infos])
    span = t_last_done - t_first
    return {"scans": n, "makespan_s": round(span, 3),
            "scans_per_hour": round(3600.0 * n / span, 1),
            "latency_ms": {"p50": round(float(np.percentile(tot, 50)), 1),
                           "p99": round(float(np.percentile(tot, 99)), 1)},
            "queue_ms_p50": round(float(np.percentile([i["queue_ms"] for i in infos], 50)), 1)}


def run_batch(args):
    jobs = configs.c4_batch(args.jobs, n_outer=args.outer)
    scans = Scans()
    inputs = [scans.get(s) for s in jobs]
    results = {}
    ref_img = None
    for slots in args.slots:
        with Server(slots=slots, autotune=args.autotune) as srv:
            if slots == args.slots[0]:   # warm the code objects outside the timed runs
                srv.wait(submit(srv, scans.get(scaled(jobs[0], n_views=8, op_iter=[1, 0]))))
            t0 = time.perf_counter()
            tickets = [submit(srv, sc) for sc in inputs]
            srv.drain()
            t1 = time.perf_counter()
            infos = [srv.info(t) for t in tickets]
            assert all(i["state"] == "done" for i in infos), infos
            imgs = [t.image for t in tickets]
        if ref_img is None:
            ref_img = imgs
            worst = 0.0
        else:
            worst = max(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
                        for a, b in zip(imgs, ref_img))
        r = stats(infos, 0.0, t1 - t0, len(jobs))
        r.update({"slots": slots, "max_rel_l2_vs_first_run": worst})
        results[slots] = r
        log(f"[serve_bench] batch slots={slots}: {json.dumps(r)}")
    base = results[args.slots[0]]["scans_per_hour"]
    out = {"mode": "batch", "jobs": args.jobs, "outer": args.outer, "autotune": args.autotune,
           "runs": list(results.values()),
           "speedup_vs_first": {s: round(r["scans_per_hour"] / base, 3)
                                for s, r in results.items()}}
    print(json.dumps(out), flush=True)


def serve_mix(n, seed, outer):
    """n requests drawn from the batch mix plus 10% large scans."""
    rng = SplitMix64(seed)
    pool = configs.c4_batch(16, n_outer=outer)
    large = scaled(configs.c5_large(), op_iter=[1, 0] * outer)
    out = []
    for k in range(n):
        u = float(rng.uniform(1, 0.0, 1.0)[0])
        if u < 0.1:
            out.append(large)
        else:
            out.append(pool[int(rng.integers(1, 0, len(pool) - 1)[0])])
    return out


def run_serve(args):
    scans = Scans()
    mix = serve_mix(args.requests, seed_for(5, 99), args.outer)
    inputs = [scans.get(s) for s in mix]
    runs = []
    for rate in args.rates:
        rng = np.random.default_rng(seed_for(5, int(rate * 100)))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(inputs)))
        row = {"rate_per_s": rate}
        for label, slots in (("plain", 1), ("batched", args.slots_served)):
            with Server(slots=slots, autotune=args.autotune) as srv:
                t0 = time.perf_counter()
                tickets = []
                for k, (a, sc) in enumerate(zip(arrivals, inputs)):
                    dt = a - (time.perf_counter() - t0)
                    if dt > 0:
                        time.sleep(dt)
                    tickets.append(submit(srv, sc))
                    log(f"[serve_bench] rate {rate}/s {label}: request {k} ({sc[0].name}) "
                        f"submitted at {time.perf_counter() - t0:.1f} s")
                for k, t in enumerate(tickets):   # in submission order, reporting progress
                    srv.wait(t)
                    log(f"[serve_bench] rate {rate}/s {label}: request {k} done by "
                        f"{time.perf_counter() - t0:.1f} s")
                srv.drain()
                t1 = time.perf_counter()
                infos = [srv.info(t) for t in tickets]
                assert all(i["state"] == "done" for i in infos), infos
            row[label] = stats(infos, 0.0, t1 - t0, len(inputs))
            row[label]["slots"] = slots
            log(f"[serve_bench] rate {rate}/s {label}: {json.dumps(row[label])}")
        row["scans_per_hour_ratio"] = round(row["batched"]["scans_per_hour"] /
                                            row["plain"]["scans_per_hour"], 3)
        runs.append(row)
    print(json.dumps({"mode": "serve", "requests": args.requests, "outer": args.outer,
                      "large_fraction": round(sum(1 for s in mix if s.nx == 512) / len(mix), 3),
                      "runs": runs}), flush=True)


KINDS = ("projector", "projector_motion", "backprojector", "warp", "warp_add", "motion")


def _work_units(spec):
    """Work per launch of each kernel kind: ray-planes (detector cells x planes) for the
    projectors, voxels for the rest."""
    rays = spec.na * spec.nb * max(spec.nx, spec.ny)
    vox = spec.nx * spec.ny * spec.nz
    return {"projector": rays, "projector_motion": rays, "backprojector": vox, "warp": vox,
            "warp_add": vox, "motion": vox}


def _launches(spec):
    """Launches of each kind in the scan's schedule (image sweep: warp, projector,
    backprojector, warp-add per view; motion sweep: projector and fused backprojector per view
    after the first)."""
    n_img = sum(1 for o in spec.op_iter if o == 1)
    n_mot = len(spec.op_iter) - n_img
    V = spec.n_views
    return {"projector": n_img * V, "backprojector": n_img * V, "warp": n_img * V,
            "warp_add": n_img * V, "projector_motion": n_mot * (V - 1), "motion": n_mot * (V - 1)}


def run_ceiling(args):
    """Upper bound on what packing the batch's jobs together (streams or one ragged launch) can
    gain over running them one at a time.  Every kernel kind is timed alone (all views of the job
    back to back, the server's automatic launch shapes) for each distinct (volume, detector)
    class of the batch and divided by its work; the lowest cost per unit over the classes is the
    device's saturated rate for that kernel.  One-at-a-time time of the batch = sum over its
    jobs' launches of the class's measured launch time; the packed bound = the same launches
    costed at the saturated rate with no gaps between kernels.  ceiling = their ratio."""
    from .driver import Context, Job
    jobs = configs.c4_batch(args.jobs, n_outer=args.outer)
    for s in jobs:
        s.noise_photons = None
    classes = {}
    for s in jobs:
        classes.setdefault((s.nx, s.ny, s.nz, s.na, s.nb), s)
    ctx = Context(0)
    per_class = {}
    for key, s in classes.items():
        sc = scaled(s, n_views=min(s.n_views, args.ceiling_views))
        d = make_scan_inputs(sc)
        p = measure(sc, d, sc.seed + 404)
        job = Job(ctx, sc.geometry, sc.angles, d.alpha, d.beta, [1, 0], lam=sc.lam,
                  motion=True, nonneg=sc.nonneg)
        job.set_projections(p)
        job.set_image(None)
        job.set_fields(d.fields)
        us = {k: job.kernel_time(k, reps=3) for k in KINDS}
        job.close()
        units = _work_units(s)
        per_class[key] = {"us": us, "us_per_munit": {k: 1e6 * us[k] / units[k] for k in KINDS}}
        log(f"[serve_bench] ceiling class {key}: " +
            ", ".join(f"{k} {us[k]:.1f} us" for k in KINDS))
    ctx.close()
    best = {k: min(c["us_per_munit"][k] for c in per_class.values()) for k in KINDS}
    alone_ms = packed_ms = 0.0
    for s in jobs:
        c = per_class[(s.nx, s.ny, s.nz, s.na, s.nb)]
        n, u = _launches(s), _work_units(s)
        alone_ms += sum(n[k] * c["us"][k] for k in KINDS) * 1e-3
        packed_ms += sum(n[k] * u[k] * best[k] * 1e-6 for k in KINDS) * 1e-3
    out = {"mode": "ceiling", "jobs": args.jobs, "outer": args.outer,
           "classes": [{"volume": list(k[:3]), "detector": list(k[3:]),
                        "us_per_launch": {q: round(v, 2) for q, v in c["us"].items()},
                        "us_per_million_units": {q: round(v, 3)
                                                 for q, v in c["us_per_munit"].items()}}
                       for k, c in per_class.items()],
           "saturated_us_per_million_units": {k: round(v, 3) for k, v in best.items()},
           "kernel_ms_one_at_a_time": round(alone_ms, 1),
           "kernel_ms_packed_bound": round(packed_ms, 1),
           "packing_ceiling": round(alone_ms / packed_ms, 3)}
    print(json.dumps(out), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["batch", "serve", "ceiling"])
    ap.add_argument("--ceiling-views", type=int, default=128,
                    help="ceiling: views of each class's timing job")
    ap.add_argument("--jobs", type=int, default=16)
    ap.add_argument("--outer", type=int, default=10, help="outer iterations (image+motion sweep)")
    ap.add_argument("--slots", type=lambda s: [int(x) for x in s.split(",")], default=[1, 2, 4])
    ap.add_argument("--slots-served", type=int, default=4)
    ap.add_argument("--rates", type=lambda s: [float(x) for x in s.split(",")],
                    default=[0.5, 1.0, 2.0, 4.0])
    ap.add_argument("--requests", type=int, default=20)
    ap.add_argument("--autotune", action="store_true")
    args = ap.parse_args()
    if args.mode == "batch":
        run_batch(args)
    elif args.mode == "ceiling":
        run_ceiling(args)
    else:
        run_serve(args)


if __name__ == "__main__":
    main()
