"""Batched and served McSART on one GPU: scans/hour and p50/p99 latency against plain
one-job-at-a-time processing of the same requests.

    python -m mcsart_amd.serve_bench batch [--jobs 16] [--outer 10] [--slots 1,2,4,8]
    python -m mcsart_amd.serve_bench serve [--rates 0.5,1,2,4] [--requests 20] [--outer 2]

``batch``: the 16-scan uneven mix (volumes 128^3 .. 320x320x160, 200-900 views, detectors
256x192 .. 512x384), all submitted at t = 0; one run per slot count (slots = 1 is plain
one-at-a-time processing through the same code).  ``serve``: Poisson arrivals at each rate drawn
from the same mix plus 10% large scans (512x512x256, 900 views, 1024x768 detector), each arrival
submitted at its arrival time; plain (1 slot) and batched runs see the same stream.  Inputs are
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["batch", "serve"])
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
    else:
        run_serve(args)


if __name__ == "__main__":
    main()
