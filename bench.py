#!/usr/bin/env python3
"""McSART benchmark on one MI355X.

Workload: the 128^3 motion-compensated McSART scan (volume 128^3 float32 at 2 mm, detector
256x192 at 1.5 mm, 200 views over 360 deg, 10 outer iterations of one image sweep and one motion
sweep, lambda 1, nonnegativity clamp); `--workload clinical256` runs the clinical-size scan
instead (256x256x128 at 1.5x1.5x2 mm, detector 512x384 at 0.78 mm, 656 views, 20 outer
iterations).  One step is one complete reconstruction of that scan,
with the scan's measured projections, initial image and 5DCT fields resident in HBM before the
timed region starts.  Synthetic data: seeded ellipsoid phantom, breathing trace and Gaussian-blob
v/f fields; the measured projections are made on the GPU with this library's warp and projector
and Poisson noise at 1e5 photons (the FP64 CPU path is far too slow to synthesise them here).

Output: one JSON line with scans/hour (value), per-scan latency p50/p99, the roofline figures of
every kernel of the sweeps and of the dominant one (the largest share of a scan; HIP events on
the stream the kernels run on), the whole-step figure (the unfused loop's algorithmic bytes over
the scan time) and the CPU baseline (the FP64 OpenMP oracle timed on a bounded sample of the same
scan on this host).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(spec, kind, layouts=None):
    """HBM bytes one launch must move at minimum (4-byte floats).
    projector, image sweep : read x (N) and p (D), write (r, P1) (2D)      -> 4N + 12D
    projector, motion sweep: read g, Q (2N) and p (D), write (r, Pq) (2D)  -> 8N + 12D
    backprojector          : read (r, P1) (2D), write u (N)                -> 4N + 8D
    warp                   : read 6 fields and x, write x_i in each layout the view's rays
                             read (1 or 2)                                 -> 28N + 4N per layout
    warp-add               : read 6 fields, u and x, write x               -> 36N
    motion step            : the motion sweep's ratio backprojection with the view's update and
                             the next view's prep fused: read (r, Pq) (2D), read and write
                             the two accumulators, read the constants (Ga, Gb, H) and x,
                             write g in each layout the next view's rays read (q is the
                             sweep's Q times a scalar, written once per sweep)
                                                                           -> 32N + 4N per layout
                                                                              + 8D
    `layouts` (per view, bit 0 natural, bit 1 transposed) averages the per-layout terms over the
    views the kernel is launched for.
    """
    g = spec.geometry
    N, D = g.nvox, g.ndet

    def mean_layouts(views):
        if layouts is None or len(views) == 0:
            return 1.0
        return float(np.mean([bin(int(layouts[i]) & 3).count("1") for i in views]))

    V = spec.n_views
    if kind == "proj_image":
        return 4 * N + 12 * D
    if kind == "proj_motion":
        return 8 * N + 12 * D
    if kind == "back":
        return 4 * N + 8 * D
    if kind == "warp":
        return 28 * N + 4 * N * mean_layouts(range(V))
    if kind == "warp_add":
        return 36 * N
    if kind == "motion":
        return 32 * N + 4 * N * mean_layouts(range(2, V)) + 8 * D   # prep of views 2..V-1
    raise KeyError(kind)


def step_bytes(spec):
    """Algorithmic bytes of one view of each sweep kind, unfused (one HBM pass per operator of
    the reference's loop): image view 84N + 28D, motion view 128N + 44D."""
    g = spec.geometry
    return 84 * g.nvox + 28 * g.ndet, 128 * g.nvox + 44 * g.ndet


def profiled(workload, kind, field="traffic_bytes"):
    """A per-launch figure of the kernel from the committed rocprofv3 counter summary of this
    workload: HBM bytes (FETCH_SIZE + WRITE_SIZE, calibrated) or the VALU-busy fraction (SQ
    counters).  None when the summary or the figure is absent."""
    path = os.path.join(ROOT, "profiles", f"{workload}_kernels.json")
    if not os.path.exists(path):
        return None
    pattern = {"projector": r"k_project_rows<1,", "projector_motion": r"k_project_rows<2,",
               "backprojector": r"k_backproject<1, \d+, \w+, 0>",
               "warp": r"k_warp_tile", "warp_add": r"k_warp_add_tile",
               "motion": r"k_backproject<1, \d+, \w+, 1>"}[kind]
    with open(path) as f:
        rows = json.load(f).get("kernels", [])
    best = [r for r in rows if re.match(pattern, r["kernel"]) and field in r]
    if not best:
        return None
    best.sort(key=lambda r: -r["launches"])
    v = best[0][field]
    return int(v) if field == "traffic_bytes" else v


def concurrent_scans(ctx, spec, g, data, proj, tuning, jobs, rounds):
    """Throughput with `jobs` copies of the scan resident on separate streams, all started
    together, `rounds` times; returns scans/hour and the per-round wall time."""
    if jobs < 2:
        return None
    from mcsart_amd.driver import Context, Job
    ctxs = [Context(0) for _ in range(jobs)]
    js = []
    for c in ctxs:
        j = Job(c, g, spec.angles, data.alpha, data.beta, spec.op_iter, lam=spec.lam,
                motion=spec.motion, nonneg=spec.nonneg)
        j.set_projections(proj)
        j.set_image(None)
        j.set_fields(data.fields)
        j.set_tuning(**tuning)
        js.append(j)
    for j in js:   # warm-up: captures every job's graphs
        j.reset()
        j.run()
    for c in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    for _ in range(rounds):
        for j in js:
            j.reset()
            j.run()
        for c in ctxs:
            c.synchronize()
    el = time.perf_counter() - t0
    for j in js:
        j.close()
    for c in ctxs:
        c.close()
    return {"jobs_in_flight": jobs, "scans_per_hour": round(3600.0 * jobs * rounds / el, 1),
            "ms_per_round": round(el * 1e3 / rounds, 2)}


def host_boundary_ms(spec, g, data, proj):
    """Wall time of one mcsart_aumisart_motion call (the host_AUMISART replacement): upload,
    schedule, download, with allocation inside the call."""
    import ctypes
    from mcsart_amd._lib import check, load
    lib = load()
    V = spec.n_views
    op = np.ascontiguousarray(spec.op_iter, np.int32)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    img0 = np.zeros(g.vol_shape, np.float32)
    out_img, out_ax = np.empty_like(img0), np.empty_like(img0)
    out_nrm = np.empty(len(op) * V, np.float32)
    out_res = np.empty((V,) + g.det_shape, np.float32)
    vol, flo, ang = f32(data.alpha), f32(data.beta), f32(spec.angles)
    ew = np.ones(V, np.float32)
    fl = [f32(f) for f in data.fields]
    t0 = time.perf_counter()
    check(lib.mcsart_aumisart_motion(
        out_img.ctypes.data, out_res.ctypes.data, out_nrm.ctypes.data, out_ax.ctypes.data,
        img0.ctypes.data, f32(proj).ctypes.data, g.nx, g.ny, g.nz, g.na, g.nb, 1, V, len(op),
        op.ctypes.data, g.da, g.db, g.ai, g.bi, g.SO, g.SD, 1.0, spec.lam, vol.ctypes.data,
        flo.ctypes.data, ew.ctypes.data, ang.ctypes.data, *[f.ctypes.data for f in fl]),
        "mcsart_aumisart_motion")
    return (time.perf_counter() - t0) * 1e3


def launches_per_scan(spec):
    n_img = sum(1 for o in spec.op_iter if o == 1)
    n_mot = len(spec.op_iter) - n_img
    stride = 1 if spec.motion else 10
    mot_views = len(range(1, spec.n_views, stride))
    return n_img * spec.n_views, n_mot * mot_views


def cpu_baseline(spec, data, proj, sample_views):
    """FP64 oracle (OpenMP) on the first `sample_views` views of the same scan: one image sweep
    and one motion sweep, scaled to a whole scan."""
    import oracle
    g = spec.geometry
    V = sample_views
    ang = spec.angles[:V]
    al, be = data.alpha[:V], data.beta[:V]
    p = proj[:V]
    zero = np.zeros(g.vol_shape)
    t0 = time.perf_counter()
    oracle.aumisart(zero, p, g, ang, al, be, [1], fields=data.fields, nonneg=True)
    t_img = (time.perf_counter() - t0) / V
    t0 = time.perf_counter()
    oracle.aumisart(zero, p, g, ang, al, be, [0], fields=data.fields, nonneg=True)
    t_mot = (time.perf_counter() - t0) / (V - 1)
    n_img_views, n_mot_views = launches_per_scan(spec)
    sec_per_scan = n_img_views * t_img + n_mot_views * t_mot
    return {
        "value": 3600.0 / sec_per_scan,
        "unit": "scans/hour",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": (f"FP64 oracle, {spec.name} geometry, first {V} views: one image sweep "
                   f"({t_img:.3f} s/view) and one motion sweep ({t_mot:.3f} s/view), scaled to "
                   f"{n_img_views} image-sweep + {n_mot_views} motion-sweep views per scan"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample-views", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--concurrent", type=int, default=4,
                    help="also measure this many scans in flight at once (0: skip)")
    ap.add_argument("--workload", default="mcsart128", choices=["mcsart128", "clinical256"],
                    help="mcsart128: 128^3, 200 views, 10 outer iterations (the headline scan); "
                         "clinical256: 256x256x128, 656 views, 20 outer iterations")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(scans are independent: more GPUs would run more scans side by side)",
              file=sys.stderr)
        sys.exit(2)

    from mcsart_amd import configs
    from mcsart_amd.configs import make_scan_inputs
    from mcsart_amd.driver import Context, Job
    from mcsart_amd.simulate import measure

    spec = {"mcsart128": configs.c2_mcsart128, "clinical256": configs.c3_clinical}[args.workload]()
    g = spec.geometry
    log(f"[bench] workload {spec.name}: {g.nx}x{g.ny}x{g.nz}, det {g.na}x{g.nb}, V={spec.n_views}, "
        f"op_iter={spec.op_iter}")
    data = make_scan_inputs(spec)
    ctx = Context(0)
    proj = measure(spec, data, spec.seed + 404)
    log("[bench] measured projections ready")

    job = Job(ctx, g, spec.angles, data.alpha, data.beta, spec.op_iter, lam=spec.lam,
              motion=spec.motion, nonneg=spec.nonneg)
    job.set_projections(proj)
    job.set_image(None)
    job.set_fields(data.fields)

    # Launch shapes of the two footprint kernels, timed on a few of this scan's views.
    tuning = job.autotune()
    log(f"[bench] autotuned launch shapes: {tuning}")

    # Pass 1 (the measurement): each sweep replays its captured HIP graph; the warm-up captures.
    for w in range(max(args.warmup, 1)):
        job.reset()
        job.run()
        ctx.synchronize()
        log(f"[bench] warmup {w + 1}/{max(args.warmup, 1)} done")
    lat = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        job.reset()
        job.run()
        ctx.synchronize()
        lat.append(time.perf_counter() - ts)
        log(f"[bench] step {s + 1}/{args.steps}: {lat[-1] * 1e3:.1f} ms")
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0

    img = job.image()
    finite = float(np.isfinite(img).mean())

    # Pass 2 (kernel roofline): the same K scans with HIP events recorded on the job's stream
    # around every launch (events cannot be timed inside a graph, so these launches are eager).
    ctx.enable_timing(True)
    ctx.reset_timing()
    t2 = time.perf_counter()
    for s in range(args.steps):
        job.reset()
        job.run()
        ctx.synchronize()
    eager_ms = (time.perf_counter() - t2) * 1e3 / args.steps
    log(f"[bench] timed eager pass: {eager_ms:.1f} ms per scan")
    times = {k: ctx.kernel_time(k) for k in ("projector", "backprojector", "warp", "motion")}
    ctx.enable_timing(False)
    K = args.steps
    # Rooflines: every view of the scan launched back to back between two HIP events on the
    # job's stream (no per-launch event overhead), for each kernel of the sweeps.  The dominant
    # kernel is the one with the largest share of a scan (average launch x launches per scan).
    layouts = job.view_layouts()
    n_img_views, n_mot_views = launches_per_scan(spec)
    n_warps = n_img_views if spec.motion else 0
    per_scan = {"projector": n_img_views, "projector_motion": n_mot_views,
                "backprojector": n_img_views, "warp": n_warps, "warp_add": n_warps,
                "motion": n_mot_views}
    byte_kind = {"projector": "proj_image", "projector_motion": "proj_motion",
                 "backprojector": "back", "warp": "warp", "warp_add": "warp_add",
                 "motion": "motion"}
    iso_us = {k: job.kernel_time(k, reps=3) for k in per_scan}
    log(f"[bench] isolated launches: {iso_us}")
    kernels_roofline = {}
    for k, us in iso_us.items():
        nbytes = algorithmic_bytes(spec, byte_kind[k], layouts)
        ach = nbytes / (us * 1e-6) / 1e9
        kernels_roofline[k] = {
            "avg_launch_us": round(us, 2), "bytes_per_launch": int(nbytes),
            "achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
            "launches_per_scan": per_scan[k],
            "ms_per_scan": round(us * per_scan[k] * 1e-3, 2),
            "traffic": profiled(spec.name, k),
            "valu_busy": profiled(spec.name, k, "valu_busy")}
    dom = max(kernels_roofline, key=lambda k: kernels_roofline[k]["ms_per_scan"])
    d = kernels_roofline[dom]
    traffic = d["traffic"]
    roofline = {
        "bound": "hbm", "kernel": dom, "achieved": d["achieved"], "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": d["frac"], "traffic": traffic,
        "avg_launch_us": d["avg_launch_us"], "bytes_per_launch": d["bytes_per_launch"],
        "timing": "HIP events around all views launched back to back, 3 repetitions",
        "traffic_source": (f"profiles/{spec.name}_kernels.json (rocprofv3 FETCH_SIZE + "
                           "WRITE_SIZE, calibrated on a 64 MiB copy)") if traffic else None,
        "valu_busy": d["valu_busy"],
        "valu_busy_source": (f"profiles/{spec.name}_kernels.json (rocprofv3 SQ_INSTS_VALU, "
                             "SQ_BUSY_CYCLES: share of SIMD cycles issuing vector "
                             "instructions)") if d["valu_busy"] is not None else None,
    }
    kernel_ms = {k: {"total_ms": round(v[0], 2), "launches": v[1],
                     "avg_us": round(v[0] * 1e3 / max(v[1], 1), 2)} for k, v in times.items()}

    # Several copies of the scan in flight at once, each a resident job on its own stream.
    conc = concurrent_scans(ctx, spec, g, data, proj, tuning, args.concurrent, K)
    if conc:
        log(f"[bench] {args.concurrent} scans in flight: {conc}")

    # The drop-in boundary (host_AUMISART: host buffers in and out, device memory allocated and
    # freed per call): one whole scan through it, PCIe transfers included.  Not the headline value.
    host_ms = host_boundary_ms(spec, g, data, proj)
    log(f"[bench] host-buffer driver call: {host_ms:.1f} ms per scan")

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] timing the FP64 CPU oracle on a bounded sample ...")
        cpu = cpu_baseline(spec, data, proj, args.cpu_sample_views)

    vol_name = f"{g.nx}^3" if g.nx == g.ny == g.nz else f"{g.nx}x{g.ny}x{g.nz}"
    lat_ms = np.array(lat) * 1e3
    value = 3600.0 * K / elapsed
    # Whole-step figure: the unfused loop's algorithmic bytes per scan over the measured scan time
    img_b, mot_b = step_bytes(spec)
    scan_bytes = n_img_views * img_b + n_mot_views * mot_b
    step_ach = scan_bytes / (elapsed / K) / 1e9
    step_roofline = {"image_view_bytes": int(img_b), "motion_view_bytes": int(mot_b),
                     "bytes_per_scan": int(scan_bytes), "achieved": round(step_ach, 1),
                     "unit": "GB/s", "frac": round(step_ach / HBM_PEAK_GBS, 4)}
    out = {
        "metric": (f"McSART scans/hour ({vol_name} motion-compensated scan, {spec.n_views} views, "
                   f"{len(spec.op_iter) // 2} outer iterations)"),
        "value": round(value, 2),
        "unit": "scans/hour",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / K, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded phantom, trace, fields; projections made on GPU + Poisson noise)",
        "config": {"workload": spec.name, "volume": [g.nx, g.ny, g.nz], "detector": [g.na, g.nb],
                   "views": spec.n_views, "op_iter": f"[1,0]x{len(spec.op_iter) // 2}",
                   "lambda": spec.lam,
                   "nonneg": spec.nonneg, "jobs_in_flight": 1},
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 2),
                       "p99": round(float(np.percentile(lat_ms, 99)), 2)},
        "kernels": kernel_ms,
        "kernels_note": "eager pass with HIP events around every launch (includes event overhead)",
        "eager_ms_per_scan_with_events": round(eager_ms, 2),
        "isolated_launch_us": {k: round(v, 2) for k, v in iso_us.items()},
        "kernels_roofline": kernels_roofline,
        "step_roofline": step_roofline,
        "host_boundary_ms_per_scan": round(host_ms, 1),
        "concurrent": conc,
        "launch": "one HIP graph per sweep",
        "tuning": tuning,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "output_finite_fraction": finite,
    }
    print(json.dumps(out), flush=True)
    job.close()
    ctx.close()


if __name__ == "__main__":
    main()
