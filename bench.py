#!/usr/bin/env python3
"""McSART benchmark on one MI355X.

Workload (default): the clinical-size motion-compensated McSART scan, the largest single-scan
workload: volume 256x256x128 float32 at 1.5x1.5x2 mm, detector 512x384 at 0.78 mm, 656 views over
360 deg, 20 outer iterations of one image sweep and one motion sweep, lambda 1, nonnegativity
clamp; `--workload mcsart128` runs the 128^3 scan (2 mm, 256x192 detector, 200 views, 10 outer
iterations).  One step is one complete reconstruction of the scan, with its measured
projections, initial image and 5DCT fields resident in HBM before the timed region starts.
Synthetic data: seeded ellipsoid phantom, breathing trace and Gaussian-blob v/f fields; the
measured projections are made on the GPU with this library's warp and projector and Poisson noise
at 1e5 photons.

Output: one JSON line with scans/hour (value), back-to-back scan wall time p50/p99, the roofline
figures of every kernel of the sweeps and of the dominant one (HIP events on the stream the
kernels run on; counter figures from profiles/<workload>_kernels.json), the 16-scan uneven batch
through the server (plain one slot against several) and a served Poisson stream with p50/p99
latency, and the CPU baseline (the FP64 OpenMP oracle on a bounded sample of the same scan,
1 warm-up + median of 5).  `--cpu-protocol` runs the CPU baseline's full protocol instead.
"""
import os

# OpenMP placement of the CPU baseline (read by the oracle's OpenMP runtime when it loads; it
# then binds this thread to one core, so the process's CPU set is read first)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
HOST_CPUS = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()

import argparse
import json
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(spec, kind, layouts=None, voxels=None):
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

    With culling (`voxels`: the per-view set sizes of mcsart_view_voxels) the warp, the warp-add
    and the motion step count only the voxels a launch processes (mcsart_amd/traffic.py): the
    warp's per-voxel terms over view v's read set R_v, the warp-add's sample and x terms over
    the voxels H_v where u can be nonzero, the motion step's accumulator terms over H_v (and
    H_v u R_{v+1}) and its constant and g terms over R_{v+1}.
    """
    g = spec.geometry
    N, D = g.nvox, g.ndet

    def n_layouts(i):
        return 1.0 if layouts is None else float(bin(int(layouts[i]) & 3).count("1"))

    def mean_layouts(views):
        if layouts is None or len(views) == 0:
            return 1.0
        return float(np.mean([n_layouts(i) for i in views]))

    V = spec.n_views
    if kind == "proj_image":
        return 4 * N + 12 * D
    if kind == "proj_motion":
        return 8 * N + 12 * D
    if kind == "back":
        return 4 * N + 8 * D
    if voxels is None:
        if kind == "warp":
            return 28 * N + 4 * N * mean_layouts(range(V))
        if kind == "warp_add":
            return 36 * N
        if kind == "motion":
            return 32 * N + 4 * N * mean_layouts(range(2, V)) + 8 * D   # prep of views 2..V-1
        raise KeyError(kind)
    from mcsart_amd import traffic
    if kind == "warp":
        return traffic.warp_bytes(voxels)
    if kind == "warp_add":
        return traffic.warp_add_bytes(voxels, N)
    if kind == "motion":   # views 1..V-2, each preparing view v + 1 (the isolated launches)
        return traffic.motion_middle_bytes(voxels, N, D)
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


def cpu_env(oracle_threads):
    """The host the CPU baseline ran on and its OpenMP settings."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": HOST_CPUS, "cpu_model": model,
            "omp_threads": oracle_threads,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "OMP_PLACES": os.environ.get("OMP_PLACES")}


def median_time(fn, reps=5):
    """One warm-up call, then the median (and all) of `reps` timed calls (steady clock)."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_operator_times(spec, data, reps=5):
    """Per-operator seconds at the scan's size (FP64 oracle): projection, backprojection, warp
    (DVF formed on the fly + trilinear), update_udvf; 1 warm-up then the median of `reps`."""
    import oracle
    g = spec.geometry
    ph = data.phantom
    ang = spec.angles[min(1, spec.n_views - 1)]
    p = oracle.projection(ph, g, ang)
    fields = data.fields or [np.zeros(g.vol_shape, np.float32)] * 6
    out = {}
    out["projection"] = median_time(lambda: oracle.projection(ph, g, ang), reps)[0]
    out["backprojection"] = median_time(lambda: oracle.backprojection(p, g, ang), reps)[0]
    out["warp"] = median_time(lambda: oracle.deform(ph, 0.7, -0.3, fields), reps)[0]
    out["update_udvf"] = median_time(
        lambda: oracle.update_udvf(fields, ph, 0.9 * ph, 0.05, -0.02), reps)[0]
    return {k: round(v, 4) for k, v in out.items()}


def cpu_view_times(spec, data, proj, reps=5, views=2):
    """Seconds per view of an image sweep and of a motion sweep (FP64 oracle driver) over the
    first `views` (image) / `views` updated (motion) views of the scan: 1 warm-up, median of
    `reps`."""
    import oracle
    g = spec.geometry
    zero = np.zeros(g.vol_shape)
    V = views + 1
    ang, al, be, p = spec.angles[:V], data.alpha[:V], data.beta[:V], proj[:V]
    t_img = median_time(lambda: oracle.aumisart(zero, p[:views], g, ang[:views], al[:views],
                                                be[:views], [1], fields=data.fields,
                                                nonneg=True), reps)[0] / views
    t_mot = median_time(lambda: oracle.motion_sweep(data.phantom, p, g, ang, al, be,
                                                    data.fields), reps)[0] / views
    return t_img, t_mot


def cpu_baseline(spec, data, proj, reps=5):
    """FP64 oracle (OpenMP, all host cores) on a bounded sample of the same scan, timed by the
    protocol 1 warm-up + median of `reps`: per-operator times, then the seconds per image-sweep
    view and per motion-sweep view over two views, scaled to the whole scan's views."""
    import oracle
    ops = cpu_operator_times(spec, data, reps)
    t_img, t_mot = cpu_view_times(spec, data, proj, reps)
    n_img_views, n_mot_views = launches_per_scan(spec)
    sec_per_scan = n_img_views * t_img + n_mot_views * t_mot
    return {
        "value": 3600.0 / sec_per_scan,
        "unit": "scans/hour",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": (f"FP64 oracle, {spec.name} scan: 1 warm-up + median of {reps}; two views of an "
                   f"image sweep ({t_img:.3f} s/view) and two of a motion sweep "
                   f"({t_mot:.3f} s/view), extrapolated to {n_img_views} image-sweep + "
                   f"{n_mot_views} motion-sweep views per scan"),
        "seconds_per_view": {"image_sweep": round(t_img, 4), "motion_sweep": round(t_mot, 4)},
        "operator_seconds": ops,
        "host": cpu_env(oracle.num_threads()),
    }


def cpu_protocol(reps=5):
    """The whole CPU-baseline protocol: per-operator times at the 64^3, 128^3 and 256x256x128
    scans, the 64^3 static reconstruction (5 SART iterations, 60 views; median of `reps` after a
    warm-up) and the 128^3 motion-compensated reconstruction (10 outer iterations, 200 views;
    one run) in full, the 256x256x128 reconstruction extrapolated from two views of each sweep."""
    import oracle
    from mcsart_amd import configs
    from mcsart_amd.configs import make_scan_inputs
    out = {"host": cpu_env(oracle.num_threads()), "reps": reps, "operators_s": {}}
    specs = {"static64": configs.c1_static64(), "mcsart128": configs.c2_mcsart128(),
             "clinical256": configs.c3_clinical()}
    for name, spec in specs.items():
        log(f"[bench] CPU protocol: operators at {name}")
        out["operators_s"][name] = cpu_operator_times(spec, make_scan_inputs(spec), reps)
    # 64^3 static: data from the oracle projector, the whole reconstruction
    spec = specs["static64"]
    g = spec.geometry
    d = make_scan_inputs(spec)
    proj = np.stack([oracle.projection(d.phantom, g, a) for a in spec.angles])
    zero = np.zeros(g.vol_shape)
    z = np.zeros(spec.n_views)
    log("[bench] CPU protocol: 64^3 static reconstruction")
    t, ts = median_time(lambda: oracle.aumisart(zero, proj, g, spec.angles, z, z, spec.op_iter,
                                                nonneg=True), reps)
    out["static64_recon_s"] = {"median": round(t, 3), "runs": [round(x, 3) for x in ts]}
    # 128^3 motion-compensated: one run of the full schedule
    spec = specs["mcsart128"]
    g = spec.geometry
    d = make_scan_inputs(spec)
    log("[bench] CPU protocol: 128^3 projections (FP64 oracle, warped phantom)")
    proj = np.stack([oracle.projection(oracle.deform(d.phantom, d.alpha[i], d.beta[i], d.fields),
                                       g, spec.angles[i]) for i in range(spec.n_views)])
    log("[bench] CPU protocol: 128^3 reconstruction, 10 outer iterations (one run)")
    # sweep by sweep (the driver's own composition: an image sweep leaves the image, the motion
    # sweep after it starts from that image and updates the six fields), with progress lines
    t0 = time.perf_counter()
    img = np.zeros(g.vol_shape)
    F = [f.astype(np.float64) for f in d.fields]
    for k, op in enumerate(spec.op_iter):
        if op == 1:
            img = oracle.aumisart(img, proj, g, spec.angles, d.alpha, d.beta, [1], fields=F,
                                  nonneg=True)[0]
        else:
            F, _ = oracle.motion_sweep(img, proj, g, spec.angles, d.alpha, d.beta, F)
        log(f"[bench] CPU protocol: 128^3 sweep {k + 1}/{len(spec.op_iter)} at "
            f"{time.perf_counter() - t0:.1f} s")
    out["mcsart128_recon_s"] = round(time.perf_counter() - t0, 2)
    # 256x256x128: extrapolated
    spec = specs["clinical256"]
    g = spec.geometry
    d = make_scan_inputs(spec)
    proj = np.stack([oracle.projection(d.phantom, g, spec.angles[i]) for i in range(3)])
    log("[bench] CPU protocol: 256x256x128 per-view sweep times")
    t_img, t_mot = cpu_view_times(spec, d, proj, reps)
    n_img, n_mot = launches_per_scan(spec)
    out["clinical256_recon_s_extrapolated"] = {
        "value": round(n_img * t_img + n_mot * t_mot, 1), "image_view_s": round(t_img, 4),
        "motion_view_s": round(t_mot, 4),
        "note": f"{n_img} image-sweep + {n_mot} motion-sweep views at the sampled per-view times"}
    return out


def batch_specs(n_jobs, outer):
    """The 16-scan uneven batch (volumes 128^3 .. 320x320x160, 200-900 views, detectors
    256x192 .. 512x384) with `outer` outer iterations, noise-free data (the kernels' timing does
    not depend on the noise)."""
    from mcsart_amd import configs
    specs = configs.c4_batch(n_jobs, n_outer=outer)
    for s in specs:
        s.noise_photons = None
    return specs


def make_inputs(specs):
    from mcsart_amd.configs import make_scan_inputs
    from mcsart_amd.simulate import measure
    out = []
    for k, s in enumerate(specs):
        d = make_scan_inputs(s)
        out.append((s, d, measure(s, d, s.seed + 404)))
        log(f"[bench] inputs {k + 1}/{len(specs)}: {s.name} {s.nx}x{s.ny}x{s.nz} det "
            f"{s.na}x{s.nb} V={s.n_views}")
    return out


def _submit(srv, sc):
    s, d, p = sc
    return srv.submit(s.geometry, s.angles, d.alpha, d.beta, s.op_iter, p, fields=d.fields,
                      lam=s.lam, motion=s.motion, nonneg=s.nonneg)


def _server_stats(infos, span_s):
    tot = np.array([i["total_ms"] for i in infos])
    return {"scans": len(infos), "makespan_s": round(span_s, 3),
            "scans_per_hour": round(3600.0 * len(infos) / span_s, 1),
            "latency_ms": {"p50": round(float(np.percentile(tot, 50)), 1),
                           "p99": round(float(np.percentile(tot, 99)), 1)}}


def batch_leg(inputs, slots):
    """All scans submitted at t = 0 to the server with 1 slot (plain one-at-a-time processing)
    and with `slots` slots; scans/hour of each, their ratio, and the largest difference between
    the two runs' images."""
    from mcsart_amd.serve import Server
    runs, imgs = {}, {}
    for n in (1, slots):
        with Server(slots=n) as srv:
            t0 = time.perf_counter()
            tickets = [_submit(srv, sc) for sc in inputs]
            srv.drain()
            span = time.perf_counter() - t0
            infos = [srv.info(t) for t in tickets]
            assert all(i["state"] == "done" for i in infos), infos
            imgs[n] = [t.image for t in tickets]
        runs[n] = _server_stats(infos, span)
        log(f"[bench] batch, {n} slot(s): {runs[n]}")
    diff = max(float(np.abs(a - b).max()) for a, b in zip(imgs[1], imgs[slots]))
    return {"plain": runs[1], "batched": dict(runs[slots], slots=slots),
            "scans_per_hour_ratio": round(runs[slots]["scans_per_hour"] /
                                          runs[1]["scans_per_hour"], 3),
            "max_abs_image_difference": diff}


def served_leg(inputs, rate, slots, seed):
    """A Poisson arrival stream (`rate` per second) of the given scans, each submitted at its
    arrival time, to the server with 1 slot (plain) and with `slots` slots: p50/p99 time from
    submission to the reconstruction in the request's buffers, and scans/hour."""
    from mcsart_amd.serve import Server
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(inputs)))
    out = {"rate_per_s": rate, "requests": len(inputs)}
    for label, n in (("plain", 1), ("batched", slots)):
        with Server(slots=n) as srv:
            t0 = time.perf_counter()
            tickets = []
            for a, sc in zip(arrivals, inputs):
                dt = a - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                tickets.append(_submit(srv, sc))
            srv.drain()
            span = time.perf_counter() - t0
            infos = [srv.info(t) for t in tickets]
            assert all(i["state"] == "done" for i in infos), infos
        out[label] = dict(_server_stats(infos, span), slots=n)
        log(f"[bench] served at {rate}/s, {n} slot(s): {out[label]}")
    out["scans_per_hour_ratio"] = round(out["batched"]["scans_per_hour"] /
                                        out["plain"]["scans_per_hour"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="clinical256", choices=["mcsart128", "clinical256"],
                    help="clinical256: 256x256x128, 656 views, 20 outer iterations (the headline "
                         "single-scan workload); mcsart128: 128^3, 200 views, 10 outer iterations")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-protocol", action="store_true",
                    help="run only the full CPU-baseline protocol (per-operator times at 64^3, "
                         "128^3 and 256x256x128, whole reconstructions) and print its JSON")
    ap.add_argument("--batch-outer", type=int, default=2,
                    help="outer iterations of the 16-scan uneven batch leg (0: skip the batch "
                         "and served legs)")
    ap.add_argument("--batch-slots", type=int, default=4)
    ap.add_argument("--serve-rates", default="1,4",
                    help="arrivals per second of the served leg, comma-separated (empty: skip)")
    ap.add_argument("--concurrent", type=int, default=0,
                    help="also time this many copies of the scan in flight at once (0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(scans are independent: more GPUs would run more scans side by side)",
              file=sys.stderr)
        sys.exit(2)
    if args.cpu_protocol:
        print(json.dumps(cpu_protocol()), flush=True)
        return

    from mcsart_amd import configs
    from mcsart_amd.configs import make_scan_inputs
    from mcsart_amd.driver import Context, Job, view_voxels
    from mcsart_amd.simulate import measure

    spec = {"mcsart128": configs.c2_mcsart128, "clinical256": configs.c3_clinical}[args.workload]()
    g = spec.geometry
    log(f"[bench] workload {spec.name}: {g.nx}x{g.ny}x{g.nz}, det {g.na}x{g.nb}, V={spec.n_views}, "
        f"op_iter={spec.op_iter}")
    data = make_scan_inputs(spec)
    ctx = Context(0)
    proj = measure(spec, data, spec.seed + 404, log=log)
    log("[bench] measured projections ready")

    job = Job(ctx, g, spec.angles, data.alpha, data.beta, spec.op_iter, lam=spec.lam,
              motion=spec.motion, nonneg=spec.nonneg)
    job.set_projections(proj)
    job.set_image(None)
    job.set_fields(data.fields)

    # Launch shapes of the two footprint kernels, timed on a few of this scan's views.
    tuning = job.autotune()
    log(f"[bench] autotuned launch shapes: {tuning}")

    # The measurement: each sweep replays its captured HIP graph; the warm-up captures them.
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
    elapsed = time.perf_counter() - t0
    K = args.steps
    img = job.image()
    finite = float(np.isfinite(img).mean())

    # One scan launched kernel by kernel with HIP events on the job's stream around every
    # launch (events cannot be timed inside a graph): per-kernel-kind totals.
    ctx.enable_timing(True)
    ctx.reset_timing()
    t2 = time.perf_counter()
    job.reset()
    job.run()
    ctx.synchronize()
    eager_ms = (time.perf_counter() - t2) * 1e3
    log(f"[bench] timed eager scan: {eager_ms:.1f} ms")
    times = {k: ctx.kernel_time(k) for k in ("projector", "backprojector", "warp", "motion")}
    ctx.enable_timing(False)
    # Rooflines: every view of the scan launched back to back between two HIP events on the
    # job's stream (no per-launch event overhead), for each kernel of the sweeps.  The dominant
    # kernel is the one with the largest share of a scan (average launch x launches per scan).
    layouts = job.view_layouts()
    voxels = view_voxels(spec.geometry, spec.angles)   # sizes of the culled sets per view
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
        nbytes = algorithmic_bytes(spec, byte_kind[k], layouts, voxels)
        ach = nbytes / (us * 1e-6) / 1e9
        kernels_roofline[k] = {
            "avg_launch_us": round(us, 2), "bytes_per_launch": int(nbytes),
            # the same launch's bytes if it processed every voxel (the culled kernels: warp,
            # warp-add, motion step); frac and achieved count only the voxels it processes
            "bytes_per_launch_uncut": int(algorithmic_bytes(spec, byte_kind[k], layouts, None)),
            "achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
            "launches_per_scan": per_scan[k],
            "ms_per_scan": round(us * per_scan[k] * 1e-3, 2),
            "traffic": profiled(spec.name, k),
            "valu_busy": profiled(spec.name, k, "valu_busy"),
            "l2_hit_rate": profiled(spec.name, k, "l2_hit_rate")}
    dom = max(kernels_roofline, key=lambda k: kernels_roofline[k]["ms_per_scan"])
    d = kernels_roofline[dom]
    traffic = d["traffic"]
    src = f"profiles/{spec.name}_kernels.json"
    roofline = {
        "bound": "hbm", "kernel": dom, "achieved": d["achieved"], "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": d["frac"], "traffic": traffic,
        "avg_launch_us": d["avg_launch_us"], "bytes_per_launch": d["bytes_per_launch"],
        "bytes_per_launch_uncut": d["bytes_per_launch_uncut"],
        "bytes_note": ("bytes_per_launch counts the voxels the culled launch processes (achieved "
                       "and frac use it); bytes_per_launch_uncut, every voxel, is what the kernel "
                       "moved before culling, not a bound on this launch"),
        "timing": "HIP events around all views launched back to back, 3 repetitions",
        "traffic_source": (f"{src} (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated on a 64 MiB "
                           "copy)") if traffic else None,
        "valu_busy": d["valu_busy"], "l2_hit_rate": d["l2_hit_rate"],
        "counter_source": f"{src} (rocprofv3 SQ and TCC counters)" if d["valu_busy"] else None,
    }
    kernel_ms = {k: {"total_ms": round(v[0], 2), "launches": v[1],
                     "avg_us": round(v[0] * 1e3 / max(v[1], 1), 2)} for k, v in times.items()}
    job.close()

    # Several copies of the scan in flight at once, each a resident job on its own stream.
    conc = None
    if args.concurrent > 1:
        conc = concurrent_scans(ctx, spec, g, data, proj, tuning, args.concurrent, 1)
        log(f"[bench] {args.concurrent} scans in flight: {conc}")

    # The drop-in boundary (host_AUMISART: host buffers in and out, device memory allocated and
    # freed per call): one whole scan through it, PCIe transfers included.  Not the headline.
    host_ms = host_boundary_ms(spec, g, data, proj)
    log(f"[bench] host-buffer driver call: {host_ms:.1f} ms per scan")

    # The 16-scan uneven batch through the server (plain = one slot) and a served Poisson stream
    # of the same scans plus one large scan (512x512x256, 900 views, 1024x768 detector).
    batch = served = None
    if args.batch_outer > 0:
        bspecs = batch_specs(16, args.batch_outer)
        log(f"[bench] batch leg: making inputs of {len(bspecs)} scans")
        binputs = make_inputs(bspecs)
        batch = batch_leg(binputs, args.batch_slots)
        batch["mix"] = (f"16 seeded scans, volumes 128^3..320x320x160, 200-900 views, detectors "
                        f"256x192..512x384, [1,0]x{args.batch_outer}, noise-free data, all "
                        f"submitted at t = 0")
        rates = [float(r) for r in args.serve_rates.split(",") if r.strip()]
        if rates:
            large = configs.scaled(configs.c5_large(), op_iter=[1, 0] * args.batch_outer)
            large.noise_photons = None
            log("[bench] served leg: making the large scan's inputs")
            linputs = make_inputs([large])
            stream = binputs[:8] + linputs + binputs[8:]
            served = {
                "stream": (f"the 16 batch scans with one 512x512x256 scan (900 views, 1024x768 "
                           f"detector) after the eighth, [1,0]x{args.batch_outer}, Poisson "
                           f"arrivals; latency = submission to the full schedule's "
                           f"reconstruction in the request's buffers"),
                "runs": [served_leg(stream, r, args.batch_slots, 20261203 + 5099) for r in rates]}

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] timing the FP64 CPU oracle on a bounded sample ...")
        cpu = cpu_baseline(spec, data, proj)
        log(f"[bench] CPU baseline: {cpu['value']:.3f} scans/hour")

    vol_name = f"{g.nx}^3" if g.nx == g.ny == g.nz else f"{g.nx}x{g.ny}x{g.nz}"
    lat_ms = np.array(lat) * 1e3
    value = 3600.0 * K / elapsed
    # The reference loop's unfused algorithmic bytes per scan over the measured scan time.  The
    # fused kernels move fewer bytes than that, so this is an equivalent bandwidth, not a
    # fraction of the HBM peak.
    img_b, mot_b = step_bytes(spec)
    scan_bytes = n_img_views * img_b + n_mot_views * mot_b
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
                   "lambda": spec.lam, "nonneg": spec.nonneg, "jobs_in_flight": 1},
        "scan_wall_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 2),
                         "p99": round(float(np.percentile(lat_ms, 99)), 2),
                         "note": "back-to-back single scans, inputs resident"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batch": batch,
        "served": served,
        "kernels_roofline": kernels_roofline,
        "kernels_eager_scan": kernel_ms,
        "eager_ms_per_scan_with_events": round(eager_ms, 2),
        "step_equivalent_bandwidth": {
            "image_view_bytes": int(img_b), "motion_view_bytes": int(mot_b),
            "bytes_per_scan": int(scan_bytes),
            "GBps": round(scan_bytes / (elapsed / K) / 1e9, 1),
            "note": "the reference loop's unfused algorithmic bytes (image view 84N+28D, motion "
                    "view 128N+44D) over the scan time; not a fraction of peak"},
        "host_boundary_ms_per_scan": round(host_ms, 1),
        "concurrent": conc,
        "launch": "one HIP graph per sweep",
        "culling": {
            "on": True,
            "note": "warp, warp-add and motion step skip the voxels no later kernel sees; "
                    "results identical bit for bit (mcsart_job_set_culling)",
            "mean_fraction_of_voxels": {
                "read_set": round(float(np.mean(voxels[:, 0])) / g.nvox, 4),
                "read_set_with_shell": round(float(np.mean(voxels[:, 1])) / g.nvox, 4),
                "u_nonzero_possible": round(float(np.mean(voxels[:, 2])) / g.nvox, 4)}},
        "tuning": tuning,
        "output_finite_fraction": finite,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
