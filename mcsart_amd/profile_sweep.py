"""Run a bounded piece of a McSART reconstruction for profilers (rocprofv3 kernel traces and
hardware-counter passes).

    python -m mcsart_amd.profile_sweep --workload mcsart128 --sweeps 1,0 --copy-mb 64

Builds the named workload's scan (synthetic inputs, measured projections made on the GPU, or
read from a file that an earlier --save-proj run wrote, so that the profile holds the sweeps'
kernels only),
autotunes the footprint kernels' launch shapes as the benchmark does, then runs the given sweeps
(1 = image sweep, 0 = motion sweep) kernel by kernel (no graphs, so every kernel is its own
dispatch), optionally the isolated launches the benchmark times (--isolated: every view of the
scan back to back for each named kernel, as mcsart_job_kernel_time launches them), and finally
one device-to-device copy of --copy-mb MiB: the profiler's FETCH_SIZE / WRITE_SIZE for that copy
calibrate the counters against a known byte count.
"""
import argparse
import sys
import time

import numpy as np

from . import configs
from .configs import make_scan_inputs
from .driver import Context, Job
from .ops import DeviceArray
from .simulate import measure
from ._lib import check, load

WORKLOADS = {"mcsart128": configs.c2_mcsart128, "clinical256": configs.c3_clinical,
             "static64": configs.c1_static64}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="mcsart128", choices=sorted(WORKLOADS))
    ap.add_argument("--sweeps", default="1,0")
    ap.add_argument("--views", type=int, default=0, help="use only the first V views (0: all)")
    ap.add_argument("--copy-mb", type=int, default=64)
    ap.add_argument("--isolated", default="",
                    help="comma-separated kernels to launch over all views back to back after "
                         "the sweeps: projector, backprojector, warp, warp_add, motion")
    ap.add_argument("--save-proj", default="",
                    help="make the scan's measured projections, save them to this .npy file and "
                         "exit (run outside the profiler: making them launches warps and "
                         "projections of their own)")
    ap.add_argument("--load-proj", default="",
                    help="measured projections from a --save-proj file instead of making them")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default="",
                    help="fixed launch shapes proj_split,back_zslab,back_stage,proj_rows"
                         "[,back_zslab_motion] "
                         "(no autotune: the profile then holds the sweep's kernels only)")
    args = ap.parse_args(argv)
    spec = WORKLOADS[args.workload]()
    if args.views:
        spec = configs.scaled(spec, n_views=args.views)
    sweeps = [int(s) for s in args.sweeps.split(",") if s != ""]
    g = spec.geometry
    data = make_scan_inputs(spec)
    if args.load_proj:
        proj = np.load(args.load_proj)
        assert proj.shape == (spec.n_views,) + g.det_shape, proj.shape
    else:
        proj = measure(spec, data, spec.seed + 404)
    if args.save_proj:
        np.save(args.save_proj, proj)
        print(f"[profile_sweep] measured projections saved to {args.save_proj}", flush=True)
        return 0
    ctx = Context(0)
    job = Job(ctx, g, spec.angles, data.alpha, data.beta, sweeps, lam=spec.lam,
              motion=spec.motion, nonneg=spec.nonneg)
    job.set_graphs(False)
    job.set_projections(proj)
    job.set_image(None)
    job.set_fields(data.fields)
    if args.tuning:
        job.set_tuning(*[int(v) for v in args.tuning.split(",")])
        print(f"[profile_sweep] launch shapes: {job.tuning()}", flush=True)
    elif not args.no_autotune:
        print(f"[profile_sweep] autotuned launch shapes: {job.autotune()}", flush=True)
    t0 = time.perf_counter()
    job.run()
    ctx.synchronize()
    print(f"[profile_sweep] {spec.name}: sweeps {sweeps} over {spec.n_views} views in "
          f"{(time.perf_counter() - t0) * 1e3:.1f} ms", flush=True)
    for kind in [k for k in args.isolated.split(",") if k]:
        us = job.kernel_time(kind, reps=1)
        print(f"[profile_sweep] isolated {kind}: {us:.2f} us per launch", flush=True)
    if args.copy_mb:
        n = args.copy_mb << 20
        a, b = DeviceArray.empty((n // 4,)), DeviceArray.empty((n // 4,))
        check(load().mcsart_memset(a.ptr, 0, n), "mcsart_memset")
        check(load().mcsart_memcpy_dtod(b.ptr, a.ptr, n), "mcsart_memcpy_dtod")
        check(load().mcsart_device_synchronize(), "sync")
        print(f"[profile_sweep] calibration copy: {n} bytes", flush=True)
    job.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
