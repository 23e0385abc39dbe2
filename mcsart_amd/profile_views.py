"""Per-view kernel durations of whole sweeps, from a rocprofv3 kernel trace of profile_sweep.

    python -m mcsart_amd.profile_views --workload clinical256 --trace <dir> --md <file>

<dir> holds the run_kernel_trace.csv that

    rocprofv3 --kernel-trace --output-format csv -d <dir> -o run -- \\
        python3 -m mcsart_amd.profile_sweep --workload W --sweeps 1,0 --views 0 ...

writes (mcsart_amd/profile_views.sh runs both steps).  profile_sweep launches kernel by kernel
with fixed launch shapes, so each kernel kind's launches come in view order: the image sweep's
warp, projector, backprojector and warp-add once per view 0 .. V-1, the motion sweep's projector
and motion step once per view 1 .. V-1.  Every launch is grouped by the layouts the rays it
serves read (mcsart_view_voxels, last column: the view's own for the warp and the projectors,
the next view's for the motion step, which prepares that view's g) and by the view angle modulo
180 degrees in 15-degree bins.
"""
import argparse
import csv
import glob
import os
import re

import numpy as np

from . import configs
from .driver import view_voxels

WORKLOADS = {"mcsart128": configs.c2_mcsart128, "clinical256": configs.c3_clinical,
             "static64": configs.c1_static64}

# kernel kind: (name pattern, first view, layouts of the view itself (0) or of the next (1))
KINDS = {
    "warp": (r"k_warp_tile<", 0, 0),
    "projector": (r"k_project_rows<1,", 0, 0),
    "backprojector": (r"k_backproject<1, \d+, \w+, 0>", 0, 0),
    "warp_add": (r"k_warp_add_tile<", 0, 0),
    "projector_motion": (r"k_project_rows<2,", 1, 0),
    "motion": (r"k_backproject<1, \d+, \w+, 1>", 1, 1),
}
LAYOUTS = {1: "natural only", 2: "transposed only", 3: "both"}


def kernel_name(full: str) -> str:
    n = full.replace("void ", "").replace("mcsart::(anonymous namespace)::", "")
    return n.replace("mcsart::", "").split("(")[0].strip()


def read_trace(trace_dir: str):
    """(kernel name, microseconds) of every dispatch, in dispatch order."""
    files = glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no kernel trace under {trace_dir}")
    rows = list(csv.DictReader(open(files[0])))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    return [(kernel_name(r["Kernel_Name"]),
             (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3) for r in rows]


def summarise(spec, dispatches):
    g = spec.geometry
    ang = spec.angles
    V = len(ang)
    layouts = view_voxels(g, ang)[:, 4]
    deg = np.degrees(ang.astype(np.float64)) % 180.0
    out = {}
    for kind, (pat, first, nxt) in KINDS.items():
        t = np.array([us for name, us in dispatches if re.match(pat, name)])
        n = V - first
        if len(t) < n:
            continue
        t = t[:n]   # one sweep of this kind
        views = np.arange(first, V)
        lay = layouts[np.minimum(views + nxt, V - 1)]
        by_layout = {LAYOUTS[k]: (float(t[lay == k].mean()), int((lay == k).sum()))
                     for k in (1, 2, 3) if (lay == k).any()}
        bins = [float(t[(deg[views] >= lo) & (deg[views] < lo + 15)].mean())
                for lo in range(0, 180, 15)]
        out[kind] = {"launches": n, "mean_us": float(t.mean()), "by_layout": by_layout,
                     "by_angle_15deg": bins}
    return out


def write_md(spec, res, path, commands):
    lines = [f"# Per-view kernel durations: {spec.name} (one image sweep, one motion sweep, "
             f"{spec.n_views} views)", "", "Commands (on the MI355X box, from the repository root):",
             ""]
    lines += [f"    {c}" for c in commands]
    lines += ["", "Kernel-by-kernel launches (rocprofv3 --kernel-trace), microseconds per launch.  "
              "Layouts: what the rays of the view a launch serves read (the next view's for the "
              "motion step, which prepares its g).  Angle bins: the view angle modulo 180 "
              "degrees.", "",
              "| kernel | launches | mean | natural only | transposed only | both |",
              "|---|---:|---:|---:|---:|---:|"]
    for kind, r in res.items():
        cells = []
        for name in LAYOUTS.values():
            v = r["by_layout"].get(name)
            cells.append(f"{v[0]:.1f} ({v[1]})" if v else "—")
        lines.append(f"| {kind} | {r['launches']} | {r['mean_us']:.1f} | " + " | ".join(cells) +
                     " |")
    lines += ["", "| kernel | " + " | ".join(f"{lo}-{lo + 15}°" for lo in range(0, 180, 15)) + " |",
              "|---|" + "---:|" * 12]
    for kind, r in res.items():
        lines.append(f"| {kind} | " + " | ".join(f"{b:.1f}" for b in r["by_angle_15deg"]) + " |")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="clinical256", choices=sorted(WORKLOADS))
    ap.add_argument("--trace", required=True, help="directory of the rocprofv3 kernel trace")
    ap.add_argument("--md", required=True)
    ap.add_argument("--command", action="append", default=[])
    args = ap.parse_args(argv)
    spec = WORKLOADS[args.workload]()
    res = summarise(spec, read_trace(args.trace))
    write_md(spec, res, args.md, args.command)
    print(f"[profile_views] wrote {args.md}")


if __name__ == "__main__":
    main()
