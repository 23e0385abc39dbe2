"""Summarise rocprofv3 output of ``mcsart_amd.profile_sweep`` into a per-kernel table.

    python -m mcsart_amd.profile_summary --workload mcsart128 \\
        --trace prof/trace_mcsart128 --fetch prof/pmc_FETCH_SIZE_mcsart128 \\
        --write prof/pmc_WRITE_SIZE_mcsart128 --md profiles/x.md --json profiles/x.json

Inputs are rocprofv3 CSV directories: a ``--kernel-trace --stats`` run and one ``--pmc`` run per
counter (FETCH_SIZE, WRITE_SIZE; KiB per dispatch).  The counters are calibrated against the last
device-to-device copy of the run (profile_sweep ends with one of --copy-mb MiB): factor =
true bytes / counted bytes for reads and for writes.  For every kernel the table gives launches,
average duration, algorithmic bytes per launch (the minimum HBM traffic of its job, 4-byte
floats, N voxels, D detector cells), achieved GB/s = algorithmic bytes / duration, the fraction
of the 8 TB/s HBM peak, and the calibrated counter traffic per launch.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import os
import re

from . import configs

HBM_PEAK_GBS = 8000.0
WORKLOADS = {"mcsart128": configs.c2_mcsart128, "clinical256": configs.c3_clinical,
             "static64": configs.c1_static64}


def short(name: str) -> str:
    n = name.replace("mcsart::(anonymous namespace)::", "").replace("mcsart::", "")
    n = n.replace("void ", "")
    return re.sub(r"\(.*", "", n).strip()


def algorithmic_bytes(kernel: str, N: int, D: int, culled=None):
    """Minimum bytes one launch must move (reads + writes, 4-byte floats).  `culled`: the
    per-launch averages of the culled kernels over the profiled sweeps (culled_bytes), which
    replace the uncut figures of the warps and the motion step."""
    if culled:
        if re.match(r"k_backproject<1, \d+, \w+, 1>", kernel):
            return culled["motion"]
        base = re.sub(r"<(true|false)(, \d+)?>$", "", kernel)
        if base in ("k_warp_tile", "k_warp_add_tile"):
            return culled[base]
    m = re.match(r"k_project(?:_rows)?<(\d)", kernel)
    if m:
        mode = int(m.group(1))
        return {0: 4 * N + 4 * D, 1: 4 * N + 12 * D, 2: 8 * N + 12 * D}[mode]
    m = re.match(r"k_backproject<(\d), \d+, \w+, (\d)>", kernel)
    if m:
        if m.group(2) == "1":   # ratio mode with the motion step: one prep layout
            return 32 * N + 4 * N + 8 * D
        if m.group(2) == "2":   # ratio mode with the static add: x read and written, one xT
            return 12 * N + 8 * D
        return 4 * N + (4 * D if m.group(1) == "0" else 8 * D)
    table = {
        "k_warp_tile": 32 * N,        # 6 fields + x read, warped image written
        "k_warp_add_tile": 36 * N,    # 6 fields + u + x read, x written
        "k_motion_step<true, true>": 72 * N,    # 6 fields read+written, u, x, pairs
        "k_motion_step<true, false>": 56 * N,
        "k_motion_step<false, true>": 40 * N,
        "k_motion_begin": 56 * N,               # fields, x read; (Ga, Gb, H, x), Q x2, g written
        "k_motion_finish": 60 * N,              # fields read+written, accumulators, x
        "k_transpose": 8 * N,
    }
    return table.get(re.sub(r"<(true|false)(, \d+)?>$", "", kernel))   # warps: <FAST, ROWS>


def load_stats(d):
    out = {}
    for r in csv.DictReader(open(os.path.join(d, "run_kernel_stats.csv"))):
        out[short(r["Name"])] = (int(r["Calls"]), float(r["AverageNs"]) / 1e3)
    return out


def load_pmc(d):
    per = collections.defaultdict(list)
    copies = []
    for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
        k = short(r["Kernel_Name"])
        v = float(r["Counter_Value"]) * 1024.0   # KiB -> bytes
        per[k].append(v)
        if "copyBuffer" in k:
            copies.append((int(r["Dispatch_Id"]), v))
    calib = max(copies)[1] if copies else None   # the run's last copy is the calibration copy
    return {k: sum(v) / len(v) for k, v in per.items()}, calib


SIMDS = 1024          # 256 CUs x 4 SIMDs
SQ_INSTANCES = 32     # SQ_BUSY_CYCLES is summed over the 32 shader engines
# SIMD cycles per wave64 VALU instruction.  Measured, not assumed: counting 4 (a 16-lane SIMD),
# the instructions counted for the 256x256x128 motion projector and image backprojector would
# need 1.37 and 1.05 times the SIMD cycles their launches had, so gfx950 retires its common
# 32-bit VALU instructions in 2; the figure is an upper-bound estimate of issue occupancy.
VALU_CYCLES = 2


def load_sq(d):
    """Per kernel, the average per dispatch of every SQ counter in a --pmc run."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
        per[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in per.items()}


def valu_busy(c):
    """Fraction of the SIMDs' cycles spent issuing VALU instructions while the kernel ran."""
    if "SQ_INSTS_VALU" not in c or not c.get("SQ_BUSY_CYCLES"):
        return None
    return VALU_CYCLES * c["SQ_INSTS_VALU"] / (SIMDS * c["SQ_BUSY_CYCLES"] / SQ_INSTANCES)


def load_tcc(d):
    """Per kernel, the L2 (TCC) hits and misses per dispatch from a --pmc run with TCC_HIT_sum and
    TCC_MISS_sum."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
        per[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = {}
    for k, cs in per.items():
        hit = cs.get("TCC_HIT_sum") or cs.get("TCC_HIT")
        miss = cs.get("TCC_MISS_sum") or cs.get("TCC_MISS")
        if hit and miss:
            out[k] = (sum(hit) / len(hit), sum(miss) / len(miss))
    return out


def culled_bytes(spec):
    """Per-launch algorithmic bytes of the culled kernels averaged over the profiled sweeps (an
    image sweep over views 0..V-1, a motion sweep over views 1..V-1), from the views' set sizes
    (mcsart_view_voxels; mcsart_amd/traffic.py)."""
    from . import traffic
    from .driver import view_voxels
    g = spec.geometry
    vox = view_voxels(g, spec.angles)
    return {"k_warp_tile": traffic.warp_bytes(vox),
            "k_warp_add_tile": traffic.warp_add_bytes(vox, g.nvox),
            "motion": traffic.motion_sweep_bytes(vox, g.nvox, g.ndet)}


def summarise(workload, trace, fetch, write, copy_mb=64, views=0, sq=None, tcc=None,
              culled=True):
    spec = WORKLOADS[workload]()
    if views:
        spec = configs.scaled(spec, n_views=views)
    g = spec.geometry
    N, D = g.nvox, g.ndet
    cb = culled_bytes(spec) if culled else None
    stats = load_stats(trace)
    fe, fcal = load_pmc(fetch) if fetch else ({}, None)
    wr, wcal = load_pmc(write) if write else ({}, None)
    sqc = load_sq(sq) if sq else {}
    tc = load_tcc(tcc) if tcc else {}
    true_bytes = copy_mb << 20
    ffac = true_bytes / fcal if fcal else None
    wfac = true_bytes / wcal if wcal else None
    rows = []
    for k, (calls, us) in sorted(stats.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        if k.startswith("__amd"):
            continue
        alg = algorithmic_bytes(k, N, D, cb)
        row = {"kernel": k, "launches": calls, "avg_us": round(us, 2),
               "algorithmic_bytes": alg}
        if alg:
            gbs = alg / (us * 1e-6) / 1e9
            row.update({"achieved_gbs": round(gbs, 1), "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)})
        if k in fe and ffac:
            row["fetch_bytes"] = int(fe[k] * ffac)
        if k in wr and wfac:
            row["write_bytes"] = int(wr[k] * wfac)
        if k in sqc:
            vb = valu_busy(sqc[k])
            if vb is not None:
                row["valu_busy"] = round(vb, 3)
            waves = sqc[k].get("SQ_WAVES")
            if waves and "SQ_INSTS_VALU" in sqc[k]:
                row["valu_per_wave"] = round(sqc[k]["SQ_INSTS_VALU"] / waves, 1)
        if k in tc:
            hit, miss = tc[k]
            row["l2_hits"], row["l2_misses"] = int(hit), int(miss)
            row["l2_hit_rate"] = round(hit / max(hit + miss, 1.0), 3)
        if "fetch_bytes" in row and "write_bytes" in row:
            row["traffic_bytes"] = row["fetch_bytes"] + row["write_bytes"]
            if alg:
                row["traffic_over_algorithmic"] = round(row["traffic_bytes"] / alg, 2)
        rows.append(row)
    return {"workload": workload, "N": N, "D": D, "views": views or spec.n_views,
            "culled": bool(culled),
            "fetch_factor": ffac and round(ffac, 4), "write_factor": wfac and round(wfac, 4),
            "kernels": rows}


def to_markdown(s, command_lines):
    out = [f"# Kernel profile: {s['workload']} (N = {s['N']} voxels, D = {s['D']} detector cells)",
           "", "Commands (on the MI355X box, from /tmp, PYTHONPATH = repository root):", ""]
    out += ["    " + c for c in command_lines]
    out += ["", f"Counter calibration against the run's final 64 MiB device copy: FETCH_SIZE x "
            f"{s['fetch_factor']}, WRITE_SIZE x {s['write_factor']}.  VALU busy = "
            "2 x SQ_INSTS_VALU / (1024 SIMDs x SQ_BUSY_CYCLES / 32): the share of the SIMDs' "
            "cycles spent issuing vector instructions (2 cycles per wave64 instruction: with 4, "
            "some kernels would exceed 1).  L2 hit rate = TCC_HIT_sum / (TCC_HIT_sum + "
            "TCC_MISS_sum) per dispatch." + (
                "  The warps and the motion step skip the voxels no later kernel sees "
                "(culling): their algorithmic bytes count the voxels they process, averaged over "
                "the profiled views (mcsart_amd/traffic.py)." if s.get("culled") else ""), "",
            "| kernel | launches | avg us | algorithmic MB | achieved GB/s | frac of 8 TB/s | "
            "fetch MB | write MB | traffic / algorithmic | VALU busy | L2 hit rate |",
            "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for r in s["kernels"]:
        mb = lambda b: "" if b is None else f"{b / 1e6:.2f}"
        out.append(f"| `{r['kernel']}` | {r['launches']} | {r['avg_us']} | "
                   f"{mb(r.get('algorithmic_bytes'))} | {r.get('achieved_gbs', '')} | "
                   f"{r.get('frac_hbm_peak', '')} | {mb(r.get('fetch_bytes'))} | "
                   f"{mb(r.get('write_bytes'))} | {r.get('traffic_over_algorithmic', '')} | "
                   f"{r.get('valu_busy', '')} | {r.get('l2_hit_rate', '')} |")
    return "\n".join(out) + "\n"


def stats_markdown(trace, title, command_lines, notes=()):
    """Plain rocprofv3 --stats table (every kernel of a run, by total time)."""
    rows = list(csv.DictReader(open(os.path.join(trace, "run_kernel_stats.csv"))))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = [f"# {title}", "", "Commands (on the MI355X box, from /tmp, PYTHONPATH = repository root):",
           ""]
    out += ["    " + c for c in command_lines]
    out += [""] + list(notes) + ["" if notes else None]
    out = [o for o in out if o is not None]
    out += ["| kernel | calls | avg us | min us | max us | total ms | % of kernel time |",
            "|---|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                   f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.1f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out) + "\n"


def lds_markdown(pmc, title, command_lines, notes=()):
    """Per-kernel LDS and issue shares from one rocprofv3 --pmc pass with SQ_LDS_BANK_CONFLICT,
    SQ_LDS_IDX_ACTIVE, SQ_INSTS_LDS, SQ_WAIT_INST_LDS, SQ_ACTIVE_INST_VALU, SQ_WAVE_CYCLES and
    SQ_WAIT_INST_ANY (per-dispatch means)."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(os.path.join(pmc, "run_counter_collection.csv"))):
        per[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = [f"# {title}", "", "Command (on the MI355X box, from /tmp, PYTHONPATH = repository root):",
           ""]
    out += ["    " + c for c in command_lines]
    out += ["", "Per-dispatch means over the run.  Bank-conflict ratio = SQ_LDS_BANK_CONFLICT / "
            "(SQ_LDS_IDX_ACTIVE - SQ_LDS_BANK_CONFLICT) (the tool's own derived metric); the "
            "wait and active columns are per-wave cycle shares (SQ_WAIT_INST_LDS, "
            "SQ_WAIT_INST_ANY, SQ_ACTIVE_INST_VALU over SQ_WAVE_CYCLES).", ""]
    out += ["| kernel | launches | LDS instructions | bank-conflict ratio | LDS wait share | "
            "any-wait share | VALU-active share |", "|---|---:|---:|---:|---:|---:|---:|"]
    for k, v in per.items():
        m = {c: sum(x) / len(x) for c, x in v.items()}
        if "SQ_WAVE_CYCLES" not in m or "rocclr" in k or not m["SQ_WAVE_CYCLES"]:
            continue
        bc, ia, wc = m["SQ_LDS_BANK_CONFLICT"], m["SQ_LDS_IDX_ACTIVE"], m["SQ_WAVE_CYCLES"]
        out.append(f"| `{k}` | {len(v['SQ_WAVE_CYCLES'])} | {m['SQ_INSTS_LDS']:.0f} | "
                   f"{bc / max(ia - bc, 1.0):.3f} | {m['SQ_WAIT_INST_LDS'] / wc:.3f} | "
                   f"{m['SQ_WAIT_INST_ANY'] / wc:.3f} | {m['SQ_ACTIVE_INST_VALU'] / wc:.3f} |")
    out += [""] + list(notes)
    return "\n".join(out) + "\n"


def occupancy_markdown(pmc, title, command_lines, notes=()):
    """Mean waves resident per CU (rocprofv3's MeanOccupancyPerCU) per kernel and grid size, from
    one --pmc pass: the grid size tells the launches of different scan sizes apart."""
    per = collections.defaultdict(list)
    for r in csv.DictReader(open(os.path.join(pmc, "run_counter_collection.csv"))):
        if r["Counter_Name"] != "MeanOccupancyPerCU":
            continue
        k = short(r["Kernel_Name"])
        if "rocclr" in k:
            continue
        per[(k, int(r["Grid_Size"]), int(r["Workgroup_Size"]))].append(float(r["Counter_Value"]))
    out = [f"# {title}", "", "Commands (on the MI355X box, from /tmp, PYTHONPATH = repository "
           "root):", ""]
    out += ["    " + c for c in command_lines]
    out += ["", "Per-dispatch means of MeanOccupancyPerCU (waves resident per CU, averaged over the "
            "dispatch) by kernel and grid size (work-items per launch).", ""]
    out += ["| kernel | grid | workgroup | wavefronts per launch | launches | waves per CU |",
            "|---|---:|---:|---:|---:|---:|"]
    for (k, grid, wg), v in sorted(per.items()):
        out.append(f"| `{k}` | {grid} | {wg} | {grid // 64} | {len(v)} | "
                   f"{sum(v) / len(v):.1f} |")
    out += [""] + list(notes)
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", help="write only the --stats table of this trace directory")
    ap.add_argument("--lds", help="write only the LDS/issue table of this --pmc directory")
    ap.add_argument("--occupancy", help="write only the occupancy table of this --pmc directory "
                                        "(MeanOccupancyPerCU)")
    ap.add_argument("--title", default="rocprofv3 kernel statistics")
    ap.add_argument("--note", action="append", default=[])
    ap.add_argument("--workload", choices=sorted(WORKLOADS))
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--sq", help="rocprofv3 --pmc run with SQ_INSTS_VALU, SQ_BUSY_CYCLES, SQ_WAVES")
    ap.add_argument("--tcc", help="rocprofv3 --pmc run with TCC_HIT_sum, TCC_MISS_sum")
    ap.add_argument("--views", type=int, default=0)
    ap.add_argument("--md")
    ap.add_argument("--json")
    ap.add_argument("--command", action="append", default=[],
                    help="command line that produced the inputs (repeatable)")
    args = ap.parse_args(argv)
    if args.lds or args.occupancy:
        md = (lds_markdown(args.lds, args.title, args.command, args.note) if args.lds else
              occupancy_markdown(args.occupancy, args.title, args.command, args.note))
        if args.md:
            with open(args.md, "w") as f:
                f.write(md)
        else:
            print(md)
        return
    if args.stats:
        md = stats_markdown(args.stats, args.title, args.command, args.note)
        if args.md:
            with open(args.md, "w") as f:
                f.write(md)
        else:
            print(md)
        return
    if not args.workload or not args.trace:
        ap.error("--workload and --trace are required without --stats")
    s = summarise(args.workload, args.trace, args.fetch, args.write, views=args.views, sq=args.sq,
                  tcc=args.tcc)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(s, f, indent=1)
    md = to_markdown(s, args.command)
    if args.md:
        with open(args.md, "w") as f:
            f.write(md)
    else:
        print(md)


if __name__ == "__main__":
    main()
