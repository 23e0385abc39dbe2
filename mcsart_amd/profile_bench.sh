#!/bin/bash
# Kernel trace of bench.py itself, reproducible from this script alone.
#
#     bash mcsart_amd/profile_bench.sh [workload] [out-dir]
#
# rocprofv3 --kernel-trace --stats around one short bench.py run (one warm-up scan, one timed
# scan, the per-kernel isolated timings; no batch, served or CPU legs) with MCSART_GRAPHS=0:
# every kernel is its own dispatch (rocprofv3 crashed on the host while tracing HIP-graph
# replays).  The summary (<out-dir>/bench_<workload>_kernels.md) lists every kernel by total time
# and states, for the bench line's dominant kernel, the average launch time bench.py measured
# with HIP events in the same run beside the trace's average.  Run from the repository root on
# the GPU box.
set -o pipefail
W=${1:-clinical256}
REPO=$(pwd)
OUT=${2:-prof/bench_trace_$W}
export PYTHONPATH=$REPO
export TMPDIR=/tmp
export MCSART_GRAPHS=0
mkdir -p "$REPO/$OUT"
BENCH="python3 $REPO/bench.py --workload $W --steps 1 --warmup 1 --batch-outer 0 --no-cpu-baseline"
CMD="rocprofv3 --kernel-trace --stats --output-format csv -d $REPO/$OUT -o run -- $BENCH"
echo "[profile_bench] $CMD" >&2
(cd /tmp && timeout -s KILL 600 $CMD > "$REPO/$OUT/bench.json") || exit $?
NOTE=$(python3 - "$REPO/$OUT" <<'PY'
import csv, json, os, re, sys
d = sys.argv[1]
b = json.loads(open(os.path.join(d, "bench.json")).read().strip().splitlines()[-1])
r = b["roofline"]
pat = {"motion": r"k_backproject<1, \d+, \w+, 1>", "warp": r"k_warp_tile",
       "warp_add": r"k_warp_add_tile", "projector": r"k_project_rows<1,",
       "projector_motion": r"k_project_rows<2,", "backprojector": r"k_backproject<1, \d+, \w+, 0>"}[r["kernel"]]
rows = [x for x in csv.DictReader(open(os.path.join(d, "run_kernel_stats.csv")))
        if re.search(pat, x["Name"])]
rows.sort(key=lambda x: -int(x["Calls"]))
t = rows[0]
print(f"Dominant kernel of the bench line ({r['kernel']}): HIP events in the same run "
      f"{r['avg_launch_us']} us per launch (all views back to back); this trace "
      f"{float(t['AverageNs']) / 1e3:.2f} us average over {t['Calls']} launches of "
      f"the most-launched instance.")
PY
)
python3 -m mcsart_amd.profile_summary --stats "$OUT" \
    --title "rocprofv3 kernel statistics of bench.py ($W, eager launches)" \
    --command "PYTHONPATH=\$REPO TMPDIR=/tmp MCSART_GRAPHS=0 bash mcsart_amd/profile_bench.sh $W $OUT" \
    --command "(cd /tmp && ${CMD//$REPO/\$REPO})" \
    --note "$NOTE" --md "$OUT/bench_${W}_kernels.md"
# the per-dispatch trace is large and the summary uses only the statistics
rm -f "$REPO/$OUT/run_kernel_trace.csv"
