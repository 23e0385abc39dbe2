#!/bin/bash
# Kernel profile of one workload, reproducible from this script alone.
#
#     bash mcsart_amd/profile_workload.sh <workload> <tuning> <views> [out-dir]
#
# <workload>: mcsart128 | clinical256 | static64 (mcsart_amd/profile_summary.py WORKLOADS);
# <tuning>: fixed launch shapes proj_split,back_zslab,back_stage,proj_rows,back_zslab_motion (the
# shapes the benchmark's autotuner picks for the workload), so every pass runs the same kernels;
# <views>: views of the scan to run (0 = all).  Run from the repository root on the GPU box.
#
# The scan's measured projections are made first, outside the profiler (making them launches
# warps and projections of their own), into /tmp.  Then five runs of mcsart_amd.profile_sweep on
# them (one image sweep and one motion sweep, kernel by kernel, then a 64 MiB device copy that
# calibrates the byte counters): a --kernel-trace --stats run and four --pmc passes (FETCH_SIZE;
# WRITE_SIZE; SQ_INSTS_VALU SQ_BUSY_CYCLES SQ_WAVES; TCC_HIT_sum TCC_MISS_sum), each within the
# per-block counter limits and under its own kill timeout; then mcsart_amd.profile_summary
# writes profiles/<workload>_kernels.md and .json with every command.
set -o pipefail
W=$1
T=$2
V=$3
REPO=$(pwd)
OUT=${4:-prof/prof_$W}
export PYTHONPATH=$REPO
export TMPDIR=/tmp
mkdir -p "$REPO/$OUT"
PROJ=/tmp/mcsart_profile_proj_$W.npy
SWEEP="python3 -m mcsart_amd.profile_sweep --workload $W --sweeps 1,0 --views $V --tuning $T --copy-mb 64 --load-proj $PROJ"
CMDS=()
run() {
    echo "[profile_workload] $*" >&2
    CMDS+=("${*//$REPO/\$REPO}")
    (cd /tmp && timeout -s KILL 240 "$@") || { rc=$?; echo "[profile_workload] failed ($rc): $*" >&2; exit $rc; }
}
run python3 -m mcsart_amd.profile_sweep --workload $W --views $V --save-proj $PROJ
run rocprofv3 --kernel-trace --stats --output-format csv -d "$REPO/$OUT/trace" -o run -- $SWEEP
run rocprofv3 --pmc FETCH_SIZE --output-format csv -d "$REPO/$OUT/fetch" -o run -- $SWEEP
run rocprofv3 --pmc WRITE_SIZE --output-format csv -d "$REPO/$OUT/write" -o run -- $SWEEP
run rocprofv3 --pmc SQ_INSTS_VALU SQ_BUSY_CYCLES SQ_WAVES --output-format csv -d "$REPO/$OUT/sq" -o run -- $SWEEP
run rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum --output-format csv -d "$REPO/$OUT/tcc" -o run -- $SWEEP
ARGS=()
for c in "${CMDS[@]}"; do ARGS+=(--command "(cd /tmp && $c)"); done
ARGS+=(--command "python3 -m mcsart_amd.profile_summary --workload $W --views $V --trace $OUT/trace --fetch $OUT/fetch --write $OUT/write --sq $OUT/sq --tcc $OUT/tcc --md profiles/${W}_kernels.md --json profiles/${W}_kernels.json")
python3 -m mcsart_amd.profile_summary --workload "$W" --views "$V" --trace "$OUT/trace" \
    --fetch "$OUT/fetch" --write "$OUT/write" --sq "$OUT/sq" --tcc "$OUT/tcc" \
    --md "$OUT/${W}_kernels.md" --json "$OUT/${W}_kernels.json" \
    --command "PYTHONPATH=\$REPO TMPDIR=/tmp bash mcsart_amd/profile_workload.sh $W $T $V $OUT" \
    "${ARGS[@]}"
# the per-dispatch trace is large and the summary uses only the statistics
rm -f "$REPO/$OUT/trace/run_kernel_trace.csv"
