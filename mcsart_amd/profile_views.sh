#!/bin/bash
# Per-view kernel durations of one image sweep and one motion sweep over every view of a
# workload, reproducible from this script alone.
#
#     bash mcsart_amd/profile_views.sh <workload> <tuning> [out-dir]
#
# <tuning>: fixed launch shapes proj_split,back_zslab,back_stage,proj_rows,back_zslab_motion (the
# benchmark autotuner's picks).  The measured projections are made first, outside the profiler,
# then a rocprofv3 kernel trace of the two sweeps (kernel by kernel, every view), and
# mcsart_amd.profile_views groups the launches by the layouts their view's rays read and by angle
# into <out-dir>/<workload>_views.md.  Run from the repository root on the GPU box.
set -o pipefail
W=$1
T=$2
REPO=$(pwd)
OUT=${3:-prof/views_$W}
export PYTHONPATH=$REPO
export TMPDIR=/tmp
mkdir -p "$REPO/$OUT"
PROJ=/tmp/mcsart_views_proj_$W.npy
CMDS=()
run() {
    echo "[profile_views] $*" >&2
    CMDS+=("${*//$REPO/\$REPO}")
    (cd /tmp && timeout -s KILL 300 "$@") || { rc=$?; echo "[profile_views] failed ($rc): $*" >&2; exit $rc; }
}
run python3 -m mcsart_amd.profile_sweep --workload $W --views 0 --save-proj $PROJ
run rocprofv3 --kernel-trace --output-format csv -d "$REPO/$OUT" -o run -- python3 -m mcsart_amd.profile_sweep --workload $W --sweeps 1,0 --views 0 --tuning $T --copy-mb 0 --load-proj $PROJ
ARGS=(--command "PYTHONPATH=\$REPO TMPDIR=/tmp bash mcsart_amd/profile_views.sh $W $T $OUT")
for c in "${CMDS[@]}"; do ARGS+=(--command "(cd /tmp && $c)"); done
python3 -m mcsart_amd.profile_views --workload "$W" --trace "$OUT" --md "$OUT/${W}_views.md" \
    "${ARGS[@]}"
# the per-dispatch trace is large; the summary holds what it uses
find "$REPO/$OUT" -name '*kernel_trace.csv' -delete
