#!/bin/bash
# Regenerates the rocprofv3 summaries in profiles/ (run from the repository root
# on an MI355X; raw output goes to $PROF_OUT, default ./prof_raw).
#   1. kernel trace + stats of a short bench run (volume-render frame)
#   2. FETCH_SIZE and WRITE_SIZE passes (one counter group per run) with a
#      256 MiB device copy as the calibration kernel
set -o pipefail
R=$(pwd)
OUT=${PROF_OUT:-$R/prof_raw}
case $OUT in /*) ;; *) OUT=$R/$OUT ;; esac
mkdir -p $OUT
cd /tmp && export TMPDIR=/tmp
BENCH="python3 $R/bench.py --steps 5 --warmup 1 --no-cpu --no-extra --no-batched --no-served --trace-config 16,256,32 --volume-layout 4 --copy-calibration"
echo "== kernel trace"
timeout -s KILL 300 rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/ktrace -o run -- $BENCH > $OUT/ktrace.log 2>&1 || exit 1
echo "== FETCH_SIZE"
timeout -s KILL 120 rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/fetch -o run -- $BENCH > $OUT/fetch.log 2>&1 || exit 1
echo "== WRITE_SIZE"
timeout -s KILL 120 rocprofv3 --pmc WRITE_SIZE --output-format csv -d $OUT/write -o run -- $BENCH > $OUT/write.log 2>&1 || exit 1
echo "== SQ counters"
timeout -s KILL 120 rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_INSTS_VMEM_RD SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_SALU --output-format csv -d $OUT/sq -o run -- $BENCH > $OUT/sq.log 2>&1 || exit 1
echo "== TCC"
timeout -s KILL 120 rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum --output-format csv -d $OUT/tcc -o run -- $BENCH > $OUT/tcc.log 2>&1 || exit 1
cd $R && python3 profile_summary.py $OUT profiles
