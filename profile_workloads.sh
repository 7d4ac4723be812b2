#!/bin/bash
# Kernel-trace summaries of the mesh frames, the batched form and the served
# form (run from the repository root on an MI355X; raw output goes to
# $PROF_OUT, default ./prof_raw; summaries to profiles/).  Launch settings are
# fixed to the values the tuners chose, so no tuning runs are profiled.
set -o pipefail
R=$(pwd)
OUT=${PROF_OUT:-$R/prof_raw}
case $OUT in /*) ;; *) OUT=$R/$OUT ;; esac
mkdir -p $OUT
cd /tmp && export TMPDIR=/tmp
FIX="--trace-config 16,256,32 --volume-layout 4 --geometry-stack 32 --batch-config 256,0,0,4,8"
MESH="python3 $R/bench.py --steps 5 --warmup 1 --no-cpu --no-batched --no-served $FIX"
BATCH="python3 $R/bench.py --steps 2 --warmup 1 --no-cpu --no-extra --no-served --batch-steps 2 $FIX"
SERVE="python3 $R/bench.py --steps 2 --warmup 1 --no-cpu --no-extra --no-batched --serve-requests 40 $FIX"
echo "== mesh frames"
timeout -s KILL 300 rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/mesh -o run -- $MESH > $OUT/mesh.log 2>&1 || exit 1
echo "== batched"
timeout -s KILL 400 rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/batched -o run -- $BATCH > $OUT/batched.log 2>&1 || exit 1
GEOM="python3 $R/bench.py --steps 2 --warmup 1 --no-cpu --no-batched --no-served --copy-calibration $FIX"
SQ="SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_INSTS_VMEM_RD SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_SALU"
echo "== mesh frames: SQ counters"
timeout -s KILL 300 rocprofv3 --pmc $SQ --output-format csv -d $OUT/mesh_sq -o run -- $GEOM > $OUT/mesh_sq.log 2>&1 || exit 1
echo "== mesh frames: FETCH_SIZE"
timeout -s KILL 300 rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/mesh_fetch -o run -- $GEOM > $OUT/mesh_fetch.log 2>&1 || exit 1
echo "== mesh frames: TCC"
timeout -s KILL 300 rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum --output-format csv -d $OUT/mesh_tcc -o run -- $GEOM > $OUT/mesh_tcc.log 2>&1 || exit 1
echo "== served"
timeout -s KILL 400 rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/served -o run -- $SERVE > $OUT/served.log 2>&1 || exit 1
cd $R
python3 profile_summary.py --stats $OUT/mesh profiles/mesh_frames_kernels.md "Kernel trace, sphere-mesh, nine-icosphere and path-line frames (with the volume frame)" "${MESH/$R\//}"
python3 profile_summary.py --stats $OUT/batched profiles/batched_lists_kernels.md "Kernel trace, batched form (256 lists, 64 datasets) and its per-list loop" "${BATCH/$R\//}"
python3 profile_summary.py --counters $OUT profiles/mesh_frames_counters.md "Counters of the geometry trace kernels (sphere-mesh, nine-icosphere and path-line frames)" "${GEOM/$R\//}"
python3 profile_summary.py --stats $OUT/served profiles/served_stream_kernels.md "Kernel trace, served form (40 requests, continuous and one at a time)" "${SERVE/$R\//}"
