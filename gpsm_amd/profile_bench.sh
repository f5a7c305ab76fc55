#!/bin/bash
# rocprofv3 runs of bench.py on one MI355X whose summaries go to profiles/
# (gpsm_amd/profile_summary.py): a kernel trace with statistics, FETCH_SIZE
# and WRITE_SIZE of the bench and of its 1 GiB copy calibration, and SQ /
# GRBM counters.  Raw output goes to $1 (default prof_out).
#
#   bash gpsm_amd/profile_bench.sh prof_out
#   python3 -m gpsm_amd.profile_summary --prof prof_out --out profiles
set -u
R=$(cd "$(dirname "$0")/.." && pwd)
O=$(cd "$R" && mkdir -p "${1:-prof_out}" && cd "${1:-prof_out}" && pwd)
cd /tmp && export TMPDIR=/tmp
B="--steps 2 --warmup 1 --no-cpu --no-plain --no-autotune"
echo "step: kernel trace"
timeout -k 10 300 rocprofv3 --kernel-trace --stats -d $O/trace -o trace --output-format csv -- python3 $R/bench.py $B > $O/trace.log 2>&1 || exit $?
echo "step: fetch"
timeout -s KILL 240 rocprofv3 --pmc FETCH_SIZE -d $O/fetch -o fetch --output-format csv -- python3 $R/bench.py $B > $O/fetch.log 2>&1 || exit $?
echo "step: write"
timeout -s KILL 240 rocprofv3 --pmc WRITE_SIZE -d $O/write -o write --output-format csv -- python3 $R/bench.py $B > $O/write.log 2>&1 || exit $?
echo "step: copy fetch"
timeout -s KILL 120 rocprofv3 --pmc FETCH_SIZE -d $O/cfetch -o cfetch --output-format csv -- python3 $R/bench.py --copy-calibration > $O/cfetch.log 2>&1 || exit $?
echo "step: copy write"
timeout -s KILL 120 rocprofv3 --pmc WRITE_SIZE -d $O/cwrite -o cwrite --output-format csv -- python3 $R/bench.py --copy-calibration > $O/cwrite.log 2>&1 || exit $?
echo "step: sq"
timeout -s KILL 240 rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_INSTS_VMEM_RD SQ_BUSY_CYCLES SQ_WAVE_CYCLES SQ_INSTS_SMEM GRBM_GUI_ACTIVE -d $O/sq -o sq --output-format csv -- python3 $R/bench.py $B > $O/sq.log 2>&1 || exit $?
echo "done"
