#!/bin/bash
# The CPU tests (-m "not gpu") with the oracle's C, the synthetic generator and the
# host C++ of libkyto_gpu.so built with AddressSanitizer and UndefinedBehaviorSanitizer
# (clang's shared runtime, preloaded into the Python process). Any report fails the run.
set -euo pipefail
cd "$(dirname "$0")/.."
python3 -m kyto_amd.build sanitize
KYTO_AMD_SANITIZED=1 make -s -C oracle sanitize
RT=$(ls /opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)
export KYTO_AMD_SANITIZED=1
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:halt_on_error=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
LD_PRELOAD="$RT" python3 -m pytest tests -m "not gpu" -x -q -p no:cacheprovider "$@"
