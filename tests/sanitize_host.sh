#!/bin/bash
# Runs the CPU test suite against AddressSanitizer + UndefinedBehaviorSanitizer
# builds of the host-side code: the oracle (oracle/mf_oracle.c), the host
# build of the kernel math headers (tests/native/host_math.cpp) and the host
# half of libmultifloats_hip.so (the C-ABI layer and every launcher; the GPU
# code objects are built without sanitizers). The builds go to a scratch
# folder (first argument, default /tmp/mf_sanitize); the in-tree libraries
# are left alone and selected away through MF_ORACLE_LIB, MF_HOSTMATH_LIB and
# MULTIFLOATS_HIP_LIB. One sanitizer runtime (clang's) serves all three, and
# it is preloaded because the Python interpreter is not instrumented.
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
OUT="${1:-/tmp/mf_sanitize}"
LLVM=/opt/rocm/llvm/bin
HIPCC=/opt/rocm/bin/hipcc
SAN="-fsanitize=address,undefined -fno-sanitize-recover=undefined -shared-libasan"
mkdir -p "$OUT"

$LLVM/clang -O1 -g -std=gnu11 -ffp-contract=off -fno-fast-math -fPIC -shared $SAN \
  -o "$OUT/liboracle.so" "$ROOT/oracle/mf_oracle.c" -lm
$LLVM/clang++ -O1 -g -std=c++17 -ffp-contract=off -fno-fast-math -fPIC -shared -Wno-unknown-pragmas $SAN \
  -o "$OUT/libhostmath.so" "$ROOT/tests/native/host_math.cpp"

# host-side sanitizers only: every -fsanitize flag goes to the host pass
HOSTSAN="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -shared-libasan"
objs=()
for f in gemm_plain gemm_ozaki batched elementwise reduce trsm capi; do
  $HIPCC --offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -ffp-contract=off -fno-fast-math \
    -Xclang -target-feature -Xclang -packed-fp32-ops -fvisibility=hidden $HOSTSAN \
    -c "$ROOT/multifloats_amd/csrc/$f.hip" -o "$OUT/$f.o" &
  objs+=("$OUT/$f.o")
done
wait
$HIPCC --offload-arch=gfx950 -shared -fPIC $HOSTSAN -o "$OUT/libmultifloats_hip.so" "${objs[@]}"

RT="$(ls $LLVM/../lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)"
cd "$ROOT"
MF_ORACLE_LIB="$OUT/liboracle.so" MF_HOSTMATH_LIB="$OUT/libhostmath.so" \
MULTIFLOATS_HIP_LIB="$OUT/libmultifloats_hip.so" \
ASAN_OPTIONS=detect_leaks=0:halt_on_error=1 UBSAN_OPTIONS=print_stacktrace=1 \
LD_PRELOAD="$RT" python -m pytest tests -q -m "not gpu" -p no:cacheprovider
