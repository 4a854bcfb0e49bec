#!/bin/bash
# Host-side AddressSanitizer check of the library on a GPU machine.
#   tests/cpp/host_asan_gpu.sh build   (where hipcc is; no GPU needed)
#   tests/cpp/host_asan_gpu.sh run     (on the machine with the GPU)
# "build" compiles the library with its host code instrumented (device code as shipped: the
# sanitizer flag follows -Xarch_host) and links tests/cpp/abi_drive.cpp and the C++ drop-in test
# against it, with clang's shared sanitizer runtime; everything goes to tests/cpp/_asan/.
set -e
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
C=$ROOT/hnsw_rabitq_amd/csrc
O=$ROOT/tests/cpp/_asan
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
# HG_SANITIZE=address,undefined adds UndefinedBehaviorSanitizer to the host code.  Its function
# and vptr checks stay off: with them the host-side kernel stubs launch nothing (every search
# returns an untouched buffer, no error), an interaction of that instrumentation with the HIP
# launch path; the library runs clean under every other check.
SAN=${HG_SANITIZE:-address}
NOSAN=-fno-sanitize=function,vptr
CLANG=${CLANG:-/opt/rocm/llvm/bin/clang++}
RT=$(dirname "$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)")
if [ "$1" = "build" ]; then
    mkdir -p "$O"
    HF="--offload-arch=gfx950 -O2 -g -std=c++17 -fPIC -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops"
    HF="$HF -Wno-unused-result -Wno-unused-command-line-argument -Xarch_host -fsanitize=$SAN -Xarch_host $NOSAN -Xarch_host -fno-omit-frame-pointer"
    pids=()
    for s in search.hip build.hip mvsearch.hip bruteforce.hip mlp.hip; do
        $HIPCC $HF -c "$C/$s" -o "$O/$s.o" & pids+=($!)
    done
    for s in capi.cpp capi_bf.cpp capi_stop.cpp; do
        $HIPCC $HF -x hip -c "$C/$s" -o "$O/$s.o" & pids+=($!)
    done
    $CLANG -O2 -g -std=c++17 -fPIC -ffp-contract=off -march=x86-64-v3 -pthread -fsanitize=$SAN $NOSAN \
        -fno-omit-frame-pointer -c "$C/host/hnsw_index.cpp" -o "$O/hnsw_index.o" & pids+=($!)
    for p in "${pids[@]}"; do wait "$p"; done
    $HIPCC --offload-arch=gfx950 -shared -fPIC -o "$O/libhnswgpu.so" "$O"/*.o -pthread -Xarch_host -fsanitize=$SAN \
        -shared-libasan
    for t in abi_drive closer_first_test; do
        $CLANG -O1 -g -std=c++17 -fsanitize=$SAN $NOSAN -shared-libasan -fno-omit-frame-pointer -I"$ROOT/include" \
            "$ROOT/tests/cpp/$t.cpp" -o "$O/$t" -L"$O" -lhnswgpu -Wl,-rpath,'$ORIGIN' -Wl,-rpath,"$RT" -pthread
    done
    echo "built $O"
elif [ "$1" = "run" ]; then
    # the GPU runtime maps memory in the sanitizer's shadow gap; its own allocations are not
    # checked for leaks
    export ASAN_OPTIONS=verify_asan_link_order=0:detect_leaks=0:protect_shadow_gap=0
    export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
    "$O/closer_first_test"
    "$O/abi_drive"
else
    echo "usage: $0 build|run" >&2
    exit 2
fi
