#!/bin/sh
# AddressSanitizer + UndefinedBehaviorSanitizer run of the host code: the
# host half of libdsim (formats, host substrate builds, schedule, tensor fit,
# writer) and the oracle, each driven over the reference's sample inputs.
# Host code only, built with g++/gcc (no device code is sanitized).
#   sh tests/sanitize/run.sh [out_dir]
set -e
here=$(cd "$(dirname "$0")" && pwd)
root=$(cd "$here/../.." && pwd)
out=${1:-$(mktemp -d)}
mkdir -p "$out"
san="-fsanitize=address,undefined -fno-omit-frame-pointer -fno-sanitize-recover=undefined -g -O1"
inc="-I$root/include -I/opt/rocm/include -D__HIP_PLATFORM_AMD__"
for f in host sim formats serve phase; do
    g++ -std=c++17 -fPIC -ffp-contract=off -pthread $san $inc -c "$root/dsim_amd/csrc/$f.cpp" -o "$out/$f.o"
done
g++ -std=c++17 $san $inc -c "$here/host_driver.cpp" -o "$out/host_driver.o"
# the device half links from the library's own (unsanitized) HIP objects
hipo="$root/dsim_amd/_build"
[ -f "$hipo/kernels.hip.o" ] || (cd "$root" && python -m dsim_amd.build)
g++ $san -o "$out/host_driver" "$out"/*.o "$hipo/kernels.hip.o" "$hipo/grid.hip.o" \
    "$hipo/nalloc.hip.o" -L/opt/rocm/lib -lamdhip64 -Wl,-rpath,/opt/rocm/lib -pthread
gcc -std=c11 $san -ffp-contract=off -I"$root/oracle" -o "$out/oracle_driver" \
    "$here/oracle_driver.c" "$root/oracle/dsim_oracle.c" -lm
export ASAN_OPTIONS=detect_leaks=1:verify_asan_link_order=0 UBSAN_OPTIONS=print_stacktrace=1
HIP_VISIBLE_DEVICES=-1 "$out/host_driver" "$root/tests/golden/reference_inputs" "$out"
"$out/oracle_driver" "$root/tests/golden/reference_inputs"
