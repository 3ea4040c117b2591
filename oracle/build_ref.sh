#!/bin/bash
# Builds oracle/_ref/libref_flux.so from the reference's own flux headers,
# read in place from the reference checkout (never copied into this tree).
# g++ 11 rejects static_assert(false) in discarded constexpr branches of
# source/common_functions.hpp, so the build uses ROCm's clang++ as host
# compiler.  No FMA contraction, as in the reference's x86-64 build.
set -euo pipefail
REF=${PAMHD_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
if [ ! -d "$REF/source/mhd" ]; then
	echo "reference checkout not found at $REF; skipping oracle/_ref" >&2
	exit 0
fi
CXX=${CXX_REF:-/opt/rocm/llvm/bin/clang++}
mkdir -p "$HERE/_ref"
"$CXX" -std=c++20 -O2 -ffp-contract=off -fno-fast-math -fPIC -shared \
	-I"$REF/source" -o "$HERE/_ref/libref_flux.so" "$HERE/ref_flux_harness.cpp"
echo "built $HERE/_ref/libref_flux.so"
