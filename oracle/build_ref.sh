#!/usr/bin/env bash
# Builds oracle/_ref/libgpis_ref_leaf.so from the reference's own leaf sources
# (hash, PCG sampler, TangentFrame, Box-Muller in Gaussian.cpp, the SDF shapes of
# SdfFunctionsQueries.cpp with Mat4f.cpp, the simplex fbm of SdfFunctions.cpp) where they
# lie under /root/reference, plus
# oracle/ref_leaf.cpp which exposes them via a C ABI.
# The reference realization (SparseConvolutionNoise.cpp) is not buildable here:
# GPFunctions.hpp needs Boost and GPFunctions.cpp needs FFTW, both absent.
# CONSTEXPR=constexpr is the definition the reference's CMakeLists.txt:54 passes.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
ref="${GPIS_REFERENCE_ROOT:-/root/reference}/src"
if [ ! -d "$ref/core" ]; then
    echo "reference tree not present; skipping oracle/_ref build" >&2
    exit 0
fi
mkdir -p "$here/_ref"
g++ -std=c++17 -O2 -fPIC -shared -DCONSTEXPR=constexpr -w \
    -I"$ref/core" -I"$ref/thirdparty" -I"$ref/thirdparty/eigen" \
    "$here/ref_leaf.cpp" "$ref/core/sampling/Gaussian.cpp" \
    "$ref/core/math/SdfFunctionsQueries.cpp" "$ref/core/math/SdfFunctions.cpp" "$ref/core/math/Mat4f.cpp" \
    -o "$here/_ref/libgpis_ref_leaf.so"
echo "built $here/_ref/libgpis_ref_leaf.so"
