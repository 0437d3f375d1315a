#!/bin/bash
# Copies the reference's own data fixtures for the SG4 path into tests/golden/.
# Run in the build container (the reference tree is not present elsewhere).
#   - Gauss-Hermite nodes/weights (50 digits): Internal_data/HermQuadra/herm{n}.txt
#   - Henon-Heiles Davidson eigenvalues: Working_tests/MPI_tests/{6D,21D}_Davidson_S0/benchmark
set -euo pipefail
REF=${1:-/root/reference}
OUT=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$OUT/hermquadra"
for n in $(seq 1 40); do cp "$REF/Internal_data/HermQuadra/herm$n.txt" "$OUT/hermquadra/"; done
cp "$REF/Working_tests/MPI_tests/6D_Davidson_S0/benchmark" "$OUT/hh6d_L3_benchmark.txt"
cp "$REF/Working_tests/MPI_tests/21D_Davidson_S0/benchmark" "$OUT/hh21d_L2_benchmark.txt"
