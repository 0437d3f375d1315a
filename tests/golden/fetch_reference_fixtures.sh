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
# initial wave packets (version-0 lines, 'end wp0' closed) held inline in the
# exa_hcn test inputs
H=$REF/Working_tests/exa_hcn-test
sed -n '71,74p' "$H/dat_2DSG_type4"            > "$OUT/psi_v0_2d_dat_2DSG_type4.txt"
sed -n '80,87p' "$H/dat_Davidson"              > "$OUT/psi_v0_3d_dat_Davidson.txt"
sed -n '67,70p' "$H/dat_Davidson_1D_WithGrid"  > "$OUT/psi_v0_1d_dat_Davidson_1D_WithGrid.txt"
sed -n '68,77p' "$H/dat_Davidson_WithGrid"     > "$OUT/psi_v0_3d_dat_Davidson_WithGrid.txt"
sed -n '65,68p' "$H/dat_Davidson_AutoContract" > "$OUT/psi_v0_2d_dat_Davidson_AutoContract.txt"
# 12D pyrazine two-surface propagation at L=1 (run_jobs: shell_run 3 1):
# autocorrelation T(au), Re C, Im C, |C| every 0.1 fs up to 6 fs
cp "$REF/Working_tests/MPI_tests/12D_propagation_S3/benchmark" "$OUT/pyrazine12d_L1_autocorr_benchmark.txt"
# 24D pyrazine at L=2 (run_jobs: shell_run 3 2; the S0 variant stores the same values)
cp "$REF/Working_tests/MPI_tests/24D_propagation_S3/benchmark" "$OUT/pyrazine24d_L2_autocorr_benchmark.txt"
