// Split type-1 path, FP64: the term kernel and the kinetic rows.
#include "sg4_split.h"

namespace sg4 {

__global__ __launch_bounds__(NT) void k_hpsi_t1s_f64(KArgs a, KMats<double> km) { split_items(a, km); }

// Kinetic part of the split path: hpsi(r, v) = sum_e K(r, col_e) psi(col_e, v)
// (CSR rows, the diagonal first); one thread per (row, vector), plain stores,
// so it also initialises the result that the term kernel then adds to.
__global__ void k_kin_rows(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                           const double* __restrict__ val, const double* __restrict__ psi,
                           double* __restrict__ hpsi, int64_t nrows, int nvec) {
  const int64_t n = nrows * nvec;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx / nvec;
    const int v = (int)(idx - r * nvec);
    const int64_t e1 = ptr[r + 1];
    double s0 = 0.0, s1 = 0.0;
    int64_t e = ptr[r];
    for (; e + 1 < e1; e += 2) {
      s0 = fma(val[e], psi[(int64_t)col[e] * nvec + v], s0);
      s1 = fma(val[e + 1], psi[(int64_t)col[e + 1] * nvec + v], s1);
    }
    if (e < e1) s0 = fma(val[e], psi[(int64_t)col[e] * nvec + v], s0);
    hpsi[idx] = s0 + s1;
  }
}

}  // namespace sg4
