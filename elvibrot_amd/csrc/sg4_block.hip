// Dense products of the device block Davidson over packed vectors: blocks of
// a few to ~100 rows of length n = nb_ba (1.26 M at 10D L=8), so every
// product is a pass over HBM with a tiny output.
//
//   gram:    C (m x q) = A B^T, A: m rows, B: q rows (stride lda, ldb), the
//            inner dimension n long.  Each workgroup takes 32-column tiles in
//            a grid-stride loop, stages the rows of a tile in LDS (sized for
//            this m and q, so that small grams keep several workgroups per
//            CU and their tile loads in flight) and keeps 4 x 4 register
//            blocks of C: with fewer blocks than threads the tile's columns
//            are dealt to slices of threads, summed in LDS at the end; the
//            workgroup partials go to a work array and a second kernel adds
//            them, one wave per output, in a fixed order (deterministic, no
//            atomics).
//   combine: Y (k x n) = beta Y + U X, U (k x d) small, X: d rows; one
//            column per thread, coalesced across the wave, U entries from
//            scalar loads (wave-uniform addresses) as FMA operands.
// Measured at 10D L=8 (n = 1.26 M): gram 8 x 8 0.065 ms, 16 x 20 0.16 ms,
// 20 x 20 0.19 ms, 4 x 84 0.32 ms (HBM floor 0.02, 0.05, 0.05, 0.11 ms; with
// LDS for 124 rows at two workgroups per CU and no column slices: 0.22, 0.28 ms
// for 8 x 8 and 20 x 20); combine 8 x 8 0.031 ms, 16 x 48 0.16 ms, 20 x 84
// 0.34 ms, 36 x 84 0.69 ms (U broadcast from LDS instead: 0.05, 0.20, 0.48,
// 0.83 ms).  The Davidson driver uses them for the small grams and the
// rotations to at most 20 rows, and the library GEMM for the others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "sg4_internal.h"

namespace sg4 {
namespace {

constexpr int GT = 32;         // columns per gram tile
constexpr int GMAX = 124;      // rows of A or B (LDS: 32 x (4 ceil(rows/4) + 1) doubles each)
constexpr int GBLK = 512;      // 4 x 4 output blocks per workgroup (two per thread)
constexpr int GNB = 2048;      // workgroups of the gram's first kernel (at most)
constexpr int CMAXU = 8192;    // entries of U (k d) at most
constexpr int CK = 16;         // outputs of the combine per pass over X (k <= 16)
constexpr int CKW = 48;        // the same for wider k (k > 48: several passes)

// column slices of the gram's workgroup for nblk 4 x 4 output blocks
__host__ __device__ constexpr int gram_slices(int nblk) {
  return nblk >= 256 ? 1 : (256 / nblk < GT ? 256 / nblk : GT);
}

__global__ __launch_bounds__(256) void k_gram_part(int m, int q, int64_t n, const double* __restrict__ A,
                                                   int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                   double* __restrict__ part) {
  // As[c pa + i]: column c of the tile, row i (rows padded to a multiple of 4
  // with zeros; pitch pa = 4 m4 + 1), Bs likewise with pitch pb = 4 q4 + 1
  extern __shared__ double gl[];
  const int m4 = (m + 3) >> 2, q4 = (q + 3) >> 2, nblk = m4 * q4;
  const int pa = 4 * m4 + 1, pb = 4 * q4 + 1;
  double* As = gl;
  double* Bs = gl + GT * pa;
  const int tid = threadIdx.x;
  // with fewer than 256 output blocks the columns of a tile are dealt to ns
  // slices of threads (block tid % nblk, slice tid / nblk: columns slice,
  // slice + ns, ...), whose sums are added in LDS at the end in slice order;
  // otherwise a thread keeps output blocks tid and tid + 256 over all columns
  const int ns = gram_slices(nblk);
  const int b0 = ns > 1 ? tid % nblk : tid, b1 = tid + 256;
  const int sl = ns > 1 ? tid / nblk : 0;
  const bool on0 = ns > 1 ? tid < ns * nblk : b0 < nblk;
  double acc0[16], acc1[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) { acc0[e] = 0.0; acc1[e] = 0.0; }
  const int64_t ntile = (n + GT - 1) / GT;
  for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
    const int64_t c0 = t * GT;
    __syncthreads();                            // previous tile consumed
    for (int e = tid; e < GT * 4 * m4; e += 256) {
      const int i = e / GT, c = e % GT;         // consecutive threads: consecutive columns
      As[c * pa + i] = (i < m && c0 + c < n) ? A[(int64_t)i * lda + c0 + c] : 0.0;
    }
    for (int e = tid; e < GT * 4 * q4; e += 256) {
      const int i = e / GT, c = e % GT;
      Bs[c * pb + i] = (i < q && c0 + c < n) ? B[(int64_t)i * ldb + c0 + c] : 0.0;
    }
    __syncthreads();
    if (on0) {
      const int bi = (b0 / q4) * 4, bj = (b0 % q4) * 4;
      for (int c = sl; c < GT; c += ns) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { a[u] = As[c * pa + bi + u]; b[u] = Bs[c * pb + bj + u]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc0[4 * u + v] = fma(a[u], b[v], acc0[4 * u + v]);
      }
    }
    if (ns == 1 && b1 < nblk) {
      const int bi = (b1 / q4) * 4, bj = (b1 % q4) * 4;
      for (int c = 0; c < GT; ++c) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { a[u] = As[c * pa + bi + u]; b[u] = Bs[c * pb + bj + u]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc1[4 * u + v] = fma(a[u], b[v], acc1[4 * u + v]);
      }
    }
  }
  // partial of output e = i q + j at part[e * nb + workgroup]
  const int64_t nb = gridDim.x;
  if (ns > 1) {
    __syncthreads();                            // the last tile consumed: LDS reused
    double* red = gl;                           // [slice][block][16]
    if (on0)
#pragma unroll
      for (int e = 0; e < 16; ++e) red[(sl * nblk + b0) * 16 + e] = acc0[e];
    __syncthreads();
    for (int o = tid; o < nblk * 16; o += 256) {
      const int b = o >> 4, e = o & 15;
      double sum = 0.0;
      for (int k = 0; k < ns; ++k) sum += red[(k * nblk + b) * 16 + e];
      const int i = (b / q4) * 4 + (e >> 2), j = (b % q4) * 4 + (e & 3);
      if (i < m && j < q) part[((int64_t)i * q + j) * nb + blockIdx.x] = sum;
    }
    return;
  }
  if (b0 < nblk) {
    const int bi = (b0 / q4) * 4, bj = (b0 % q4) * 4;
    for (int u = 0; u < 4; ++u)
      for (int v = 0; v < 4; ++v)
        if (bi + u < m && bj + v < q) part[((int64_t)(bi + u) * q + bj + v) * nb + blockIdx.x] = acc0[4 * u + v];
  }
  if (b1 < nblk) {
    const int bi = (b1 / q4) * 4, bj = (b1 % q4) * 4;
    for (int u = 0; u < 4; ++u)
      for (int v = 0; v < 4; ++v)
        if (bi + u < m && bj + v < q) part[((int64_t)(bi + u) * q + bj + v) * nb + blockIdx.x] = acc1[4 * u + v];
  }
}

// C[e] = sum over the nb workgroup partials of output e: one wave per output,
// lane l adding partials l, l + 64, ..., then a fixed shuffle tree (the same
// order on every call)
__global__ __launch_bounds__(256) void k_gram_sum(int mq, int nb, const double* __restrict__ part,
                                                  double* __restrict__ C) {
  const int lane = threadIdx.x & 63;
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= mq) return;                          // uniform over the wave
  double s = 0.0;
  for (int b = lane; b < nb; b += 64) s += part[(int64_t)e * nb + b];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (lane == 0) C[e] = s;
}

// Y rows i0..i0+CK-1 for the columns of this thread; U entries are read
// straight from global memory at wave-uniform addresses (scalar loads, one
// SGPR operand per FMA), X once per pass.
template <int CK>
__global__ __launch_bounds__(256) void k_combine(int k, int d, int64_t n, const double* __restrict__ U,
                                                 const double* __restrict__ X, int64_t ldx, double beta,
                                                 double* __restrict__ Y, int64_t ldy) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    for (int i0 = 0; i0 < k; i0 += CK) {
      double acc[CK];
#pragma unroll
      for (int u = 0; u < CK; ++u) acc[u] = 0.0;
      int j = 0;
      for (; j + 4 <= d; j += 4) {               // four loads in flight
        double x[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) x[t] = X[(int64_t)(j + t) * ldx + r];
#pragma unroll
        for (int u = 0; u < CK; ++u)
          if (i0 + u < k) {
            const double* ur = U + (int64_t)(i0 + u) * d + j;
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[u] = fma(ur[t], x[t], acc[u]);
          }
      }
      for (; j < d; ++j) {
        const double x = X[(int64_t)j * ldx + r];
#pragma unroll
        for (int u = 0; u < CK; ++u)
          if (i0 + u < k) acc[u] = fma(U[(int64_t)(i0 + u) * d + j], x, acc[u]);
      }
#pragma unroll
      for (int u = 0; u < CK; ++u)
        if (i0 + u < k) {
          double* y = Y + (int64_t)(i0 + u) * ldy + r;
          *y = (beta == 0.0) ? acc[u] : fma(beta, *y, acc[u]);
        }
    }
  }
}

#define BK_TRY(expr)                                                              \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr); \
      return SG4_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

}  // namespace
}  // namespace sg4

using namespace sg4;

extern "C" {

int sg4_block_gram_work(int32_t m, int32_t q, int64_t* nwork) {
  if (!nwork) { set_error("null argument"); return SG4_ERR_ARG; }
  if (m < 1 || q < 1 || m > GMAX || q > GMAX || ((m + 3) / 4) * ((q + 3) / 4) > GBLK) {
    set_error("sg4_block_gram: m, q <= 124 with ceil(m/4) ceil(q/4) <= 512 expected");
    return SG4_ERR_ARG;
  }
  *nwork = (int64_t)GNB * m * q;
  return SG4_OK;
}

int sg4_block_gram(sg4_plan* p, int32_t m, int32_t q, int64_t n, const double* A, int64_t lda,
                   const double* B, int64_t ldb, double* C, double* work, int64_t nwork) {
  if (!p || !A || !B || !C || !work) { set_error("null argument"); return SG4_ERR_ARG; }
  int64_t need;
  int rc;
  if ((rc = sg4_block_gram_work(m, q, &need))) return rc;
  if (n < 1 || lda < n || ldb < n) { set_error("sg4_block_gram: bad n or row strides"); return SG4_ERR_ARG; }
  if (nwork < (int64_t)m * q) { set_error("sg4_block_gram: work array too small"); return SG4_ERR_ARG; }
  if ((rc = device_block(p, A, "A")) || (rc = device_block(p, B, "B")) ||
      (rc = device_block(p, C, "C")) || (rc = device_block(p, work, "work")))
    return rc;
  const hipStream_t st = (hipStream_t)plan_stream(p);
  const int64_t ntile = (n + GT - 1) / GT;
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)GNB, ntile, nwork / ((int64_t)m * q)}));
  // LDS: the tile's rows, or the slice sums at the end if larger (<= 32 KiB)
  const int m4 = (m + 3) / 4, q4 = (q + 3) / 4, ns = gram_slices(m4 * q4);
  const size_t tile = sizeof(double) * GT * ((size_t)(4 * m4 + 1) + (4 * q4 + 1));
  const size_t red = ns > 1 ? sizeof(double) * 16 * (size_t)ns * m4 * q4 : 0;
  const size_t lds = std::max(tile, red);
  hipLaunchKernelGGL(k_gram_part, dim3(nb), dim3(256), lds, st, m, q, n, A, lda, B, ldb, work);
  BK_TRY(hipGetLastError());
  const int mq = m * q;
  hipLaunchKernelGGL(k_gram_sum, dim3((mq + 3) / 4), dim3(256), 0, st, mq, nb, work, C);
  BK_TRY(hipGetLastError());
  return SG4_OK;
}

int sg4_block_combine(sg4_plan* p, int32_t k, int32_t d, int64_t n, const double* U, const double* X,
                      int64_t ldx, double beta, double* Y, int64_t ldy) {
  if (!p || !U || !X || !Y) { set_error("null argument"); return SG4_ERR_ARG; }
  if (k < 1 || d < 1 || (int64_t)k * d > CMAXU || n < 1 || ldx < n || ldy < n) {
    set_error("sg4_block_combine: k d <= 8192, n >= 1 and row strides >= n expected");
    return SG4_ERR_ARG;
  }
  int rc;
  if ((rc = device_block(p, U, "U")) || (rc = device_block(p, X, "X")) || (rc = device_block(p, Y, "Y")))
    return rc;
  const char* xb = (const char*)X;
  const char* yb = (const char*)Y;
  const int64_t xe = 8 * ((int64_t)(d - 1) * ldx + n), ye = 8 * ((int64_t)(k - 1) * ldy + n);
  if (xb < yb + ye && yb < xb + xe) { set_error("sg4_block_combine: X and Y overlap"); return SG4_ERR_ARG; }
  const hipStream_t st = (hipStream_t)plan_stream(p);
  const unsigned nb = (unsigned)std::min<int64_t>((n + 255) / 256, 65535);
  if (k <= CK)
    hipLaunchKernelGGL(k_combine<CK>, dim3(nb), dim3(256), 0, st, k, d, n, U, X, ldx, beta, Y, ldy);
  else
    hipLaunchKernelGGL(k_combine<CKW>, dim3(nb), dim3(256), 0, st, k, d, n, U, X, ldx, beta, Y, ldy);
  BK_TRY(hipGetLastError());
  return SG4_OK;
}

}  // extern "C"
