// Device-side types, constants and templates shared by the H.psi kernels
// (sg4_k_split_*.hip, sg4_k_grid_*.hip) and the plan code (sg4_hpsi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sg4_internal.h"
#include "sg4_shapes.h"

namespace sg4 {

constexpr int MAXACT = 16;
constexpr int NT = 512;                 // maximum threads per workgroup (launch bound)
constexpr int LDS_SMALL = 64 * 1024;    // LDS budget of the common class (2 WG/CU)
constexpr int LDS_MAX = 160 * 1024;     // gfx950 LDS per workgroup
// slab products along modes of >= this many points run row by row (one matrix
// row live in scalar registers; whole 7x7 and larger matrices spilled them)
constexpr int SLAB_ROW_OUTER = 7;

struct MatDev {          // offsets (in doubles) of the 1-D matrices of (k,l)
  int64_t B, BtW, D1, D2;
  int32_t nq, nb;
  double b00, btw00, d200;  // entries used when nq = nb = 1
};

struct TermDev {
  int64_t boff, qoff;
  double wgt;      // W_t times the 1x1 transforms of the inactive modes
  double vshift;   // type 1: sum over inactive modes of c_k d2_k(0,0)
  int32_t nb, nq;  // product sizes
  int32_t nact, pad;
  int16_t mat[MAXACT];   // MatDev index of each active mode
  int16_t kmode[MAXACT]; // index of the mode's kinetic coefficient (ck)
  // register-blocked passes: groups of consecutive active modes
  int32_t ngrp;          // 0: use the one-mode-per-pass path
  int8_t gfirst[MAXACT]; // first active mode of group g
  int8_t gshape[MAXACT]; // shape id (sg4_shapes.h)
};

struct Item {
  int32_t term;
  int16_t v0;
  int8_t lgnv;
  int8_t big;     // second buffer in global scratch (slot) instead of LDS
  int32_t slot;
  int32_t pad;
};

struct KArgs {
  const Item* items;
  const TermDev* terms;
  const MatDev* md;
  const int32_t* map;
  const double* psi;
  double* hpsi;
  double* scratch;
  int64_t slot_size;
  int nvec;
  int nitems;
  // deterministic accumulation (split path): 128-bit fixed-point accumulator
  // per result element (two words) and an overflow counter; nullptr: FP64 atomics
  unsigned long long* fx = nullptr;
  unsigned long long* fxflag = nullptr;
};

// Deterministic scatter-add: c * 2^64 as a 128-bit two's complement integer
// (the magnitude truncated below 2^-64) added to acc[0] (low word) and acc[1]
// (high word) with 64-bit integer atomics; the carry out of the low word is
// the one this addition produced, so the final (high, low) pair is the exact
// integer sum whatever the order of the additions.  |c| >= 2^63 is counted
// in *flag instead (the result is then reported as NaN).
__device__ __forceinline__ void fx_add(unsigned long long* acc, unsigned long long* flag, double c) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(c);
  const int ex = (int)((bits >> 52) & 0x7ff);
  if (ex == 0) return;                       // zero or below 2^-1022
  if (ex == 0x7ff) { atomicAdd(flag, 1ULL); return; }
  const unsigned long long m = (bits & 0xfffffffffffffULL) | (1ULL << 52);
  const int sh = ex - 1075 + 64;             // c 2^64 = m 2^sh
  unsigned long long lo, hi;
  if (sh >= 75) { atomicAdd(flag, 1ULL); return; }
  if (sh >= 64) { hi = m << (sh - 64); lo = 0; }
  else if (sh > 0) { lo = m << sh; hi = m >> (64 - sh); }
  else if (sh > -53) { lo = m >> (-sh); hi = 0; }
  else return;                               // below 2^-64
  if (bits >> 63) {                          // negate (two's complement, 128 bits)
    lo = ~lo + 1ULL;
    hi = ~hi + (lo == 0 ? 1ULL : 0ULL);
  }
  const unsigned long long old = atomicAdd(&acc[0], lo);
  const unsigned long long h = hi + ((old + lo < old) ? 1ULL : 0ULL);
  if (h) atomicAdd(&acc[1], h);
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int fdiv(int n, int d, float inv) {
  int q = (int)((float)n * inv);
  int r = n - q * d;
  while (r < 0) { --q; r += d; }
  while (r >= d) { ++q; r -= d; }
  return q;
}

enum { IN_LDS = 0, IN_GATHER = 1 };
enum { OUT_STORE = 0, OUT_ADD = 1, OUT_VSTORE = 2, OUT_SCATTER = 3 };

template <class T>
using cptr = const __attribute__((address_space(4))) T*;   // scalar (constant) loads

// the value is the same in every lane (it derives from the work item);
// readfirstlane states that so the pointer can live in scalar registers
template <class T>
__device__ __forceinline__ const T* uniform_ptr(const T* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (const T*)(((uint64_t)hi << 32) | lo);
}

template <class T>
struct PassCtx {
  const T* MR;   // NO x NI, row-major: MR[i*NI + j]
  const T* in;
  T* out;
  int AV;        // A * nv
  float invAV;
  int A;
  int nfib;
  int lgnv;
  T coef;
  // gather / scatter
  const int32_t* __restrict__ map;  // term's map slice
  const double* __restrict__ psi;
  double* hpsi;
  int nvec, v0;
  double wgt;
  // grid potential
  const T* __restrict__ V;  // term's grid slice
  T vshift;
};

__device__ __forceinline__ double fmat(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmat(float a, float b, float c) { return fmaf(a, b, c); }

template <class T, int KOUT>
__device__ __forceinline__ void emit(const PassCtx<T>& c, int av, int b, int i, int NO, T y, T xi) {
  if constexpr (KOUT == OUT_STORE) {
    c.out[av + c.AV * (i + NO * b)] = y;
  } else if constexpr (KOUT == OUT_ADD) {
    T* o = c.out + av + c.AV * (i + NO * b);
    *o = fmat(c.coef, y, *o);
  } else if constexpr (KOUT == OUT_VSTORE) {
    const int e = (av >> c.lgnv) + c.A * (i + NO * b);
    c.out[av + c.AV * (i + NO * b)] = fmat(c.coef, y, (c.V[e] + c.vshift) * xi);
  } else {
    // last active mode: nothing after it, b == 0
    const int e = (av >> c.lgnv) + c.A * i;
    const int m = c.map[e];
    if (m >= 0)
      atomicAdd(&c.hpsi[(int64_t)m * c.nvec + c.v0 + (av & ((1 << c.lgnv) - 1))],
                c.wgt * (double)y);
  }
}

// One mode product over all fibres of the current tensor, one fibre per
// thread: y(i) = sum_j M(i,j) x(j).  KIN selects where the fibre comes from
// (LDS, or the packed block through the map), KOUT what is done with y.
// The matrix rows are read with scalar loads (wave-uniform); the pointer is
// made opaque in every iteration so that the rows are not hoisted into
// vector registers across the fibre loop.
template <class T, int NI, int NO, int KIN, int KOUT>
__device__ __forceinline__ void mode_pass(const PassCtx<T>& c) {
  for (int f = threadIdx.x; f < c.nfib; f += blockDim.x) {
    const int b = fdiv(f, c.AV, c.invAV);
    const int av = f - b * c.AV;
    T x[NI];
    if constexpr (KIN == IN_GATHER) {
      // first active mode: A = 1 so av is the vector index
      const int32_t* mp = c.map + NI * b;
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        const int m = mp[j];
        x[j] = (m >= 0) ? (T)c.psi[(int64_t)m * c.nvec + c.v0 + av] : (T)0;
      }
    } else {
      const T* src = c.in + av + c.AV * NI * b;
#pragma unroll
      for (int j = 0; j < NI; ++j) x[j] = src[c.AV * j];
    }
    cptr<T> M = (cptr<T>)uniform_ptr(c.MR);
    asm volatile("" : "+s"(M));
#pragma unroll
    for (int i = 0; i < NO; ++i) {
      T y = 0;
#pragma unroll
      for (int j = 0; j < NI; ++j) y = fmat(M[i * NI + j], x[j], y);
      emit<T, KOUT>(c, av, b, i, NO, y, (KOUT == OUT_VSTORE) ? x[i < NI ? i : 0] : (T)0);
    }
  }
}

// Any sizes (rectangular, > 24, or the rare terms whose second buffer is in
// global scratch): the fibre is re-read for every output instead of being
// held in registers, which keeps this path's register use small.
template <class T, int KIN, int KOUT>
__device__ __noinline__ void mode_pass_generic(const PassCtx<T>& c, int NI, int NO) {
  for (int f = threadIdx.x; f < c.nfib; f += blockDim.x) {
    const int b = fdiv(f, c.AV, c.invAV);
    const int av = f - b * c.AV;
    for (int i = 0; i < NO; ++i) {
      T y = 0, xi = 0;
      for (int j = 0; j < NI; ++j) {
        T xj;
        if constexpr (KIN == IN_GATHER) {
          const int m = c.map[NI * b + j];
          xj = (m >= 0) ? (T)c.psi[(int64_t)m * c.nvec + c.v0 + av] : (T)0;
        } else {
          xj = c.in[av + c.AV * (j + NI * b)];
        }
        if (j == i) xi = xj;
        y = fmat(c.MR[i * NI + j], xj, y);
      }
      emit<T, KOUT>(c, av, b, i, NO, y, xi);
    }
  }
}

template <class T, bool FAST, int KIN, int KOUT>
__device__ __forceinline__ void dispatch_pass(const PassCtx<T>& c, int NI, int NO) {
  if (FAST && NI == NO) {
    switch (NI) {
#define SG4_CASE(n) case n: mode_pass<T, n, n, KIN, KOUT>(c); return;
      SG4_CASE(1) SG4_CASE(2) SG4_CASE(3) SG4_CASE(4) SG4_CASE(5) SG4_CASE(6) SG4_CASE(7)
      SG4_CASE(8) SG4_CASE(9) SG4_CASE(10) SG4_CASE(11) SG4_CASE(12) SG4_CASE(13)
      SG4_CASE(14) SG4_CASE(15) SG4_CASE(16) SG4_CASE(17) SG4_CASE(18) SG4_CASE(19)
      SG4_CASE(20) SG4_CASE(21) SG4_CASE(22) SG4_CASE(23) SG4_CASE(24)
#undef SG4_CASE
      default: break;
    }
  }
  mode_pass_generic<T, KIN, KOUT>(c, NI, NO);
}

// ---------------------------------------------------------------------------
// Register-blocked group passes.  A slab = all grid values of a group of up to
// three consecutive active modes (sizes N0, N1, N2) for fixed indices of the
// other modes and one vector; one thread holds a slab in registers and applies
// every 1-D matrix of the group before writing it back, so a term costs one
// pass per group instead of one per mode.
//   KIND_CHAIN: x <- M2 x3 (M1 x2 (M0 x1 x))          (basis<->grid, in place)
//   KIND_GRID : y <- [V x | y] + sum_i c_i (D2_i xi x)  (potential + kinetic)
// Slab element J (first index fastest) of slab f lives at base + AV*J with
// base = (f mod AV) + AV*P*(f div AV), P = N0 N1 N2, AV = A*nv.
// ---------------------------------------------------------------------------
enum { KIND_CHAIN = 0, KIND_GRID = 1 };

template <class T>
struct KMats {
  const T* mats;
  const T* V;
  const T* ck;
};


template <class T>
struct GCtx {
  const T* M[3];   // row-major n x n matrices of the group's modes
  T c[3];          // kinetic coefficients (grid kind)
  const T* in;
  T* out;
  int AV;
  float invAV;
  int A;
  int nfib;
  int first;       // grid kind: out holds V + vshift (staged by the gather); start from V x
};

// y = M x along one mode of the slab (stride S, N entries), in place
template <class T, int N, int S, int P>
__device__ __forceinline__ void slab_mode(T (&x)[P], cptr<T> M) {
  if constexpr (N >= SLAB_ROW_OUTER) {
    // row by row: one row of M (2N scalar registers) is live at a time
    // instead of the whole matrix, at the price of a second slab in VGPRs
    T y[P];
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
      for (int o = 0; o < P / N; ++o) {
        const int base = (o % S) + (o / S) * S * N;
        T acc = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) acc = fmat(M[i * N + j], x[base + S * j], acc);
        y[base + S * i] = acc;
      }
    }
#pragma unroll
    for (int J = 0; J < P; ++J) x[J] = y[J];
    return;
  }
#pragma unroll
  for (int o = 0; o < P / N; ++o) {
    const int base = (o % S) + (o / S) * S * N;
    T y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T acc = 0;
#pragma unroll
      for (int j = 0; j < N; ++j) acc = fmat(M[i * N + j], x[base + S * j], acc);
      y[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) x[base + S * i] = y[i];
  }
}

// acc += c (M x) along one mode of the slab
template <class T, int N, int S, int P>
__device__ __forceinline__ void slab_mode_acc(const T (&x)[P], T (&acc)[P], cptr<T> M, T c) {
#pragma unroll
  for (int o = 0; o < P / N; ++o) {
    const int base = (o % S) + (o / S) * S * N;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T y = 0;
#pragma unroll
      for (int j = 0; j < N; ++j) y = fmat(M[i * N + j], x[base + S * j], y);
      acc[base + S * i] = fmat(c, y, acc[base + S * i]);
    }
  }
}

template <class T, int N0, int N1, int N2, int KIND>
__device__ __forceinline__ void group_pass(const GCtx<T>& c) {
  constexpr int P = N0 * N1 * N2;
  for (int f = threadIdx.x; f < c.nfib; f += blockDim.x) {
    const int b = fdiv(f, c.AV, c.invAV);
    const int av = f - b * c.AV;
    const int base = av + c.AV * P * b;
    T x[P];
#pragma unroll
    for (int J = 0; J < P; ++J) x[J] = c.in[base + c.AV * J];
    cptr<T> M0 = (cptr<T>)uniform_ptr(c.M[0]);
    cptr<T> M1 = (cptr<T>)uniform_ptr(c.M[1]);
    cptr<T> M2 = (cptr<T>)uniform_ptr(c.M[2]);
    asm volatile("" : "+s"(M0), "+s"(M1), "+s"(M2));
    if constexpr (KIND == KIND_CHAIN) {
      slab_mode<T, N0, 1, P>(x, M0);
      if constexpr (N1 > 1) slab_mode<T, N1, N0, P>(x, M1);
      if constexpr (N2 > 1) slab_mode<T, N2, N0 * N1, P>(x, M2);
#pragma unroll
      for (int J = 0; J < P; ++J) c.out[base + c.AV * J] = x[J];
    } else {
      T acc[P];
      if (c.first) {
#pragma unroll
        for (int J = 0; J < P; ++J) acc[J] = c.out[base + c.AV * J] * x[J];
      } else {
#pragma unroll
        for (int J = 0; J < P; ++J) acc[J] = c.out[base + c.AV * J];
      }
      slab_mode_acc<T, N0, 1, P>(x, acc, M0, c.c[0]);
      if constexpr (N1 > 1) slab_mode_acc<T, N1, N0, P>(x, acc, M1, c.c[1]);
      if constexpr (N2 > 1) slab_mode_acc<T, N2, N0 * N1, P>(x, acc, M2, c.c[2]);
#pragma unroll
      for (int J = 0; J < P; ++J) c.out[base + c.AV * J] = acc[J];
    }
  }
}

template <class T, int KIND>
__device__ __forceinline__ void dispatch_group(const GCtx<T>& c, int shape) {
  switch (shape) {
#define SG4_GCASE(id, a, b, d) case id: group_pass<T, a, b, d, KIND>(c); return;
    SG4_SHAPES(SG4_GCASE)
#undef SG4_GCASE
    default: return;
  }
}

// Gather and scatter walk the item in batches of GK elements per thread so
// that the map loads of a batch, then its psi loads (or atomics), are all in
// flight together instead of one dependent load pair per element.
constexpr int GK = 8;

// One term with every active mode square (nq = nb), through group passes:
// gather (psi rows into X, V + vshift into Y) -> basis->grid groups (in
// place on X) -> grid groups (X -> Y, the first one multiplying by the staged
// potential) -> grid->basis groups (in place on Y) -> weighted scatter-add.
template <class T>
__device__ __forceinline__ void term_t1_grouped(const KArgs& a, const KMats<T>& km,
                                                const Item& it, T* lds) {
  const TermDev* Td = &a.terms[it.term];
  const int lg = it.lgnv, nv = 1 << lg;
  const int32_t* map = a.map + Td->boff;
  const T* V = km.V + Td->qoff;
  const T vshift = (T)Td->vshift;
  const int N = Td->nq;
  const int NE = N * nv;
  const int nt = blockDim.x;
  T* X = lds;
  T* Y = lds + NE;
  for (int e0 = threadIdx.x; e0 < NE; e0 += GK * nt) {
    int m[GK];
    T v[GK], x[GK];
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      const int e = e0 + k * nt;
      m[k] = (e < NE) ? map[e >> lg] : -1;
      v[k] = (e < NE) ? V[e >> lg] : (T)0;
    }
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      const int e = e0 + k * nt;
      x[k] = (m[k] >= 0) ? (T)a.psi[(int64_t)m[k] * a.nvec + it.v0 + (e & (nv - 1))] : (T)0;
    }
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      const int e = e0 + k * nt;
      if (e < NE) {
        X[e] = x[k];
        Y[e] = v[k] + vshift;
      }
    }
  }
  __syncthreads();
  GCtx<T> c;
  for (int stage = 0; stage < 3; ++stage) {
    int A = 1;
    for (int g = 0; g < Td->ngrp; ++g) {
      const int p0 = Td->gfirst[g];
      const int p1 = (g + 1 < Td->ngrp) ? Td->gfirst[g + 1] : Td->nact;
      int P = 1;
      for (int q = 0; q < 3; ++q) {
        const int pm = p0 + q;
        if (pm < p1) {
          const MatDev* md = &a.md[Td->mat[pm]];
          P *= md->nq;
          c.M[q] = km.mats + (stage == 0 ? md->B : stage == 1 ? md->D2 : md->BtW);
          c.c[q] = km.ck[Td->kmode[pm]];
        } else {
          c.M[q] = c.M[0];
          c.c[q] = 0;
        }
      }
      c.AV = A * nv;
      c.invAV = 1.0f / (float)c.AV;
      c.A = A;
      c.nfib = NE / P;
      if (stage == 1) {
        c.in = X; c.out = Y; c.first = (g == 0);
        dispatch_group<T, KIND_GRID>(c, Td->gshape[g]);
      } else {
        T* buf = (stage == 0) ? X : Y;
        c.in = buf; c.out = buf; c.first = 0;
        dispatch_group<T, KIND_CHAIN>(c, Td->gshape[g]);
      }
      __syncthreads();
      A *= P;
    }
  }
  const double w = Td->wgt;
  for (int e0 = threadIdx.x; e0 < NE; e0 += GK * nt) {
    int m[GK];
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      const int e = e0 + k * nt;
      m[k] = (e < NE) ? map[e >> lg] : -1;
    }
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      const int e = e0 + k * nt;
      if (m[k] >= 0)
        atomicAdd(&a.hpsi[(int64_t)m[k] * a.nvec + it.v0 + (e & (nv - 1))], w * (double)Y[e]);
    }
  }
}

constexpr int NT32 = 768;

struct KMats10 {
  const double* mats;
  const double* V;
  const double* G;
  const double* Jac;
  const double* sqRJ;
};

// general type-1 operator: derivative terms (two mode indices each, -1 =
// none) and their grid coefficients F[e * sum_nq + point]
struct KMats1G {
  const double* mats;
  const double* F;
  const int32_t* der;
  int nder;
  int64_t sum_nq;
};

struct KGrad {
  const double* mats;
  const double* qw;        // 1-D quadrature weights
  const int64_t* qw_off;   // (k*(Lmax+1)+l) -> offset
  const double* phi;
  double* grad;
};

// Kernels (defined in sg4_k_split_*.hip, sg4_k_grid_*.hip) that the plan code
// launches.
__global__ void k_hpsi_t1s_f64(KArgs a, KMats<double> km);
__global__ void k_hpsi_t1s_f32(KArgs a, KMats<float> km);
__global__ void k_kin_rows(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                           const double* __restrict__ val, const double* __restrict__ psi,
                           double* __restrict__ hpsi, int64_t nrows, int nvec);
__global__ void k_hpsi_t1_f64(KArgs a, KMats<double> km);
__global__ void k_hpsi_t1_f32(KArgs a, KMats<float> km);
__global__ void k_hpsi_t10_f64(KArgs a, KMats10 km);
__global__ void k_grad_v(KArgs a, KGrad kg);
__global__ void k_hpsi_t1g_f64(KArgs a, KMats1G km);

}  // namespace sg4
