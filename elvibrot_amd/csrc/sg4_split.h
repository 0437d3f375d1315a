// Split type-1 path of the SG4 H.psi: device templates of the term kernel
// (potential part), instantiated by sg4_k_split_f64.hip / sg4_k_split_f32.hip.
#pragma once
#include "sg4_device.h"

namespace sg4 {

// ---------------------------------------------------------------------------
// Split type-1 path.  When the quadrature of every 1-D basis integrates the
// products of two basis functions exactly (B^T W B = I, checked when the plan
// is built), the kinetic part of a term needs no grid: along mode k it is
// c_k T_k(l_k) with T_k(l) = BtW D2 B (nb x nb), the identity in the other
// modes.  Summed over the Smolyak terms with the combination weights, the
// kinetic part of H becomes one sparse matrix K between packed rows that
// differ in one mode index (built on the host, applied by k_kin_rows), and a
// term only needs  W_t P_t^T B^T W V_t B P_t :  basis->grid products, the
// potential, grid->basis products.  The grid->basis groups run in reverse
// order so that the last basis->grid group, the potential and the first
// grid->basis group are one register pass.
// LDS: X (N nv values of the item), Vs (the N potential values), Ms (the N
// map entries, so that the psi gathers and the scatter-adds only wait for
// LDS, not for a global map load).
// ---------------------------------------------------------------------------
template <class T>
struct SCtx {
  const T* M[3];   // basis->grid (or, not fused, the pass's) matrices of the group
  const T* Mt[3];  // fused pass: grid->basis matrices
  T* X;
  const T* Vs;
  const int32_t* Ms;     // the term's map, staged in LDS
  const double* psi;     // gather: slabs read from the packed block (group 0 only)
  double* hpsi;          // scatter: slabs added into the packed result (group 0 only)
  double w;
  int nvec, v0;
  int gather, scatter;
  int AV;
  float invAV;
  int nfib;
  int lgnv;
};

// A pass over group 0 (A = 1, AV = nv) may read its slabs straight from the
// packed block (gather) and add them straight into the result (scatter):
// element J of slab (b, av) is grid point P b + J, vector av.
template <class T, int N0, int N1, int N2, bool FUSED>
__device__ __forceinline__ void split_pass(const SCtx<T>& c) {
  constexpr int P = N0 * N1 * N2;
  for (int f = threadIdx.x; f < c.nfib; f += blockDim.x) {
    const int b = fdiv(f, c.AV, c.invAV);
    const int av = f - b * c.AV;
    const int base = av + c.AV * P * b;
    T x[P];
    if (c.gather) {
#pragma unroll
      for (int J = 0; J < P; ++J) {
        const int m = c.Ms[P * b + J];
        x[J] = (m >= 0) ? (T)c.psi[(int64_t)m * c.nvec + c.v0 + av] : (T)0;
      }
    } else {
#pragma unroll
      for (int J = 0; J < P; ++J) x[J] = c.X[base + c.AV * J];
    }
    {
      cptr<T> M0 = (cptr<T>)uniform_ptr(c.M[0]);
      cptr<T> M1 = (cptr<T>)uniform_ptr(c.M[1]);
      cptr<T> M2 = (cptr<T>)uniform_ptr(c.M[2]);
      asm volatile("" : "+s"(M0), "+s"(M1), "+s"(M2));
      slab_mode<T, N0, 1, P>(x, M0);
      if constexpr (N1 > 1) slab_mode<T, N1, N0, P>(x, M1);
      if constexpr (N2 > 1) slab_mode<T, N2, N0 * N1, P>(x, M2);
    }
    if constexpr (FUSED) {
#pragma unroll
      for (int J = 0; J < P; ++J) x[J] *= c.Vs[(base + c.AV * J) >> c.lgnv];
      cptr<T> Q0 = (cptr<T>)uniform_ptr(c.Mt[0]);
      cptr<T> Q1 = (cptr<T>)uniform_ptr(c.Mt[1]);
      cptr<T> Q2 = (cptr<T>)uniform_ptr(c.Mt[2]);
      asm volatile("" : "+s"(Q0), "+s"(Q1), "+s"(Q2));
      slab_mode<T, N0, 1, P>(x, Q0);
      if constexpr (N1 > 1) slab_mode<T, N1, N0, P>(x, Q1);
      if constexpr (N2 > 1) slab_mode<T, N2, N0 * N1, P>(x, Q2);
    }
    if (c.scatter) {
#pragma unroll
      for (int J = 0; J < P; ++J) {
        const int m = c.Ms[P * b + J];
        if (m >= 0) atomicAdd(&c.hpsi[(int64_t)m * c.nvec + c.v0 + av], c.w * (double)x[J]);
      }
    } else {
#pragma unroll
      for (int J = 0; J < P; ++J) c.X[base + c.AV * J] = x[J];
    }
  }
}

template <class T, bool FUSED>
__device__ __forceinline__ void dispatch_split(const SCtx<T>& c, int shape) {
  switch (shape) {
#define SG4_SCASE(id, a, b, d) case id: split_pass<T, a, b, d, FUSED>(c); return;
    SG4_SHAPES(SG4_SCASE)
#undef SG4_SCASE
    default: return;
  }
}

// Potential part of one term for one chunk of vectors (every active mode
// square and grouped): stage map and V -> basis->grid groups 0..G-2 (group 0
// gathering from the packed block) -> fused group G-1 -> grid->basis groups
// G-2..0 (group 0 scatter-adding into the result).  A term with one group
// does all of it in one register pass.
template <class T>
__device__ __forceinline__ void term_t1_split(const KArgs& a, const KMats<T>& km, const Item& it,
                                              T* lds) {
  const TermDev* Td = &a.terms[it.term];
  const int lg = it.lgnv, nv = 1 << lg;
  const int32_t* map = a.map + Td->boff;
  const T* V = km.V + Td->qoff;
  const double w = Td->wgt;
  if (Td->nact == 0) {  // the one-point term
    const int m = map[0];
    if (m >= 0)
      for (int v = threadIdx.x; v < nv; v += blockDim.x) {
        const int64_t r = (int64_t)m * a.nvec + it.v0 + v;
        const double val = w * (double)(V[0] * (T)a.psi[r]);
        if (a.fx) fx_add(a.fx + 2 * r, a.fxflag, val);
        else atomicAdd(&a.hpsi[r], val);
      }
    return;
  }
  const int N = Td->nq;
  const int NE = N * nv;
  const int nt = blockDim.x;
  T* X = lds;
  T* Vs = lds + NE;
  int32_t* Ms = (int32_t*)(Vs + N);
  for (int p0 = threadIdx.x; p0 < N; p0 += GK * nt) {
    int m[GK];
    T v[GK];
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      const int p = p0 + k * nt;
      m[k] = (p < N) ? map[p] : -1;
      v[k] = (p < N) ? V[p] : (T)0;
    }
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      const int p = p0 + k * nt;
      if (p < N) {
        Ms[p] = m[k];
        Vs[p] = v[k];
      }
    }
  }
  __syncthreads();
  const int G = Td->ngrp;
  SCtx<T> c;
  c.X = X;
  c.Vs = Vs;
  c.Ms = Ms;
  c.psi = a.psi;
  c.hpsi = a.hpsi;
  c.w = w;
  c.nvec = a.nvec;
  c.v0 = it.v0;
  c.lgnv = lg;
  int Ag[MAXACT];
  {
    int A = 1;
    for (int g = 0; g < G; ++g) {
      Ag[g] = A;
      const int p1 = (g + 1 < G) ? Td->gfirst[g + 1] : Td->nact;
      for (int pm = Td->gfirst[g]; pm < p1; ++pm) A *= a.md[Td->mat[pm]].nq;
    }
  }
  auto setup = [&](int g, bool fused, bool btw) {
    const int p0 = Td->gfirst[g];
    const int p1 = (g + 1 < G) ? Td->gfirst[g + 1] : Td->nact;
    int P = 1;
    for (int q = 0; q < 3; ++q) {
      const int pm = p0 + q;
      if (pm < p1) {
        const MatDev* md = &a.md[Td->mat[pm]];
        P *= md->nq;
        c.M[q] = km.mats + (btw ? md->BtW : md->B);
        c.Mt[q] = km.mats + md->BtW;
      } else {
        c.M[q] = c.M[0];
        c.Mt[q] = c.Mt[0];
      }
    }
    (void)fused;
    c.AV = Ag[g] * nv;
    c.invAV = 1.0f / (float)c.AV;
    c.nfib = NE / P;
  };
  // group 0 gathers; the last grid->basis pass (group 0) scatter-adds (FP64
  // atomics), or, with the deterministic accumulator, leaves its output in X
  // for one fixed-point pass below (kept out of the unrolled shape passes)
  const int scat = (a.fx == nullptr);
  for (int g = 0; g + 1 < G; ++g) {
    setup(g, false, false);
    c.gather = (g == 0);
    c.scatter = 0;
    dispatch_split<T, false>(c, Td->gshape[g]);
    __syncthreads();
  }
  setup(G - 1, true, false);
  c.gather = (G == 1);
  c.scatter = (G == 1) && scat;
  dispatch_split<T, true>(c, Td->gshape[G - 1]);
  for (int g = G - 2; g >= 0; --g) {
    __syncthreads();
    setup(g, false, true);
    c.gather = 0;
    c.scatter = (g == 0) && scat;
    dispatch_split<T, false>(c, Td->gshape[g]);
  }
  if (!scat) {
    // group 0's layout: X[av + nv p] is grid point p, vector av
    __syncthreads();
    for (int e = threadIdx.x; e < NE; e += nt) {
      const int m = Ms[e >> lg];
      if (m >= 0)
        fx_add(a.fx + 2 * ((int64_t)m * a.nvec + it.v0 + (e & (nv - 1))), a.fxflag, w * (double)X[e]);
    }
  }
}

template <class T>
__device__ __forceinline__ void split_items(const KArgs& a, const KMats<T>& km) {
  extern __shared__ double lds_d[];
  T* lds = (T*)lds_d;
  for (int w = blockIdx.x; w < a.nitems; w += gridDim.x) {
    if (w != (int)blockIdx.x) __syncthreads();   // LDS reuse between items
    const Item it = a.items[w];
    if (it.term < 0) continue;
    term_t1_split<T>(a, km, it, lds);
  }
}

}  // namespace sg4
