// Type-1 grid path of the SG4 H.psi (kinetic derivatives on the grid): the
// per-mode term template, instantiated by sg4_k_grid_f64.hip /
// sg4_k_grid_f32.hip.
#pragma once
#include "sg4_device.h"

namespace sg4 {

// ---------------------------------------------------------------------------
// Type-1 operator: one workgroup per work item.  T is the type of the mode
// contractions (double: parity mode; float: the FP32 mode); gather and
// scatter stay in FP64.
// ---------------------------------------------------------------------------

template <class T, bool BIG>
__device__ __forceinline__ void term_t1(const KArgs& a, const KMats<T>& km, const Item& it,
                                        T* lds) {
  const TermDev* Td = &a.terms[it.term];
  const int lg = it.lgnv, nv = 1 << lg;
  const int nact = Td->nact;
  const int32_t* map = a.map + Td->boff;
  const T* V = km.V + Td->qoff;

  if (nact == 0) {  // the one-point term
    const int m = map[0];
    if (m >= 0)
      for (int v = threadIdx.x; v < nv; v += blockDim.x) {
        const int64_t r = (int64_t)m * a.nvec + it.v0 + v;
        atomicAdd(&a.hpsi[r], Td->wgt * (double)(((T)V[0] + (T)Td->vshift) * (T)a.psi[r]));
      }
    return;
  }

  T* buf0 = lds;
  T* buf1 = BIG ? (T*)a.scratch + (int64_t)it.slot * a.slot_size : lds + (size_t)Td->nq * nv;

  PassCtx<T> c;
  c.map = map; c.psi = a.psi; c.hpsi = a.hpsi; c.nvec = a.nvec; c.v0 = it.v0;
  c.wgt = Td->wgt; c.V = V; c.vshift = (T)Td->vshift; c.lgnv = lg; c.coef = 1;

  // ---- basis -> grid ----
  int A = 1;             // product of already transformed sizes
  int Bafter = Td->nb;   // product of untransformed sizes (incl. current)
  T* cur = nullptr;
  T* nxt = buf0;
  for (int p = 0; p < nact; ++p) {
    const MatDev* md = &a.md[Td->mat[p]];
    const int nq = md->nq, nb = md->nb;
    Bafter /= nb;
    c.MR = km.mats + md->B;
    c.in = cur; c.out = nxt;
    c.A = A; c.AV = A * nv; c.invAV = 1.0f / (float)c.AV;
    c.nfib = A * nv * Bafter;
    if (p == 0) dispatch_pass<T, !BIG, IN_GATHER, OUT_STORE>(c, nb, nq);
    else dispatch_pass<T, !BIG, IN_LDS, OUT_STORE>(c, nb, nq);
    __syncthreads();
    A *= nq;
    cur = nxt;
    nxt = (nxt == buf0) ? buf1 : buf0;
  }
  // ---- grid: V psi + sum_k c_k d2_k psi ----
  T* pg = cur;
  T* acc = nxt;
  {
    int Ab = 1, Aa = Td->nq;
    for (int p = 0; p < nact; ++p) {
      const MatDev* md = &a.md[Td->mat[p]];
      const int nq = md->nq;
      Aa /= nq;
      c.MR = km.mats + md->D2;
      c.in = pg; c.out = acc;
      c.A = Ab; c.AV = Ab * nv; c.invAV = 1.0f / (float)c.AV;
      c.nfib = Ab * nv * Aa;
      c.coef = km.ck[Td->kmode[p]];
      if (p == 0) dispatch_pass<T, !BIG, IN_LDS, OUT_VSTORE>(c, nq, nq);
      else dispatch_pass<T, !BIG, IN_LDS, OUT_ADD>(c, nq, nq);
      __syncthreads();
      Ab *= nq;
    }
  }
  // ---- grid -> basis, last pass scatters ----
  c.coef = 1;
  cur = acc;
  nxt = pg;
  A = 1;
  int Aq = Td->nq;
  for (int p = 0; p < nact; ++p) {
    const MatDev* md = &a.md[Td->mat[p]];
    const int nq = md->nq, nb = md->nb;
    Aq /= nq;
    c.MR = km.mats + md->BtW;
    c.in = cur; c.out = nxt;
    c.A = A; c.AV = A * nv; c.invAV = 1.0f / (float)c.AV;
    c.nfib = A * nv * Aq;
    if (p == nact - 1) {
      dispatch_pass<T, !BIG, IN_LDS, OUT_SCATTER>(c, nq, nb);
    } else {
      dispatch_pass<T, !BIG, IN_LDS, OUT_STORE>(c, nq, nb);
      __syncthreads();
    }
    A *= nb;
    T* t = cur; cur = nxt; nxt = t;
  }
}

}  // namespace sg4
