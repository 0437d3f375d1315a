// Grid paths, FP64: the type-1 kernel with the kinetic derivatives on the
// grid, the type-10 kernel and the grid-value gradient kernel.
#include "sg4_grid.h"

namespace sg4 {

// The grid is either one workgroup per item or a persistent grid that walks
// the (cost-sorted) item list with stride gridDim.x.
__global__ __launch_bounds__(NT) void k_hpsi_t1_f64(KArgs a, KMats<double> km) {
  extern __shared__ double lds[];
  for (int w = blockIdx.x; w < a.nitems; w += gridDim.x) {
    if (w != (int)blockIdx.x) __syncthreads();   // LDS reuse between items
    const Item it = a.items[w];
    if (it.term < 0) continue;
    if (it.big) term_t1<double, true>(a, km, it, lds);
    else if (a.terms[it.term].ngrp > 0) term_t1_grouped<double>(a, km, it, lds);
    else term_t1<double, false>(a, km, it, lds);
  }
}

// ---------------------------------------------------------------------------
// Type-10 operator (curvilinear kinetic d1 G d1), FP64
// (sub_TabOpPsi_OF_ONEDP_FOR_SGtype4 case 10, SRC/sub_Operator/sub_OpPsi_SG4.f90
// :1546-1649):
//   u = sqRJ psi ; u_j = d1_j u ; w_i = Jac sum_j G(j,i) u_j ;
//   H psi = -1/2 sum_i d1_i w_i / (Jac sqRJ) + V psi
// Every mode takes part (a one-point mode has a 1x1 d1).  LDS holds D + 3
// buffers of nq_t x nv: two for the transforms, the kinetic accumulator, and
// one per derivative u_j.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(NT) void k_hpsi_t10_f64(KArgs a, KMats10 km) {
  extern __shared__ double lds[];
  const Item it = a.items[blockIdx.x];
  const TermDev* Td = &a.terms[it.term];
  const int lg = it.lgnv, nv = 1 << lg;
  const int D = Td->nact;
  const size_t SZ = (size_t)Td->nq * nv;
  double* b0 = lds;
  double* b1 = lds + SZ;
  double* O = lds + 2 * SZ;
  double* U = lds + 3 * SZ;
  const int64_t q0 = Td->qoff;

  PassCtx<double> c;
  c.map = a.map + Td->boff; c.psi = a.psi; c.hpsi = a.hpsi; c.nvec = a.nvec; c.v0 = it.v0;
  c.wgt = Td->wgt; c.V = km.V + q0; c.vshift = 0.0; c.lgnv = lg; c.coef = 1.0;

  // ---- basis -> grid ----
  int A = 1, Bafter = Td->nb;
  double* cur = nullptr;
  double* nxt = b0;
  for (int p = 0; p < D; ++p) {
    const MatDev* md = &a.md[Td->mat[p]];
    const int nq = md->nq, nb = md->nb;
    Bafter /= nb;
    c.MR = km.mats + md->B;
    c.in = cur; c.out = nxt;
    c.A = A; c.AV = A * nv; c.invAV = 1.0f / (float)c.AV;
    c.nfib = A * nv * Bafter;
    if (p == 0) dispatch_pass<double, true, IN_GATHER, OUT_STORE>(c, nb, nq);
    else dispatch_pass<double, true, IN_LDS, OUT_STORE>(c, nb, nq);
    __syncthreads();
    A *= nq;
    cur = nxt;
    nxt = (nxt == b0) ? b1 : b0;
  }
  double* pg = cur;   // psi on the grid
  double* S = nxt;    // scratch: u, then w_i
  for (size_t e = threadIdx.x; e < SZ; e += blockDim.x) S[e] = km.sqRJ[q0 + (e >> lg)] * pg[e];
  __syncthreads();
  // ---- u_j = d1_j u ----
  {
    int Ab = 1, Aa = Td->nq;
    for (int j = 0; j < D; ++j) {
      const MatDev* md = &a.md[Td->mat[j]];
      const int nq = md->nq;
      Aa /= nq;
      c.MR = km.mats + md->D1;
      c.in = S; c.out = U + (size_t)j * SZ;
      c.A = Ab; c.AV = Ab * nv; c.invAV = 1.0f / (float)c.AV;
      c.nfib = Ab * nv * Aa;
      dispatch_pass<double, true, IN_LDS, OUT_STORE>(c, nq, nq);
      Ab *= nq;
    }
  }
  __syncthreads();
  // ---- sum_i d1_i (Jac sum_j G(j,i) u_j) ----
  {
    int Ab = 1, Aa = Td->nq;
    for (int i = 0; i < D; ++i) {
      for (size_t e = threadIdx.x; e < SZ; e += blockDim.x) {
        const int64_t q = q0 + (int64_t)(e >> lg);
        const double* Gq = km.G + (int64_t)D * D * q;
        double w = 0.0;
        for (int j = 0; j < D; ++j) w = fma(Gq[j + D * i], U[(size_t)j * SZ + e], w);
        S[e] = km.Jac[q] * w;
      }
      __syncthreads();
      const MatDev* md = &a.md[Td->mat[i]];
      const int nq = md->nq;
      Aa /= nq;
      c.MR = km.mats + md->D1;
      c.in = S; c.out = O;
      c.A = Ab; c.AV = Ab * nv; c.invAV = 1.0f / (float)c.AV;
      c.nfib = Ab * nv * Aa;
      if (i == 0) dispatch_pass<double, true, IN_LDS, OUT_STORE>(c, nq, nq);
      else dispatch_pass<double, true, IN_LDS, OUT_ADD>(c, nq, nq);
      __syncthreads();
      Ab *= nq;
    }
  }
  for (size_t e = threadIdx.x; e < SZ; e += blockDim.x) {
    const int64_t q = q0 + (int64_t)(e >> lg);
    O[e] = -0.5 * O[e] / (km.Jac[q] * km.sqRJ[q]) + km.V[q] * pg[e];
  }
  __syncthreads();
  // ---- grid -> basis, last pass scatters ----
  cur = O;
  nxt = S;
  A = 1;
  int Aq = Td->nq;
  for (int p = 0; p < D; ++p) {
    const MatDev* md = &a.md[Td->mat[p]];
    const int nq = md->nq, nb = md->nb;
    Aq /= nq;
    c.MR = km.mats + md->BtW;
    c.in = cur; c.out = nxt;
    c.A = A; c.AV = A * nv; c.invAV = 1.0f / (float)c.AV;
    c.nfib = A * nv * Aq;
    if (p == D - 1) {
      dispatch_pass<double, true, IN_LDS, OUT_SCATTER>(c, nq, nb);
    } else {
      dispatch_pass<double, true, IN_LDS, OUT_STORE>(c, nq, nb);
      __syncthreads();
    }
    A *= nb;
    double* t = cur; cur = nxt; nxt = t;
  }
}

// ---------------------------------------------------------------------------
// Grid-value gradient of <phi|H|psi>: both vectors transformed basis -> grid
// (the first half of the H.psi pipeline), then the weighted product.
// LDS: four buffers of nq_t x nv.
// ---------------------------------------------------------------------------

__device__ void bg_transform(const KArgs& a, const double* mats, const TermDev* Td, const Item& it,
                             const double* src, double* b0, double* b1, double** result) {
  const int lg = it.lgnv, nv = 1 << lg;
  PassCtx<double> c;
  c.map = a.map + Td->boff; c.psi = src; c.hpsi = nullptr; c.nvec = a.nvec; c.v0 = it.v0;
  c.wgt = 1.0; c.V = nullptr; c.vshift = 0.0; c.lgnv = lg; c.coef = 1.0;
  int A = 1, Bafter = Td->nb;
  double* cur = nullptr;
  double* nxt = b0;
  for (int p = 0; p < Td->nact; ++p) {
    const MatDev* md = &a.md[Td->mat[p]];
    const int nq = md->nq, nb = md->nb;
    Bafter /= nb;
    c.MR = mats + md->B;
    c.in = cur; c.out = nxt;
    c.A = A; c.AV = A * nv; c.invAV = 1.0f / (float)c.AV;
    c.nfib = A * nv * Bafter;
    if (p == 0) dispatch_pass<double, true, IN_GATHER, OUT_STORE>(c, nb, nq);
    else dispatch_pass<double, true, IN_LDS, OUT_STORE>(c, nb, nq);
    __syncthreads();
    A *= nq;
    cur = nxt;
    nxt = (nxt == b0) ? b1 : b0;
  }
  *result = cur;
}

__global__ __launch_bounds__(NT) void k_grad_v(KArgs a, KGrad kg) {
  extern __shared__ double lds[];
  const Item it = a.items[blockIdx.x];
  const TermDev* Td = &a.terms[it.term];
  const int lg = it.lgnv, nv = 1 << lg;
  const size_t SZ = (size_t)Td->nq * nv;
  double* gphi = nullptr;
  double* gpsi = nullptr;
  if (Td->nact == 0) {   // one point, no transform
    for (int v = threadIdx.x; v < nv; v += blockDim.x) {
      const int m = a.map[Td->boff];
      const int64_t r = (int64_t)m * a.nvec + it.v0 + v;
      kg.grad[Td->qoff * a.nvec + it.v0 + v] = (m >= 0) ? Td->wgt * kg.phi[r] * a.psi[r] : 0.0;
    }
    return;
  }
  bg_transform(a, kg.mats, Td, it, kg.phi, lds, lds + SZ, &gphi);
  bg_transform(a, kg.mats, Td, it, a.psi, lds + 2 * SZ, lds + 3 * SZ, &gpsi);
  for (size_t e = threadIdx.x; e < SZ; e += blockDim.x) {
    int pt = (int)(e >> lg);
    double w = Td->wgt;
    for (int p = 0; p < Td->nact; ++p) {
      const MatDev* md = &a.md[Td->mat[p]];
      const int i = pt % md->nq;
      pt /= md->nq;
      w *= kg.qw[kg.qw_off[Td->mat[p]] + i];
    }
    kg.grad[(Td->qoff + (int64_t)(e >> lg)) * a.nvec + it.v0 + (e & (nv - 1))] = w * gphi[e] * gpsi[e];
  }
}

}  // namespace sg4

namespace sg4 {

// ---------------------------------------------------------------------------
// General type-1 operator, FP64 (sub_TabOpPsi_OF_ONEDP_FOR_SGtype4 case 1,
// SRC/sub_Operator/sub_OpPsi_SG4.f90:1488-1544):  on the grid of a term
//   H psi = sum_e F_e(q) D_e psi,
// D_e from the two mode indices of term e (-1 = none): the identity, dnRGG
// d1 along i, d2 along i, or d1 along j then along i (the 1-D matrices of
// DerivOp_TO_RDP_OF_SmolaykRep).  Every mode takes part.  LDS: four buffers
// of nq_t x nv -- the two transform buffers (the grid psi X and a free one W),
// the accumulator Y and a temporary T for the mixed derivatives.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void term_t1g(const KArgs& a, const KMats1G& km, const Item& it,
                                         double* lds) {
  const TermDev* Td = &a.terms[it.term];
  const int lg = it.lgnv, nv = 1 << lg;
  const int nact = Td->nact;
  const int64_t SZ = (int64_t)Td->nq * nv;
  double* b0 = lds;
  double* b1 = b0 + SZ;
  double* Y = b1 + SZ;
  double* T = Y + SZ;
  const int nt = blockDim.x;
  PassCtx<double> c;
  c.map = a.map + Td->boff; c.psi = a.psi; c.hpsi = a.hpsi; c.nvec = a.nvec; c.v0 = it.v0;
  c.wgt = Td->wgt; c.V = nullptr; c.vshift = 0.0; c.lgnv = lg; c.coef = 1.0;
  // basis -> grid
  int A = 1, Bafter = Td->nb;
  double* cur = nullptr;
  double* nxt = b0;
  for (int p = 0; p < nact; ++p) {
    const MatDev* md = &a.md[Td->mat[p]];
    const int nq = md->nq, nb = md->nb;
    Bafter /= nb;
    c.MR = km.mats + md->B;
    c.in = cur; c.out = nxt;
    c.A = A; c.AV = A * nv; c.invAV = 1.0f / (float)c.AV;
    c.nfib = A * nv * Bafter;
    if (p == 0) dispatch_pass<double, true, IN_GATHER, OUT_STORE>(c, nb, nq);
    else dispatch_pass<double, true, IN_LDS, OUT_STORE>(c, nb, nq);
    __syncthreads();
    A *= nq;
    cur = nxt;
    nxt = (nxt == b0) ? b1 : b0;
  }
  double* X = cur;
  double* W = nxt;
  for (int64_t e = threadIdx.x; e < SZ; e += nt) Y[e] = 0.0;
  // one derivative along mode p of the grid tensor: in -> out
  auto deriv = [&](int p, int64_t moff, const double* in, double* out) {
    int Ab = 1, Aa = Td->nq;
    for (int q = 0; q < p; ++q) Ab *= a.md[Td->mat[q]].nq;
    const MatDev* md = &a.md[Td->mat[p]];
    Aa /= Ab * md->nq;
    c.MR = km.mats + moff;
    c.in = in; c.out = out;
    c.A = Ab; c.AV = Ab * nv; c.invAV = 1.0f / (float)c.AV;
    c.nfib = Ab * nv * Aa;
    dispatch_pass<double, true, IN_LDS, OUT_STORE>(c, md->nq, md->nq);
  };
  __syncthreads();
  for (int e = 0; e < km.nder; ++e) {
    int i = km.der[2 * e], j = km.der[2 * e + 1];
    if (i < 0 && j >= 0) { i = j; j = -1; }
    const double* src = X;
    if (i >= 0) {
      const MatDev* mi = &a.md[Td->mat[i]];
      if (j < 0 || j == i) {
        deriv(i, j < 0 ? mi->D1 : mi->D2, X, W);
        src = W;
      } else {
        deriv(j, a.md[Td->mat[j]].D1, X, W);
        __syncthreads();
        deriv(i, mi->D1, W, T);
        src = T;
      }
      __syncthreads();
    }
    const double* Fe = km.F + (int64_t)e * km.sum_nq + Td->qoff;
    for (int64_t idx = threadIdx.x; idx < SZ; idx += nt) Y[idx] += Fe[idx >> lg] * src[idx];
    __syncthreads();
  }
  // grid -> basis, last pass scatters
  cur = Y;
  nxt = W;
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
      dispatch_pass<double, true, IN_LDS, OUT_SCATTER>(c, nq, nb);
    } else {
      dispatch_pass<double, true, IN_LDS, OUT_STORE>(c, nq, nb);
      __syncthreads();
    }
    A *= nb;
    double* t = cur; cur = nxt; nxt = t;
  }
}

__global__ __launch_bounds__(NT) void k_hpsi_t1g_f64(KArgs a, KMats1G km) {
  extern __shared__ double lds[];
  for (int w = blockIdx.x; w < a.nitems; w += gridDim.x) {
    if (w != (int)blockIdx.x) __syncthreads();
    const Item it = a.items[w];
    if (it.term < 0) continue;
    term_t1g(a, km, it, lds);
  }
}

}  // namespace sg4
