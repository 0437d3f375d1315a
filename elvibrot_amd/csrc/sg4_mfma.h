// Split-precision MFMA path of the SG4 H.psi (SG4_PREC_F32X2): device
// templates of the term kernel, instantiated by sg4_k_mfma.hip.
//
// The potential part of a term, W_t P_t^T (x)B^T W V_t (x)B P_t, is computed
// with the matrix cores.  The active modes of a term are cut into groups of
// consecutive modes whose Kronecker size K = prod n_k is at most 32; a group
// contraction y = (B_a (x) B_b (x) ...) x along the group's index is a GEMM
// with a K x K matrix (padded to 16 or 32) and N/K x nv columns, done as
// 16x16 output tiles of v_mfma_f32_16x16x{16,32}_bf16.
//
// Split precision: every FP32 operand x is written as three bf16 limbs
// x = hi + mid + lo (round to nearest at each step; the residuals x - hi and
// x - hi - mid are exact in FP32), and a product is the sum of the six limb
// products whose order is at most 2^-16 relative (lo*Hi, mid*Mid, hi*Lo,
// mid*Hi, hi*Mid, hi*Hi, smallest first), accumulated in FP32 by the MFMA.
// The representation error of three limbs is below 2^-24 relative, the
// dropped products are below 2^-25: each contraction is as accurate as an
// FP32 one.  The group matrices are split once on the host.
//
// LDS holds the item's tensor in FP32, in place, element (point p, vector v)
// at X[p*RS + v] (points first-mode-fastest, as the map; RS >= nv pads the
// rows against bank conflicts).  A pass over group
// g reads, for each 16-column tile, the B operand straight from X (lane l:
// column l%16, group indices k of its K slots), multiplies, and writes the
// output rows back to the same addresses (output row m of a lane is the
// group index of one of its input slots: the C layout of the 16x16 MFMA
// holds rows 4(l/16)..4(l/16)+3 of column l%16, and the K slots are ordered
// to match).  Tiles of one pass touch disjoint columns, so a pass needs no
// barrier inside; passes are separated by one.  Basis->grid groups 0..G-2,
// then the last group fused in registers (basis->grid, times V, grid->basis:
// the C tile is the next B operand), then grid->basis groups G-2..0.  Group 0
// gathers its B operands from the packed block through the term's map and
// the last pass adds W_t times its output into H.psi with FP64 atomics (the
// partial sums of the Smolyak combination are up to ~1e3 times larger than
// the result, so the sum over terms is accumulated in FP64).
#pragma once
#include "sg4_device.h"

namespace sg4 {

constexpr int MG_MAX = 16;          // groups per term
constexpr int NT_MFMA = 1024;       // largest workgroup of the MFMA term kernel
constexpr int MAT_STRIDE = 3072;    // bf16 entries per group matrix (3 limbs, 32 x 32)

// A term of the MFMA path.
struct TermM {
  int64_t boff, qoff;   // offsets of the term in the map and in the grids
  double wgt;           // W_t times the 1 x 1 factors of the inactive modes
  int32_t N;            // grid points (= basis functions: square modes)
  int32_t G;            // groups (0: the one-point term)
  int32_t A[MG_MAX];    // point stride of group g (product of the earlier groups' K)
  int32_t K[MG_MAX];    // Kronecker size of group g
  int32_t mB[MG_MAX];   // group-matrix index of B_g; grid->basis BtW_g is mB[g] + 1
};

struct MArgs {
  const Item* items;
  const TermM* terms;
  const uint16_t* mats;   // group matrices, MAT_STRIDE bf16 each, in A-operand lane order
  const int32_t* map;
  const float* V;         // FP32 potential grids
  const double* psi;      // rows layout psi[row*nvec + v]
  double* hpsi;           // rows layout H.psi (holds the kinetic part; terms are added)
  unsigned long long* fx;       // deterministic accumulator (or nullptr), see fx_add
  unsigned long long* fxflag;
  int nvec;
  int nitems;
};

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef short s16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));

// two floats -> two bf16 (round to nearest even), packed
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  f32x2_t f = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f, bf16x2_t));
}

// Operand of S bf16 values per lane (S = 4: x16 instruction, S = 8: x32).
template <int S> struct Limbs;
template <> struct Limbs<4> { u32x2_t h, m, l; };
template <> struct Limbs<8> { u32x4_t h, m, l; };

// x = hi + mid + lo, three bf16 limbs, pairs packed
template <int S>
__device__ __forceinline__ void split3(const float (&x)[S], Limbs<S>& o) {
#pragma unroll
  for (int i = 0; i < S / 2; ++i) {
    const float a = x[2 * i], b = x[2 * i + 1];
    const unsigned hp = pack_bf16(a, b);
    const float ra = a - __uint_as_float(hp << 16), rb = b - __uint_as_float(hp & 0xffff0000u);
    const unsigned mp = pack_bf16(ra, rb);
    const float sa = ra - __uint_as_float(mp << 16), sb = rb - __uint_as_float(mp & 0xffff0000u);
    o.h[i] = hp;
    o.m[i] = mp;
    o.l[i] = pack_bf16(sa, sb);
  }
}

template <int S> struct MF;
template <> struct MF<4> {
  static __device__ __forceinline__ f32x4_t mma(u32x2_t a, u32x2_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s16x4_t, a),
                                                      __builtin_bit_cast(s16x4_t, b), c, 0, 0, 0);
  }
};
template <> struct MF<8> {
  static __device__ __forceinline__ f32x4_t mma(u32x4_t a, u32x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                    __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
  }
};

// acc += A (3 limbs) * B (3 limbs), six products, smallest first
template <int S>
__device__ __forceinline__ f32x4_t mma3(const Limbs<S>& A, const Limbs<S>& B, f32x4_t c) {
  c = MF<S>::mma(A.h, B.l, c);
  c = MF<S>::mma(A.m, B.m, c);
  c = MF<S>::mma(A.l, B.h, c);
  c = MF<S>::mma(A.h, B.m, c);
  c = MF<S>::mma(A.m, B.h, c);
  c = MF<S>::mma(A.h, B.h, c);
  return c;
}

// A operand of a group matrix for this lane: limb-major, M tile, lane, slot
template <int S>
__device__ __forceinline__ void load_amat(const uint16_t* __restrict__ mat, int lane,
                                          Limbs<S> (&A)[S / 4]) {
  constexpr int MT = S / 4;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    if constexpr (S == 8) {
      const u32x4_t* p = (const u32x4_t*)mat;
      A[mt].h = p[(0 * MT + mt) * 64 + lane];
      A[mt].m = p[(1 * MT + mt) * 64 + lane];
      A[mt].l = p[(2 * MT + mt) * 64 + lane];
    } else {
      const u32x2_t* p = (const u32x2_t*)mat;
      A[mt].h = p[(0 * MT + mt) * 64 + lane];
      A[mt].m = p[(1 * MT + mt) * 64 + lane];
      A[mt].l = p[(2 * MT + mt) * 64 + lane];
    }
  }
}

// group index held by slot s of a lane in lane group j = lane/16: for the x16
// instruction 4j + s; for x32 the first four slots are rows 4j..4j+3 of the
// first 16-row output tile and the last four rows 16+4j.. of the second one,
// so that an output tile is also the next product's B operand
template <int S>
__device__ __forceinline__ int slot_k(int j, int s) {
  return (S == 4) ? 4 * j + s : (s < 4 ? 4 * j + s : 12 + 4 * j + s);
}

enum { MIN_LDS = 0, MIN_GATHER = 1 };
enum { MOUT_LDS = 0, MOUT_SCATTER = 1 };

// LDS row stride (floats) of an item with nv vectors: nv + 4 from 16 vectors
// on (20, 36, 68), so that the four lane groups of a B-operand read (rows 4A
// apart, A odd for odd mode sizes) fall in distinct banks; nv below 16.
__host__ __device__ constexpr int mrow_stride(int nv) { return nv >= 16 ? nv + 4 : nv; }

struct MCtx {
  float* X;                 // LDS tensor, element (p, v) at X[p*RS + v]
  const float* Vs;          // LDS potential (N)
  const int32_t* Ms;        // LDS map (N)
  const uint16_t* matA;     // this pass's matrix (fused pass: basis->grid)
  const uint16_t* matB;     // fused pass: grid->basis matrix
  const double* psi;
  double* hpsi;
  unsigned long long* fx;
  unsigned long long* fxflag;
  uint32_t dummy;           // byte address of a per-lane scratch word (masked-out rows)
  int nvec, v0, lgnv, RS;
  int N, K, A;
  float invA;
  double w;
};

// n / d for 0 <= n < 2^22, 1 <= d < 2^16: float quotient, one exact correction
__device__ __forceinline__ int qdiv(int n, int d, float inv) {
  int q = (int)((float)n * inv);
  const int r = n - q * d;
  q += (r >= d) ? 1 : 0;
  q -= (r < 0) ? 1 : 0;
  return q;
}

template <int S, int IN, int OUT, bool FUSED>
__device__ __forceinline__ void mpass(const MCtx& c) {
  constexpr int MT = S / 4;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6, nwav = blockDim.x >> 6;
  const int j = lane >> 4, n = lane & 15;
  const int nv = 1 << c.lgnv;
  Limbs<S> Am[MT], Bm[MT];
  load_amat<S>(c.matA, lane, Am);
  if constexpr (FUSED) load_amat<S>(c.matB, lane, Bm);
  // per-slot offsets, invariant over the tiles: point offset k A and LDS byte
  // offset; a slot beyond K reads point 0 of its column (finite, multiplied
  // by the zero padding of the matrix) and writes to the lane's scratch word
  int koff[S];
  uint32_t kofb[S];
  bool kval[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int k = slot_k<S>(j, s);
    kval[s] = k < c.K;
    koff[s] = kval[s] ? k * c.A : 0;
    kofb[s] = (uint32_t)(koff[s] * c.RS) * 4u;
  }
  const int ncol = (c.N / c.K) << c.lgnv;
  const int ntile = (ncol + 15) >> 4;
  const int KA = c.K * c.A;
  const uint32_t xb = (uint32_t)(uintptr_t)c.X;
  for (int t = wave; t < ntile; t += nwav) {
    const int col0 = t * 16 + n;
    const bool cv = col0 < ncol;
    const int col = cv ? col0 : ncol - 1;   // a partial tile repeats its last column
    const int v = col & (nv - 1);
    const int rest = col >> c.lgnv;
    const int ihi = qdiv(rest, c.A, c.invA);
    const int pb = (rest - ihi * c.A) + ihi * KA;
    const uint32_t base = xb + (uint32_t)(pb * c.RS + v) * 4u;
    float x[S];
    if constexpr (IN == MIN_GATHER) {
      const uint32_t rowb = (uint32_t)c.nvec * 8u, vb = (uint32_t)(c.v0 + v) * 8u;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int m = c.Ms[pb + koff[s]];
        const uint32_t off = (uint32_t)(m < 0 ? 0 : m) * rowb + vb;
        const double val = *(const double*)((const char*)c.psi + off);
        x[s] = (m >= 0) ? (float)val : 0.0f;
      }
    } else {
#pragma unroll
      for (int s = 0; s < S; ++s)
        x[s] = *(const __attribute__((address_space(3))) float*)(uintptr_t)(base + kofb[s]);
    }
    Limbs<S> B;
    split3<S>(x, B);
    f32x4_t y[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      y[mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
      y[mt] = mma3<S>(Am[mt], B, y[mt]);
    }
    if constexpr (FUSED) {
      // rows beyond K are exactly zero (zero matrix rows)
#pragma unroll
      for (int s = 0; s < S; ++s) x[s] = y[s >> 2][s & 3] * c.Vs[pb + koff[s]];
      split3<S>(x, B);
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        y[mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        y[mt] = mma3<S>(Bm[mt], B, y[mt]);
      }
    }
    if constexpr (OUT == MOUT_LDS) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const uint32_t ad = kval[s] ? base + kofb[s] : c.dummy;
        *(__attribute__((address_space(3))) float*)(uintptr_t)ad = y[s >> 2][s & 3];
      }
    } else {
      const uint32_t rowb = (uint32_t)c.nvec * 8u, vb = (uint32_t)(c.v0 + v) * 8u;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int m = c.Ms[pb + koff[s]];
        if (cv && kval[s] && m >= 0) {
          const uint32_t off = (uint32_t)m * rowb + vb;   // bytes of the FP64 element
          if (c.fx) fx_add(c.fx + 2 * (off >> 3), c.fxflag, c.w * (double)y[s >> 2][s & 3]);
          else atomicAdd((double*)((char*)c.hpsi + off), c.w * (double)y[s >> 2][s & 3]);
        }
      }
    }
  }
}

template <int IN, int OUT, bool FUSED>
__device__ __forceinline__ void mpass_k(const MCtx& c) {
  if (c.K <= 16) mpass<4, IN, OUT, FUSED>(c);
  else mpass<8, IN, OUT, FUSED>(c);
}

// The potential part of one term for one chunk of vectors.
__device__ __forceinline__ void term_mfma(const MArgs& a, const Item& it, float* lds) {
  const TermM* T = &a.terms[it.term];
  const int lg = it.lgnv, nv = 1 << lg;
  const int32_t* map = a.map + T->boff;
  const float* V = a.V + T->qoff;
  const double w = T->wgt;
  const int G = T->G;
  if (G == 0) {   // the one-point term
    const int m = map[0];
    if (m >= 0)
      for (int v = threadIdx.x; v < nv; v += blockDim.x) {
        const int64_t r = (int64_t)m * a.nvec + it.v0 + v;
        const double val = w * (double)(V[0] * (float)a.psi[r]);
        if (a.fx) fx_add(a.fx + 2 * r, a.fxflag, val);
        else atomicAdd(&a.hpsi[r], val);
      }
    return;
  }
  const int N = T->N;
  const int RS = mrow_stride(nv);
  float* X = lds;
  float* Vs = lds + (int64_t)N * RS;
  int32_t* Ms = (int32_t*)(Vs + N);
  float* dummy = (float*)(Ms + N);
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    Ms[p] = map[p];
    Vs[p] = V[p];
  }
  __syncthreads();
  MCtx c;
  c.X = X;
  c.Vs = Vs;
  c.Ms = Ms;
  c.dummy = (uint32_t)(uintptr_t)(dummy + (threadIdx.x & 63));
  c.psi = a.psi;
  c.hpsi = a.hpsi;
  c.fx = a.fx;
  c.fxflag = a.fxflag;
  c.nvec = a.nvec;
  c.v0 = it.v0;
  c.lgnv = lg;
  c.RS = RS;
  c.N = N;
  c.w = w;
  auto set = [&](int g, int mat, int mat2) {
    c.K = T->K[g];
    c.A = T->A[g];
    c.invA = 1.0f / (float)c.A;
    c.matA = a.mats + (int64_t)mat * MAT_STRIDE;
    c.matB = a.mats + (int64_t)mat2 * MAT_STRIDE;
  };
  for (int g = 0; g + 1 < G; ++g) {
    set(g, T->mB[g], T->mB[g]);
    if (g == 0) mpass_k<MIN_GATHER, MOUT_LDS, false>(c);
    else mpass_k<MIN_LDS, MOUT_LDS, false>(c);
    __syncthreads();
  }
  set(G - 1, T->mB[G - 1], T->mB[G - 1] + 1);
  if (G == 1) mpass_k<MIN_GATHER, MOUT_SCATTER, true>(c);
  else mpass_k<MIN_LDS, MOUT_LDS, true>(c);
  for (int g = G - 2; g >= 0; --g) {
    __syncthreads();
    set(g, T->mB[g] + 1, T->mB[g] + 1);
    if (g == 0) mpass_k<MIN_LDS, MOUT_SCATTER, false>(c);
    else mpass_k<MIN_LDS, MOUT_LDS, false>(c);
  }
}

__device__ __forceinline__ void mfma_items(const MArgs& a) {
  extern __shared__ double lds_d[];
  float* lds = (float*)lds_d;
  for (int w = blockIdx.x; w < a.nitems; w += gridDim.x) {
    if (w != (int)blockIdx.x) __syncthreads();   // LDS reuse between items
    const Item it = a.items[w];
    if (it.term < 0) continue;
    term_mfma(a, it, lds);
  }
}

__global__ void k_hpsi_mfma(MArgs a);

}  // namespace sg4
