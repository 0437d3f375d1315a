// MI355X (gfx950) Smolyak SG4 H.psi: device plan, kernels and C ABI.
//
// Counterpart of sub_TabOpPsi_FOR_SGtype4 (SRC/sub_Operator/sub_OpPsi_SG4.f90
// :681-977) and of the per-term action sub_TabOpPsi_OF_ONEDP_FOR_SGtype4
// (:1352-1666).  The reference loops over Smolyak terms with OpenMP, gathers
// each term's coefficients, transforms basis->grid mode by mode, applies the
// potential and the kinetic derivatives on the grid, transforms back and
// scatter-adds W_t times the result.  Here one workgroup handles one
// (term, chunk of vectors) work item in LDS, in passes over groups of up to
// three modes whose slabs a thread holds in registers while it multiplies
// them by the small 1-D matrices (entries wave-uniform, read through the
// scalar cache).  Three type-1 paths:
//   split path (k_kin_rows + k_hpsi_t1s_*): the kinetic part as one sparse
//     matrix over the packed rows, the terms keep only basis->grid, the
//     potential and grid->basis, with the gather and the weighted scatter-add
//     folded into the first and last passes;
//   grouped grid path (k_hpsi_t1_*, term_t1_grouped): the reference's
//     structure with the kinetic derivatives on the grid;
//   per-mode grid path (k_hpsi_t1_*, term_t1): one pass per mode, any sizes.
// Type 10 (d1 G d1): k_hpsi_t10_f64.  Eigenvalue gradients: k_grad_v.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "sg4_mfma.h"

namespace sg4 {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

#define HIP_TRY(expr)                                                             \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr); \
      return SG4_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

// V = V0 + eps * mu on every grid point (piecewise-constant field segment)
__global__ void k_axpy_grid(const double* __restrict__ v0, const double* __restrict__ mu,
                            double eps, double* __restrict__ v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = mu ? fma(eps, mu[i], v0[i]) : v0[i];
}

// deterministic accumulation: clear the 128-bit accumulators and the overflow
// counter; then add their values (hi * 2^64 + lo) * 2^-64 to hpsi, or NaN if
// some contribution overflowed the fixed-point range
__global__ void k_fx_clear(unsigned long long* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ULL;
}
__global__ void k_fx_finish(const unsigned long long* __restrict__ fx,
                            const unsigned long long* __restrict__ flag, double* __restrict__ hpsi,
                            int64_t n) {
  const bool bad = *flag != 0ULL;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = (double)(long long)fx[2 * i + 1] + ldexp((double)fx[2 * i], -64);
    hpsi[i] = bad ? __builtin_nan("") : hpsi[i] + v;
  }
}

__global__ void k_to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
// Layout changes between the column-major block (v * nrow + r) and the rows
// layout (r * W + v) of width W >= nvec (columns nvec..W-1 of the rows layout
// are zero going in and ignored coming out), through 64-row x 16-vector LDS
// tiles so that both the reads and the writes are contiguous.
// grid = (ceil(nrow/64), ceil(W/16)), 256 threads.
constexpr int TT_R = 64, TT_V = 16;
__global__ __launch_bounds__(256) void k_cols_to_rows(const double* __restrict__ in,
                                                      double* __restrict__ out, int64_t nrow,
                                                      int nvec, int W) {
  __shared__ double tile[TT_V][TT_R + 1];
  const int64_t r0 = (int64_t)blockIdx.x * TT_R;
  const int v0 = blockIdx.y * TT_V;
  for (int k = threadIdx.x; k < TT_R * TT_V; k += blockDim.x) {
    const int vv = k / TT_R, rr = k % TT_R;   // read columns: rows contiguous
    if (r0 + rr < nrow && v0 + vv < W)
      tile[vv][rr] = (v0 + vv < nvec) ? in[(int64_t)(v0 + vv) * nrow + r0 + rr] : 0.0;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < TT_R * TT_V; k += blockDim.x) {
    const int rr = k / TT_V, vv = k % TT_V;   // write rows: vectors contiguous
    if (r0 + rr < nrow && v0 + vv < W) out[(r0 + rr) * W + v0 + vv] = tile[vv][rr];
  }
}
__global__ __launch_bounds__(256) void k_rows_to_cols(const double* __restrict__ in,
                                                      double* __restrict__ out, int64_t nrow,
                                                      int nvec, int W) {
  __shared__ double tile[TT_V][TT_R + 1];
  const int64_t r0 = (int64_t)blockIdx.x * TT_R;
  const int v0 = blockIdx.y * TT_V;
  for (int k = threadIdx.x; k < TT_R * TT_V; k += blockDim.x) {
    const int rr = k / TT_V, vv = k % TT_V;   // read rows: vectors contiguous
    if (r0 + rr < nrow && v0 + vv < nvec) tile[vv][rr] = in[(r0 + rr) * W + v0 + vv];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < TT_R * TT_V; k += blockDim.x) {
    const int vv = k / TT_R, rr = k % TT_R;   // write columns: rows contiguous
    if (r0 + rr < nrow && v0 + vv < nvec) out[(int64_t)(v0 + vv) * nrow + r0 + rr] = tile[vv][rr];
  }
}
static dim3 tile_grid(int64_t nrow, int W) {
  return dim3((unsigned)((nrow + TT_R - 1) / TT_R), (unsigned)((W + TT_V - 1) / TT_V));
}
__global__ void k_fill(double* __restrict__ p, int64_t n, double val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = val;
}

// Henon-Heiles grid of every term (calcN_op of
// Working_tests/MPI_tests/6D_Davidson_S0/sub_system_HenonHeiles.f:14-43):
// V = sum 1/2 q_i^2 + lambda sum (q_i^2 q_{i+1} - q_{i+1}^3/3)
struct HHArgs {
  const int32_t* levels;  // D x nterms
  const int64_t* qoff;
  const double* nodes;    // for k, for l: nq nodes
  const int64_t* node_off;  // (k*(Lmax+1)+l) -> offset
  const int32_t* nqtab;     // (k*(Lmax+1)+l) -> nq
  int D, Lmax;
  int64_t nterms;
  double lambda;
  double* V;
};
__global__ void k_grid_hh(HHArgs a) {
  for (int64_t t = blockIdx.x; t < a.nterms; t += gridDim.x) {
    const int64_t q0 = a.qoff[t], nqt = a.qoff[t + 1] - q0;
    for (int64_t e = threadIdx.x; e < nqt; e += blockDim.x) {
      int64_t r = e;
      double v = 0.0, qprev = 0.0;
      for (int k = 0; k < a.D; ++k) {
        const int l = a.levels[(int64_t)t * a.D + k];
        const int idx = k * (a.Lmax + 1) + l;
        const int n = a.nqtab[idx];
        const int i = (int)(r % n);
        r /= n;
        const double q = a.nodes[a.node_off[idx] + i];
        v += 0.5 * q * q;
        if (k > 0) v += a.lambda * (qprev * qprev * q - q * q * q / 3.0);
        qprev = q;
      }
      a.V[q0 + e] = v;
    }
  }
}

// Dipole mu = sum_k q_k at every grid point of every term (first index fastest).
__global__ void k_grid_dipole(HHArgs a) {
  for (int64_t t = blockIdx.x; t < a.nterms; t += gridDim.x) {
    const int64_t q0 = a.qoff[t], nqt = a.qoff[t + 1] - q0;
    for (int64_t e = threadIdx.x; e < nqt; e += blockDim.x) {
      int64_t r = e;
      double m = 0.0;
      for (int k = 0; k < a.D; ++k) {
        const int l = a.levels[(int64_t)t * a.D + k];
        const int idx = k * (a.Lmax + 1) + l;
        const int n = a.nqtab[idx];
        m += a.nodes[a.node_off[idx] + (int)(r % n)];
        r /= n;
      }
      a.V[q0 + e] = m;
    }
  }
}

// Triatomic model in Jacobi coordinates (R, r, theta), J = 0, type-10
// operator: V = D(1 - e^{-a(R-Re)})^2 + D(1 - e^{-a(r-re)})^2 + c cos^2 theta,
// G = diag(1/mu_R, 1/mu_r, 1/(mu_R R^2) + 1/(mu_r r^2)), Jac = sin theta,
// sqRhoOVERJac = 1, on every term grid (first index fastest; modes 0, 1, 2 =
// R, r, theta).  The first-call evaluation of get_OpGrid_type10_OF_ONEDP_FOR_SG4
// (SRC/sub_Operator/sub_OpPsi_SG4.f90:2536-2814) for this model.
struct JacArgs {
  const int32_t* levels;
  const int64_t* qoff;
  const double* nodes;
  const int64_t* node_off;
  const int32_t* nqtab;
  int Lmax;
  int64_t nterms;
  double muR, mur, D, a, Re, re, c;
  double *V, *G, *Jac, *sqRJ;
};
__global__ void k_grid_jacobi(JacArgs a) {
  for (int64_t t = blockIdx.x; t < a.nterms; t += gridDim.x) {
    const int64_t q0 = a.qoff[t], nqt = a.qoff[t + 1] - q0;
    for (int64_t e = threadIdx.x; e < nqt; e += blockDim.x) {
      int64_t r = e;
      double q[3];
      for (int k = 0; k < 3; ++k) {
        const int l = a.levels[t * 3 + k];
        const int idx = k * (a.Lmax + 1) + l;
        const int n = a.nqtab[idx];
        q[k] = a.nodes[a.node_off[idx] + (int)(r % n)];
        r /= n;
      }
      const double mR = 1.0 - exp(-a.a * (q[0] - a.Re));
      const double mr = 1.0 - exp(-a.a * (q[1] - a.re));
      const double ct = cos(q[2]);
      const int64_t p = q0 + e;
      a.V[p] = a.D * mR * mR + a.D * mr * mr + a.c * ct * ct;
      double* g = a.G + 9 * p;
      for (int i = 0; i < 9; ++i) g[i] = 0.0;
      g[0] = 1.0 / a.muR;
      g[4] = 1.0 / a.mur;
      g[8] = 1.0 / (a.muR * q[0] * q[0]) + 1.0 / (a.mur * q[1] * q[1]);
      a.Jac[p] = sin(q[2]);
      a.sqRJ[p] = 1.0;
    }
  }
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  ~DevBuf() { if (p) (void)hipFree(p); }
  int alloc(size_t bytes) {
    if (p) { (void)hipFree(p); p = nullptr; }
    n = bytes;
    if (bytes == 0) return SG4_OK;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of " + std::to_string(bytes) + " bytes failed");
      p = nullptr;
      return SG4_ERR_OOM;
    }
    return SG4_OK;
  }
  template <class T> T* as() const { return (T*)p; }
};

struct ItemPlan {
  DevBuf items, scratch;
  int64_t nitems = 0, nbig = 0, slot_size = 0;
  int lds = 0;
  // split path: items whose LDS need exceeds the budget, in a launch of their own
  DevBuf items2;
  int64_t nitems2 = 0;
  int lds2 = 0;
  double flops_main = 0;   // algorithmic flops of the main H.psi launch
  double flops_mfma = 0;   // MFMA path: flops the matrix cores execute in that launch
                           // (padded Kronecker tiles, six limb products)
};

// Launch configuration of the H.psi kernel; chosen by sg4_plan_autotune.
struct LaunchCfg {
  int nthreads = 256;          // 256 or 512
  int lds_budget = LDS_SMALL;  // bytes of work buffers per workgroup
  int grid = 0;                // 0: one workgroup per item, else persistent grid
  int xcd = 0;                 // 1: keep the items of a term on one XCD; k >= 2: keep a
                               //    class of terms with equal levels in modes k..D-1 on one XCD
};

}  // namespace sg4

struct sg4_plan {
  // one or several independent SG4 instances (a batch): their terms share the
  // work-item list and their packed rows are concatenated
  struct Inst { int D, Lmax; int64_t nterms, nb_ba, term0, row0, md0, ck0, lev0; };
  std::vector<Inst> inst;
  int nck = 0;          // kinetic coefficients: sum of the instances' D
  int D = 0, Lmax = 0;  // of the (first) instance; type 10 needs a single instance
  int64_t nterms = 0, nb_ba = 0, sum_nb = 0, sum_nq = 0;
  int max_nvec = 0;
  int type_op = 0;
  int device = 0;       // the GPU current when the plan was created
  hipStream_t stream = nullptr;
  std::vector<int32_t> h_levels, h_nq, h_nb;
  std::vector<int64_t> h_qoff, h_boff;
  std::vector<double> h_weights;
  std::vector<sg4::MatDev> h_md;
  std::vector<sg4::TermDev> h_terms;
  std::vector<double> h_ck;
  bool have_nodes = false;
  sg4::DevBuf d_mats, d_md, d_terms, d_map, d_V, d_ck, d_nodes, d_node_off, d_nqtab,
      d_levels, d_qoff;
  sg4::DevBuf d_mats32, d_V32, d_ck32;   // FP32 copies for the FP32 modes
  sg4::DevBuf d_G, d_Jac, d_sqRJ;        // type-10 grids
  sg4::DevBuf d_qw;                      // 1-D quadrature weights (offsets as nodes)
  sg4::DevBuf d_V0;                      // field-free potential (kept by sg4_plan_set_field)
  bool have_qw = false;
  bool f32_valid = false;
  int64_t nmats = 0;
  sg4::DevBuf d_rows_in, d_rows_out, d_cols_in, d_cols_out;
  std::map<int64_t, std::unique_ptr<sg4::ItemPlan>> items;  // per (nvec, size, budget, order)
  std::map<int, sg4::LaunchCfg> tuned;                       // per (nvec, precision)
  hipStream_t own_stream = nullptr;                          // for graph capture on the null stream
  struct GraphEntry { int nvec, prec; const double* psi; double* hpsi; hipGraphExec_t exec; };
  std::vector<GraphEntry> graphs;
  double flops_per_vec = 0, bytes_fixed = 0;
  int64_t last_items = -1;   // work items of the last apply
  double last_flops = 0;     // algorithmic flops of the last apply's main H.psi launch
  double last_flops_mfma = 0;  // MFMA path: flops the matrix cores executed in that launch
  bool grouping = true;      // register-blocked group passes for square terms
  bool no_alloc = false;     // inside sg4_apply_noalloc: use prepared state only
  // SG4_OPT_DETERMINISTIC: term contributions summed exactly in 128-bit fixed
  // point (order-independent); d_fx holds 2 words per result element and the
  // overflow counter at the end
  bool deterministic = false;
  sg4::DevBuf d_fx;
  // general type-1 operator (type_op 11): derivative terms and their grids
  int nder = 0;
  sg4::DevBuf d_der, d_F;
  // split type-1 path: kinetic sparse matrix K (CSR over packed rows) and the
  // host copies of the map and matrices it is assembled from; split_ok: K is
  // built (the algebra holds), split_fp: the FP64/FP32 term kernels also
  // have a register-blocked shape for every term and fit its LDS
  bool split_ok = false;
  bool split_fp = false;
  int64_t kin_nnz = 0;
  std::vector<int32_t> h_map;
  std::vector<double> h_mats;
  sg4::DevBuf d_kptr, d_kcol, d_kval;
  // split-precision MFMA path (SG4_PREC_F32X2): terms cut into mode groups,
  // group matrices in A-operand lane order
  bool mfma_valid = false;
  std::vector<sg4::TermM> h_termm;
  sg4::DevBuf d_termm, d_mtab;
  int64_t nmtab = 0;
  // optional kernel timing: HIP events around the H.psi launch of each apply
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pairs;
  size_t ev_used = 0;
  ~sg4_plan() {
    for (auto& e : ev_pairs) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }
};

namespace sg4 {

void* plan_stream(const sg4_plan* p) { return p->stream; }
int64_t plan_rows(const sg4_plan* p) { return p->nb_ba; }
int plan_max_nvec(const sg4_plan* p) { return p->max_nvec; }

// The allocation-free entry point found no prepared state for its call.
static int not_prepared() {
  set_error("sg4_apply_noalloc: the plan holds no work items (or FP32 / matrix-core tables) "
            "for this block size, precision and launch configuration; one sg4_apply call "
            "with them prepares the plan");
  return SG4_ERR_STATE;
}

// Captured graphs hold kernel arguments that point at plan buffers (item
// lists, the term and matrix tables, the kinetic matrix, the FP32 copies) and
// the launch configuration of their capture: every call that reallocates
// those buffers or changes the configuration destroys them, and the next
// sg4_apply_graph captures again.
static void drop_graphs(sg4_plan* P) {
  if (P->graphs.empty()) return;
  if (P->own_stream) (void)hipStreamSynchronize(P->own_stream);
  for (auto& g : P->graphs) (void)hipGraphExecDestroy(g.exec);
  P->graphs.clear();
}

// The operator or the term set changed: everything derived from them.
static void invalidate(sg4_plan* P) {
  drop_graphs(P);
  P->items.clear();
  P->f32_valid = false;
  P->mfma_valid = false;
}

// Greedy grouping of consecutive active modes into the shapes of
// sg4_shapes.h (triples, then pairs, then singles); 0 if some mode is not
// square or has no single-mode shape (the term then uses one pass per mode).
static int make_groups(const sg4_plan* P, TermDev& T) {
  static const int shapes[][3] = {
#define SG4_ROW(id, a, b, d) {a, b, d},
      SG4_SHAPES(SG4_ROW)
#undef SG4_ROW
  };
  auto find = [&](int a, int b, int d) {
    for (int i = 0; i < SG4_NSHAPES; ++i)
      if (shapes[i][0] == a && shapes[i][1] == b && shapes[i][2] == d) return i;
    return -1;
  };
  int n[MAXACT];
  for (int p = 0; p < T.nact; ++p) {
    const MatDev& m = P->h_md[T.mat[p]];
    if (m.nq != m.nb) return 0;
    n[p] = m.nq;
  }
  int g = 0, p = 0;
  while (p < T.nact) {
    int id = -1, len = 0;
    if (p + 2 < T.nact && (id = find(n[p], n[p + 1], n[p + 2])) >= 0) len = 3;
    else if (p + 1 < T.nact && (id = find(n[p], n[p + 1], 1)) >= 0) len = 2;
    else if ((id = find(n[p], 1, 1)) >= 0) len = 1;
    else return 0;
    T.gfirst[g] = (int8_t)p;
    T.gshape[g] = (int8_t)id;
    ++g;
    p += len;
  }
  return g;
}

static double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// Kinetic matrix of the split type-1 path (see the device comment above
// split_pass).  For a packed row r with mode indices i_j, let lambda_j be the
// smallest level whose basis holds i_j and Lambda_k = sum_{j != k} lambda_j.
// The terms holding r and a row r' that differs in mode k only are those with
// l_j >= lambda_j (j != k) and l_k >= the levels of i_k and i'_k; their
// combination weights W(|l|) = (-1)^(L-|l|) C(D-1, L-|l|) sum, for fixed l_k,
// to sum_e C(e+D-2, D-2) W(l_k + Lambda_k + e), the coefficient of z^u in
// (1-z)^-(D-1) (1-z)^(D-1) = 1 with u = L - l_k - Lambda_k, i.e. to 1 for
// l_k = L - Lambda_k and 0 otherwise.  Hence
//   K[r][r'] = c_k T_k(L - Lambda_k)[i_k][i'_k],  T_k(m) = BtW D2 B at level m,
// read off the term l* = (lambda_1 .. L - Lambda_k .. lambda_D) (|l*| = L),
// whose map holds both rows.  Every (row, mode) pair is visited once: term l*
// with |l*| = L, mode k, mode indices j != k in the top level band of l*_j.
// Entries of T_k(m) below 1e-12 max|T_k(m)| are quadrature rounding of
// integrals that vanish exactly (odd parity, and |i - j| > 2 for HO bases;
// they are ~1e-16 relative, the others >= 1e-2) and are not stored.
// The split path is enabled only when every 1-D basis has B^T W B = I to
// 1e-11 and the terms are exactly the Smolyak set {L-D+1 <= |l| <= L} with
// these weights; the FP64/FP32 split kernels need in addition a shape for
// every term's mode groups (split_fp), the matrix-core path does not.
static int build_kinetic(sg4_plan* P) {
  P->split_ok = false;
  P->split_fp = false;
  if (P->type_op != 1 || !P->grouping) return SG4_OK;
  bool fp = true;
  for (const TermDev& T : P->h_terms) {
    if (T.nact > 0 && T.ngrp == 0) fp = false;
    if (20.0 * T.nq > (double)LDS_MAX) fp = false;   // FP64: one vector + potential + map
  }
  const size_t nmd = P->h_md.size();
  std::vector<int64_t> toff(nmd);
  std::vector<double> tm, tcut(nmd);
  for (size_t i = 0; i < nmd; ++i) {
    const MatDev& m = P->h_md[i];
    if (m.nq != m.nb) return SG4_OK;
    const int n = m.nq;
    const double* B = &P->h_mats[m.B];     // n x n, row-major
    const double* Bt = &P->h_mats[m.BtW];
    const double* D2 = &P->h_mats[m.D2];
    std::vector<long double> DB((size_t)n * n);
    for (int a = 0; a < n; ++a)
      for (int j = 0; j < n; ++j) {
        long double sI = 0, sD = 0;
        for (int b = 0; b < n; ++b) {
          sI += (long double)Bt[a * n + b] * B[b * n + j];
          sD += (long double)D2[a * n + b] * B[b * n + j];
        }
        if (std::fabs((double)sI - (a == j ? 1.0 : 0.0)) > 1e-11) return SG4_OK;
        DB[(size_t)a * n + j] = sD;
      }
    toff[i] = (int64_t)tm.size();
    double tmax = 0.0;
    for (int a = 0; a < n; ++a)
      for (int j = 0; j < n; ++j) {
        long double t = 0;
        for (int b = 0; b < n; ++b) t += (long double)Bt[a * n + b] * DB[(size_t)b * n + j];
        tm.push_back((double)t);
        tmax = std::max(tmax, std::fabs((double)t));
      }
    tcut[i] = 1e-12 * tmax;
  }
  // Smolyak structure of every instance
  std::vector<int> Linst(P->inst.size());
  for (size_t ii = 0; ii < P->inst.size(); ++ii) {
    const sg4_plan::Inst& I = P->inst[ii];
    const int D = I.D;
    int L = 0;
    for (int64_t tl = 0; tl < I.nterms; ++tl) {
      int sl = 0;
      for (int k = 0; k < D; ++k) sl += P->h_levels[I.lev0 + tl * D + k];
      L = std::max(L, sl);
    }
    if (L > I.Lmax) return SG4_OK;
    double cnt = 0;
    for (int sl = std::max(0, L - D + 1); sl <= L; ++sl) cnt += binom(sl + D - 1, D - 1);
    if (cnt != (double)I.nterms) return SG4_OK;
    for (int64_t tl = 0; tl < I.nterms; ++tl) {
      int sl = 0;
      for (int k = 0; k < D; ++k) sl += P->h_levels[I.lev0 + tl * D + k];
      const double wexp = (((L - sl) & 1) ? -1.0 : 1.0) * binom(D - 1, L - sl);
      if (std::fabs(P->h_weights[I.term0 + tl] - wexp) > 1e-12 * std::fabs(wexp)) return SG4_OK;
    }
    Linst[ii] = L;
  }
  // two sweeps over (row, mode) pairs: count, then fill
  const int64_t nrows = P->nb_ba;
  std::vector<int64_t> ptr(nrows + 1, 0);
  std::vector<int32_t> col;
  std::vector<double> val;
  std::vector<double> diag;
  for (int sweep = 0; sweep < 2; ++sweep) {
    std::vector<int64_t> pos;
    if (sweep == 1) {
      for (int64_t r = 0; r < nrows; ++r) ptr[r + 1] += ptr[r] + 1;   // + the diagonal
      col.assign(ptr[nrows], 0);
      val.assign(ptr[nrows], 0.0);
      diag.assign(nrows, 0.0);
      pos.resize(nrows);
      for (int64_t r = 0; r < nrows; ++r) pos[r] = ptr[r] + 1;
    }
    for (size_t ii = 0; ii < P->inst.size(); ++ii) {
      const sg4_plan::Inst& I = P->inst[ii];
      const int D = I.D, L1 = I.Lmax + 1, L = Linst[ii];
      auto nbk = [&](int k, int l) { return l < 0 ? 0 : P->h_nb[I.md0 + k * L1 + l]; };
      int lv[SG4_MAX_D], lo[SG4_MAX_D], hi[SG4_MAX_D], dig[SG4_MAX_D];
      int64_t stride[SG4_MAX_D + 1];
      for (int64_t tl = 0; tl < I.nterms; ++tl) {
        int sl = 0;
        for (int k = 0; k < D; ++k) sl += (lv[k] = P->h_levels[I.lev0 + tl * D + k]);
        if (sl != L) continue;
        const int64_t t = I.term0 + tl;
        const int32_t* mp = &P->h_map[P->h_boff[t]];
        stride[0] = 1;
        for (int k = 0; k < D; ++k) stride[k + 1] = stride[k] * nbk(k, lv[k]);
        for (int k = 0; k < D; ++k) {
          bool empty = false;
          for (int j = 0; j < D; ++j) {
            lo[j] = (j == k) ? 0 : nbk(j, lv[j] - 1);
            hi[j] = nbk(j, lv[j]);
            if (lo[j] >= hi[j]) empty = true;
            dig[j] = lo[j];
          }
          if (empty) continue;
          const int nk = hi[k];
          const int64_t mi = I.md0 + k * L1 + lv[k];
          const double* Tk = &tm[toff[mi]];
          const double cut = tcut[mi];
          const double ck = P->h_ck[I.ck0 + k];
          int64_t e = 0;
          for (int j = 0; j < D; ++j) e += dig[j] * stride[j];
          while (true) {
            const int32_t r = mp[e];
            if (r >= 0) {
              const int ik = dig[k];
              for (int i2 = 0; i2 < nk; ++i2) {
                if (i2 == ik || std::fabs(Tk[ik * nk + i2]) <= cut) continue;
                const int32_t r2 = mp[e + (int64_t)(i2 - ik) * stride[k]];
                if (r2 < 0) continue;
                if (sweep == 0) {
                  ++ptr[r + 1];
                } else {
                  col[pos[r]] = r2;
                  val[pos[r]] = ck * Tk[ik * nk + i2];
                  ++pos[r];
                }
              }
              if (sweep == 1) diag[r] += ck * Tk[ik * nk + ik];
            }
            int j = 0;
            for (; j < D; ++j) {
              ++dig[j];
              e += stride[j];
              if (dig[j] < hi[j]) break;
              e -= (int64_t)(dig[j] - lo[j]) * stride[j];
              dig[j] = lo[j];
            }
            if (j == D) break;
          }
        }
      }
    }
    if (sweep == 1)
      for (int64_t r = 0; r < nrows; ++r) {
        col[ptr[r]] = (int32_t)r;
        val[ptr[r]] = diag[r];
        if (pos[r] != ptr[r + 1]) { set_error("internal: kinetic matrix row count"); return SG4_ERR_STATE; }
      }
  }
  int rc;
  if ((rc = P->d_kptr.alloc(sizeof(int64_t) * (nrows + 1)))) return rc;
  if ((rc = P->d_kcol.alloc(sizeof(int32_t) * col.size()))) return rc;
  if ((rc = P->d_kval.alloc(sizeof(double) * val.size()))) return rc;
  HIP_TRY(hipMemcpy(P->d_kptr.p, ptr.data(), sizeof(int64_t) * (nrows + 1), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(P->d_kcol.p, col.data(), sizeof(int32_t) * col.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(P->d_kval.p, val.data(), sizeof(double) * val.size(), hipMemcpyHostToDevice));
  P->kin_nnz = (int64_t)col.size();
  P->split_ok = true;
  P->split_fp = fp;
  return SG4_OK;
}

static int build_terms(sg4_plan* P) {
  P->mfma_valid = false;
  P->h_terms.resize(P->nterms);
  std::vector<double> ck = P->h_ck;
  if ((int)ck.size() < P->nck) ck.assign(P->nck, 0.0);
  for (const sg4_plan::Inst& I : P->inst) {
    const int D = I.D, L1 = I.Lmax + 1;
    for (int64_t tl = 0; tl < I.nterms; ++tl) {
      const int64_t t = I.term0 + tl;
      TermDev T{};
      T.boff = P->h_boff[t];
      T.qoff = P->h_qoff[t];
      T.nb = (int32_t)(P->h_boff[t + 1] - P->h_boff[t]);
      T.nq = (int32_t)(P->h_qoff[t + 1] - P->h_qoff[t]);
      double w = P->h_weights[t], vs = 0.0;
      int nact = 0;
      for (int k = 0; k < D; ++k) {
        const int l = P->h_levels[(size_t)I.lev0 + (size_t)tl * D + k];
        const int64_t idx = I.md0 + k * L1 + l;
        const MatDev& m = P->h_md[idx];
        if (m.nq == 1 && m.nb == 1 && P->type_op == 1) {
          w *= m.b00 * m.btw00;
          vs += ck[I.ck0 + k] * m.d200;
        } else {
          if (nact >= MAXACT) {
            set_error("more than 16 modes with more than one basis function in a term");
            return SG4_ERR_UNSUPPORTED;
          }
          if (idx > INT16_MAX) { set_error("too many 1-D bases in the batch"); return SG4_ERR_UNSUPPORTED; }
          T.mat[nact] = (int16_t)idx;
          T.kmode[nact] = (int16_t)(I.ck0 + k);
          ++nact;
        }
      }
      T.nact = nact;
      T.wgt = w;
      T.vshift = vs;
      T.ngrp = (P->grouping && P->type_op == 1 && nact > 0) ? make_groups(P, T) : 0;
      P->h_terms[t] = T;
    }
  }
  int rc = P->d_terms.alloc(sizeof(TermDev) * P->h_terms.size());
  if (rc) return rc;
  HIP_TRY(hipMemcpy(P->d_terms.p, P->h_terms.data(), sizeof(TermDev) * P->h_terms.size(),
                    hipMemcpyHostToDevice));
  return build_kinetic(P);
}

struct HItem { Item it; double cost; bool large; };

// Item order of a launch: by decreasing cost, or (xcd >= 1) dealt to the
// eight XCDs (workgroups go round-robin to the XCDs) so that the items of a
// term (xcd = 1) or of a class of terms with equal levels in modes
// xcd..D-1 (xcd >= 2) run on one XCD and share its L2 for their gathers.
static std::vector<Item> order_items(const sg4_plan* P, const std::vector<HItem>& all, bool large,
                                     int xcd) {
  std::vector<Item> h;
  if (!xcd) {
    for (auto& x : all)
      if (x.large == large) h.push_back(x.it);
    return h;
  }
  std::vector<std::vector<Item>> lane(8);
  double acc[8] = {0};
  if (xcd >= 2) {
    // terms with equal levels in modes xcd..D-1 share the rows that agree in
    // those modes: put such a class on one XCD, its items next to each other
    // (classes by decreasing largest item, items by decreasing cost)
    std::map<std::vector<int32_t>, int> cls_of;
    std::vector<std::vector<const HItem*>> cls;
    for (auto& x : all) {
      if (x.large != large) continue;
      int64_t t = x.it.term;
      size_t ii = 0;
      while (ii + 1 < P->inst.size() && t >= P->inst[ii + 1].term0) ++ii;
      const sg4_plan::Inst& I = P->inst[ii];
      std::vector<int32_t> key(1, (int32_t)ii);
      for (int k = xcd; k < I.D; ++k) key.push_back(P->h_levels[I.lev0 + (t - I.term0) * I.D + k]);
      auto f = cls_of.find(key);
      int c;
      if (f == cls_of.end()) {
        c = (int)cls.size();
        cls_of[key] = c;
        cls.emplace_back();
      } else {
        c = f->second;
      }
      cls[c].push_back(&x);
    }
    for (auto& members : cls) {
      double w = 0;
      for (const HItem* x : members) w += x->cost;
      const int l = (int)(std::min_element(acc, acc + 8) - acc);
      for (const HItem* x : members) lane[l].push_back(x->it);
      acc[l] += w;
    }
  } else {
    std::map<int32_t, int> owner;
    for (auto& x : all) {
      if (x.large != large) continue;
      auto o = owner.find(x.it.term);
      int l;
      if (o == owner.end()) {
        l = (int)(std::min_element(acc, acc + 8) - acc);
        owner[x.it.term] = l;
      } else {
        l = o->second;
      }
      lane[l].push_back(x.it);
      acc[l] += x.cost;
    }
  }
  size_t m = 0;
  for (auto& v : lane) m = std::max(m, v.size());
  Item pad{};
  pad.term = -1;
  for (size_t i = 0; i < m; ++i)
    for (int l = 0; l < 8; ++l) h.push_back(i < lane[l].size() ? lane[l][i] : pad);
  return h;
}

// Work items for a block of nvec vectors: for every term the vectors are cut
// into power-of-two chunks so that the work buffers fit the LDS budget; a
// term too large for that even with one vector keeps its second buffer in a
// global scratch slot.  Items are ordered by decreasing work so that the long
// ones start first.  With xcd = 1 the items of one term are placed at
// positions congruent modulo 8 (workgroups are dealt round-robin to the eight
// XCDs), so the chunks of a term share one L2 for their gathers; terms are
// given to the XCD with the least accumulated work.  With xcd = k >= 2 the unit kept on one XCD is a
// class of terms with equal levels in modes k..D-1 (they share many rows), its
// items consecutive in the lane.
// Split path (split = true): an item needs (nv + 1) N values of LDS (the
// chunk and the potential) and N map entries; a term that does not fit the budget with one
// vector goes to the second list, launched with its own LDS size.
static int get_items(sg4_plan* P, int nvec, int esize, ItemPlan** out, int nbuf_force = 0,
                     int budget = LDS_SMALL, int xcd = 0, bool split = false) {
  const int64_t key =
      ((((int64_t)nvec * 16 + esize + (nbuf_force ? 1 : 0)) * 1024 + budget / 1024) * 32 + xcd) * 2 +
      (split ? 1 : 0);
  auto itp = P->items.find(key);
  if (itp != P->items.end()) { *out = itp->second.get(); return SG4_OK; }
  if (P->no_alloc) return not_prepared();
  std::vector<HItem> all;
  double flops_main = 0;
  all.reserve(P->nterms * 2);
  int64_t nbig = 0, slot = 0;
  int lds = 0, lds2 = 0;
  if (split) {
    for (int64_t t = 0; t < P->nterms; ++t) {
      const TermDev& T = P->h_terms[t];
      const int64_t n = (T.nact == 0) ? 0 : T.nq;
      const bool large = 2 * n * esize + 4 * n > budget;
      const int cap = large ? LDS_MAX : budget;
      int gmax = 0;
      while ((1 << (gmax + 1)) <= nvec && ((1 << (gmax + 1)) + 1) * n * esize + 4 * n <= cap)
        ++gmax;
      double work = 0;
      for (int q = 0; q < T.nact; ++q) work += P->h_md[T.mat[q]].nq;
      work = (work + 1.0) * (double)std::max<int64_t>(n, 1);
      for (int v0 = 0; v0 < nvec;) {
        int g = gmax;
        while ((1 << g) > nvec - v0) --g;
        HItem h{};
        h.it.term = (int32_t)t;
        h.it.v0 = (int16_t)v0;
        h.it.lgnv = (int8_t)g;
        h.cost = work * (1 << g);
        h.large = large;
        if (!large) {   // basis->grid + grid->basis products and the potential
          double f = 0;
          for (int q = 0; q < T.nact; ++q) f += 4.0 * (double)n * P->h_md[T.mat[q]].nq;
          flops_main += (f + (double)std::max<int64_t>(n, 1)) * (1 << g);
        }
        int& l = large ? lds2 : lds;
        l = std::max<int>(l, (int)(((1 << g) + 1) * n * esize + 4 * n));
        all.push_back(h);
        v0 += 1 << g;
      }
    }
  }
  for (int64_t t = 0; t < (split ? 0 : P->nterms); ++t) {
    const TermDev& T = P->h_terms[t];
    const int64_t nmax = std::max(T.nq, T.nb);
    const int nbuf = nbuf_force ? nbuf_force : ((P->type_op == 10) ? P->D + 3 : 2);
    const int64_t per = nbuf * nmax * esize;  // bytes per vector (work buffers)
    const bool big = per > budget;
    if (big && (P->type_op == 10 || nbuf_force)) {
      set_error("term too large for the LDS work buffers of this kernel");
      return SG4_ERR_UNSUPPORTED;
    }
    if (big && nmax * esize > budget) {
      set_error("term with " + std::to_string(nmax) + " grid points exceeds the LDS budget");
      return SG4_ERR_UNSUPPORTED;
    }
    int gmax = 0;
    if (!big)
      while ((1 << (gmax + 1)) <= nvec && per * (1 << (gmax + 1)) <= budget) ++gmax;
    double work = 0;
    for (int p = 0; p < T.nact; ++p) work += P->h_md[T.mat[p]].nq;
    work = (work + 1.0) * (double)nmax;
    int v0 = 0;
    while (v0 < nvec) {
      int g = gmax;
      while ((1 << g) > nvec - v0) --g;
      HItem h{};
      h.it.term = (int32_t)t;
      h.it.v0 = (int16_t)v0;
      h.it.lgnv = (int8_t)g;
      h.it.big = big ? 1 : 0;
      h.it.slot = big ? (int32_t)(nbig++) : 0;
      lds = std::max<int>(lds, (int)(big ? nmax * esize : per << g));
      slot = std::max<int64_t>(slot, big ? nmax : 0);
      h.cost = work * (1 << g);
      all.push_back(h);
      v0 += 1 << g;
    }
  }
  std::stable_sort(all.begin(), all.end(), [](const HItem& x, const HItem& y) { return x.cost > y.cost; });
  const std::vector<Item> h = order_items(P, all, false, xcd);
  const std::vector<Item> h2 = order_items(P, all, true, xcd);
  auto ip = std::make_unique<ItemPlan>();
  ip->nitems = (int64_t)h.size();
  ip->nbig = nbig;
  ip->slot_size = slot;
  ip->lds = lds;
  ip->nitems2 = (int64_t)h2.size();
  ip->lds2 = lds2;
  ip->flops_main = split ? flops_main
                         : (P->type_op == 10 ? 4.0 : 3.0) * P->flops_per_vec * nvec;
  int rc = ip->items.alloc(sizeof(Item) * h.size());
  if (rc) return rc;
  HIP_TRY(hipMemcpy(ip->items.p, h.data(), sizeof(Item) * h.size(), hipMemcpyHostToDevice));
  if (!h2.empty()) {
    if ((rc = ip->items2.alloc(sizeof(Item) * h2.size()))) return rc;
    HIP_TRY(hipMemcpy(ip->items2.p, h2.data(), sizeof(Item) * h2.size(), hipMemcpyHostToDevice));
  }
  if (nbig > 0 && (rc = ip->scratch.alloc((size_t)esize * nbig * slot))) return rc;
  *out = ip.get();
  P->items[key] = std::move(ip);
  return SG4_OK;
}

static int grid_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 8192);
}

static bool use_split(const sg4_plan* P);
static bool use_split_alg(const sg4_plan* P);

// bf16 limbs of the group matrices (round to nearest even)
static uint16_t to_bf16(double d) {
  const float f = (float)d;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static double from_bf16(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Terms and group matrices of the split-precision MFMA path.  The active
// modes of a term are cut into groups of consecutive modes with Kronecker
// size K <= 32 (the cut with the least padded work); the K x K matrices B_g = B_a (x) B_b (x) ... and
// BtW_g (first index fastest, as the in-term index) are split into three
// bf16 limbs and stored in the A-operand order of the lanes (limb, 16-row
// tile, lane, slot: row 16 mt + lane%16, group index slot_k(lane/16, slot)),
// zero-padded to 16 or 32.  Modes with equal 1-D matrices share group
// matrices.
static int build_mfma(sg4_plan* P) {
  if (P->mfma_valid) return SG4_OK;
  if (P->no_alloc) return not_prepared();
  if (!use_split_alg(P)) {
    set_error("the split-precision MFMA mode needs the split type-1 path (square HO/Legendre "
              "modes with exact quadrature, complete Smolyak term set)");
    return SG4_ERR_UNSUPPORTED;
  }
  const size_t nmd = P->h_md.size();
  std::vector<int> canon(nmd);
  for (size_t i = 0; i < nmd; ++i) {
    canon[i] = (int)i;
    const MatDev& a = P->h_md[i];
    const size_t nn = (size_t)a.nq * a.nb;
    for (size_t q = 0; q < i; ++q) {
      if (canon[q] != (int)q) continue;
      const MatDev& b = P->h_md[q];
      if (a.nq != b.nq || a.nb != b.nb) continue;
      if (std::equal(&P->h_mats[a.B], &P->h_mats[a.B] + nn, &P->h_mats[b.B]) &&
          std::equal(&P->h_mats[a.BtW], &P->h_mats[a.BtW] + nn, &P->h_mats[b.BtW])) {
        canon[i] = (int)q;
        break;
      }
    }
  }
  std::map<std::vector<int>, int> gid;
  std::vector<uint16_t> tab;
  auto put = [&](const std::vector<int>& key, bool btw) {
    int K = 1;
    for (int q : key) K *= P->h_md[q].nq;
    const int S = (K <= 16) ? 4 : 8, MT = S / 4;
    const size_t base = tab.size();
    tab.resize(base + MAT_STRIDE, 0);
    for (int mt = 0; mt < MT; ++mt)
      for (int lane = 0; lane < 64; ++lane)
        for (int sl = 0; sl < S; ++sl) {
          const int row = 16 * mt + (lane & 15), j = lane >> 4;
          const int k = (S == 4) ? 4 * j + sl : (sl < 4 ? 4 * j + sl : 12 + 4 * j + sl);
          double a = 0.0;
          if (row < K && k < K) {
            a = 1.0;
            int mr = row, kr = k;
            for (int q : key) {
              const MatDev& md = P->h_md[q];
              const int n = md.nq;
              a *= P->h_mats[(btw ? md.BtW : md.B) + (int64_t)(mr % n) * n + kr % n];
              mr /= n;
              kr /= n;
            }
          }
          const uint16_t h = to_bf16(a);
          const double r = a - from_bf16(h);
          const uint16_t m = to_bf16(r);
          const uint16_t l = to_bf16(r - from_bf16(m));
          const uint16_t limb[3] = {h, m, l};
          for (int li = 0; li < 3; ++li) tab[base + ((size_t)(li * MT + mt) * 64 + lane) * S + sl] = limb[li];
        }
  };
  std::vector<TermM> tm(P->nterms);
  for (int64_t t = 0; t < P->nterms; ++t) {
    const TermDev& T = P->h_terms[t];
    TermM M{};
    M.boff = T.boff;
    M.qoff = T.qoff;
    M.wgt = T.wgt;
    M.N = T.nq;
    // cut the active modes into consecutive groups with K <= 32, minimising
    // the per-point vector work of the passes: a 16-column tile costs about
    // 47 vector instructions for K <= 16 (x16 products) and 73 for K <= 32
    // (x32), spread over the K valid rows of the tile
    const int na = T.nact;
    std::vector<double> best(na + 1, 1e300);
    std::vector<int> from(na + 1, -1);
    best[0] = 0.0;
    for (int i = 1; i <= na; ++i) {
      int K = 1;
      for (int q = i - 1; q >= 0; --q) {
        K *= P->h_md[T.mat[q]].nq;
        if (K > 32) break;
        const double c = best[q] + (K <= 16 ? 47.0 : 73.0) / K;
        if (c < best[i]) { best[i] = c; from[i] = q; }
      }
    }
    if (na > 0 && from[na] < 0) {
      set_error("a mode with more than 32 grid points has no MFMA tile");
      return SG4_ERR_UNSUPPORTED;
    }
    std::vector<int> cut;
    for (int i = na; i > 0; i = from[i]) cut.push_back(from[i]);
    std::reverse(cut.begin(), cut.end());
    cut.push_back(na);
    int G = 0, A = 1;
    for (size_t gi = 0; gi + 1 < cut.size(); ++gi) {
      std::vector<int> key;
      int K = 1;
      for (int p = cut[gi]; p < cut[gi + 1]; ++p) {
        K *= P->h_md[T.mat[p]].nq;
        key.push_back(canon[T.mat[p]]);
      }
      if (G >= MG_MAX) { set_error("too many mode groups in a term"); return SG4_ERR_UNSUPPORTED; }
      auto f = gid.find(key);
      int idx;
      if (f == gid.end()) {
        idx = (int)(tab.size() / MAT_STRIDE);
        put(key, false);
        put(key, true);
        gid[key] = idx;
      } else {
        idx = f->second;
      }
      M.A[G] = A;
      M.K[G] = K;
      M.mB[G] = idx;
      ++G;
      A *= K;
    }
    M.G = G;
    tm[t] = M;
  }
  int rc;
  if ((rc = P->d_termm.alloc(sizeof(TermM) * tm.size()))) return rc;
  HIP_TRY(hipMemcpy(P->d_termm.p, tm.data(), sizeof(TermM) * tm.size(), hipMemcpyHostToDevice));
  if ((rc = P->d_mtab.alloc(sizeof(uint16_t) * std::max<size_t>(tab.size(), 1)))) return rc;
  if (!tab.empty())
    HIP_TRY(hipMemcpy(P->d_mtab.p, tab.data(), sizeof(uint16_t) * tab.size(), hipMemcpyHostToDevice));
  P->nmtab = (int64_t)(tab.size() / MAT_STRIDE);
  P->h_termm.swap(tm);
  P->mfma_valid = true;
  return SG4_OK;
}

// Work items of the MFMA path: per term, chunks of nv = 2^g <= 64 vectors
// with (4 RS + 8) N + 256 bytes of LDS (the FP32 tensor with row stride RS,
// the FP32 potential, the map, scratch words) within the budget; a term that needs more than the budget with one
// vector goes to the second list, launched with the full LDS.
static int get_items_mfma(sg4_plan* P, int nvec, int budget, int xcd, ItemPlan** out) {
  const int64_t key = -(((((int64_t)nvec * 1024 + budget / 1024) * 32 + xcd) * 2) + 1);
  auto itp = P->items.find(key);
  if (itp != P->items.end()) { *out = itp->second.get(); return SG4_OK; }
  if (P->no_alloc) return not_prepared();
  std::vector<HItem> all;
  double flops_main = 0, flops_mfma = 0;
  int lds = 0, lds2 = 0;
  for (int64_t t = 0; t < P->nterms; ++t) {
    const TermM& M = P->h_termm[t];
    const TermDev& T = P->h_terms[t];
    const int64_t N = M.G ? M.N : 0;
    // LDS bytes: tensor (row stride RS >= nv floats), potential, map, 64 scratch words
    auto need = [&](int g) { return (int64_t)(4 * mrow_stride(1 << g) + 8) * N + 256; };
    if (need(0) > LDS_MAX) {
      set_error("term with " + std::to_string(N) + " grid points exceeds the LDS of the MFMA path");
      return SG4_ERR_UNSUPPORTED;
    }
    const bool large = need(0) > budget;
    const int cap = large ? LDS_MAX : budget;
    int gmax = 0;
    while ((1 << (gmax + 1)) <= std::min(nvec, 64) && need(gmax + 1) <= cap) ++gmax;
    double c1 = 0, fm1 = 0;   // per vector: cost model, MFMA flops
    for (int g = 0; g < M.G; ++g) {
      const double kp = M.K[g] <= 16 ? 16.0 : 32.0;
      c1 += (double)N * (2.0 + 0.05 * kp * kp / M.K[g]);
      fm1 += 2.0 * 6.0 * 2.0 * kp * kp * ((double)N / M.K[g]);
    }
    c1 = std::max(c1, 1.0);
    double fa1 = 0;
    for (int q = 0; q < T.nact; ++q) fa1 += 4.0 * (double)N * P->h_md[T.mat[q]].nq;
    fa1 += (double)std::max<int64_t>(N, 1);
    for (int v0 = 0; v0 < nvec;) {
      int g = gmax;
      while ((1 << g) > nvec - v0) --g;
      HItem h{};
      h.it.term = (int32_t)t;
      h.it.v0 = (int16_t)v0;
      h.it.lgnv = (int8_t)g;
      h.cost = c1 * (1 << g);
      h.large = large;
      if (!large) {
        flops_main += fa1 * (1 << g);
        flops_mfma += fm1 * (1 << g);
      }
      int& l = large ? lds2 : lds;
      l = std::max<int>(l, (int)need(g));
      all.push_back(h);
      v0 += 1 << g;
    }
  }
  std::stable_sort(all.begin(), all.end(), [](const HItem& x, const HItem& y) { return x.cost > y.cost; });
  const std::vector<Item> h = order_items(P, all, false, xcd);
  const std::vector<Item> h2 = order_items(P, all, true, xcd);
  auto ip = std::make_unique<ItemPlan>();
  ip->nitems = (int64_t)h.size();
  ip->lds = std::max(lds, 16);
  ip->nitems2 = (int64_t)h2.size();
  ip->lds2 = lds2;
  ip->flops_main = flops_main;
  ip->flops_mfma = flops_mfma;
  int rc = ip->items.alloc(sizeof(Item) * std::max<size_t>(h.size(), 1));
  if (rc) return rc;
  if (!h.empty()) HIP_TRY(hipMemcpy(ip->items.p, h.data(), sizeof(Item) * h.size(), hipMemcpyHostToDevice));
  if (!h2.empty()) {
    if ((rc = ip->items2.alloc(sizeof(Item) * h2.size()))) return rc;
    HIP_TRY(hipMemcpy(ip->items2.p, h2.data(), sizeof(Item) * h2.size(), hipMemcpyHostToDevice));
  }
  *out = ip.get();
  P->items[key] = std::move(ip);
  return SG4_OK;
}

static int ensure_f32(sg4_plan* P) {
  if (P->f32_valid) return SG4_OK;
  if (P->no_alloc && (!P->d_mats32.p || !P->d_V32.p || !P->d_ck32.p)) return not_prepared();
  int rc;
  if (!P->d_mats32.p && (rc = P->d_mats32.alloc(sizeof(float) * P->nmats))) return rc;
  if (!P->d_V32.p && (rc = P->d_V32.alloc(sizeof(float) * P->sum_nq))) return rc;
  if (!P->d_ck32.p && (rc = P->d_ck32.alloc(sizeof(float) * P->nck))) return rc;
  hipLaunchKernelGGL(k_to_f32, dim3(grid_blocks(P->nmats)), dim3(256), 0, P->stream,
                     P->d_mats.as<double>(), P->d_mats32.as<float>(), P->nmats);
  hipLaunchKernelGGL(k_to_f32, dim3(grid_blocks(P->sum_nq)), dim3(256), 0, P->stream,
                     P->d_V.as<double>(), P->d_V32.as<float>(), P->sum_nq);
  hipLaunchKernelGGL(k_to_f32, dim3(1), dim3(64), 0, P->stream, P->d_ck.as<double>(),
                     P->d_ck32.as<float>(), (int64_t)P->nck);
  HIP_TRY(hipGetLastError());
  P->f32_valid = true;
  return SG4_OK;
}

static int check_apply(sg4_plan* P, int nvec, int prec) {
  if (P->type_op == 0) { set_error("no operator set on the plan"); return SG4_ERR_STATE; }
  if (nvec < 1 || nvec > P->max_nvec) { set_error("nvec out of range"); return SG4_ERR_ARG; }
  if (prec != SG4_PREC_FP64 && prec != SG4_PREC_F32 && prec != SG4_PREC_F32X2) {
    set_error("unknown precision mode");
    return SG4_ERR_ARG;
  }
  if ((P->type_op == 10 || P->type_op == 11) && prec != SG4_PREC_FP64) {
    set_error("the type-10 and general type-1 operators are available in FP64 only");
    return SG4_ERR_UNSUPPORTED;
  }
  return SG4_OK;
}

// The split algebra (kinetic part as K, terms keep the potential part): the
// matrix-core path needs only this; the FP64/FP32 term kernels also need
// their shapes (use_split).
static bool use_split_alg(const sg4_plan* P) {
  return P->type_op == 1 && P->grouping && P->split_ok;
}

static bool use_split(const sg4_plan* P) { return use_split_alg(P) && P->split_fp; }

static bool use_split_for(const sg4_plan* P, int prec) {
  return prec == SG4_PREC_F32X2 ? use_split_alg(P) : use_split(P);
}

// Launch configuration for (nvec, prec): the one tuned or set for it, else
// the one set for the nearest block width of the same precision (the
// narrowest wider one first): the best launch shape depends on the term set
// far more than on the block width, so a block of 8 vectors in a Davidson
// iteration runs with the launch tuned for 16.  Defaults when none is set.
static LaunchCfg cfg_for(const sg4_plan* P, int nvec, int prec) {
  auto it = P->tuned.find(nvec * 16 + prec);
  if (it != P->tuned.end()) return it->second;
  const LaunchCfg* best = nullptr;
  int bw = 0;
  for (const auto& kv : P->tuned) {
    if (kv.first % 16 != prec) continue;
    const int w = kv.first / 16;
    const bool better = !best || (w > nvec && (bw < nvec || w < bw)) || (w < nvec && bw < nvec && w > bw);
    if (better) { best = &kv.second; bw = w; }
  }
  return best ? *best : LaunchCfg{};
}

// Everything an apply needs that may allocate or copy (FP32 copies, item
// lists); done before any stream capture.
static int prepare(sg4_plan* P, int nvec, int prec, const LaunchCfg& cfg, ItemPlan** ip) {
  int rc;
  if (P->deterministic) {
    if (!use_split_for(P, prec)) {
      set_error("deterministic accumulation is available on the split type-1 path only");
      return SG4_ERR_UNSUPPORTED;
    }
    const size_t need = 8 * (2 * (size_t)P->nb_ba * P->max_nvec + 2);
    if (P->d_fx.n < need) {
      if (P->no_alloc) return not_prepared();
      if ((rc = P->d_fx.alloc(need))) return rc;
    }
  }
  const int esize = (prec == SG4_PREC_FP64) ? 8 : 4;
  if (esize == 4 && (rc = ensure_f32(P))) return rc;
  if (prec == SG4_PREC_F32X2) {
    if ((rc = build_mfma(P))) return rc;
    return get_items_mfma(P, nvec, cfg.lds_budget, cfg.xcd, ip);
  }
  if (P->type_op == 11) return get_items(P, nvec, esize, ip, 4, LDS_MAX, 0, false);
  const bool t10 = P->type_op == 10;
  return get_items(P, nvec, esize, ip, 0, t10 ? LDS_SMALL : cfg.lds_budget, t10 ? 0 : cfg.xcd,
                   use_split(P));
}

// Width of the rows-layout staging block of a column-major apply of nvec
// vectors.  Work items carry 2^g vectors (as many as the LDS budget allows)
// and a term's item costs little more with 16 vectors than with 8 (its
// passes are latency-bound), so the
// block is padded with zero vectors to the next power of two (nvec <= 16) or
// multiple of 16 when the plan's max_nvec allows: 9 vectors, say, then run as
// one 16-vector item per term instead of an 8- and a 1-vector item (10D L=8:
// 5.3 instead of 7.8 ms).
static int pad_width(const sg4_plan* p, int nvec) {
  int W = 1;
  if (nvec <= 16) {
    while (W < nvec) W <<= 1;
  } else {
    W = 16 * ((nvec + 15) / 16);
  }
  return W <= p->max_nvec ? W : nvec;
}

static int apply_rows(sg4_plan* P, int nvec, const double* psi, double* hpsi, int prec,
                      const LaunchCfg* cfgp = nullptr) {
  int rc;
  if ((rc = check_apply(P, nvec, prec))) return rc;
  const LaunchCfg cfg = cfgp ? *cfgp : cfg_for(P, nvec, prec);
  ItemPlan* ip = nullptr;
  if ((rc = prepare(P, nvec, prec, cfg, &ip))) return rc;
  const int esize = (prec == SG4_PREC_FP64) ? 8 : 4;
  const int64_t n = P->nb_ba * nvec;
  const bool split = use_split_for(P, prec);
  if (!split) hipLaunchKernelGGL(k_fill, dim3(grid_blocks(n)), dim3(256), 0, P->stream, hpsi, n, 0.0);
  KArgs a;
  a.terms = P->d_terms.as<TermDev>();
  a.md = P->d_md.as<MatDev>();
  a.map = P->d_map.as<int32_t>();
  a.psi = psi;
  a.hpsi = hpsi;
  a.nvec = nvec;
  a.nitems = (int)ip->nitems;
  a.items = ip->items.as<Item>();
  a.scratch = ip->scratch.as<double>();
  a.slot_size = ip->slot_size;
  unsigned long long* fx = nullptr;
  unsigned long long* fxflag = nullptr;
  if (P->deterministic) {
    fx = P->d_fx.as<unsigned long long>();
    fxflag = fx + 2 * n;
    hipLaunchKernelGGL(k_fx_clear, dim3(grid_blocks(2 * n + 1)), dim3(256), 0, P->stream, fx,
                       2 * n + 1);
  }
  a.fx = fx;
  a.fxflag = fxflag;
  P->last_items = ip->nitems + ip->nitems2;
  P->last_flops = ip->flops_main;
  P->last_flops_mfma = ip->flops_mfma;
  if (P->timing) {
    if (P->ev_used == P->ev_pairs.size()) {
      hipEvent_t e0, e1;
      HIP_TRY(hipEventCreate(&e0));
      HIP_TRY(hipEventCreate(&e1));
      P->ev_pairs.emplace_back(e0, e1);
    }
    P->ev0 = P->ev_pairs[P->ev_used].first;
    P->ev1 = P->ev_pairs[P->ev_used].second;
    ++P->ev_used;
    if (!split) HIP_TRY(hipEventRecord(P->ev0, P->stream));
  }
  const unsigned grid = (unsigned)((cfg.grid > 0 && cfg.grid < ip->nitems) ? cfg.grid : ip->nitems);
  const dim3 blk(P->type_op == 10 ? 256 : (split ? cfg.nthreads : std::min(cfg.nthreads, NT)));
  if (prec == SG4_PREC_F32X2) {
    // kinetic rows in FP64 (initialise hpsi), then the terms on the matrix cores
    hipLaunchKernelGGL(k_kin_rows, dim3(grid_blocks(n)), dim3(256), 0, P->stream,
                       P->d_kptr.as<int64_t>(), P->d_kcol.as<int32_t>(), P->d_kval.as<double>(),
                       psi, hpsi, P->nb_ba, nvec);
    MArgs m;
    m.terms = P->d_termm.as<TermM>();
    m.mats = P->d_mtab.as<uint16_t>();
    m.map = P->d_map.as<int32_t>();
    m.V = P->d_V32.as<float>();
    m.psi = psi;
    m.hpsi = hpsi;
    m.fx = fx;
    m.fxflag = fxflag;
    m.nvec = nvec;
    for (int list = 0; list < 2; ++list) {
      const int64_t ni = list ? ip->nitems2 : ip->nitems;
      if (P->timing && list == 0) HIP_TRY(hipEventRecord(P->ev0, P->stream));
      if (ni > 0) {
        m.nitems = (int)ni;
        m.items = list ? ip->items2.as<Item>() : ip->items.as<Item>();
        const unsigned gr = (unsigned)((cfg.grid > 0 && cfg.grid < ni) ? cfg.grid : ni);
        hipLaunchKernelGGL(k_hpsi_mfma, dim3(gr), dim3(std::min(cfg.nthreads, NT_MFMA)),
                           list ? ip->lds2 : ip->lds, P->stream, m);
      }
      if (P->timing && list == 0) HIP_TRY(hipEventRecord(P->ev1, P->stream));
    }
    if (fx)
      hipLaunchKernelGGL(k_fx_finish, dim3(grid_blocks(n)), dim3(256), 0, P->stream, fx, fxflag,
                         hpsi, n);
    HIP_TRY(hipGetLastError());
    return SG4_OK;
  }
  if (split) {
    // kinetic rows (initialise hpsi), then the potential part of the terms
    hipLaunchKernelGGL(k_kin_rows, dim3(grid_blocks(n)), dim3(256), 0, P->stream,
                       P->d_kptr.as<int64_t>(), P->d_kcol.as<int32_t>(), P->d_kval.as<double>(),
                       psi, hpsi, P->nb_ba, nvec);
    for (int list = 0; list < 2; ++list) {
      const int64_t ni = list ? ip->nitems2 : ip->nitems;
      if (ni == 0) {
        if (P->timing && list == 0) {
          HIP_TRY(hipEventRecord(P->ev0, P->stream));
          HIP_TRY(hipEventRecord(P->ev1, P->stream));
        }
        continue;
      }
      KArgs b = a;
      b.nitems = (int)ni;
      b.items = list ? ip->items2.as<Item>() : ip->items.as<Item>();
      const unsigned gr = (unsigned)((cfg.grid > 0 && cfg.grid < ni) ? cfg.grid : ni);
      const int lds = list ? ip->lds2 : ip->lds;
      if (P->timing && list == 0) HIP_TRY(hipEventRecord(P->ev0, P->stream));
      if (esize == 8) {
        KMats<double> km{P->d_mats.as<double>(), P->d_V.as<double>(), P->d_ck.as<double>()};
        hipLaunchKernelGGL(k_hpsi_t1s_f64, dim3(gr), blk, lds, P->stream, b, km);
      } else {
        KMats<float> km{P->d_mats32.as<float>(), P->d_V32.as<float>(), P->d_ck32.as<float>()};
        hipLaunchKernelGGL(k_hpsi_t1s_f32, dim3(gr), blk, lds, P->stream, b, km);
      }
      if (P->timing && list == 0) HIP_TRY(hipEventRecord(P->ev1, P->stream));
    }
    if (fx)
      hipLaunchKernelGGL(k_fx_finish, dim3(grid_blocks(n)), dim3(256), 0, P->stream, fx, fxflag,
                         hpsi, n);
  } else if (P->type_op == 11) {
    KMats1G km{P->d_mats.as<double>(), P->d_F.as<double>(), P->d_der.as<int32_t>(), P->nder,
               P->sum_nq};
    hipLaunchKernelGGL(k_hpsi_t1g_f64, dim3(grid), dim3(256), ip->lds, P->stream, a, km);
  } else if (P->type_op == 10) {
    KMats10 km{P->d_mats.as<double>(), P->d_V.as<double>(), P->d_G.as<double>(),
               P->d_Jac.as<double>(), P->d_sqRJ.as<double>()};
    hipLaunchKernelGGL(k_hpsi_t10_f64, dim3((unsigned)ip->nitems), blk, ip->lds, P->stream, a, km);
  } else if (esize == 8) {
    KMats<double> km{P->d_mats.as<double>(), P->d_V.as<double>(), P->d_ck.as<double>()};
    hipLaunchKernelGGL(k_hpsi_t1_f64, dim3(grid), blk, ip->lds, P->stream, a, km);
  } else {
    KMats<float> km{P->d_mats32.as<float>(), P->d_V32.as<float>(), P->d_ck32.as<float>()};
    hipLaunchKernelGGL(k_hpsi_t1_f32, dim3(grid), blk, ip->lds, P->stream, a, km);
  }
  if (P->timing && !split) HIP_TRY(hipEventRecord(P->ev1, P->stream));
  HIP_TRY(hipGetLastError());
  return SG4_OK;
}

}  // namespace sg4

using namespace sg4;

namespace sg4 {
// The kernels add into their outputs with hardware floating-point atomics
// compiled with -munsafe-fp-atomics; those are correct on coarse-grained
// device memory only.  On host (pinned), managed or fine-grained memory the
// adds may be lost, so every block handed over by pointer must be ordinary
// device memory (hipMalloc, hipMallocAsync, a torch CUDA tensor) on the plan's
// GPU.  hipPointerGetAttributes looks the pointer up in the runtime's table
// of allocations; it allocates nothing and enqueues nothing.
int device_block(const sg4_plan* p, const void* ptr, const char* name) {
  hipPointerAttribute_t at{};
  const hipError_t e = hipPointerGetAttributes(&at, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();   // the lookup failure is this call's answer, not a sticky error
    set_error(std::string(name) + " is not memory allocated by HIP: device pointers are required");
    return SG4_ERR_ARG;
  }
  if (at.type != hipMemoryTypeDevice || at.isManaged) {
    set_error(std::string(name) + " is host or managed memory: device pointers are required");
    return SG4_ERR_ARG;
  }
  // hipDeviceMallocFinegrained (1), hipMallocSignalMemory (2) and
  // hipDeviceMallocUncached (3) all set one of the two low flag bits
  if (at.allocationFlags & 3u) {
    set_error(std::string(name) + " is fine-grained or uncached device memory: "
              "floating-point atomics need coarse-grained memory");
    return SG4_ERR_ARG;
  }
  if (at.device != p->device) {
    set_error(std::string(name) + " is on GPU " + std::to_string(at.device) +
              ", the plan on GPU " + std::to_string(p->device));
    return SG4_ERR_ARG;
  }
  return SG4_OK;
}
}  // namespace sg4

extern "C" {

const char* sg4_last_error(void) { return g_last_error.c_str(); }

// The kernels read psi while they write or scatter-add into hpsi: the two
// nb_ba x nvec blocks must be distinct memory.

static int device_blocks(const sg4_plan* p, const double* psi, const double* hpsi) {
  int rc = sg4::device_block(p, psi, "psi");
  return rc ? rc : sg4::device_block(p, hpsi, "hpsi");
}

static bool blocks_overlap(const sg4_plan* p, int32_t nvec, const double* psi, const double* hpsi) {
  if (nvec < 1) return false;
  const uintptr_t a = (uintptr_t)psi, b = (uintptr_t)hpsi;
  const uintptr_t n = (uintptr_t)p->nb_ba * (uintptr_t)nvec * sizeof(double);
  if (a < b + n && b < a + n) {
    set_error("psi and hpsi overlap");
    return true;
  }
  return false;
}

int sg4_setup_create(int32_t D, int32_t L, int32_t LB, const sg4_mode_desc* modes,
                     sg4_setup** out) {
  if (!modes || !out) { set_error("null argument"); return SG4_ERR_ARG; }
  try {
    auto* s = new sg4_setup;
    s->s = new Setup(D, L, LB, modes);
    *out = s;
    return SG4_OK;
  } catch (const std::exception& e) {
    set_error(e.what());
    return SG4_ERR_ARG;
  }
}

int sg4_setup_sizes(const sg4_setup* s, int64_t* nterms, int64_t* nb_ba, int64_t* sum_nb,
                    int64_t* sum_nq) {
  if (!s) { set_error("null setup"); return SG4_ERR_ARG; }
  if (nterms) *nterms = s->s->nterms;
  if (nb_ba) *nb_ba = s->s->nb_ba;
  if (sum_nb) *sum_nb = s->s->boff.back();
  if (sum_nq) *sum_nq = s->s->qoff.back();
  return SG4_OK;
}

int sg4_setup_tables(const sg4_setup* s, sg4_tables* out) {
  if (!s || !out) { set_error("null argument"); return SG4_ERR_ARG; }
  s->s->fill_tables(out);
  return SG4_OK;
}

int sg4_setup_term_grid(const sg4_setup* s, int64_t t, double* q) {
  if (!s || !q || t < 0 || t >= s->s->nterms) { set_error("bad term"); return SG4_ERR_ARG; }
  s->s->term_grid(t, q);
  return SG4_OK;
}

int sg4_setup_basis1d(const sg4_setup* s, int32_t k, int32_t l, int32_t* nq, int32_t* nb,
                      double* x, double* w, double* d0, double* d1, double* d2) {
  if (!s || k < 0 || k >= s->s->D || l < 0 || l > s->s->L) { set_error("bad mode/level"); return SG4_ERR_ARG; }
  const Basis1D& b = s->s->b1[k][l];
  if (nq) *nq = b.nq;
  if (nb) *nb = b.nb;
  if (x) std::memcpy(x, b.x.data(), sizeof(double) * b.nq);
  if (w) std::memcpy(w, b.w.data(), sizeof(double) * b.nq);
  size_t qb = (size_t)b.nq * b.nb;
  if (d0) std::memcpy(d0, b.d0.data(), sizeof(double) * qb);
  if (d1) std::memcpy(d1, b.d1.data(), sizeof(double) * qb);
  if (d2) std::memcpy(d2, b.d2.data(), sizeof(double) * qb);
  return SG4_OK;
}

int sg4_setup_packed(const sg4_setup* s, int32_t* idx) {
  if (!s || !idx) { set_error("null argument"); return SG4_ERR_ARG; }
  const Setup& S = *s->s;
  const int D = S.D, L1 = S.L + 1;
  std::vector<char> seen((size_t)S.nb_ba, 0);
  int64_t nseen = 0;
  int64_t dims[SG4_MAX_D];
  for (int64_t t = 0; t < S.nterms && nseen < S.nb_ba; ++t) {
    int64_t n = 1;
    for (int k = 0; k < D; ++k) {
      dims[k] = S.nb_tab[(size_t)k * L1 + S.levels[(size_t)t * D + k]];
      n *= dims[k];
    }
    const int32_t* mp = &S.map[S.boff[t]];
    for (int64_t e = 0; e < n; ++e) {
      const int32_t r = mp[e];
      if (r < 0 || seen[r]) continue;
      int64_t q = e;
      for (int k = 0; k < D; ++k) {   // in-term index, first index fastest
        idx[(size_t)r * D + k] = (int32_t)(q % dims[k]);
        q /= dims[k];
      }
      seen[r] = 1;
      ++nseen;
    }
  }
  if (nseen != S.nb_ba) { set_error("some packed rows occur in no term"); return SG4_ERR_STATE; }
  return SG4_OK;
}

int sg4_setup_destroy(sg4_setup* s) {
  if (s) { delete s->s; delete s; }
  return SG4_OK;
}

int sg4_plan_create_batch(int32_t ninst, const sg4_tables* tabs, int32_t max_nvec,
                          sg4_plan** out) {
  if (!tabs || !out || ninst < 1) { set_error("null table pointer"); return SG4_ERR_ARG; }
  if (max_nvec < 1 || max_nvec > 4096) { set_error("max_nvec out of range"); return SG4_ERR_ARG; }
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    (void)hipFuncSetAttribute((const void*)k_hpsi_t1_f64,
                              hipFuncAttributeMaxDynamicSharedMemorySize, LDS_MAX);
    (void)hipFuncSetAttribute((const void*)k_hpsi_t1_f32,
                              hipFuncAttributeMaxDynamicSharedMemorySize, LDS_MAX);
    (void)hipFuncSetAttribute((const void*)k_hpsi_t1s_f64,
                              hipFuncAttributeMaxDynamicSharedMemorySize, LDS_MAX);
    (void)hipFuncSetAttribute((const void*)k_hpsi_t1s_f32,
                              hipFuncAttributeMaxDynamicSharedMemorySize, LDS_MAX);
    (void)hipFuncSetAttribute((const void*)k_hpsi_t10_f64,
                              hipFuncAttributeMaxDynamicSharedMemorySize, LDS_MAX);
    (void)hipFuncSetAttribute((const void*)k_grad_v,
                              hipFuncAttributeMaxDynamicSharedMemorySize, LDS_MAX);
    (void)hipFuncSetAttribute((const void*)k_hpsi_mfma,
                              hipFuncAttributeMaxDynamicSharedMemorySize, LDS_MAX);
    (void)hipFuncSetAttribute((const void*)k_hpsi_t1g_f64,
                              hipFuncAttributeMaxDynamicSharedMemorySize, LDS_MAX);
  });
  std::unique_ptr<sg4_plan> P(new sg4_plan);
  HIP_TRY(hipGetDevice(&P->device));
  P->max_nvec = max_nvec;
  bool nodes = true, qws = true;
  int64_t off = 0, noff = 0;
  std::vector<int64_t> node_off;
  std::vector<double> matsR, nodev, qwv;
  std::vector<int32_t> mapv;
  P->h_boff.push_back(0);
  P->h_qoff.push_back(0);
  for (int ii = 0; ii < ninst; ++ii) {
    const sg4_tables* t = &tabs[ii];
    if (!t->levels || !t->weights || !t->map || !t->nq || !t->nb || !t->mats) {
      set_error("null table pointer");
      return SG4_ERR_ARG;
    }
    if (t->D < 1 || t->D > SG4_MAX_D || t->Lmax < 0 || t->Lmax > SG4_MAX_L || t->nterms < 1 ||
        t->nb_ba < 1 || t->nb_ba > INT32_MAX) {
      set_error("table sizes out of range");
      return SG4_ERR_ARG;
    }
    nodes = nodes && t->nodes;
    qws = qws && t->qweights;
    sg4_plan::Inst I;
    I.D = t->D; I.Lmax = t->Lmax; I.nterms = t->nterms; I.nb_ba = t->nb_ba;
    I.term0 = P->nterms; I.row0 = P->nb_ba; I.md0 = (int64_t)P->h_md.size(); I.ck0 = P->nck;
    I.lev0 = (int64_t)P->h_levels.size();
    const int D = t->D, L1 = t->Lmax + 1;
    const int64_t noff0 = noff;
    for (int k = 0; k < D; ++k)
      for (int l = 0; l < L1; ++l) {
        const int nq = t->nq[k * L1 + l], nb = t->nb[k * L1 + l];
        if (nq < 1 || nb < 1 || nb > nq || nq > SG4_MAX_N) {
          set_error("1-D sizes out of range (need 1 <= nb <= nq <= 64)");
          return SG4_ERR_ARG;
        }
        MatDev m{};
        m.nq = nq; m.nb = nb;
        m.B = off; off += (int64_t)nq * nb;
        m.BtW = off; off += (int64_t)nq * nb;
        m.D1 = off; off += (int64_t)nq * nq;
        m.D2 = off; off += (int64_t)nq * nq;
        P->h_md.push_back(m);
        P->h_nq.push_back(nq);
        P->h_nb.push_back(nb);
        node_off.push_back(noff);
        noff += nq;
      }
    // row-major copies of this instance's matrices (the kernels read rows)
    const int64_t base = P->h_md[I.md0].B;
    matsR.resize(off);
    for (int64_t i = I.md0; i < (int64_t)P->h_md.size(); ++i) {
      MatDev& m = P->h_md[i];
      auto tr = [&](int64_t o, int rows, int cols) {
        for (int r = 0; r < rows; ++r)
          for (int c = 0; c < cols; ++c)
            matsR[o + (int64_t)r * cols + c] = t->mats[o - base + r + (int64_t)rows * c];
      };
      tr(m.B, m.nq, m.nb);
      tr(m.BtW, m.nb, m.nq);
      tr(m.D1, m.nq, m.nq);
      tr(m.D2, m.nq, m.nq);
      m.b00 = t->mats[m.B - base]; m.btw00 = t->mats[m.BtW - base]; m.d200 = t->mats[m.D2 - base];
    }
    if (t->nodes) nodev.insert(nodev.end(), t->nodes, t->nodes + (noff - noff0));
    if (t->qweights) qwv.insert(qwv.end(), t->qweights, t->qweights + (noff - noff0));
    P->h_levels.insert(P->h_levels.end(), t->levels, t->levels + (size_t)D * t->nterms);
    P->h_weights.insert(P->h_weights.end(), t->weights, t->weights + t->nterms);
    int64_t nbsum = 0;
    for (int64_t tt = 0; tt < t->nterms; ++tt) {
      int64_t nbt = 1, nqt = 1;
      for (int k = 0; k < D; ++k) {
        const int l = t->levels[(size_t)tt * D + k];
        if (l < 0 || l > t->Lmax) { set_error("term level out of range"); return SG4_ERR_ARG; }
        nbt *= t->nb[k * L1 + l];
        nqt *= t->nq[k * L1 + l];
      }
      if (nqt > INT32_MAX / 64) { set_error("term too large"); return SG4_ERR_UNSUPPORTED; }
      P->h_boff.push_back(P->h_boff.back() + nbt);
      P->h_qoff.push_back(P->h_qoff.back() + nqt);
      nbsum += nbt;
    }
    for (int64_t e = 0; e < nbsum; ++e) {
      const int32_t m = t->map[e];
      if (m < -1 || m >= t->nb_ba) { set_error("map entry out of range"); return SG4_ERR_ARG; }
      if (I.row0 + t->nb_ba > INT32_MAX) { set_error("batch exceeds int32 rows"); return SG4_ERR_UNSUPPORTED; }
      mapv.push_back(m < 0 ? -1 : (int32_t)(m + I.row0));
    }
    P->nterms += t->nterms;
    P->nb_ba += t->nb_ba;
    P->nck += D;
    P->inst.push_back(I);
  }
  P->D = tabs[0].D;
  P->Lmax = tabs[0].Lmax;
  P->sum_nb = P->h_boff.back();
  P->sum_nq = P->h_qoff.back();
  P->nmats = off;
  int rc;
  if ((rc = P->d_mats.alloc(sizeof(double) * off))) return rc;
  HIP_TRY(hipMemcpy(P->d_mats.p, matsR.data(), sizeof(double) * off, hipMemcpyHostToDevice));
  P->h_mats = matsR;
  if ((rc = P->d_md.alloc(sizeof(MatDev) * P->h_md.size()))) return rc;
  HIP_TRY(hipMemcpy(P->d_md.p, P->h_md.data(), sizeof(MatDev) * P->h_md.size(), hipMemcpyHostToDevice));
  if ((rc = P->d_map.alloc(sizeof(int32_t) * P->sum_nb))) return rc;
  HIP_TRY(hipMemcpy(P->d_map.p, mapv.data(), sizeof(int32_t) * P->sum_nb, hipMemcpyHostToDevice));
  P->h_map.swap(mapv);
  if ((rc = P->d_levels.alloc(sizeof(int32_t) * P->h_levels.size()))) return rc;
  HIP_TRY(hipMemcpy(P->d_levels.p, P->h_levels.data(), sizeof(int32_t) * P->h_levels.size(), hipMemcpyHostToDevice));
  if ((rc = P->d_qoff.alloc(sizeof(int64_t) * P->h_qoff.size()))) return rc;
  HIP_TRY(hipMemcpy(P->d_qoff.p, P->h_qoff.data(), sizeof(int64_t) * P->h_qoff.size(), hipMemcpyHostToDevice));
  if ((rc = P->d_nqtab.alloc(sizeof(int32_t) * P->h_nq.size()))) return rc;
  HIP_TRY(hipMemcpy(P->d_nqtab.p, P->h_nq.data(), sizeof(int32_t) * P->h_nq.size(), hipMemcpyHostToDevice));
  if ((rc = P->d_node_off.alloc(sizeof(int64_t) * node_off.size()))) return rc;
  HIP_TRY(hipMemcpy(P->d_node_off.p, node_off.data(), sizeof(int64_t) * node_off.size(), hipMemcpyHostToDevice));
  if (nodes) {
    if ((rc = P->d_nodes.alloc(sizeof(double) * noff))) return rc;
    HIP_TRY(hipMemcpy(P->d_nodes.p, nodev.data(), sizeof(double) * noff, hipMemcpyHostToDevice));
    P->have_nodes = true;
  }
  if (qws) {
    if ((rc = P->d_qw.alloc(sizeof(double) * noff))) return rc;
    HIP_TRY(hipMemcpy(P->d_qw.p, qwv.data(), sizeof(double) * noff, hipMemcpyHostToDevice));
    P->have_qw = true;
  }
  if ((rc = P->d_V.alloc(sizeof(double) * P->sum_nq))) return rc;
  // work buffers for the column-major entry point
  const size_t blk = sizeof(double) * (size_t)P->nb_ba * max_nvec;
  if ((rc = P->d_rows_in.alloc(blk))) return rc;
  if ((rc = P->d_rows_out.alloc(blk))) return rc;
  // algorithmic flop / byte counts: per vector sum_t S_t with
  // S_t = sum_k 2 N_t n_k (x3 for type 1, x4 for type 10); bytes: V grid + map
  double S = 0;
  for (const sg4_plan::Inst& I : P->inst)
    for (int64_t tl = 0; tl < I.nterms; ++tl) {
      const int64_t tt = I.term0 + tl;
      const double Nt = (double)(P->h_qoff[tt + 1] - P->h_qoff[tt]);
      for (int k = 0; k < I.D; ++k)
        S += 2.0 * Nt * P->h_nq[I.md0 + k * (I.Lmax + 1) + P->h_levels[I.lev0 + tl * I.D + k]];
    }
  P->flops_per_vec = S;
  P->bytes_fixed = 8.0 * P->sum_nq + 4.0 * P->sum_nb;
  *out = P.release();
  return SG4_OK;
}

int sg4_plan_create(const sg4_tables* t, int32_t max_nvec, sg4_plan** out) {
  return sg4_plan_create_batch(1, t, max_nvec, out);
}

int sg4_plan_set_stream(sg4_plan* p, void* s) {
  if (!p) { set_error("null plan"); return SG4_ERR_ARG; }
  p->stream = (hipStream_t)s;
  return SG4_OK;
}

int sg4_plan_set_operator(sg4_plan* p, const sg4_operator* op) {
  if (!p || !op || !op->V) { set_error("null operator"); return SG4_ERR_ARG; }
  invalidate(p);
  int rc;
  if (op->type_op == 1) {
    if (!op->ck) { set_error("type 1 needs ck"); return SG4_ERR_ARG; }
    p->h_ck.assign(op->ck, op->ck + p->nck);
    if ((rc = p->d_ck.alloc(sizeof(double) * p->nck))) return rc;
    HIP_TRY(hipMemcpy(p->d_ck.p, op->ck, sizeof(double) * p->nck, hipMemcpyHostToDevice));
  } else if (op->type_op == 10) {
    if (!op->G || !op->Jac || !op->sqRJ) { set_error("type 10 needs G, Jac and sqRJ"); return SG4_ERR_ARG; }
    if (p->inst.size() != 1) { set_error("type 10 needs a single-instance plan"); return SG4_ERR_UNSUPPORTED; }
    const size_t n = (size_t)p->sum_nq;
    if ((rc = p->d_G.alloc(sizeof(double) * n * p->D * p->D))) return rc;
    if ((rc = p->d_Jac.alloc(sizeof(double) * n))) return rc;
    if ((rc = p->d_sqRJ.alloc(sizeof(double) * n))) return rc;
    HIP_TRY(hipMemcpy(p->d_G.p, op->G, sizeof(double) * n * p->D * p->D, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_Jac.p, op->Jac, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_sqRJ.p, op->sqRJ, sizeof(double) * n, hipMemcpyHostToDevice));
    p->h_ck.assign(p->nck, 0.0);
  } else {
    set_error("type_op must be 1 or 10");
    return SG4_ERR_UNSUPPORTED;
  }
  p->type_op = op->type_op;
  HIP_TRY(hipMemcpy(p->d_V.p, op->V, sizeof(double) * p->sum_nq, hipMemcpyHostToDevice));
  (void)p->d_V0.alloc(0);
  return build_terms(p);
}

int sg4_plan_set_operator_type1(sg4_plan* p, int32_t nder, const int32_t* der, const double* F) {
  if (!p || !der || !F || nder < 1) { set_error("null argument"); return SG4_ERR_ARG; }
  if (p->inst.size() != 1) {
    set_error("the general type-1 operator needs a single-instance plan");
    return SG4_ERR_UNSUPPORTED;
  }
  if (p->D > MAXACT) { set_error("more than 16 modes"); return SG4_ERR_UNSUPPORTED; }
  for (int e = 0; e < 2 * nder; ++e)
    if (der[e] < -1 || der[e] >= p->D) { set_error("derivative mode index out of range"); return SG4_ERR_ARG; }
  invalidate(p);
  int rc;
  if ((rc = p->d_der.alloc(sizeof(int32_t) * 2 * nder))) return rc;
  if ((rc = p->d_F.alloc(sizeof(double) * (size_t)nder * p->sum_nq))) return rc;
  HIP_TRY(hipMemcpy(p->d_der.p, der, sizeof(int32_t) * 2 * nder, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(p->d_F.p, F, sizeof(double) * (size_t)nder * p->sum_nq, hipMemcpyHostToDevice));
  p->nder = nder;
  p->type_op = 11;
  p->h_ck.assign(p->nck, 0.0);
  (void)p->d_V0.alloc(0);
  return build_terms(p);
}

int sg4_plan_set_henon_heiles(sg4_plan* p, double lambda) {
  if (!p) { set_error("null plan"); return SG4_ERR_ARG; }
  if (!p->have_nodes) { set_error("plan was created without grid nodes"); return SG4_ERR_STATE; }
  invalidate(p);
  for (const sg4_plan::Inst& I : p->inst) {
    HHArgs a;
    a.levels = p->d_levels.as<int32_t>() + I.lev0;
    a.qoff = p->d_qoff.as<int64_t>() + I.term0;
    a.nodes = p->d_nodes.as<double>();
    a.node_off = p->d_node_off.as<int64_t>() + I.md0;
    a.nqtab = p->d_nqtab.as<int32_t>() + I.md0;
    a.D = I.D; a.Lmax = I.Lmax; a.nterms = I.nterms; a.lambda = lambda;
    a.V = p->d_V.as<double>();
    hipLaunchKernelGGL(k_grid_hh, dim3((unsigned)std::min<int64_t>(I.nterms, 65535)), dim3(256), 0,
                       p->stream, a);
  }
  HIP_TRY(hipGetLastError());
  p->type_op = 1;
  (void)p->d_V0.alloc(0);
  p->h_ck.assign(p->nck, -0.5);
  int rc;
  if ((rc = p->d_ck.alloc(sizeof(double) * p->nck))) return rc;
  HIP_TRY(hipMemcpy(p->d_ck.p, p->h_ck.data(), sizeof(double) * p->nck, hipMemcpyHostToDevice));
  rc = build_terms(p);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(p->stream));
  return SG4_OK;
}

int sg4_plan_grid_dipole(sg4_plan* p, double* mu) {
  if (!p || !mu) { set_error("null argument"); return SG4_ERR_ARG; }
  if (!p->have_nodes) { set_error("plan was created without grid nodes"); return SG4_ERR_STATE; }
  for (const sg4_plan::Inst& I : p->inst) {
    HHArgs a;
    a.levels = p->d_levels.as<int32_t>() + I.lev0;
    a.qoff = p->d_qoff.as<int64_t>() + I.term0;
    a.nodes = p->d_nodes.as<double>();
    a.node_off = p->d_node_off.as<int64_t>() + I.md0;
    a.nqtab = p->d_nqtab.as<int32_t>() + I.md0;
    a.D = I.D; a.Lmax = I.Lmax; a.nterms = I.nterms; a.lambda = 0.0;
    a.V = mu;
    hipLaunchKernelGGL(k_grid_dipole, dim3((unsigned)std::min<int64_t>(I.nterms, 65535)), dim3(256),
                       0, p->stream, a);
  }
  HIP_TRY(hipGetLastError());
  return SG4_OK;
}

int sg4_plan_set_jacobi_triatomic(sg4_plan* p, const double* prm) {
  if (!p || !prm) { set_error("null argument"); return SG4_ERR_ARG; }
  if (!p->have_nodes) { set_error("plan was created without grid nodes"); return SG4_ERR_STATE; }
  if (p->inst.size() != 1 || p->D != 3) {
    set_error("the Jacobi triatomic model needs a single 3-mode instance (R, r, theta)");
    return SG4_ERR_UNSUPPORTED;
  }
  if (!(prm[0] > 0) || !(prm[1] > 0)) { set_error("masses must be positive"); return SG4_ERR_ARG; }
  invalidate(p);
  int rc;
  const size_t n = (size_t)p->sum_nq;
  if ((rc = p->d_G.alloc(sizeof(double) * n * 9))) return rc;
  if ((rc = p->d_Jac.alloc(sizeof(double) * n))) return rc;
  if ((rc = p->d_sqRJ.alloc(sizeof(double) * n))) return rc;
  const sg4_plan::Inst& I = p->inst[0];
  JacArgs a;
  a.levels = p->d_levels.as<int32_t>();
  a.qoff = p->d_qoff.as<int64_t>();
  a.nodes = p->d_nodes.as<double>();
  a.node_off = p->d_node_off.as<int64_t>() + I.md0;
  a.nqtab = p->d_nqtab.as<int32_t>() + I.md0;
  a.Lmax = I.Lmax;
  a.nterms = I.nterms;
  a.muR = prm[0]; a.mur = prm[1]; a.D = prm[2]; a.a = prm[3]; a.Re = prm[4]; a.re = prm[5];
  a.c = prm[6];
  a.V = p->d_V.as<double>(); a.G = p->d_G.as<double>(); a.Jac = p->d_Jac.as<double>();
  a.sqRJ = p->d_sqRJ.as<double>();
  hipLaunchKernelGGL(k_grid_jacobi, dim3((unsigned)std::min<int64_t>(I.nterms, 65535)), dim3(256), 0,
                     p->stream, a);
  HIP_TRY(hipGetLastError());
  p->type_op = 10;
  (void)p->d_V0.alloc(0);
  p->h_ck.assign(p->nck, 0.0);
  rc = build_terms(p);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(p->stream));
  return SG4_OK;
}

int sg4_apply_rows(sg4_plan* p, int32_t nvec, const double* psi, double* hpsi, int32_t prec) {
  if (!p || !psi || !hpsi) { set_error("null argument"); return SG4_ERR_ARG; }
  if (int rc0 = device_blocks(p, psi, hpsi)) return rc0;
  if (blocks_overlap(p, nvec, psi, hpsi)) return SG4_ERR_ARG;
  return apply_rows(p, nvec, psi, hpsi, prec);
}

int sg4_apply(sg4_plan* p, int32_t nvec, const double* psi, double* hpsi, int32_t prec) {
  if (!p || !psi || !hpsi) { set_error("null argument"); return SG4_ERR_ARG; }
  if (int rc0 = device_blocks(p, psi, hpsi)) return rc0;
  if (nvec < 1 || nvec > p->max_nvec) { set_error("nvec out of range"); return SG4_ERR_ARG; }
  if (blocks_overlap(p, nvec, psi, hpsi)) return SG4_ERR_ARG;
  double* ri = p->d_rows_in.as<double>();
  double* ro = p->d_rows_out.as<double>();
  if (nvec == 1) {
    return apply_rows(p, 1, psi, hpsi, prec);
  }
  const int W = pad_width(p, nvec);
  hipLaunchKernelGGL(k_cols_to_rows, tile_grid(p->nb_ba, W), dim3(256), 0, p->stream, psi, ri, p->nb_ba, nvec, W);
  int rc = apply_rows(p, W, ri, ro, prec);
  if (rc) return rc;
  hipLaunchKernelGGL(k_rows_to_cols, tile_grid(p->nb_ba, W), dim3(256), 0, p->stream, ro, hpsi, p->nb_ba, nvec, W);
  HIP_TRY(hipGetLastError());
  return SG4_OK;
}

}  // extern "C"

// Allocation-free entry point: the two rows-layout staging blocks of
// sg4_apply live in this file-scope device array (one copy per GPU, held
// until the process ends).  In host code `scratch` only names the symbol:
// its device address on the current GPU comes from hipGetSymbolAddress,
// looked up on the first call on that GPU and kept per device.
static constexpr size_t SG4_SCRATCH_BYTES = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[SG4_SCRATCH_BYTES];
static constexpr int SG4_MAX_GPUS = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[SG4_MAX_GPUS] = {};

static int scratch_address(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= SG4_MAX_GPUS) { set_error("device index out of range"); return SG4_ERR_ARG; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return SG4_OK;
}

extern "C" {

int sg4_apply_noalloc(sg4_plan* p, int32_t nvec, const double* psi, double* hpsi, int32_t prec) {
  if (!p || !psi || !hpsi) { set_error("null argument"); return SG4_ERR_ARG; }
  if (int rc0 = device_blocks(p, psi, hpsi)) return rc0;
  if (nvec < 1 || nvec > p->max_nvec) { set_error("nvec out of range"); return SG4_ERR_ARG; }
  if (blocks_overlap(p, nvec, psi, hpsi)) return SG4_ERR_ARG;
  const int W = pad_width(p, nvec);
  const size_t n = (size_t)p->nb_ba * W;
  if (nvec > 1 && 2 * sizeof(double) * n > SG4_SCRATCH_BYTES) {
    set_error("sg4_apply_noalloc: nb_ba x " + std::to_string(W) + " = " + std::to_string(n) +
              " exceeds its device scratch (two FP64 blocks in 256 MiB: nb_ba x nvec <= 16777216, "
              "nvec counted after rounding up to a power of two or a multiple of 16)");
    return SG4_ERR_ARG;
  }
  double* base = nullptr;
  int rc;
  if (nvec > 1 && (rc = scratch_address((void**)&base))) return rc;
  // prepared state only, and no timing events (creating them allocates)
  struct Guard {
    sg4_plan* p;
    bool timing;
    explicit Guard(sg4_plan* q) : p(q), timing(q->timing) { p->no_alloc = true; p->timing = false; }
    ~Guard() { p->no_alloc = false; p->timing = timing; }
  } guard(p);
  if (nvec == 1) return apply_rows(p, 1, psi, hpsi, prec);
  double* ri = base;
  double* ro = base + n;
  hipLaunchKernelGGL(k_cols_to_rows, tile_grid(p->nb_ba, W), dim3(256), 0, p->stream, psi, ri, p->nb_ba, nvec, W);
  if ((rc = apply_rows(p, W, ri, ro, prec))) return rc;
  hipLaunchKernelGGL(k_rows_to_cols, tile_grid(p->nb_ba, W), dim3(256), 0, p->stream, ro, hpsi, p->nb_ba, nvec, W);
  HIP_TRY(hipGetLastError());
  return SG4_OK;
}

int sg4_apply_host(sg4_plan* p, int32_t nvec, const double* psi, double* hpsi, int32_t prec) {
  if (!p || !psi || !hpsi) { set_error("null argument"); return SG4_ERR_ARG; }
  if (nvec < 1 || nvec > p->max_nvec) { set_error("nvec out of range"); return SG4_ERR_ARG; }
  const size_t bytes = sizeof(double) * (size_t)p->nb_ba * nvec;
  int rc;
  if (p->d_cols_in.n < bytes) {
    if ((rc = p->d_cols_in.alloc(sizeof(double) * (size_t)p->nb_ba * p->max_nvec))) return rc;
    if ((rc = p->d_cols_out.alloc(sizeof(double) * (size_t)p->nb_ba * p->max_nvec))) return rc;
  }
  HIP_TRY(hipMemcpyAsync(p->d_cols_in.p, psi, bytes, hipMemcpyHostToDevice, p->stream));
  rc = sg4_apply(p, nvec, p->d_cols_in.as<double>(), p->d_cols_out.as<double>(), prec);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(hpsi, p->d_cols_out.p, bytes, hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return SG4_OK;
}

int sg4_plan_stats(const sg4_plan* p, int32_t nvec, double* flops, double* bytes, int64_t* items) {
  if (!p) { set_error("null plan"); return SG4_ERR_ARG; }
  if (flops)
    *flops = p->last_items >= 0 ? p->last_flops
                                : (p->type_op == 10 ? 4.0 : 3.0) * p->flops_per_vec * nvec;
  if (bytes) *bytes = p->bytes_fixed + 16.0 * (double)p->nb_ba * nvec;
  if (items) {
    *items = p->last_items;
  }
  return SG4_OK;
}

int sg4_plan_mfma_stats(const sg4_plan* p, double* mfma_flops, int64_t* group_mats) {
  if (!p) { set_error("null plan"); return SG4_ERR_ARG; }
  if (mfma_flops) *mfma_flops = p->last_flops_mfma;
  if (group_mats) *group_mats = p->mfma_valid ? p->nmtab : -1;
  return SG4_OK;
}

int sg4_plan_split_info(const sg4_plan* p, int32_t* split, int64_t* kin_nnz) {
  if (!p) { set_error("null plan"); return SG4_ERR_ARG; }
  if (split) *split = use_split(p) ? 1 : 0;
  if (kin_nnz) *kin_nnz = p->kin_nnz;
  return SG4_OK;
}

int sg4_plan_set_field(sg4_plan* p, const double* mu, double eps) {
  if (!p) { set_error("null plan"); return SG4_ERR_ARG; }
  if (p->type_op == 0) { set_error("no operator set on the plan"); return SG4_ERR_STATE; }
  if (p->type_op == 11) {
    set_error("field terms of the general type-1 operator go into its identity term");
    return SG4_ERR_UNSUPPORTED;
  }
  int rc;
  if (mu && (rc = sg4::device_block(p, mu, "mu"))) return rc;
  if (!p->d_V0.p) {
    if ((rc = p->d_V0.alloc(sizeof(double) * p->sum_nq))) return rc;
    HIP_TRY(hipMemcpyAsync(p->d_V0.p, p->d_V.p, sizeof(double) * p->sum_nq,
                           hipMemcpyDeviceToDevice, p->stream));
  }
  hipLaunchKernelGGL(k_axpy_grid, dim3(grid_blocks(p->sum_nq)), dim3(256), 0, p->stream,
                     p->d_V0.as<double>(), mu, eps, p->d_V.as<double>(), p->sum_nq);
  HIP_TRY(hipGetLastError());
  p->f32_valid = false;
  return SG4_OK;
}

int sg4_grad_V(sg4_plan* p, int32_t nvec, const double* phi, const double* psi, double* grad) {
  if (!p || !phi || !psi || !grad) { set_error("null argument"); return SG4_ERR_ARG; }
  if (p->type_op == 0) { set_error("no operator set on the plan"); return SG4_ERR_STATE; }
  if (!p->have_qw) { set_error("plan was created without quadrature weights"); return SG4_ERR_STATE; }
  if (nvec < 1 || nvec > p->max_nvec) { set_error("nvec out of range"); return SG4_ERR_ARG; }
  if (p->type_op == 11) {
    set_error("grid-value gradients are available for the type-1 (V, c_k) and type-10 operators");
    return SG4_ERR_UNSUPPORTED;
  }
  int rc;
  if ((rc = device_blocks(p, phi, psi)) || (rc = sg4::device_block(p, grad, "grad"))) return rc;
  ItemPlan* ip = nullptr;
  rc = get_items(p, nvec, 8, &ip, 4);
  if (rc) return rc;
  KArgs a;
  a.terms = p->d_terms.as<TermDev>();
  a.md = p->d_md.as<MatDev>();
  a.map = p->d_map.as<int32_t>();
  a.psi = psi;
  a.hpsi = nullptr;
  a.nvec = nvec;
  a.items = ip->items.as<Item>();
  a.nitems = (int)ip->nitems;
  a.scratch = nullptr;
  a.slot_size = 0;
  KGrad kg{p->d_mats.as<double>(), p->d_qw.as<double>(), p->d_node_off.as<int64_t>(), phi, grad};
  hipLaunchKernelGGL(k_grad_v, dim3((unsigned)ip->nitems), dim3(NT), ip->lds, p->stream, a, kg);
  HIP_TRY(hipGetLastError());
  return SG4_OK;
}

int sg4_plan_autotune(sg4_plan* p, int32_t nvec, int32_t prec, float* best_ms) {
  if (!p) { set_error("null plan"); return SG4_ERR_ARG; }
  int rc;
  if ((rc = check_apply(p, nvec, prec))) return rc;
  if (p->type_op == 10) { if (best_ms) *best_ms = 0.0f; return SG4_OK; }
  DevBuf x, y;
  const size_t bytes = sizeof(double) * (size_t)p->nb_ba * nvec;
  if ((rc = x.alloc(bytes)) || (rc = y.alloc(bytes))) return rc;
  hipLaunchKernelGGL(k_fill, dim3(grid_blocks(p->nb_ba * nvec)), dim3(256), 0, p->stream,
                     x.as<double>(), p->nb_ba * nvec, 1.0);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const bool timing = p->timing;
  p->timing = false;
  LaunchCfg best;
  float bms = 1e30f;
  int cus = 256;
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cus = prop.multiProcessorCount;
  }
  const int threads3[] = {256, 512, NT32};
  std::vector<int> threads(threads3, threads3 + ((use_split(p) && prec == SG4_PREC_F32) ? 3 : 2));
  if (prec == SG4_PREC_F32X2) threads.push_back(NT_MFMA);
  const int budgets[] = {32 * 1024, 48 * 1024, 64 * 1024, 80 * 1024, 112 * 1024, 160 * 1024};
  const int grids[] = {0, 4, 8};   // 0: one per item; k: k workgroups per CU, persistent
  for (int nt : threads)
    for (int bu : budgets)
      for (int gk : grids)
        for (int xc = 0; xc < 5; ++xc) {
          LaunchCfg c;
          c.nthreads = nt; c.lds_budget = bu; c.grid = gk ? gk * cus : 0; c.xcd = xc;
          ItemPlan* ip = nullptr;
          if (prepare(p, nvec, prec, c, &ip) != SG4_OK) { (void)hipGetLastError(); continue; }
          if ((rc = apply_rows(p, nvec, x.as<double>(), y.as<double>(), prec, &c))) break;
          HIP_TRY(hipEventRecord(e0, p->stream));
          for (int r = 0; r < 3; ++r)
            if ((rc = apply_rows(p, nvec, x.as<double>(), y.as<double>(), prec, &c))) break;
          HIP_TRY(hipEventRecord(e1, p->stream));
          HIP_TRY(hipEventSynchronize(e1));
          float ms = 0;
          HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
          ms /= 3;
          if (std::getenv("SG4_AUTOTUNE_VERBOSE"))
            fprintf(stderr, "sg4 autotune nvec=%d prec=%d threads=%d lds=%d grid=%d xcd=%d: %.3f ms\n",
                    nvec, prec, c.nthreads, c.lds_budget, c.grid, c.xcd, ms);
          if (ms < bms) { bms = ms; best = c; }
        }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->timing = timing;
  if (rc) return rc;
  drop_graphs(p);
  p->tuned[nvec * 16 + prec] = best;
  if (best_ms) *best_ms = bms;
  return SG4_OK;
}

int sg4_plan_config(const sg4_plan* p, int32_t nvec, int32_t prec, int32_t* cfg4) {
  if (!p || !cfg4) { set_error("null argument"); return SG4_ERR_ARG; }
  const LaunchCfg c = cfg_for(p, nvec, prec);
  cfg4[0] = c.nthreads; cfg4[1] = c.lds_budget; cfg4[2] = c.grid; cfg4[3] = c.xcd;
  return SG4_OK;
}

int sg4_plan_set_config(sg4_plan* p, int32_t nvec, int32_t prec, const int32_t* cfg4) {
  if (!p || !cfg4) { set_error("null argument"); return SG4_ERR_ARG; }
  const bool t768 = cfg4[0] == NT32 && use_split(p) && prec == SG4_PREC_F32;
  const bool t1024 = cfg4[0] == NT_MFMA && prec == SG4_PREC_F32X2;
  if ((cfg4[0] != 256 && cfg4[0] != 512 && !t768 && !t1024) || cfg4[1] < 1024 || cfg4[1] > LDS_MAX ||
      cfg4[2] < 0 || cfg4[3] < 0 || cfg4[3] > 8) {
    set_error("launch configuration out of range (threads 256/512, 768 for the FP32 split "
              "path, 1024 for the split-precision mode; 1 KiB <= LDS budget <= 160 KiB, "
              "grid >= 0, 0 <= xcd <= 8)");
    return SG4_ERR_ARG;
  }
  LaunchCfg c;
  c.nthreads = cfg4[0]; c.lds_budget = cfg4[1]; c.grid = cfg4[2]; c.xcd = cfg4[3];
  drop_graphs(p);
  p->tuned[nvec * 16 + prec] = c;
  return SG4_OK;
}

int sg4_apply_graph(sg4_plan* p, int32_t nvec, const double* psi, double* hpsi, int32_t prec) {
  if (!p || !psi || !hpsi) { set_error("null argument"); return SG4_ERR_ARG; }
  int rc;
  if ((rc = device_blocks(p, psi, hpsi))) return rc;
  if ((rc = check_apply(p, nvec, prec))) return rc;
  if (blocks_overlap(p, nvec, psi, hpsi)) return SG4_ERR_ARG;
  hipStream_t user = p->stream;
  if (!p->own_stream) HIP_TRY(hipStreamCreateWithFlags(&p->own_stream, hipStreamNonBlocking));
  sg4_plan::GraphEntry* g = nullptr;
  for (auto& e : p->graphs)
    if (e.nvec == nvec && e.prec == prec && e.psi == psi && e.hpsi == hpsi) g = &e;
  if (!g) {
    ItemPlan* ip = nullptr;
    if ((rc = prepare(p, nvec, prec, cfg_for(p, nvec, prec), &ip))) return rc;
    HIP_TRY(hipStreamSynchronize(user));
    const bool timing = p->timing;
    p->timing = false;
    p->stream = p->own_stream;
    hipGraph_t graph;
    HIP_TRY(hipStreamBeginCapture(p->own_stream, hipStreamCaptureModeThreadLocal));
    rc = apply_rows(p, nvec, psi, hpsi, prec);
    hipError_t e = hipStreamEndCapture(p->own_stream, &graph);
    p->stream = user;
    p->timing = timing;
    if (rc) return rc;
    if (e != hipSuccess) { set_error("stream capture failed"); return SG4_ERR_HIP; }
    hipGraphExec_t exec;
    HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    p->graphs.push_back({nvec, prec, psi, hpsi, exec});
    g = &p->graphs.back();
  } else if (prec != SG4_PREC_FP64 && (rc = ensure_f32(p))) {
    // the FP32 potential is refreshed outside the graph after sg4_plan_set_field
    return rc;
  }
  // order the replay after the caller's stream and the caller's stream after it
  hipEvent_t ev;
  HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(ev, user));
  HIP_TRY(hipStreamWaitEvent(p->own_stream, ev, 0));
  HIP_TRY(hipGraphLaunch(g->exec, p->own_stream));
  HIP_TRY(hipEventRecord(ev, p->own_stream));
  HIP_TRY(hipStreamWaitEvent(user, ev, 0));
  (void)hipEventDestroy(ev);
  return SG4_OK;
}

int sg4_plan_set_option(sg4_plan* p, int32_t option, int32_t value) {
  if (!p) { set_error("null plan"); return SG4_ERR_ARG; }
  if (option == SG4_OPT_DETERMINISTIC) {
    drop_graphs(p);
    p->deterministic = value != 0;
    return SG4_OK;
  }
  if (option != SG4_OPT_GROUPING) { set_error("unknown option"); return SG4_ERR_ARG; }
  p->grouping = value != 0;
  invalidate(p);
  if (p->type_op == 0) return SG4_OK;
  return build_terms(p);
}

int sg4_plan_timing(sg4_plan* p, int32_t enable) {
  if (!p) { set_error("null plan"); return SG4_ERR_ARG; }
  p->timing = enable != 0;
  p->ev_used = 0;
  return SG4_OK;
}

int sg4_plan_kernel_times(sg4_plan* p, float* ms, int32_t max, int32_t* n) {
  if (!p || !n) { set_error("null argument"); return SG4_ERR_ARG; }
  int32_t k = 0;
  for (size_t i = 0; i < p->ev_used && k < max; ++i, ++k) {
    HIP_TRY(hipEventSynchronize(p->ev_pairs[i].second));
    HIP_TRY(hipEventElapsedTime(&ms[k], p->ev_pairs[i].first, p->ev_pairs[i].second));
  }
  *n = k;
  p->ev_used = 0;
  return SG4_OK;
}

int sg4_plan_destroy(sg4_plan* p) {
  delete p;
  return SG4_OK;
}

}  // extern "C"
