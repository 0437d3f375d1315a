// Chebyshev propagation of complex states on the device under a
// piecewise-constant control field, and the exact gradient of a final-state
// overlap with respect to the field amplitudes with checkpointed
// recomputation.
//
// Propagator: march_cheby (SRC/sub_propagation/sub_module_propa_march.f90
// :4145-4436; recurrence :4297-4309).  With Hs = (H - E0)/Esc and the
// coefficients c_0 = J_0(alpha), c_k = 2 J_k(alpha), alpha = Esc dt (cof,
// :5739-5760, supplied by the caller):
//   w_0 = psi, w_1 = -i Hs w_0, w_{k+1} = w_{k-1} - 2i Hs w_k,
//   psi(t + dt) = exp(-i E0 dt) sum_k c_k w_k.
// The adjoint step U^dagger is the same series with +i (H symmetric).  A
// complex block of ns states is held in the rows layout with 2 ns values per
// packed row (the real parts of the states, then the imaginary parts), so
// that one H.psi application (sg4_apply_rows, nvec = 2 ns) serves both, as
// the reference applies a complex psi as two real vectors
// (SRC/sub_Operator/sub_OpPsi.f90:397-411).  The recurrence, the series sum
// and the phase run as element-wise kernels on the plan's stream: the number
// of orders is fixed by the caller, so a step needs no host synchronisation.
//
// Field: H(t) = H0 + eps_s mu on segment s (sg4_plan_set_field).  Gradient of
// J = sum_i |a_i|^2, a_i = <target|psi(T)> over the rows of block i (the
// instances of a batch): lambda = a_i target on block i, propagated backwards
// with U^dagger, and dJ/deps_s = 2 Re <lambda_s | dpsi_s>, dpsi_s being the
// forward tangent of segment s (dHs/deps = mu/Esc, E0 and Esc fixed), exact
// for the discrete map.  States at the start of every ckpt_every-th segment
// are kept; the backward sweep recomputes the others block by block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "sg4_internal.h"

namespace sg4 {

namespace {

#define CB_TRY(expr)                                                              \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr); \
      return SG4_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

struct Buf {
  double* p = nullptr;
  ~Buf() { if (p) (void)hipFree(p); }
  int alloc(size_t n) {
    if (hipMalloc(&p, sizeof(double) * std::max<size_t>(n, 1)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      set_error("device allocation of the propagation work buffers failed");
      return SG4_ERR_OOM;
    }
    return SG4_OK;
  }
};

inline unsigned blocks_for(int64_t n) {
  return (unsigned)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 16384);
}

// A complex value of state s in a row of width W (values): re at s, im at ns + s.
// r = sgn * i multiplies: r (zr + i zi) = (-sgn zi) + i (sgn zr).

// acc = c x (complex block of width 2 ns, contiguous)
__global__ void k_cb_scale(const double* __restrict__ x, double* __restrict__ acc, double c, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc[i] = c * x[i];
}

__global__ void k_cb_copy(const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}

// One order of the series on a block of width W = 2 ns (plain) per row:
//   first (k = 1): wn = r z,            z = (y - E0 wk) / Esc, wk = w_0
//   next  (k > 1): wn = wm + 2 r z,     z = (y - E0 wk) / Esc
// wn is written over wm (first: over wm as well); acc += c wn.
__global__ void k_cb_order(double* __restrict__ wm, const double* __restrict__ wk,
                           const double* __restrict__ y, double* __restrict__ acc, double c,
                           double E0, double sc, double sgn, int first, int64_t rows, int ns) {
  const int64_t n = rows * ns;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ns;
    const int s = (int)(i - r * ns);
    const int64_t re = r * 2 * ns + s, im = re + ns;
    const double zr = (y[re] - E0 * wk[re]) * sc, zi = (y[im] - E0 * wk[im]) * sc;
    const double f = first ? 1.0 : 2.0;
    const double nr = (first ? 0.0 : wm[re]) - f * sgn * zi;
    const double ni = (first ? 0.0 : wm[im]) + f * sgn * zr;
    wm[re] = nr;
    wm[im] = ni;
    acc[re] += c * nr;
    acc[im] += c * ni;
  }
}

// The same for the tangent block of width 4 ns: [w_re, w_im, t_re, t_im]
// (w the state, t its derivative), y = H applied to the whole block, mw = mu
// applied to the w part (width 2 ns):  t's z = (y_t - E0 t_k + mw) / Esc.
__global__ void k_cb_order_tan(double* __restrict__ wm, const double* __restrict__ wk,
                               const double* __restrict__ y, const double* __restrict__ mw,
                               double* __restrict__ acc, double* __restrict__ tacc, double c,
                               double E0, double sc, double sgn, int first, int64_t rows, int ns) {
  const int64_t n = rows * ns;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ns;
    const int s = (int)(i - r * ns);
    const int64_t wr = r * 4 * ns + s, wi = wr + ns, tr = wr + 2 * ns, ti = wr + 3 * ns;
    const int64_t mr = r * 2 * ns + s, mi = mr + ns;
    const double f = first ? 1.0 : 2.0;
    const double zr = (y[wr] - E0 * wk[wr]) * sc, zi = (y[wi] - E0 * wk[wi]) * sc;
    const double ur = (y[tr] - E0 * wk[tr] + mw[mr]) * sc, ui = (y[ti] - E0 * wk[ti] + mw[mi]) * sc;
    const double nwr = (first ? 0.0 : wm[wr]) - f * sgn * zi;
    const double nwi = (first ? 0.0 : wm[wi]) + f * sgn * zr;
    const double ntr = (first ? 0.0 : wm[tr]) - f * sgn * ui;
    const double nti = (first ? 0.0 : wm[ti]) + f * sgn * ur;
    wm[wr] = nwr; wm[wi] = nwi; wm[tr] = ntr; wm[ti] = nti;
    acc[mr] += c * nwr;
    acc[mi] += c * nwi;
    tacc[mr] += c * ntr;
    tacc[mi] += c * nti;
  }
}

// pack [x | t] (width 2 ns each) into a width-4 ns block; unpack its w part
__global__ void k_cb_pack(const double* __restrict__ x, const double* __restrict__ t,
                          double* __restrict__ w4, int64_t rows, int ns) {
  const int64_t n = rows * 2 * ns;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / (2 * ns);
    const int c = (int)(i - r * 2 * ns);
    w4[r * 4 * ns + c] = x[i];
    w4[r * 4 * ns + 2 * ns + c] = t ? t[i] : 0.0;
  }
}
__global__ void k_cb_wpart(const double* __restrict__ w4, double* __restrict__ x, int64_t rows, int ns) {
  const int64_t n = rows * 2 * ns;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / (2 * ns);
    x[i] = w4[r * 4 * ns + (i - r * 2 * ns)];
  }
}

// out = acc * (pr + i pi)
__global__ void k_cb_phase(const double* __restrict__ acc, double* __restrict__ out, double pr,
                           double pi, int64_t rows, int ns) {
  const int64_t n = rows * ns;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ns;
    const int s = (int)(i - r * ns);
    const int64_t re = r * 2 * ns + s, im = re + ns;
    const double ar = acc[re], ai = acc[im];
    out[re] = ar * pr - ai * pi;
    out[im] = ar * pi + ai * pr;
  }
}

// sum over rows [r0, r1) and states of conj(a) b: per-workgroup partial sums
// (re, im) into part[2 * blockIdx.x], summed on the host in a fixed order
constexpr int RED_BLOCKS = 512;
__global__ __launch_bounds__(256) void k_cb_cdot(const double* __restrict__ a,
                                                 const double* __restrict__ b, int64_t r0,
                                                 int64_t r1, int ns, double* __restrict__ part) {
  __shared__ double sr[256], si[256];
  double accr = 0.0, acci = 0.0;
  const int64_t n = (r1 - r0) * ns;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = r0 + i / ns;
    const int s = (int)(i % ns);
    const int64_t re = r * 2 * ns + s, im = re + ns;
    accr += a[re] * b[re] + a[im] * b[im];
    acci += a[re] * b[im] - a[im] * b[re];
  }
  sr[threadIdx.x] = accr;
  si[threadIdx.x] = acci;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sr[threadIdx.x] += sr[threadIdx.x + o];
      si[threadIdx.x] += si[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr[0];
    part[2 * blockIdx.x + 1] = si[0];
  }
}

// y[rows r0..r1) = (ar + i ai) x
__global__ void k_cb_cscale_rows(const double* __restrict__ x, double* __restrict__ y, double ar,
                                 double ai, int64_t r0, int64_t r1, int ns) {
  const int64_t n = (r1 - r0) * ns;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = r0 + i / ns;
    const int s = (int)(i % ns);
    const int64_t re = r * 2 * ns + s, im = re + ns;
    const double xr = x[re], xi = x[im];
    y[re] = ar * xr - ai * xi;
    y[im] = ar * xi + ai * xr;
  }
}

// One order of the Taylor series (march_noD_field) on a block of width 2 ns:
//   wn = (-i a) (y - E0 wk),  a = dt / (j + 1),  psi += wn
// with (-i a)(zr + i zi) = a zi - i a zr.
__global__ void k_tl_order(const double* __restrict__ wk, const double* __restrict__ y,
                           double* __restrict__ wn, double* __restrict__ psi, double a, double E0,
                           int64_t rows, int ns) {
  const int64_t n = rows * ns;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ns;
    const int s = (int)(i - r * ns);
    const int64_t re = r * 2 * ns + s, im = re + ns;
    const double zr = y[re] - E0 * wk[re], zi = y[im] - E0 * wk[im];
    const double nr = a * zi, ni = -a * zr;
    wn[re] = nr;
    wn[im] = ni;
    psi[re] += nr;
    psi[im] += ni;
  }
}

// psi <- (pr + i pi) psi in place
__global__ void k_tl_phase(double* psi, double pr, double pi, int64_t rows, int ns) {
  const int64_t n = rows * ns;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ns;
    const int s = (int)(i - r * ns);
    const int64_t re = r * 2 * ns + s, im = re + ns;
    const double ar = psi[re], ai = psi[im];
    psi[re] = ar * pr - ai * pi;
    psi[im] = ar * pi + ai * pr;
  }
}

// squared norm of every state of a width-2 ns block: partial sums of state s
// (grid row blockIdx.y) into part[s * gridDim.x + blockIdx.x]
constexpr int NRM_BLOCKS = 64;
__global__ __launch_bounds__(256) void k_tl_norm2(const double* __restrict__ w, int64_t rows, int ns,
                                                  double* __restrict__ part) {
  __shared__ double sh[256];
  const int s = blockIdx.y;
  double acc = 0.0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
    const double xr = w[r * 2 * ns + s], xi = w[r * 2 * ns + ns + s];
    acc += xr * xr + xi * xi;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[s * gridDim.x + blockIdx.x] = sh[0];
}

struct Prop {
  sg4_plan* H;
  sg4_plan* Mu;          // the dipole operator (potential mu, no kinetic part)
  const double* mu;      // mu on the grid points (device), for sg4_plan_set_field
  int ns;
  int64_t rows, n;       // n = rows * 2 ns
  hipStream_t st;
  const sg4_cheby_params* p;
  Buf A, B, Y, ACC;      // plain step
  Buf A4, B4, Y4, E, MW, TACC;   // tangent step
  Buf part;
  std::vector<double> hpart;

  int init(bool tangent) {
    int rc;
    if ((rc = A.alloc(n)) || (rc = B.alloc(n)) || (rc = Y.alloc(n)) || (rc = ACC.alloc(n)) ||
        (rc = part.alloc(2 * RED_BLOCKS)))
      return rc;
    if (tangent && ((rc = A4.alloc(2 * n)) || (rc = B4.alloc(2 * n)) || (rc = Y4.alloc(2 * n)) ||
                    (rc = E.alloc(n)) || (rc = MW.alloc(n)) || (rc = TACC.alloc(n))))
      return rc;
    hpart.resize(2 * RED_BLOCKS);
    return SG4_OK;
  }

  // psi <- U psi (sgn = -1) or U^dagger psi (sgn = +1)
  int step(double* psi, double sgn) {
    const dim3 g(blocks_for(rows * ns)), b(256);
    const double sc = 1.0 / p->Esc;
    int rc;
    hipLaunchKernelGGL(k_cb_copy, dim3(blocks_for(n)), b, 0, st, psi, A.p, n);
    hipLaunchKernelGGL(k_cb_scale, dim3(blocks_for(n)), b, 0, st, psi, ACC.p, p->coef[0], n);
    double* wm = A.p;    // w_{k-1}
    double* wk = B.p;    // w_k
    if (p->ncoef > 1) {
      if ((rc = sg4_apply_rows(H, 2 * ns, A.p, Y.p, p->prec))) return rc;
      hipLaunchKernelGGL(k_cb_order, g, b, 0, st, B.p, A.p, Y.p, ACC.p, p->coef[1], p->E0, sc, sgn, 1,
                         rows, ns);
    }
    for (int k = 2; k < p->ncoef; ++k) {
      if ((rc = sg4_apply_rows(H, 2 * ns, wk, Y.p, p->prec))) return rc;
      hipLaunchKernelGGL(k_cb_order, g, b, 0, st, wm, wk, Y.p, ACC.p, p->coef[k], p->E0, sc, sgn, 0,
                         rows, ns);
      std::swap(wm, wk);
    }
    const double ph = sgn * p->E0 * p->dt;   // exp(-i E0 dt) forward, its conjugate backward
    hipLaunchKernelGGL(k_cb_phase, g, b, 0, st, ACC.p, psi, std::cos(ph), std::sin(ph), rows, ns);
    CB_TRY(hipGetLastError());
    return SG4_OK;
  }

  // (psi, dpsi) <- (U psi, U dpsi + (dU/deps) psi)
  int step_tangent(double* psi, double* dpsi) {
    const dim3 g(blocks_for(rows * ns)), b(256);
    const double sc = 1.0 / p->Esc, sgn = -1.0;
    int rc;
    hipLaunchKernelGGL(k_cb_pack, dim3(blocks_for(n)), b, 0, st, psi, dpsi, A4.p, rows, ns);
    hipLaunchKernelGGL(k_cb_scale, dim3(blocks_for(n)), b, 0, st, psi, ACC.p, p->coef[0], n);
    hipLaunchKernelGGL(k_cb_scale, dim3(blocks_for(n)), b, 0, st, dpsi, TACC.p, p->coef[0], n);
    double* wm = A4.p;
    double* wk = B4.p;
    for (int k = 1; k < p->ncoef; ++k) {
      const double* src = (k == 1) ? A4.p : wk;
      if ((rc = sg4_apply_rows(H, 4 * ns, src, Y4.p, p->prec))) return rc;
      hipLaunchKernelGGL(k_cb_wpart, dim3(blocks_for(n)), b, 0, st, src, E.p, rows, ns);
      if ((rc = sg4_apply_rows(Mu, 2 * ns, E.p, MW.p, p->prec))) return rc;
      if (k == 1) {
        hipLaunchKernelGGL(k_cb_order_tan, g, b, 0, st, B4.p, A4.p, Y4.p, MW.p, ACC.p, TACC.p,
                           p->coef[1], p->E0, sc, sgn, 1, rows, ns);
      } else {
        hipLaunchKernelGGL(k_cb_order_tan, g, b, 0, st, wm, wk, Y4.p, MW.p, ACC.p, TACC.p, p->coef[k],
                           p->E0, sc, sgn, 0, rows, ns);
        std::swap(wm, wk);
      }
    }
    const double ph = -p->E0 * p->dt;
    hipLaunchKernelGGL(k_cb_phase, g, b, 0, st, ACC.p, psi, std::cos(ph), std::sin(ph), rows, ns);
    hipLaunchKernelGGL(k_cb_phase, g, b, 0, st, TACC.p, dpsi, std::cos(ph), std::sin(ph), rows, ns);
    CB_TRY(hipGetLastError());
    return SG4_OK;
  }

  // sum over rows [r0, r1) of conj(a) b (synchronises the stream)
  int cdot(const double* a, const double* b, int64_t r0, int64_t r1, double* re, double* im) {
    hipLaunchKernelGGL(k_cb_cdot, dim3(RED_BLOCKS), dim3(256), 0, st, a, b, r0, r1, ns, part.p);
    CB_TRY(hipGetLastError());
    CB_TRY(hipMemcpyAsync(hpart.data(), part.p, sizeof(double) * 2 * RED_BLOCKS, hipMemcpyDeviceToHost, st));
    CB_TRY(hipStreamSynchronize(st));
    double sr = 0, si = 0;
    for (int i = 0; i < RED_BLOCKS; ++i) { sr += hpart[2 * i]; si += hpart[2 * i + 1]; }
    *re = sr;
    *im = si;
    return SG4_OK;
  }

  int field(int s) { return sg4_plan_set_field(H, mu, s < 0 ? 0.0 : p->eps[s]); }
};

// Restores the field-free potential of H when a propagation leaves on an
// error path (the normal end restores it itself and clears H).
struct FieldReset {
  sg4_plan* H;
  ~FieldReset() { if (H) (void)sg4_plan_set_field(H, nullptr, 0.0); }
};

int check_params(const sg4_plan* H, int32_t ns, const sg4_cheby_params* p, int width) {
  if (!H || !p || !p->coef || !p->eps) { set_error("null argument"); return SG4_ERR_ARG; }
  if (ns < 1 || width * ns > plan_max_nvec(H)) {
    set_error("ns out of range: the plan's max_nvec must hold " + std::to_string(width) +
              " ns vectors");
    return SG4_ERR_ARG;
  }
  if (p->ncoef < 1 || p->nseg < 1 || p->steps_per_seg < 1 || !(p->Esc > 0) || !(p->dt > 0)) {
    set_error("bad propagation parameters");
    return SG4_ERR_ARG;
  }
  return SG4_OK;
}

}  // namespace
}  // namespace sg4

using namespace sg4;

extern "C" {

int sg4_cheby_propagate(sg4_plan* H, const double* mu, int32_t ns, double* psi,
                        const sg4_cheby_params* prm) {
  int rc;
  if ((rc = check_params(H, ns, prm, 2))) return rc;
  if (!psi) { set_error("null argument"); return SG4_ERR_ARG; }
  if ((rc = device_block(H, psi, "psi")) || (mu && (rc = device_block(H, mu, "mu")))) return rc;
  Prop P{};
  P.H = H; P.Mu = nullptr; P.mu = mu; P.ns = ns; P.rows = plan_rows(H);
  P.n = P.rows * 2 * ns; P.st = (hipStream_t)plan_stream(H); P.p = prm;
  if ((rc = P.init(false))) return rc;
  FieldReset reset{H};
  for (int s = 0; s < prm->nseg; ++s) {
    if ((rc = P.field(s))) return rc;
    for (int j = 0; j < prm->steps_per_seg; ++j)
      if ((rc = P.step(psi, -1.0))) return rc;
  }
  reset.H = nullptr;
  if ((rc = sg4_plan_set_field(H, nullptr, 0.0))) return rc;
  CB_TRY(hipStreamSynchronize(P.st));
  return SG4_OK;
}

int sg4_cheby_field_gradient(sg4_plan* H, sg4_plan* Mu, const double* mu, int32_t ns,
                             const double* psi0, const double* target, int32_t nblk,
                             const int64_t* blk_rows, const sg4_cheby_params* prm, double* J,
                             double* grad) {
  int rc;
  if ((rc = check_params(H, ns, prm, 4))) return rc;
  if (!Mu || !psi0 || !target || !J || !grad || nblk < 1 || !blk_rows || !mu) {
    set_error("null argument");
    return SG4_ERR_ARG;
  }
  if (plan_rows(Mu) != plan_rows(H) || plan_max_nvec(Mu) < 2 * ns) {
    set_error("the mu operator must have the rows of H and hold 2 ns vectors");
    return SG4_ERR_ARG;
  }
  if ((rc = device_block(H, psi0, "psi0")) || (rc = device_block(H, target, "target")) ||
      (rc = device_block(H, mu, "mu")) || (rc = device_block(Mu, mu, "mu")))
    return rc;
  const int64_t rows = plan_rows(H);
  for (int i = 0; i < nblk; ++i)
    if (blk_rows[2 * i] < 0 || blk_rows[2 * i + 1] > rows || blk_rows[2 * i] >= blk_rows[2 * i + 1]) {
      set_error("block row range out of bounds");
      return SG4_ERR_ARG;
    }
  const int S = prm->nseg, C = std::max(1, prm->ckpt_every);
  Prop P{};
  P.H = H; P.Mu = Mu; P.mu = mu; P.ns = ns; P.rows = rows;
  P.n = rows * 2 * ns; P.st = (hipStream_t)plan_stream(H); P.p = prm;
  void* mu_stream = plan_stream(Mu);
  if ((rc = sg4_plan_set_stream(Mu, P.st))) return rc;
  struct Restore {
    sg4_plan* m; void* s;
    ~Restore() { (void)sg4_plan_set_stream(m, s); }
  } restore{Mu, mu_stream};
  if ((rc = P.init(true))) return rc;
  FieldReset reset{H};
  const dim3 b(256);
  // forward, keeping the state at the start of every C-th segment
  std::vector<Buf> ckpt((S + C - 1) / C);
  Buf cur;
  if ((rc = cur.alloc(P.n))) return rc;
  hipLaunchKernelGGL(k_cb_copy, dim3(blocks_for(P.n)), b, 0, P.st, psi0, cur.p, P.n);
  for (int s = 0; s < S; ++s) {
    if (s % C == 0) {
      if ((rc = ckpt[s / C].alloc(P.n))) return rc;
      hipLaunchKernelGGL(k_cb_copy, dim3(blocks_for(P.n)), b, 0, P.st, cur.p, ckpt[s / C].p, P.n);
    }
    if ((rc = P.field(s))) return rc;
    for (int j = 0; j < prm->steps_per_seg; ++j)
      if ((rc = P.step(cur.p, -1.0))) return rc;
  }
  // overlaps and the adjoint seed
  Buf lam, st, dpsi;
  if ((rc = lam.alloc(P.n)) || (rc = st.alloc(P.n)) || (rc = dpsi.alloc(P.n))) return rc;
  hipLaunchKernelGGL(k_cb_scale, dim3(blocks_for(P.n)), b, 0, P.st, target, lam.p, 0.0, P.n);
  double Jv = 0.0;
  for (int i = 0; i < nblk; ++i) {
    double ar, ai;
    if ((rc = P.cdot(target, cur.p, blk_rows[2 * i], blk_rows[2 * i + 1], &ar, &ai))) return rc;
    Jv += ar * ar + ai * ai;
    const int64_t nr = (blk_rows[2 * i + 1] - blk_rows[2 * i]) * ns;
    hipLaunchKernelGGL(k_cb_cscale_rows, dim3(blocks_for(nr)), b, 0, P.st, target, lam.p, ar, ai,
                       blk_rows[2 * i], blk_rows[2 * i + 1], ns);
  }
  *J = Jv;
  // backward sweep: recompute the states of a checkpoint block, tangent of
  // each segment against the adjoint state, then the adjoint step
  std::vector<Buf> seg(C);
  int have = -1;   // checkpoint block whose states are in seg
  for (int s = S - 1; s >= 0; --s) {
    const int cb = s / C, c0 = cb * C;
    if (have != cb) {
      hipLaunchKernelGGL(k_cb_copy, dim3(blocks_for(P.n)), b, 0, P.st, ckpt[cb].p, st.p, P.n);
      for (int u = c0; u <= s; ++u) {
        if (!seg[u - c0].p && (rc = seg[u - c0].alloc(P.n))) return rc;
        hipLaunchKernelGGL(k_cb_copy, dim3(blocks_for(P.n)), b, 0, P.st, st.p, seg[u - c0].p, P.n);
        if (u == s) break;
        if ((rc = P.field(u))) return rc;
        for (int j = 0; j < prm->steps_per_seg; ++j)
          if ((rc = P.step(st.p, -1.0))) return rc;
      }
      have = cb;
    }
    if ((rc = P.field(s))) return rc;
    hipLaunchKernelGGL(k_cb_copy, dim3(blocks_for(P.n)), b, 0, P.st, seg[s - c0].p, st.p, P.n);
    hipLaunchKernelGGL(k_cb_scale, dim3(blocks_for(P.n)), b, 0, P.st, st.p, dpsi.p, 0.0, P.n);
    for (int j = 0; j < prm->steps_per_seg; ++j)
      if ((rc = P.step_tangent(st.p, dpsi.p))) return rc;
    double gr, gi;
    if ((rc = P.cdot(lam.p, dpsi.p, 0, rows, &gr, &gi))) return rc;
    grad[s] = 2.0 * gr;
    for (int j = 0; j < prm->steps_per_seg; ++j)
      if ((rc = P.step(lam.p, +1.0))) return rc;
  }
  reset.H = nullptr;
  if ((rc = sg4_plan_set_field(H, nullptr, 0.0))) return rc;
  CB_TRY(hipStreamSynchronize(P.st));
  return SG4_OK;
}

int sg4_taylor_propagate(sg4_plan* H, const double* mu, int32_t ns, double* psi,
                         const sg4_taylor_params* prm, int32_t* orders) {
  if (!H || !prm || !prm->eps || !psi) { set_error("null argument"); return SG4_ERR_ARG; }
  if (ns < 1 || 2 * ns > plan_max_nvec(H)) {
    set_error("ns out of range: the plan's max_nvec must hold 2 ns vectors");
    return SG4_ERR_ARG;
  }
  if (prm->npoly < 1 || prm->nseg < 1 || prm->steps_per_seg < 1 || !(prm->dt > 0) ||
      !(prm->poly_tol > 0) || !(prm->max_norm2 >= 0)) {
    set_error("bad propagation parameters");
    return SG4_ERR_ARG;
  }
  int rc;
  if ((rc = device_block(H, psi, "psi")) || (mu && (rc = device_block(H, mu, "mu")))) return rc;
  const int64_t rows = plan_rows(H), n = rows * 2 * ns;
  const hipStream_t st = (hipStream_t)plan_stream(H);
  Buf A, B, Y, part;
  if ((rc = A.alloc(n)) || (rc = B.alloc(n)) || (rc = Y.alloc(n)) ||
      (rc = part.alloc((size_t)NRM_BLOCKS * ns)))
    return rc;
  std::vector<double> hpart((size_t)NRM_BLOCKS * ns);
  // largest squared norm over the states of a width-2 ns block (synchronises)
  auto max_norm2 = [&](const double* w, double* out) -> int {
    hipLaunchKernelGGL(k_tl_norm2, dim3(NRM_BLOCKS, ns), dim3(256), 0, st, w, rows, ns, part.p);
    CB_TRY(hipGetLastError());
    CB_TRY(hipMemcpyAsync(hpart.data(), part.p, sizeof(double) * hpart.size(), hipMemcpyDeviceToHost, st));
    CB_TRY(hipStreamSynchronize(st));
    double m = 0.0;
    for (int s = 0; s < ns; ++s) {
      double a = 0.0;
      for (int i = 0; i < NRM_BLOCKS; ++i) a += hpart[(size_t)s * NRM_BLOCKS + i];
      m = (a > m || a != a) ? a : m;
    }
    *out = m;
    return SG4_OK;
  };
  const dim3 g(blocks_for(rows * ns)), b(256);
  const double ph = -prm->E0 * prm->dt;
  FieldReset reset{H};
  int64_t step = 0;
  for (int s = 0; s < prm->nseg; ++s) {
    if ((rc = sg4_plan_set_field(H, mu, prm->eps[s]))) return rc;
    for (int j = 0; j < prm->steps_per_seg; ++j, ++step) {
      hipLaunchKernelGGL(k_cb_copy, dim3(blocks_for(n)), b, 0, st, psi, A.p, n);
      double* wk = A.p;
      double* wn = B.p;
      double m = 0.0;
      int used = 0;
      for (int k = 0; k < prm->npoly; ++k) {
        if ((rc = sg4_apply_rows(H, 2 * ns, wk, Y.p, prm->prec))) return rc;
        hipLaunchKernelGGL(k_tl_order, g, b, 0, st, wk, Y.p, wn, psi, prm->dt / (k + 1), prm->E0, rows, ns);
        if ((rc = max_norm2(wn, &m))) return rc;
        used = k + 1;
        if (m < prm->poly_tol) break;
        if (!(m <= 1e15)) {
          set_error("Taylor propagation: the squared norm of term " + std::to_string(used) + " is " +
                    std::to_string(m) + " (> 1e15): reduce the time step");
          return SG4_ERR_STATE;
        }
        std::swap(wk, wn);
      }
      if (!(m < prm->poly_tol)) {
        set_error("Taylor propagation: the series did not reach poly_tol within npoly = " +
                  std::to_string(prm->npoly) + " orders (last squared norm " + std::to_string(m) + ")");
        return SG4_ERR_STATE;
      }
      hipLaunchKernelGGL(k_tl_phase, g, b, 0, st, psi, std::cos(ph), std::sin(ph), rows, ns);
      CB_TRY(hipGetLastError());
      if (prm->max_norm2 > 0) {
        double pn;
        if ((rc = max_norm2(psi, &pn))) return rc;
        if (!(pn <= prm->max_norm2)) {
          set_error("Taylor propagation: a state's squared norm " + std::to_string(pn) +
                    " exceeds max_norm2 after step " + std::to_string(step + 1));
          return SG4_ERR_STATE;
        }
      }
      if (orders) orders[step] = used;
    }
  }
  reset.H = nullptr;
  if ((rc = sg4_plan_set_field(H, nullptr, 0.0))) return rc;
  CB_TRY(hipStreamSynchronize(st));
  return SG4_OK;
}

}  // extern "C"
