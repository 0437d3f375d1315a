// Host-side set-up of the Smolyak SG4 basis: 1-D primitive bases, grid
// derivative matrices, term set, combination weights, packed basis and the
// packed<->term map.  These are the tables the H.psi kernels consume; in
// ElVibRot they are produced by RecSparseGrid_ForDP_type4
// (SRC/sub_Basis/sub_quadra_SparseBasis.f90:922-1434).  A Fortran host that
// already owns them passes them straight to sg4_plan_create instead.
#include "sg4_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace sg4 {

// ---------------------------------------------------------------------------
// Symmetric tridiagonal eigenvalues (implicit QL with Wilkinson shifts),
// used for the Golub-Welsch nodes; polished afterwards by Newton steps.
// ---------------------------------------------------------------------------
static void tridiag_ql(std::vector<double>& d, std::vector<double> e) {
  const int n = (int)d.size();
  e.push_back(0.0);
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-17 * dd) break;
      }
      if (m != l) {
        if (++iter > 200) throw std::runtime_error("QL did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; break; }
          s = f / r; c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p; e[l] = g; e[m] = 0.0;
      }
    } while (m != l);
  }
  std::sort(d.begin(), d.end());
}

// Orthonormal Hermite polynomials (weight e^{-x^2}) evaluated up to degree n.
static void hermite_table(long double x, int n, long double* p) {
  p[0] = std::pow(std::acos(-1.0L), -0.25L);
  if (n >= 1) p[1] = std::sqrt(2.0L) * x * p[0];
  for (int k = 2; k <= n; ++k)
    p[k] = std::sqrt(2.0L / k) * x * p[k - 1] - std::sqrt((k - 1.0L) / k) * p[k - 2];
}

// Gauss-Hermite rule; weights returned multiplied by e^{x^2} (the quantity
// ElVibRot reads from Internal_data/HermQuadra/herm{n}.txt column 4,
// sub_quadra_herm.f90:160-178).
void gauss_hermite(int n, std::vector<double>& x, std::vector<double>& wexp) {
  std::vector<double> d(n, 0.0), e(n > 1 ? n - 1 : 0);
  for (int i = 0; i + 1 < n; ++i) e[i] = std::sqrt(0.5 * (i + 1));
  tridiag_ql(d, e);
  x.assign(n, 0.0);
  wexp.assign(n, 0.0);
  std::vector<long double> p(n + 2);
  for (int i = 0; i < n; ++i) {
    long double xx = d[i];
    for (int it = 0; it < 10; ++it) {
      hermite_table(xx, n, p.data());
      long double dx = p[n] / (std::sqrt(2.0L * n) * p[n - 1]);
      xx -= dx;
      if (std::fabs(dx) < 1e-22L * (1 + std::fabs(xx))) break;
    }
    hermite_table(xx, n - 1, p.data());
    long double s = 0, ex = std::exp(-xx * xx);
    for (int k = 0; k < n; ++k) s += p[k] * p[k] * ex;
    x[i] = (double)xx;
    wexp[i] = (double)(1.0L / s);
  }
  for (int i = 0; i < n / 2; ++i) {
    double xs = 0.5 * (x[n - 1 - i] - x[i]), ws = 0.5 * (wexp[i] + wexp[n - 1 - i]);
    x[i] = -xs; x[n - 1 - i] = xs; wexp[i] = ws; wexp[n - 1 - i] = ws;
  }
  if (n % 2) x[n / 2] = 0.0;
}

void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  std::vector<double> d(n, 0.0), e(n > 1 ? n - 1 : 0);
  for (int i = 0; i + 1 < n; ++i) { double k = i + 1; e[i] = k / std::sqrt(4 * k * k - 1); }
  tridiag_ql(d, e);
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    long double xx = d[i], pn = 0, pm = 0;
    for (int it = 0; it < 12; ++it) {
      long double a = 1, b = xx;
      for (int j = 2; j <= n; ++j) { long double c = ((2 * j - 1) * xx * b - (j - 1) * a) / j; a = b; b = c; }
      if (n == 1) { a = 1; b = xx; }
      pn = b; pm = a;
      long double dp = n * (xx * pn - pm) / (xx * xx - 1.0L);
      long double dx = pn / dp;
      xx -= dx;
      if (std::fabs(dx) < 1e-22L) break;
    }
    long double a = 1, b = xx;
    for (int j = 2; j <= n; ++j) { long double c = ((2 * j - 1) * xx * b - (j - 1) * a) / j; a = b; b = c; }
    if (n == 1) { a = 1; b = xx; }
    long double dp = n * (xx * b - a) / (xx * xx - 1.0L);
    x[i] = (double)xx;
    w[i] = (double)(2.0L / ((1.0L - xx * xx) * dp * dp));
  }
  for (int i = 0; i < n / 2; ++i) {
    double xs = 0.5 * (x[n - 1 - i] - x[i]), ws = 0.5 * (w[i] + w[n - 1 - i]);
    x[i] = -xs; x[n - 1 - i] = xs; w[i] = ws; w[n - 1 - i] = ws;
  }
  if (n % 2) x[n / 2] = 0.0;
}

// ---------------------------------------------------------------------------
// 1-D primitive bases, column-major (nq x nb).
// ---------------------------------------------------------------------------

// HO: normalised Hermite functions with first and second derivative
// (d0d1d2poly_Hermite_exp, Ext_Lib/FOR_EVRT_loc/SRC/sub_communf90/sub_math/
// sub_polyortho.f90:1090-1153), scaled by Q = Q0 + x/scaleQ
// (sub_scale_basis, SRC/sub_Basis/sub_module_basis.f90:1499).
static void ho_basis(Basis1D& b, double Q0, double scaleQ) {
  std::vector<double> xr, wr;
  gauss_hermite(b.nq, xr, wr);
  b.x.resize(b.nq); b.w.resize(b.nq);
  b.d0.assign((size_t)b.nq * b.nb, 0); b.d1 = b.d0; b.d2 = b.d0;
  std::vector<long double> p(b.nb + 1);
  const long double sc = scaleQ, s0 = std::sqrt(sc), s1 = s0 * sc, s2 = s1 * sc;
  for (int iq = 0; iq < b.nq; ++iq) {
    long double xx = xr[iq], ex = std::exp(-0.5L * xx * xx);
    hermite_table(xx, b.nb, p.data());
    for (int n = 0; n < b.nb; ++n) {
      long double f = p[n] * ex;
      long double fd = ((n > 0 ? std::sqrt(2.0L * n) * p[n - 1] : 0.0L) - xx * p[n]) * ex;
      long double fdd = (xx * xx - (2.0L * n + 1.0L)) * f;
      b.d0[iq + (size_t)b.nq * n] = (double)(f * s0);
      b.d1[iq + (size_t)b.nq * n] = (double)(fd * s1);
      b.d2[iq + (size_t)b.nq * n] = (double)(fdd * s2);
    }
    b.x[iq] = (double)(Q0 + xx / sc);
    b.w[iq] = (double)(wr[iq] / sc);
  }
}

// Legendre basis in theta: sqrt((2l+1)/2) P_l(cos theta) with derivatives in
// theta; Gauss-Legendre weights in cos(theta) (SRC/sub_Basis/
// sub_quadra_legendre.f90:48).  Points are in ascending theta.
static void legendre_basis(Basis1D& b) {
  std::vector<double> u, wu;
  gauss_legendre(b.nq, u, wu);
  b.x.resize(b.nq); b.w.resize(b.nq);
  b.d0.assign((size_t)b.nq * b.nb, 0); b.d1 = b.d0; b.d2 = b.d0;
  for (int iq = 0; iq < b.nq; ++iq) {
    int ju = b.nq - 1 - iq;
    long double uu = u[ju], th = std::acos(uu), s = std::sin(th), c = std::cos(th);
    long double P0 = 1, P1 = uu, dP0 = 0, dP1 = 1, ddP0 = 0, ddP1 = 0;
    for (int l = 0; l < b.nb; ++l) {
      long double P, dP, ddP;
      if (l == 0) { P = 1; dP = 0; ddP = 0; }
      else if (l == 1) { P = uu; dP = 1; ddP = 0; }
      else {
        P = ((2 * l - 1) * uu * P1 - (l - 1) * P0) / l;
        dP = l * P1 + uu * dP1;
        ddP = (l + 1) * dP1 + uu * ddP1;
        P0 = P1; P1 = P; dP0 = dP1; dP1 = dP; ddP0 = ddP1; ddP1 = ddP;
      }
      long double nrm = std::sqrt((2.0L * l + 1.0L) / 2.0L);
      b.d0[iq + (size_t)b.nq * l] = (double)(nrm * P);
      b.d1[iq + (size_t)b.nq * l] = (double)(-nrm * s * dP);
      b.d2[iq + (size_t)b.nq * l] = (double)(nrm * (s * s * ddP - c * dP));
    }
    (void)dP0; (void)ddP0;
    b.x[iq] = (double)th;
    b.w[iq] = wu[ju];
  }
}

// dnRGG%dk = dnRGB%dk . pinv(d0b), pinv = (d0b^T d0b)^-1 d0b^T
// (sub_dnGB_TO_dnGG, SRC/sub_Basis/sub_module_basis.f90:2048-2329).
// Cholesky of the (SPD) normal matrix in long double.
static void grid_derivatives(Basis1D& b) {
  const int nq = b.nq, nb = b.nb;
  std::vector<long double> G((size_t)nb * nb), P((size_t)nb * nq);
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) {
      long double s = 0;
      for (int q = 0; q < nq; ++q) s += (long double)b.d0[q + (size_t)nq * i] * b.d0[q + (size_t)nq * j];
      G[i + (size_t)nb * j] = s;
    }
  for (int j = 0; j < nb; ++j) {  // Cholesky, lower
    long double s = G[j + (size_t)nb * j];
    for (int k = 0; k < j; ++k) s -= G[j + (size_t)nb * k] * G[j + (size_t)nb * k];
    if (s <= 0) throw std::runtime_error("basis matrix is singular");
    G[j + (size_t)nb * j] = std::sqrt(s);
    for (int i = j + 1; i < nb; ++i) {
      long double t = G[i + (size_t)nb * j];
      for (int k = 0; k < j; ++k) t -= G[i + (size_t)nb * k] * G[j + (size_t)nb * k];
      G[i + (size_t)nb * j] = t / G[j + (size_t)nb * j];
    }
  }
  for (int q = 0; q < nq; ++q) {  // solve G G^T p = d0b(q,:)^T
    long double* col = &P[(size_t)nb * q];
    for (int i = 0; i < nb; ++i) col[i] = b.d0[q + (size_t)nq * i];
    for (int i = 0; i < nb; ++i) {
      long double t = col[i];
      for (int k = 0; k < i; ++k) t -= G[i + (size_t)nb * k] * col[k];
      col[i] = t / G[i + (size_t)nb * i];
    }
    for (int i = nb - 1; i >= 0; --i) {
      long double t = col[i];
      for (int k = i + 1; k < nb; ++k) t -= G[k + (size_t)nb * i] * col[k];
      col[i] = t / G[i + (size_t)nb * i];
    }
  }
  auto prod = [&](const std::vector<double>& src, std::vector<double>& dst) {
    dst.assign((size_t)nq * nq, 0.0);
    for (int q = 0; q < nq; ++q)
      for (int r = 0; r < nq; ++r) {
        long double s = 0;
        for (int i = 0; i < nb; ++i) s += (long double)src[q + (size_t)nq * i] * P[i + (size_t)nb * r];
        dst[q + (size_t)nq * r] = (double)s;
      }
  };
  prod(b.d1, b.gg1);
  prod(b.d2, b.gg2);
  b.BtW.assign((size_t)nb * nq, 0.0);  // dnRBGwrho%d0 (sub_dnGB_TO_dnBG :2001-2003)
  for (int i = 0; i < nb; ++i)
    for (int q = 0; q < nq; ++q) b.BtW[i + (size_t)nb * q] = b.d0[q + (size_t)nq * i] * b.w[q];
}

// ---------------------------------------------------------------------------
// Smolyak structure
// ---------------------------------------------------------------------------
static inline int nrule(int A, int B, int l) { return A + B * l; }

Setup::Setup(int D_, int L_, int LB_, const sg4_mode_desc* m) : D(D_), L(L_), LB(LB_) {
  if (D < 1 || D > SG4_MAX_D || L < 0 || L > SG4_MAX_L || LB < 0 || LB > L)
    throw std::invalid_argument("bad D, L or LB");
  modes.assign(m, m + D);
  Lmin = std::max(0, L - D + 1);
  b1.resize(D);
  for (int k = 0; k < D; ++k) {
    const sg4_mode_desc& md = modes[k];
    if (md.kind != SG4_BASIS_HO && md.kind != SG4_BASIS_LEGENDRE)
      throw std::invalid_argument("unknown basis kind");
    if (md.scaleQ <= 0) throw std::invalid_argument("scaleQ must be > 0");
    b1[k].resize(L + 1);
    for (int l = 0; l <= L; ++l) {
      Basis1D& b = b1[k][l];
      b.nq = nrule(md.nq_A, md.nq_B, l);
      b.nb = nrule(md.nb_A, md.nb_B, l);
      if (b.nb < 1 || b.nb > b.nq || b.nq > SG4_MAX_N)
        throw std::invalid_argument("need 1 <= nb(l) <= nq(l) <= 64");
      if (md.kind == SG4_BASIS_HO) ho_basis(b, md.Q0, md.scaleQ); else legendre_basis(b);
      grid_derivatives(b);
    }
  }
  // level of each 1-D basis index: smallest l with nb(l) > i
  // (Get_L_FROM_Basis_L_TO_n, SRC/sub_Basis/sub_module_Basis_LTO_n.f90:512)
  const int nmax1 = SG4_MAX_N;
  lam.assign((size_t)D * nmax1, 1 << 20);
  for (int k = 0; k < D; ++k)
    for (int i = 0; i < nmax1; ++i) {
      int l = 0;
      while (l <= L && nrule(modes[k].nb_A, modes[k].nb_B, l) <= i) ++l;
      if (l <= L) lam[(size_t)k * nmax1 + i] = l;
    }
  build_terms();
  build_packed();
  build_map();
}

// Term set {l : Lmin <= |l| <= L}, first index fastest
// (ADD_ONE_TO_nDindex_type5m, Ext_Lib/nDindex_loc/SRC/sub_nDindex/
// sub_module_nDindex.f90:2737-2790), weights (-1)^(L-|l|) C(D-1, L-|l|)
// (calc_Weight_OF_SRep, SRC/sub_Basis/sub_module_param_SGType2.f90:795-806).
void Setup::build_terms() {
  std::vector<int> l(D, 0);
  auto binom = [](int n, int k) -> double {
    if (k < 0 || k > n) return 0.0;
    double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return std::round(r);
  };
  int sum = 0;
  boff.push_back(0); qoff.push_back(0);
  for (;;) {
    if (sum >= Lmin) {
      int64_t nbt = 1, nqt = 1;
      for (int k = 0; k < D; ++k) { nbt *= b1[k][l[k]].nb; nqt *= b1[k][l[k]].nq; }
      levels.insert(levels.end(), l.begin(), l.end());
      int dl = L - sum;
      weights.push_back(((dl & 1) ? -1.0 : 1.0) * binom(D - 1, dl));
      boff.push_back(boff.back() + nbt);
      qoff.push_back(qoff.back() + nqt);
    }
    int k = 0;  // odometer, first index fastest, |l| <= L
    for (; k < D; ++k) {
      if (sum < L) { ++l[k]; ++sum; break; }
      sum -= l[k]; l[k] = 0;
    }
    if (k == D) break;
  }
  nterms = (int64_t)weights.size();
}

// Packed basis {i : sum_k lam_k(i_k) <= LB}, lexicographic with the last index
// fastest (ADD_ONE_TO_nDindex_type5p, sub_module_nDindex.f90:2670-2736).
// Ranks are computed with suffix counts: cnt[k][b] = number of index tuples
// (i_k..i_{D-1}) with level sum <= b.
void Setup::build_packed() {
  const int nmax1 = SG4_MAX_N;
  cnt.assign((size_t)(D + 1) * (LB + 1), 0);
  for (int b = 0; b <= LB; ++b) cnt[(size_t)D * (LB + 1) + b] = 1;
  for (int k = D - 1; k >= 0; --k)
    for (int b = 0; b <= LB; ++b) {
      int64_t s = 0;
      for (int i = 0; i < nmax1; ++i) {
        int li = lam[(size_t)k * nmax1 + i];
        if (li > b) break;
        s += cnt[(size_t)(k + 1) * (LB + 1) + (b - li)];
      }
      cnt[(size_t)k * (LB + 1) + b] = s;
    }
  nb_ba = cnt[LB];
  // prefix[k][b][i] = sum_{j<i} cnt[k+1][b - lam_k(j)]
  prefix.assign((size_t)D * (LB + 1) * (nmax1 + 1), -1);
  for (int k = 0; k < D; ++k)
    for (int b = 0; b <= LB; ++b) {
      int64_t s = 0;
      int64_t* p = &prefix[((size_t)k * (LB + 1) + b) * (nmax1 + 1)];
      for (int i = 0; i <= nmax1; ++i) {
        p[i] = s;
        if (i == nmax1) break;
        int li = lam[(size_t)k * nmax1 + i];
        if (li > b) { for (int j = i + 1; j <= nmax1; ++j) p[j] = -1; break; }
        s += cnt[(size_t)(k + 1) * (LB + 1) + (b - li)];
      }
    }
}

int64_t Setup::rank(const int* idx) const {
  const int nmax1 = SG4_MAX_N;
  int b = LB;
  int64_t r = 0;
  for (int k = 0; k < D; ++k) {
    int li = lam[(size_t)k * nmax1 + idx[k]];
    if (li > b) return -1;
    r += prefix[((size_t)k * (LB + 1) + b) * (nmax1 + 1) + idx[k]];
    b -= li;
  }
  return r;
}

// packed row of every in-term basis function, first index fastest; -1 where
// the level sum exceeds LB (Set_tables_FOR_SmolyakRepBasis_TO_tabPackedBasis,
// SRC/sub_Basis/sub_Basis_SG4/sub_module_basis_BtoG_GtoB_SG4.f90:630-954).
void Setup::build_map() {
  map.assign((size_t)boff.back(), -1);
  if (nb_ba > INT32_MAX) throw std::runtime_error("packed basis exceeds int32");
  for (int64_t t = 0; t < nterms; ++t) {
    int idx[SG4_MAX_D] = {0}, nk[SG4_MAX_D];
    for (int k = 0; k < D; ++k) nk[k] = b1[k][levels[(size_t)t * D + k]].nb;
    int32_t* mp = &map[(size_t)boff[t]];
    const int64_t nbt = boff[t + 1] - boff[t];
    for (int64_t e = 0; e < nbt; ++e) {
      mp[e] = (int32_t)rank(idx);
      for (int k = 0; k < D; ++k) { if (++idx[k] < nk[k]) break; idx[k] = 0; }
    }
  }
  // concatenated matrices in the C-ABI order: for k, for l: B, BtW, D1, D2
  nq_tab.assign((size_t)D * (L + 1), 0);
  nb_tab = nq_tab;
  for (int k = 0; k < D; ++k)
    for (int l = 0; l <= L; ++l) {
      const Basis1D& b = b1[k][l];
      nq_tab[(size_t)k * (L + 1) + l] = b.nq;
      nb_tab[(size_t)k * (L + 1) + l] = b.nb;
      mats.insert(mats.end(), b.d0.begin(), b.d0.end());
      mats.insert(mats.end(), b.BtW.begin(), b.BtW.end());
      mats.insert(mats.end(), b.gg1.begin(), b.gg1.end());
      mats.insert(mats.end(), b.gg2.begin(), b.gg2.end());
      nodes.insert(nodes.end(), b.x.begin(), b.x.end());
      qweights.insert(qweights.end(), b.w.begin(), b.w.end());
    }
}

void Setup::fill_tables(sg4_tables* t) const {
  std::memset(t, 0, sizeof(*t));
  t->D = D;
  t->Lmax = L;
  t->nterms = nterms;
  t->nb_ba = nb_ba;
  t->levels = levels.data();
  t->weights = weights.data();
  t->map = map.data();
  t->nq = nq_tab.data();
  t->nb = nb_tab.data();
  t->mats = mats.data();
  t->nodes = nodes.data();
  t->qweights = qweights.data();
}

// grid point coordinates of term t (first index fastest),
// Rec_Qact_SG4_with_Tab_iq (SRC/sub_Basis/sub_module_basis.f90:3979)
void Setup::term_grid(int64_t t, double* q) const {
  int idx[SG4_MAX_D] = {0};
  const int* lv = &levels[(size_t)t * D];
  const int64_t nqt = qoff[t + 1] - qoff[t];
  for (int64_t e = 0; e < nqt; ++e) {
    for (int k = 0; k < D; ++k) q[(size_t)e * D + k] = b1[k][lv[k]].x[idx[k]];
    for (int k = 0; k < D; ++k) { if (++idx[k] < b1[k][lv[k]].nq) break; idx[k] = 0; }
  }
}

}  // namespace sg4
