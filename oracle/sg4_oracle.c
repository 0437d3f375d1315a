/*
 * sg4_oracle.c -- CPU FP64 restatement of ElVibRot's Smolyak (SG type 4)
 * Hamiltonian action H.psi, used ONLY as the checker of the HIP library
 * (by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg).
 * Nothing in the product links, loads or calls this file.
 *
 * Every function cites the reference routine whose behaviour it restates.
 * The code is written from the mathematics, in plain loops, for clarity
 * rather than speed.  OpenMP (when compiled with -fopenmp) splits the term
 * loop, like the reference (sub_OpPsi_SG4.f90:820-864); the scatter then
 * goes through per-thread accumulators that are summed in thread order.
 *
 * Pins (each checked by tests/test_oracle.py):
 *   - Gauss-Hermite / Gauss-Legendre nodes and weights: the 50-digit tables
 *     Internal_data/HermQuadra/herm{n}.txt (fixtures under tests/golden/).
 *   - HO basis values and derivatives: orthonormality under the quadrature,
 *     Hermite ODE phi'' = (x^2 - 2n - 1) phi.
 *   - Grid derivative matrices: exact on the basis functions themselves.
 *   - Smolyak weights: sum of weights == 1; counts against closed forms.
 *   - Whole operator: the 28 (6D, L=3) and 22 (21D, L=2) Henon-Heiles
 *     eigenvalues of Working_tests/MPI_tests/{6D,21D}_Davidson_S0/benchmark.
 *   - Type-10 operator with constant G equals the type-1 operator
 *     (closed-form identity on a finite grid: d1 G d1 vs d2 differ only by
 *     the quadrature; checked on eigenvalues), rigid-rotor l(l+1) spectrum.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXD 32
#define ORC_MAXL 40

/* ------------------------------------------------------------------------ */
/* 1-D quadratures                                                           */
/* ------------------------------------------------------------------------ */

/* Number of eigenvalues of the symmetric tridiagonal matrix (diag a, off b)
 * that are < x (Sturm sequence count). */
static int sturm_count(int n, const double *a, const double *b, double x) {
  int c = 0;
  long double q = 1.0L;
  for (int i = 0; i < n; i++) {
    long double bb = (i > 0) ? (long double)b[i - 1] * b[i - 1] : 0.0L;
    q = (long double)a[i] - x - (i > 0 ? bb / q : 0.0L);
    if (q == 0.0L) q = -1e-300L;
    if (q < 0) c++;
  }
  return c;
}

/* k-th (0-based, ascending) eigenvalue by bisection. */
static double tridiag_eig(int n, const double *a, const double *b, int k,
                          double lo, double hi) {
  for (int it = 0; it < 200; it++) {
    double mid = 0.5 * (lo + hi);
    if (mid == lo || mid == hi) break;
    if (sturm_count(n, a, b, mid) > k) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

/* Orthonormal Hermite polynomials p_0..p_{n-1} at x (weight e^{-x^2}),
 * in long double.  Returns p_{n-1}, p_{n-2} via out args; sum of squares of
 * the Hermite functions p_k e^{-x^2/2}, k<m, in *s2f (if non-null). */
static void herm_orthonormal(int n, long double x, long double *pn,
                             long double *pnm1, int m, long double *s2f) {
  long double p0 = powl(acosl(-1.0L), -0.25L), p1 = 0, s = 0;
  long double e = expl(-0.5L * x * x);
  if (m > 0) s += (p0 * e) * (p0 * e);
  if (n == 1) { *pn = p0; *pnm1 = 0; if (s2f) *s2f = s; return; }
  p1 = sqrtl(2.0L) * x * p0;
  if (m > 1) s += (p1 * e) * (p1 * e);
  long double pm2 = p0, pm1 = p1;
  for (int k = 2; k < n; k++) {
    long double pk = sqrtl(2.0L / k) * x * pm1 - sqrtl((k - 1.0L) / k) * pm2;
    if (k < m) s += (pk * e) * (pk * e);
    pm2 = pm1; pm1 = pk;
  }
  *pn = pm1; *pnm1 = pm2;
  if (s2f) *s2f = s;
}

/* Gauss-Hermite nodes x and weights multiplied by e^{x^2} (the column the
 * reference reads: sub_quadra_herm.f90:160-178, herm{nq}.txt column 4).
 * Golub-Welsch Jacobi matrix, eigenvalues by bisection, Newton polish in
 * long double on p_n, weights = 1 / sum_k phi_k(x)^2. */
int orc_gauss_hermite(int n, double *x, double *wexp) {
  if (n < 1 || n > 200) return -1;
  double a[256], b[256];
  for (int i = 0; i < n; i++) a[i] = 0.0;
  for (int i = 0; i < n - 1; i++) b[i] = sqrt(0.5 * (i + 1));
  double R = sqrt(2.0 * n + 1.0) + 1.0;
  for (int k = 0; k < n; k++) {
    long double xx = tridiag_eig(n, a, b, k, -R, R);
    for (int it = 0; it < 8; it++) {
      long double pn, pnm1;
      /* p_n at xx needs n+1 polynomials */
      herm_orthonormal(n + 1, xx, &pn, &pnm1, 0, NULL);
      long double dpn = sqrtl(2.0L * n) * pnm1;
      long double dx = pn / dpn;
      xx -= dx;
      if (fabsl(dx) < 1e-21L * (1.0L + fabsl(xx))) break;
    }
    long double pn, pnm1, s2;
    herm_orthonormal(n, xx, &pn, &pnm1, n, &s2);
    x[k] = (double)xx;
    wexp[k] = (double)(1.0L / s2);
  }
  /* exact symmetry of the rule */
  for (int k = 0; k < n / 2; k++) {
    double xs = 0.5 * (x[n - 1 - k] - x[k]);
    x[k] = -xs; x[n - 1 - k] = xs;
    double ws = 0.5 * (wexp[k] + wexp[n - 1 - k]);
    wexp[k] = ws; wexp[n - 1 - k] = ws;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
  return 0;
}

/* Gauss-Legendre nodes on [-1,1] and weights (reference: gauleg128 used by
 * sub_quadra_legendre, SRC/sub_Basis/sub_quadra_legendre.f90:48). */
int orc_gauss_legendre(int n, double *x, double *w) {
  if (n < 1 || n > 200) return -1;
  double a[256], b[256];
  for (int i = 0; i < n; i++) a[i] = 0.0;
  for (int i = 0; i < n - 1; i++) { double k = i + 1; b[i] = k / sqrt(4 * k * k - 1); }
  for (int k = 0; k < n; k++) {
    long double xx = tridiag_eig(n, a, b, k, -1.0000001, 1.0000001);
    for (int it = 0; it < 8; it++) {
      long double p0 = 1, p1 = xx;
      if (n == 1) { p1 = xx; p0 = 1; }
      else {
        for (int j = 2; j <= n; j++) { long double p2 = ((2 * j - 1) * xx * p1 - (j - 1) * p0) / j; p0 = p1; p1 = p2; }
      }
      /* p1 = P_n, p0 = P_{n-1} */
      long double dp = n * (xx * p1 - p0) / (xx * xx - 1.0L);
      long double dx = p1 / dp;
      xx -= dx;
      if (fabsl(dx) < 1e-21L) break;
    }
    long double p0 = 1, p1 = xx;
    for (int j = 2; j <= n; j++) { long double p2 = ((2 * j - 1) * xx * p1 - (j - 1) * p0) / j; p0 = p1; p1 = p2; }
    if (n == 1) { p1 = xx; p0 = 1; }
    long double dp = n * (xx * p1 - p0) / (xx * xx - 1.0L);
    x[k] = (double)xx;
    w[k] = (double)(2.0L / ((1.0L - xx * xx) * dp * dp));
  }
  for (int k = 0; k < n / 2; k++) {
    double xs = 0.5 * (x[n - 1 - k] - x[k]);
    x[k] = -xs; x[n - 1 - k] = xs;
    double ws = 0.5 * (w[k] + w[n - 1 - k]);
    w[k] = ws; w[n - 1 - k] = ws;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* 1-D primitive bases (column-major nq x nb arrays)                        */
/* ------------------------------------------------------------------------ */

/* Harmonic-oscillator basis: normalised Hermite functions and their first
 * and second derivatives at the Gauss-Hermite nodes
 * (d0d1d2poly_Hermite_exp, Ext_Lib/FOR_EVRT_loc/SRC/sub_communf90/sub_math/
 * sub_polyortho.f90:1090-1153; sub_quadra_hermite, SRC/sub_Basis/
 * sub_quadra_herm.f90:48-250), then the affine change Q = Q0 + x/scaleQ
 * (sub_scale_basis, SRC/sub_Basis/sub_module_basis.f90:1499-1560). */
int orc_ho_basis(int nq, int nb, double Q0, double scaleQ, double *x,
                 double *w, double *d0, double *d1, double *d2) {
  if (nq < 1 || nb < 1 || nb > nq) return -1;
  double xr[256], wr[256];
  if (orc_gauss_hermite(nq, xr, wr)) return -1;
  long double sc = scaleQ;
  for (int iq = 0; iq < nq; iq++) {
    long double xx = xr[iq];
    long double e = expl(-0.5L * xx * xx);
    long double pm2 = 0, pm1 = 0, p = powl(acosl(-1.0L), -0.25L);
    for (int n = 0; n < nb; n++) {
      if (n == 1) { pm1 = p; p = sqrtl(2.0L) * xx * pm1; }
      else if (n >= 2) { pm2 = pm1; pm1 = p; p = sqrtl(2.0L / n) * xx * pm1 - sqrtl((n - 1.0L) / n) * pm2; }
      /* (p e)' = (sqrt(2n) p_{n-1} - x p) e ; (p e)'' = (x^2 - 2n - 1) p e */
      long double dp = (n > 0) ? sqrtl(2.0L * n) * pm1 : 0.0L;
      long double f0 = p * e;
      long double f1 = (dp - xx * p) * e;
      long double f2 = (xx * xx - 2.0L * n - 1.0L) * p * e;
      d0[iq + (size_t)nq * n] = (double)(f0 * sqrtl(sc));
      d1[iq + (size_t)nq * n] = (double)(f1 * sqrtl(sc) * sc);
      d2[iq + (size_t)nq * n] = (double)(f2 * sqrtl(sc) * sc * sc);
    }
    x[iq] = (double)(Q0 + xx / sc);
    w[iq] = (double)(wr[iq] / sc);
  }
  return 0;
}

/* Legendre basis in theta (reference 'Pl0' basis: normalised Legendre
 * polynomials sqrt((2l+1)/2) P_l(cos theta), quadrature in u=cos(theta),
 * weights w(u); the basis functions are tabulated on theta = acos(u) and
 * the derivatives are with respect to theta, so that d/dtheta acts on the
 * grid).  SRC/sub_Basis/sub_quadra_legendre.f90:48. */
int orc_legendre_basis(int nq, int nb, double *x, double *w, double *d0,
                       double *d1, double *d2) {
  if (nq < 1 || nb < 1 || nb > nq) return -1;
  double u[256], wu[256];
  if (orc_gauss_legendre(nq, u, wu)) return -1;
  for (int iq = 0; iq < nq; iq++) {
    /* ascending theta: u descending */
    int ju = nq - 1 - iq;
    long double uu = u[ju];
    long double th = acosl(uu), s = sinl(th), c = cosl(th);
    long double P[256], dP[256], ddP[256];
    P[0] = 1; dP[0] = 0; ddP[0] = 0;
    if (nb > 1) { P[1] = uu; dP[1] = 1; ddP[1] = 0; }
    for (int l = 2; l < nb; l++) {
      P[l] = ((2 * l - 1) * uu * P[l - 1] - (l - 1) * P[l - 2]) / l;
      dP[l] = l * P[l - 1] + uu * dP[l - 1];
      ddP[l] = (l + 1) * dP[l - 1] + uu * ddP[l - 1];
    }
    for (int l = 0; l < nb; l++) {
      long double nrm = sqrtl((2.0L * l + 1.0L) / 2.0L);
      /* f(theta) = P(cos theta): f' = -sin P'(u); f'' = sin^2 P''(u) - cos P'(u) */
      d0[iq + (size_t)nq * l] = (double)(nrm * P[l]);
      d1[iq + (size_t)nq * l] = (double)(-nrm * s * dP[l]);
      d2[iq + (size_t)nq * l] = (double)(nrm * (s * s * ddP[l] - c * dP[l]));
    }
    x[iq] = (double)th;
    w[iq] = wu[ju];
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Grid derivative matrices: dnRGG%dk = dnRGB%dk . pinv(d0b)                 */
/* (sub_dnGB_TO_dnGG, SRC/sub_Basis/sub_module_basis.f90:2048-2329).  The    */
/* pseudo-inverse is (d0b^T d0b)^-1 d0b^T, i.e. the Moore-Penrose inverse   */
/* the reference forms through d0b^T (d0b d0b^T)^+ (:2114-2189).             */
/* ------------------------------------------------------------------------ */
static int solve_ld(int n, long double *A, long double *B, int nrhs) {
  /* A (n x n col-major) X = B (n x nrhs), Gaussian elimination w/ pivoting */
  for (int k = 0; k < n; k++) {
    int p = k; long double mx = fabsl(A[k + n * k]);
    for (int i = k + 1; i < n; i++) if (fabsl(A[i + n * k]) > mx) { mx = fabsl(A[i + n * k]); p = i; }
    if (mx == 0) return -1;
    if (p != k) {
      for (int j = 0; j < n; j++) { long double t = A[k + n * j]; A[k + n * j] = A[p + n * j]; A[p + n * j] = t; }
      for (int j = 0; j < nrhs; j++) { long double t = B[k + n * j]; B[k + n * j] = B[p + n * j]; B[p + n * j] = t; }
    }
    for (int i = k + 1; i < n; i++) {
      long double f = A[i + n * k] / A[k + n * k];
      if (f == 0) continue;
      for (int j = k; j < n; j++) A[i + n * j] -= f * A[k + n * j];
      for (int j = 0; j < nrhs; j++) B[i + n * j] -= f * B[k + n * j];
    }
  }
  for (int j = 0; j < nrhs; j++)
    for (int i = n - 1; i >= 0; i--) {
      long double s = B[i + n * j];
      for (int k = i + 1; k < n; k++) s -= A[i + n * k] * B[k + n * j];
      B[i + n * j] = s / A[i + n * i];
    }
  return 0;
}

int orc_grid_deriv(int nq, int nb, const double *d0, const double *d1,
                   const double *d2, double *gg0, double *gg1, double *gg2) {
  long double *M = calloc((size_t)nb * nb, sizeof(long double));
  long double *P = calloc((size_t)nb * nq, sizeof(long double));
  if (!M || !P) { free(M); free(P); return -1; }
  for (int i = 0; i < nb; i++)
    for (int j = 0; j < nb; j++) {
      long double s = 0;
      for (int q = 0; q < nq; q++) s += (long double)d0[q + nq * i] * d0[q + nq * j];
      M[i + nb * j] = s;
    }
  for (int i = 0; i < nb; i++)
    for (int q = 0; q < nq; q++) P[i + nb * q] = d0[q + nq * i];
  int rc = solve_ld(nb, M, P, nq);
  if (rc == 0) {
    const double *src[3] = {d0, d1, d2};
    double *dst[3] = {gg0, gg1, gg2};
    for (int m = 0; m < 3; m++) {
      if (!dst[m]) continue;
      for (int q = 0; q < nq; q++)
        for (int r = 0; r < nq; r++) {
          long double s = 0;
          for (int i = 0; i < nb; i++) s += (long double)src[m][q + nq * i] * P[i + nb * r];
          dst[m][q + (size_t)nq * r] = (double)s;
        }
    }
  }
  free(M); free(P);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Smolyak SG4 structure                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
  int kind;            /* 0 = HO (Hermite), 1 = Legendre in theta */
  int nq_A, nq_B;      /* nq(l) = nq_A + nq_B * l   (sub_module_Basis_LTO_n.f90:313-316) */
  int nb_A, nb_B;      /* nb(l) = nb_A + nb_B * l   (defaults = nq rule, sub_read_data.f90:800-801) */
  double Q0, scaleQ;
} orc_mode;

typedef struct {
  int nq, nb;
  double *x, *w, *d0, *d1, *d2;   /* dnRGB (nq x nb) */
  double *BtW;                    /* dnRBGwrho (nb x nq) = d0^T w */
  double *gg1, *gg2;              /* dnRGG%d1, %d2 (nq x nq) */
} orc_b1;

typedef struct {
  int D, L, LB, Lmin;
  orc_mode mode[ORC_MAXD];
  orc_b1 b1[ORC_MAXD][ORC_MAXL + 1];
  int64_t nterms;
  int *lev;              /* D x nterms */
  double *weight;        /* nterms */
  int64_t *boff, *qoff;  /* nterms + 1 */
  int64_t nb_ba;
  int *packed;           /* D x nb_ba (0-based 1-D indices), lex order */
  int *map;              /* sum nb_t, packed row or -1 */
} orc_sg;

static int n_of(int A, int B, int l) { return A + B * l; }
/* level of 1-D basis index i (0-based): smallest l with n(l) > i
 * (Get_L_FROM_Basis_L_TO_n, sub_module_Basis_LTO_n.f90:512) */
static int lam_of(const orc_mode *m, int i) {
  int l = 0;
  while (n_of(m->nb_A, m->nb_B, l) <= i) l++;
  return l;
}

/* term set {l : Lmin <= |l| <= L} in the reference's order, first index
 * fastest (init_nDindex_type5m + ADD_ONE_TO_nDindex_type5m,
 * Ext_Lib/nDindex_loc/SRC/sub_nDindex/sub_module_nDindex.f90:1099,2737). */
static int next_term(int D, int L, int Lmin, int *l) {
  for (;;) {
    int i = 0;
    for (;;) {
      l[i]++;
      int s = 0; for (int k = 0; k < D; k++) s += l[k];
      if (s <= L) break;
      l[i] = 0; i++;
      if (i == D) return 0;
    }
    int s = 0; for (int k = 0; k < D; k++) s += l[k];
    if (s >= Lmin) return 1;
  }
}

static long long binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  long long r = 1;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return r;
}

static int cmp_lex(const int *a, const int *b, int D) {
  for (int k = 0; k < D; k++) { if (a[k] != b[k]) return a[k] < b[k] ? -1 : 1; }
  return 0;
}

static void free_sg(orc_sg *s) {
  if (!s) return;
  for (int k = 0; k < s->D; k++)
    for (int l = 0; l <= s->L; l++) {
      orc_b1 *b = &s->b1[k][l];
      free(b->x); free(b->w); free(b->d0); free(b->d1); free(b->d2);
      free(b->BtW); free(b->gg1); free(b->gg2);
    }
  free(s->lev); free(s->weight); free(s->boff); free(s->qoff);
  free(s->packed); free(s->map);
  free(s);
}

void orc_sg_destroy(orc_sg *s) { free_sg(s); }

/* Build the SG4 basis (RecSparseGrid_ForDP_type4,
 * SRC/sub_Basis/sub_quadra_SparseBasis.f90:922-1434). */
orc_sg *orc_sg_create(int D, int L, int LB, const int *kind, const int *nqA,
                      const int *nqB, const int *nbA, const int *nbB,
                      const double *Q0, const double *scaleQ) {
  if (D < 1 || D > ORC_MAXD || L < 0 || L > ORC_MAXL || LB < 0 || LB > L) return NULL;
  orc_sg *s = calloc(1, sizeof(orc_sg));
  s->D = D; s->L = L; s->LB = LB;
  s->Lmin = L - D + 1 > 0 ? L - D + 1 : 0;
  for (int k = 0; k < D; k++) {
    orc_mode *m = &s->mode[k];
    m->kind = kind[k]; m->nq_A = nqA[k]; m->nq_B = nqB[k];
    m->nb_A = nbA[k]; m->nb_B = nbB[k];
    m->Q0 = Q0[k]; m->scaleQ = scaleQ[k];
    for (int l = 0; l <= L; l++) {
      orc_b1 *b = &s->b1[k][l];
      int nq = n_of(m->nq_A, m->nq_B, l), nb = n_of(m->nb_A, m->nb_B, l);
      if (nb > nq || nb < 1) { free_sg(s); return NULL; }
      b->nq = nq; b->nb = nb;
      b->x = malloc(sizeof(double) * nq); b->w = malloc(sizeof(double) * nq);
      b->d0 = malloc(sizeof(double) * nq * nb); b->d1 = malloc(sizeof(double) * nq * nb);
      b->d2 = malloc(sizeof(double) * nq * nb); b->BtW = malloc(sizeof(double) * nq * nb);
      b->gg1 = malloc(sizeof(double) * nq * nq); b->gg2 = malloc(sizeof(double) * nq * nq);
      int rc = (m->kind == 0) ? orc_ho_basis(nq, nb, m->Q0, m->scaleQ, b->x, b->w, b->d0, b->d1, b->d2)
                              : orc_legendre_basis(nq, nb, b->x, b->w, b->d0, b->d1, b->d2);
      if (rc || orc_grid_deriv(nq, nb, b->d0, b->d1, b->d2, NULL, b->gg1, b->gg2)) { free_sg(s); return NULL; }
      for (int i = 0; i < nb; i++)
        for (int q = 0; q < nq; q++) b->BtW[i + (size_t)nb * q] = b->d0[q + (size_t)nq * i] * b->w[q];
    }
  }
  /* terms */
  int l[ORC_MAXD];
  int64_t cnt = 0;
  memset(l, 0, sizeof l); l[0] = -1;
  while (next_term(D, L, s->Lmin, l)) cnt++;
  s->nterms = cnt;
  s->lev = malloc(sizeof(int) * D * cnt);
  s->weight = malloc(sizeof(double) * cnt);
  s->boff = malloc(sizeof(int64_t) * (cnt + 1));
  s->qoff = malloc(sizeof(int64_t) * (cnt + 1));
  memset(l, 0, sizeof l); l[0] = -1;
  int64_t t = 0, sb = 0, sq = 0;
  while (next_term(D, L, s->Lmin, l)) {
    int tot = 0; int64_t nbt = 1, nqt = 1;
    for (int k = 0; k < D; k++) {
      s->lev[k + (size_t)D * t] = l[k]; tot += l[k];
      nbt *= s->b1[k][l[k]].nb; nqt *= s->b1[k][l[k]].nq;
    }
    /* W = (-1)^(L-|l|) C(D-1, L-|l|)  (calc_Weight_OF_SRep,
       SRC/sub_Basis/sub_module_param_SGType2.f90:795-806) */
    int dl = L - tot;
    s->weight[t] = (dl % 2 ? -1.0 : 1.0) * (double)binom(D - 1, dl);
    s->boff[t] = sb; s->qoff[t] = sq;
    sb += nbt; sq += nqt; t++;
  }
  s->boff[cnt] = sb; s->qoff[cnt] = sq;
  /* packed basis: sum_k lam(i_k) <= LB, lex order with the last index
     fastest (init_nDindex_type5p + ADD_ONE_TO_nDindex_type5p,
     sub_module_nDindex.f90:987,2670) */
  int maxi[ORC_MAXD];
  for (int k = 0; k < D; k++) maxi[k] = n_of(s->mode[k].nb_A, s->mode[k].nb_B, L);
  int64_t cap = 1024, np = 0;
  int *pk = malloc(sizeof(int) * D * cap);
  int idx[ORC_MAXD]; memset(idx, 0, sizeof idx);
  for (;;) {
    int sl = 0; for (int k = 0; k < D; k++) sl += lam_of(&s->mode[k], idx[k]);
    if (sl <= LB) {
      if (np == cap) { cap *= 2; pk = realloc(pk, sizeof(int) * D * cap); }
      memcpy(pk + (size_t)D * np, idx, sizeof(int) * D); np++;
    }
    /* advance, last index fastest; prune when the level sum exceeds LB */
    int k = D - 1;
    for (;;) {
      idx[k]++;
      int s2 = 0; for (int j = 0; j <= k; j++) s2 += lam_of(&s->mode[j], idx[j]);
      if (idx[k] < maxi[k] && s2 <= LB) break;
      idx[k] = 0; k--;
      if (k < 0) break;
    }
    if (k < 0) break;
  }
  s->nb_ba = np; s->packed = pk;
  /* map (Set_tables_FOR_SmolyakRepBasis_TO_tabPackedBasis,
     SRC/sub_Basis/sub_Basis_SG4/sub_module_basis_BtoG_GtoB_SG4.f90:630-954):
     in-term multi-index first index fastest; packed row by binary search;
     -1 when the level sum exceeds LB (map 0 in the reference, :917-923) */
  s->map = malloc(sizeof(int) * sb);
  for (t = 0; t < cnt; t++) {
    int nk[ORC_MAXD]; int64_t nbt = s->boff[t + 1] - s->boff[t];
    for (int k = 0; k < D; k++) nk[k] = s->b1[k][s->lev[k + (size_t)D * t]].nb;
    int ii[ORC_MAXD]; memset(ii, 0, sizeof ii);
    for (int64_t e = 0; e < nbt; e++) {
      int64_t lo = 0, hi = np - 1, found = -1;
      while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        int c = cmp_lex(pk + (size_t)D * mid, ii, D);
        if (c == 0) { found = mid; break; }
        if (c < 0) lo = mid + 1; else hi = mid - 1;
      }
      s->map[s->boff[t] + e] = (int)found;
      for (int k = 0; k < D; k++) { if (++ii[k] < nk[k]) break; ii[k] = 0; }
    }
  }
  return s;
}

int64_t orc_sg_nterms(const orc_sg *s) { return s->nterms; }
int64_t orc_sg_nb_ba(const orc_sg *s) { return s->nb_ba; }
int64_t orc_sg_sum_nb(const orc_sg *s) { return s->boff[s->nterms]; }
int64_t orc_sg_sum_nq(const orc_sg *s) { return s->qoff[s->nterms]; }
void orc_sg_tables(const orc_sg *s, int *lev, double *w, int64_t *boff,
                   int64_t *qoff, int *map, int *packed) {
  if (lev) memcpy(lev, s->lev, sizeof(int) * s->D * s->nterms);
  if (w) memcpy(w, s->weight, sizeof(double) * s->nterms);
  if (boff) memcpy(boff, s->boff, sizeof(int64_t) * (s->nterms + 1));
  if (qoff) memcpy(qoff, s->qoff, sizeof(int64_t) * (s->nterms + 1));
  if (map) memcpy(map, s->map, sizeof(int) * s->boff[s->nterms]);
  if (packed) memcpy(packed, s->packed, sizeof(int) * s->D * s->nb_ba);
}
/* 1-D basis of (mode k, level l): returns nq, nb; arrays may be NULL */
int orc_sg_basis1d(const orc_sg *s, int k, int l, int *nq, int *nb, double *x,
                   double *w, double *d0, double *d1, double *d2, double *gg1,
                   double *gg2) {
  if (k < 0 || k >= s->D || l < 0 || l > s->L) return -1;
  const orc_b1 *b = &s->b1[k][l];
  *nq = b->nq; *nb = b->nb;
  size_t qb = (size_t)b->nq * b->nb, qq = (size_t)b->nq * b->nq;
  if (x) memcpy(x, b->x, sizeof(double) * b->nq);
  if (w) memcpy(w, b->w, sizeof(double) * b->nq);
  if (d0) memcpy(d0, b->d0, sizeof(double) * qb);
  if (d1) memcpy(d1, b->d1, sizeof(double) * qb);
  if (d2) memcpy(d2, b->d2, sizeof(double) * qb);
  if (gg1) memcpy(gg1, b->gg1, sizeof(double) * qq);
  if (gg2) memcpy(gg2, b->gg2, sizeof(double) * qq);
  return 0;
}

/* grid coordinates of term t, point iq (first index fastest)
 * (Rec_Qact_SG4_with_Tab_iq, SRC/sub_Basis/sub_module_basis.f90:3979) */
void orc_sg_term_grid(const orc_sg *s, int64_t t, double *q /* D x nq_t */) {
  int D = s->D;
  const int *lv = s->lev + (size_t)D * t;
  int64_t nqt = s->qoff[t + 1] - s->qoff[t];
  int ii[ORC_MAXD]; memset(ii, 0, sizeof ii);
  for (int64_t e = 0; e < nqt; e++) {
    for (int k = 0; k < D; k++) q[k + (size_t)D * e] = s->b1[k][lv[k]].x[ii[k]];
    for (int k = 0; k < D; k++) { if (++ii[k] < s->b1[k][lv[k]].nq) break; ii[k] = 0; }
  }
}

/* ------------------------------------------------------------------------ */
/* Potentials (first-call grid evaluation, get_OpGrid_type1_OF_ONEDP_FOR_SG4 */
/* SRC/sub_Operator/sub_OpPsi_SG4.f90:2818-3082 -> calcN_op)                 */
/* ------------------------------------------------------------------------ */

/* Henon-Heiles: V = sum 1/2 q_i^2 + lambda sum (q_i^2 q_{i+1} - q_{i+1}^3/3)
 * (Working_tests/MPI_tests/6D_Davidson_S0/sub_system_HenonHeiles.f:14-43) */
double orc_henon_heiles(int D, const double *q, double lambda) {
  double v = 0;
  for (int i = 0; i < D; i++) v += 0.5 * q[i] * q[i];
  for (int i = 0; i < D - 1; i++) v += lambda * (q[i] * q[i] * q[i + 1] - q[i + 1] * q[i + 1] * q[i + 1] / 3.0);
  return v;
}

/* V grid of every term, Henon-Heiles */
void orc_sg_grid_hh(const orc_sg *s, double lambda, double *V /* sum nq */) {
  int D = s->D;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t t = 0; t < s->nterms; t++) {
    int64_t nqt = s->qoff[t + 1] - s->qoff[t];
    double *q = malloc(sizeof(double) * D * nqt);
    orc_sg_term_grid(s, t, q);
    for (int64_t e = 0; e < nqt; e++) V[s->qoff[t] + e] = orc_henon_heiles(D, q + (size_t)D * e, lambda);
    free(q);
  }
}

/* ------------------------------------------------------------------------ */
/* Per-term operator                                                        */
/* ------------------------------------------------------------------------ */

/* mode-k product on an array with dims n[0..D-1] (first fastest):
 * out(a, i, b) = sum_j M(i, j) in(a, j, b); M is (m_out x m_in) col-major.
 * (RB_TO_RG_basis / RG_TO_RB_basis, SRC/sub_Basis/sub_module_basis_set_alloc.f90
 *  :2148-2287, applied fibre by fibre as in BDP_TO_GDP_OF_SmolyakRep
 *  sub_module_basis_BtoG_GtoB_SG4.f90:2430-2469) */
static void mode_product(const double *in, double *out, const int *dims, int D,
                         int k, const double *M, int m_out) {
  int64_t A = 1, B = 1;
  for (int j = 0; j < k; j++) A *= dims[j];
  for (int j = k + 1; j < D; j++) B *= dims[j];
  int m_in = dims[k];
  for (int64_t b = 0; b < B; b++)
    for (int i = 0; i < m_out; i++)
      for (int64_t a = 0; a < A; a++) {
        double s = 0;
        for (int j = 0; j < m_in; j++) s += M[i + (size_t)m_out * j] * in[a + A * (j + (int64_t)m_in * b)];
        out[a + A * (i + (int64_t)m_out * b)] = s;
      }
}

/* Type-1 operator (sub_TabOpPsi_OF_ONEDP_FOR_SGtype4 case 1,
 * SRC/sub_Operator/sub_OpPsi_SG4.f90:1488-1544): on the grid,
 *   H psi = V psi + sum_k c_k d2_k psi
 * where d2_k is dnRGG%d2 of mode k (DerivOp_TO_RDP_OF_SmolaykRep,
 * sub_module_basis_BtoG_GtoB_SG4.f90:2670-2775).  c_k = -1/2 G_kk.
 *
 * Type-10 operator (case 10, :1546-1649):
 *   u = sqRhoOVERJac psi ; u_j = d1_j u ;
 *   w_i = Jac sum_j G(:,j,i) u_j ; H psi = -1/2 sum_i d1_i w_i / (Jac sqRhoOVERJac) + V psi
 */
typedef struct {
  int type;            /* 1 or 10 */
  const double *V;     /* sum nq */
  const double *ck;    /* type 1: D constants */
  const double *G;     /* type 10: D*D per grid point (point-major: G[(i + D j) + D*D*iq]) */
  const double *Jac;   /* type 10 */
  const double *sqRJ;  /* type 10 */
} orc_op;

static void term_apply(const orc_sg *s, int64_t t, const orc_op *op,
                       const double *psi_t, double *r_t, double *w1, double *w2,
                       double *w3, double *w4) {
  int D = s->D;
  const int *lv = s->lev + (size_t)D * t;
  int dims[ORC_MAXD];
  int64_t nqt = s->qoff[t + 1] - s->qoff[t];
  int64_t nbt = s->boff[t + 1] - s->boff[t];
  /* B -> G, mode by mode, k = 1..D (BDP_TO_GDP_OF_SmolyakRep :2430-2469) */
  for (int k = 0; k < D; k++) dims[k] = s->b1[k][lv[k]].nb;
  memcpy(w1, psi_t, sizeof(double) * nbt);
  for (int k = 0; k < D; k++) {
    const orc_b1 *b = &s->b1[k][lv[k]];
    mode_product(w1, w2, dims, D, k, b->d0, b->nq);
    dims[k] = b->nq;
    memcpy(w1, w2, sizeof(double) * nqt);  /* sizes only grow (nq >= nb) */
  }
  const double *V = op->V + s->qoff[t];
  if (op->type == 1) {
    for (int64_t e = 0; e < nqt; e++) w3[e] = V[e] * w1[e];
    for (int k = 0; k < D; k++) {
      if (op->ck[k] == 0.0) continue;
      const orc_b1 *b = &s->b1[k][lv[k]];
      mode_product(w1, w2, dims, D, k, b->gg2, b->nq);
      for (int64_t e = 0; e < nqt; e++) w3[e] += op->ck[k] * w2[e];
    }
  } else {
    const double *G = op->G + (size_t)D * D * s->qoff[t];
    const double *Jac = op->Jac + s->qoff[t];
    const double *sq = op->sqRJ + s->qoff[t];
    double *uj = w4; /* D x nqt */
    for (int64_t e = 0; e < nqt; e++) w2[e] = w1[e] * sq[e];
    for (int j = 0; j < D; j++) {
      const orc_b1 *b = &s->b1[j][lv[j]];
      mode_product(w2, uj + (size_t)j * nqt, dims, D, j, b->gg1, b->nq);
    }
    for (int64_t e = 0; e < nqt; e++) w3[e] = 0;
    for (int i = 0; i < D; i++) {
      for (int64_t e = 0; e < nqt; e++) {
        double acc = 0;
        for (int j = 0; j < D; j++) acc += G[(i * D + j) + (size_t)D * D * e] * uj[(size_t)j * nqt + e];
        w2[e] = acc * Jac[e];
      }
      const orc_b1 *b = &s->b1[i][lv[i]];
      mode_product(w2, r_t, dims, D, i, b->gg1, b->nq); /* r_t used as scratch (size >= nqt) */
      for (int64_t e = 0; e < nqt; e++) w3[e] += r_t[e];
    }
    for (int64_t e = 0; e < nqt; e++) w3[e] = -0.5 * w3[e] / (Jac[e] * sq[e]) + V[e] * w1[e];
  }
  /* G -> B, mode by mode (GDP_TO_BDP_OF_SmolyakRep :2514-2551) */
  for (int k = 0; k < D; k++) {
    const orc_b1 *b = &s->b1[k][lv[k]];
    mode_product(w3, w2, dims, D, k, b->BtW, b->nb);
    dims[k] = b->nb;
    int64_t n = 1; for (int j = 0; j < D; j++) n *= dims[j];
    memcpy(w3, w2, sizeof(double) * n);
  }
  memcpy(r_t, w3, sizeof(double) * nbt);
}

/* Block H.psi (sub_TabOpPsi_FOR_SGtype4, SRC/sub_Operator/sub_OpPsi_SG4.f90
 * :681-977): OpPsi zeroed, then for every term gather (tabPackedBasis_TO_
 * tabR_AT_iG, sub_module_basis_BtoG_GtoB_SG4.f90:1181-1251), apply, and
 * scatter-add W_t * r_t (tabR_AT_iG_TO_tabPackedBasis :1255-1294).
 * psi, hpsi: nb_ba x nvec, column-major.  Terms in [t0, t1) only (t1<0: all);
 * the whole output is still zeroed first. */
static int hpsi_impl(const orc_sg *s, int type, const double *V, const double *ck,
                     const double *G, const double *Jac, const double *sqRJ,
                     int nvec, const double *psi, double *hpsi, int64_t t0,
                     int64_t t1, int64_t tstride, int absmode) {
  orc_op op = {type, V, ck, G, Jac, sqRJ};
  if (t1 < 0 || t1 > s->nterms) t1 = s->nterms;
  if (tstride < 1) tstride = 1;
  int64_t nb_ba = s->nb_ba;
  memset(hpsi, 0, sizeof(double) * nb_ba * nvec);
  int64_t maxq = 1;
  for (int64_t t = 0; t < s->nterms; t++) {
    int64_t n = s->qoff[t + 1] - s->qoff[t];
    if (n > maxq) maxq = n;
  }
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  double *acc = calloc((size_t)nth * nb_ba * nvec, sizeof(double));
  if (!acc) return -1;
#pragma omp parallel
  {
    int th = 0;
#ifdef _OPENMP
    th = omp_get_thread_num();
#endif
    double *mine = acc + (size_t)th * nb_ba * nvec;
    double *w1 = malloc(sizeof(double) * maxq), *w2 = malloc(sizeof(double) * maxq);
    double *w3 = malloc(sizeof(double) * maxq), *w4 = malloc(sizeof(double) * maxq * s->D);
    double *pt = malloc(sizeof(double) * maxq), *rt = malloc(sizeof(double) * maxq);
#pragma omp for schedule(static)
    for (int64_t t = t0; t < t1; t += tstride) {
      int64_t nbt = s->boff[t + 1] - s->boff[t];
      const int *mp = s->map + s->boff[t];
      for (int v = 0; v < nvec; v++) {
        const double *p = psi + (size_t)nb_ba * v;
        for (int64_t e = 0; e < nbt; e++) pt[e] = mp[e] >= 0 ? p[mp[e]] : 0.0;
        term_apply(s, t, &op, pt, rt, w1, w2, w3, w4);
        double *h = mine + (size_t)nb_ba * v;
        for (int64_t e = 0; e < nbt; e++)
          if (mp[e] >= 0) h[mp[e]] += absmode ? fabs(s->weight[t] * rt[e]) : s->weight[t] * rt[e];
      }
    }
    free(w1); free(w2); free(w3); free(w4); free(pt); free(rt);
  }
  for (int th = 0; th < nth; th++) {
    const double *mine = acc + (size_t)th * nb_ba * nvec;
    for (int64_t e = 0; e < nb_ba * nvec; e++) hpsi[e] += mine[e];
  }
  free(acc);
  return 0;
}

int orc_hpsi_range(const orc_sg *s, int type, const double *V, const double *ck,
                   const double *G, const double *Jac, const double *sqRJ,
                   int nvec, const double *psi, double *hpsi, int64_t t0,
                   int64_t t1, int64_t tstride) {
  return hpsi_impl(s, type, V, ck, G, Jac, sqRJ, nvec, psi, hpsi, t0, t1, tstride, 0);
}

/* Condition of the weighted scatter-sum: sum_t |W_t r_t| per packed element
 * (same term results as orc_hpsi).  The rounding error of any summation
 * order of H psi is bounded by a small multiple of eps times this. */
int orc_hpsi_absbound(const orc_sg *s, int type, const double *V, const double *ck,
                      const double *G, const double *Jac, const double *sqRJ,
                      int nvec, const double *psi, double *bound) {
  return hpsi_impl(s, type, V, ck, G, Jac, sqRJ, nvec, psi, bound, 0, -1, 1, 1);
}

int orc_hpsi(const orc_sg *s, int type, const double *V, const double *ck,
             const double *G, const double *Jac, const double *sqRJ, int nvec,
             const double *psi, double *hpsi) {
  return orc_hpsi_range(s, type, V, ck, G, Jac, sqRJ, nvec, psi, hpsi, 0, -1, 1);
}

/* General type-1 operator (Type_HamilOp = 1 with grid-dependent
 * coefficients; sub_TabOpPsi_OF_ONEDP_FOR_SGtype4 case 1,
 * SRC/sub_Operator/sub_OpPsi_SG4.f90:1488-1544): on the grid of a term,
 *   H psi = sum_e F_e(q) D_e psi,
 * D_e = derivative term e given by two mode indices (derive_termQdyn(:,e),
 * 0-based here, -1 = none): (-1,-1) the identity, (i,-1) dnRGG%d1 along
 * mode i, (i,i) dnRGG%d2 along mode i, (i,j) i != j dnRGG%d1 along j then
 * along i (DerivOp_TO_RDP_OF_SmolaykRep, sub_module_basis_BtoG_GtoB_SG4.f90
 * :2670-2775, one 1-D matrix per mode from Get3_MatdnRGG).  F: nder x sum_nq,
 * F[e * sum_nq + point].  Block action as orc_hpsi (fixed-order sum). */
static void deriv_apply(const orc_sg *s, const int *lv, const int *dims, int i, int j,
                        const double *in, double *out, double *tmp) {
  const int D = s->D;
  if (i < 0 && j >= 0) { i = j; j = -1; }
  int64_t n = 1;
  for (int k = 0; k < D; k++) n *= dims[k];
  if (i < 0) { memcpy(out, in, sizeof(double) * n); return; }
  const orc_b1 *bi = &s->b1[i][lv[i]];
  if (j < 0) { mode_product(in, out, dims, D, i, bi->gg1, bi->nq); return; }
  if (i == j) { mode_product(in, out, dims, D, i, bi->gg2, bi->nq); return; }
  const orc_b1 *bj = &s->b1[j][lv[j]];
  mode_product(in, tmp, dims, D, j, bj->gg1, bj->nq);
  mode_product(tmp, out, dims, D, i, bi->gg1, bi->nq);
}

int orc_hpsi_t1g(const orc_sg *s, int nder, const int *der, const double *F, int nvec,
                 const double *psi, double *hpsi) {
  const int D = s->D;
  const int64_t nb_ba = s->nb_ba, sum_nq = s->qoff[s->nterms];
  int64_t maxq = 1;
  for (int64_t t = 0; t < s->nterms; t++) {
    int64_t n = s->qoff[t + 1] - s->qoff[t];
    if (n > maxq) maxq = n;
  }
  memset(hpsi, 0, sizeof(double) * nb_ba * nvec);
  double *w1 = malloc(sizeof(double) * maxq), *w2 = malloc(sizeof(double) * maxq);
  double *w3 = malloc(sizeof(double) * maxq), *w4 = malloc(sizeof(double) * maxq);
  double *w5 = malloc(sizeof(double) * maxq);
  int dims[ORC_MAXD];
  for (int64_t t = 0; t < s->nterms; t++) {
    const int *lv = s->lev + (size_t)D * t;
    const int64_t nqt = s->qoff[t + 1] - s->qoff[t], nbt = s->boff[t + 1] - s->boff[t];
    const int *mp = s->map + s->boff[t];
    for (int v = 0; v < nvec; v++) {
      const double *p = psi + (size_t)nb_ba * v;
      for (int k = 0; k < D; k++) dims[k] = s->b1[k][lv[k]].nb;
      for (int64_t e = 0; e < nbt; e++) w1[e] = mp[e] >= 0 ? p[mp[e]] : 0.0;
      for (int k = 0; k < D; k++) {           /* basis -> grid */
        const orc_b1 *b = &s->b1[k][lv[k]];
        mode_product(w1, w2, dims, D, k, b->d0, b->nq);
        dims[k] = b->nq;
        memcpy(w1, w2, sizeof(double) * nqt);
      }
      for (int64_t e = 0; e < nqt; e++) w3[e] = 0.0;
      for (int e = 0; e < nder; e++) {        /* sum_e F_e D_e psi on the grid */
        const double *Fe = F + (size_t)e * sum_nq + s->qoff[t];
        deriv_apply(s, lv, dims, der[2 * e], der[2 * e + 1], w1, w4, w5);
        for (int64_t q = 0; q < nqt; q++) w3[q] += Fe[q] * w4[q];
      }
      for (int k = 0; k < D; k++) {           /* grid -> basis */
        const orc_b1 *b = &s->b1[k][lv[k]];
        mode_product(w3, w2, dims, D, k, b->BtW, b->nb);
        dims[k] = b->nb;
        int64_t n = 1;
        for (int j = 0; j < D; j++) n *= dims[j];
        memcpy(w3, w2, sizeof(double) * n);
      }
      double *h = hpsi + (size_t)nb_ba * v;
      for (int64_t e = 0; e < nbt; e++)
        if (mp[e] >= 0) h[mp[e]] += s->weight[t] * w3[e];
    }
  }
  free(w1); free(w2); free(w3); free(w4); free(w5);
  return 0;
}

/* The type-1 block action in extended precision: the same operator (the
 * double 1-D matrices, grids and weights of orc_sg) with every mode product,
 * the grid operator and the scatter-sum carried in long double (64-bit
 * mantissa).  Its rounding, ~5e-20 times the Smolyak amplification, is far
 * below FP64's, so a comparison with it measures the other side's error
 * alone.  Test infrastructure like the rest of this file. */
static void mode_product_ld(const long double *in, long double *out, const int *dims, int D,
                            int k, const double *M, int m_out) {
  int64_t A = 1, B = 1;
  for (int j = 0; j < k; j++) A *= dims[j];
  for (int j = k + 1; j < D; j++) B *= dims[j];
  int m_in = dims[k];
  for (int64_t b = 0; b < B; b++)
    for (int i = 0; i < m_out; i++)
      for (int64_t a = 0; a < A; a++) {
        long double sum = 0;
        for (int j = 0; j < m_in; j++)
          sum += (long double)M[i + (size_t)m_out * j] * in[a + A * (j + (int64_t)m_in * b)];
        out[a + A * (i + (int64_t)m_out * b)] = sum;
      }
}

int orc_hpsi_ld(const orc_sg *s, const double *V, const double *ck, int nvec, const double *psi,
                double *hpsi) {
  const int D = s->D;
  const int64_t nb_ba = s->nb_ba;
  int64_t maxq = 1;
  for (int64_t t = 0; t < s->nterms; t++) {
    int64_t n = s->qoff[t + 1] - s->qoff[t];
    if (n > maxq) maxq = n;
  }
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  long double *acc = calloc((size_t)nth * nb_ba * nvec, sizeof(long double));
  if (!acc) return -1;
#pragma omp parallel
  {
    int th = 0;
#ifdef _OPENMP
    th = omp_get_thread_num();
#endif
    long double *mine = acc + (size_t)th * nb_ba * nvec;
    long double *w1 = malloc(sizeof(long double) * maxq), *w2 = malloc(sizeof(long double) * maxq);
    long double *w3 = malloc(sizeof(long double) * maxq);
    int dims[ORC_MAXD];
#pragma omp for schedule(dynamic, 16)
    for (int64_t t = 0; t < s->nterms; t++) {
      const int *lv = s->lev + (size_t)D * t;
      const int64_t nqt = s->qoff[t + 1] - s->qoff[t], nbt = s->boff[t + 1] - s->boff[t];
      const int *mp = s->map + s->boff[t];
      const double *Vt = V + s->qoff[t];
      for (int v = 0; v < nvec; v++) {
        const double *p = psi + (size_t)nb_ba * v;
        for (int k = 0; k < D; k++) dims[k] = s->b1[k][lv[k]].nb;
        for (int64_t e = 0; e < nbt; e++) w1[e] = mp[e] >= 0 ? (long double)p[mp[e]] : 0.0L;
        for (int k = 0; k < D; k++) {
          const orc_b1 *b = &s->b1[k][lv[k]];
          mode_product_ld(w1, w2, dims, D, k, b->d0, b->nq);
          dims[k] = b->nq;
          memcpy(w1, w2, sizeof(long double) * nqt);
        }
        for (int64_t e = 0; e < nqt; e++) w3[e] = (long double)Vt[e] * w1[e];
        for (int k = 0; k < D; k++) {
          if (ck[k] == 0.0) continue;
          const orc_b1 *b = &s->b1[k][lv[k]];
          mode_product_ld(w1, w2, dims, D, k, b->gg2, b->nq);
          for (int64_t e = 0; e < nqt; e++) w3[e] += (long double)ck[k] * w2[e];
        }
        for (int k = 0; k < D; k++) {
          const orc_b1 *b = &s->b1[k][lv[k]];
          mode_product_ld(w3, w2, dims, D, k, b->BtW, b->nb);
          dims[k] = b->nb;
          int64_t n = 1;
          for (int j = 0; j < D; j++) n *= dims[j];
          memcpy(w3, w2, sizeof(long double) * n);
        }
        long double *h = mine + (size_t)nb_ba * v;
        for (int64_t e = 0; e < nbt; e++)
          if (mp[e] >= 0) h[mp[e]] += (long double)s->weight[t] * w3[e];
      }
    }
    free(w1); free(w2); free(w3);
  }
  for (int64_t e = 0; e < nb_ba * nvec; e++) {
    long double sum = 0;
    for (int th = 0; th < nth; th++) sum += acc[(size_t)th * nb_ba * nvec + e];
    hpsi[e] = (double)sum;
  }
  free(acc);
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
