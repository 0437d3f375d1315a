/*
 * sg4hpsi.h -- C ABI of the MI355X (gfx950) Smolyak SG4 Hamiltonian action.
 *
 * Drop-in boundary for ElVibRot's SG type 4 operator action.  The Fortran
 * side keeps its OpPsi / Davidson / propagation drivers and binds these entry
 * points through iso_c_binding (module mod_sg4_gpu in
 * elvibrot_amd/fortran/mod_sg4_gpu.f90).  All arguments are plain pointers
 * and sizes; no C++ or torch types cross this boundary.
 *
 * Conventions
 *   - integers are int32 unless noted; sizes that can exceed 2^31 are int64;
 *   - arrays are column-major, first index fastest (Fortran order);
 *   - packed vectors are nb_ba x nvec blocks (Psi(v)%RvecB(:) of a block of
 *     param_psi, SRC/sub_WP/sub_module_psi_set_alloc.f90:56-169);
 *   - every call returns SG4_OK (0) or a negative status; sg4_last_error()
 *     gives the message.  The Fortran wrapper maps a nonzero status to STOP,
 *     as the reference does (sub_OpPsi_SG4.f90:741-753).
 *   - one host thread per plan; kernels run on the plan's HIP stream.
 */
#ifndef SG4HPSI_H
#define SG4HPSI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SG4_OK 0
#define SG4_ERR_ARG (-1)         /* bad argument or size */
#define SG4_ERR_HIP (-2)         /* HIP runtime error */
#define SG4_ERR_OOM (-3)         /* device allocation failed */
#define SG4_ERR_UNSUPPORTED (-4) /* valid input outside what the kernels handle */
#define SG4_ERR_STATE (-5)       /* call out of order (e.g. no operator set) */

#define SG4_BASIS_HO 0       /* 1-D harmonic oscillator ("Hm"), Gauss-Hermite grid */
#define SG4_BASIS_LEGENDRE 1 /* 1-D Legendre in theta ("Pl0"), Gauss-Legendre grid */

/* precision modes of sg4_apply */
#define SG4_PREC_FP64 0      /* FP64 throughout (parity mode, 1e-12 relative) */
#define SG4_PREC_F32 1       /* mode contractions in FP32, FP64 gather/scatter */
#define SG4_PREC_F32X2 2     /* split-precision contractions on the matrix cores (split
                                 type-1 algebra only, modes of <= 32 points):
                                 FP32 operands as three bf16 limbs
                                 (hi + mid + lo), six limb products per product on
                                 v_mfma_f32_16x16x{16,32}_bf16 with FP32 accumulation:
                                 FP32-accurate contractions; FP64 kinetic part, term sum
                                 and result */

typedef struct sg4_setup sg4_setup; /* host tables built by this library */
typedef struct sg4_plan sg4_plan;   /* device state of one SG4 operator */

/* ---- 1-D basis description (one per mode), as in &basis_nD -------------- */
typedef struct {
  int32_t kind;          /* SG4_BASIS_HO or SG4_BASIS_LEGENDRE */
  int32_t nq_A, nq_B;    /* nq(l) = nq_A + nq_B*l   (L_TO_nq_A/B) */
  int32_t nb_A, nb_B;    /* nb(l) = nb_A + nb_B*l   (defaults: = nq rule) */
  double Q0, scaleQ;     /* Q = Q0 + x/scaleQ (HO only) */
} sg4_mode_desc;

/* ---- tables consumed by the device plan --------------------------------
 * Field-by-field counterpart of what BasisnD holds after
 * RecSparseGrid_ForDP_type4 (SRC/sub_Basis/sub_quadra_SparseBasis.f90:922):
 *   levels  = para_SGType2%nDind_SmolyakRep%Tab_nDval   (D x nterms)
 *   weights = WeightSG                                   (nterms)
 *   map     = para_SGType2%tab_iB_OF_SRep_TO_iB - 1      (sum_t nb_t; -1 = absent)
 *   nq, nb  = sizes of tab_basisPrimSG(l,k)              ((Lmax+1) x D)
 *   mats    = for k = 1..D, for l = 0..Lmax, concatenated:
 *               dnRGB%d0 (nq x nb), dnRBGwrho%d0 (nb x nq),
 *               dnRGG%d1 (nq x nq), dnRGG%d2 (nq x nq)
 * Term t occupies map[boff(t) : boff(t)+nb_t) with boff the running sum of
 * nb_t = prod_k nb(levels(k,t),k); its grid occupies [qoff(t), qoff(t)+nq_t)
 * of every grid array, with the in-term index first-index-fastest.        */
typedef struct {
  int32_t D;
  int32_t Lmax;
  int64_t nterms;
  int64_t nb_ba;
  const int32_t *levels;
  const double *weights;
  const int32_t *map;
  const int32_t *nq;
  const int32_t *nb;
  const double *mats;
  const double *nodes;    /* optional: for k, for l: the nq grid nodes (basis%x) */
  const double *qweights; /* optional: for k, for l: the nq quadrature weights (basis%wrho) */
} sg4_tables;

/* ---- operator on the grids (get_OpGrid_type{1,10}_OF_ONEDP_FOR_SG4,
 * SRC/sub_Operator/sub_OpPsi_SG4.f90:2536-3082) ------------------------------
 * type_op = 1  : H = V(q) + sum_k ck[k] d^2/dq_k^2  (Type_HamilOp=1, constant
 *                diagonal G: ck = -G_kk/2)
 * type_op = 10 : H = -1/2 (1/(Jac sqRJ)) sum_ij d_i Jac G_ij d_j sqRJ + V
 * Grid arrays have sum_t nq_t points; G is D x D per point (G[i + D*j + D*D*p]). */
typedef struct {
  int32_t type_op;
  const double *V;       /* required */
  const double *ck;      /* type 1: D coefficients */
  const double *G;       /* type 10 */
  const double *Jac;     /* type 10 */
  const double *sqRJ;    /* type 10: sqRhoOVERJac */
} sg4_operator;

const char *sg4_last_error(void);

/* ---- set-up (replaces RecSparseGrid_ForDP_type4 + sub_dnGB_TO_dnGG when the
 * host does not build the tables itself; SRC/sub_Basis/
 * sub_quadra_SparseBasis.f90:922-1434, SRC/sub_Basis/sub_module_basis.f90:2048) */
int sg4_setup_create(int32_t D, int32_t L, int32_t LB, const sg4_mode_desc *modes,
                     sg4_setup **out);
/* nterms, nb_ba, sum_t nb_t, sum_t nq_t */
int sg4_setup_sizes(const sg4_setup *s, int64_t *nterms, int64_t *nb_ba, int64_t *sum_nb,
                    int64_t *sum_nq);
/* pointers into setup-owned memory, valid until sg4_setup_destroy */
int sg4_setup_tables(const sg4_setup *s, sg4_tables *out);
/* grid coordinates of term t: q[k + D*p], p < nq_t (Rec_Qact_SG4_with_Tab_iq,
 * SRC/sub_Basis/sub_module_basis.f90:3979) */
int sg4_setup_term_grid(const sg4_setup *s, int64_t t, double *q);
/* 1-D rule and basis of (mode k, level l): x, w (nq), d0, d1, d2 (nq x nb) */
int sg4_setup_basis1d(const sg4_setup *s, int32_t k, int32_t l, int32_t *nq, int32_t *nb,
                      double *x, double *w, double *d0, double *d1, double *d2);
/* 0-based mode indices of every packed row, idx[r*D + k] (r < nb_ba): the
 * nD index of packed basis function r (Rec_ndim_index, SG type 4 branch,
 * SRC/sub_Basis/sub_module_basis.f90:3055-3065, 0-based here) */
int sg4_setup_packed(const sg4_setup *s, int32_t *idx);
int sg4_setup_destroy(sg4_setup *s);

/* ---- device plan ----------------------------------------------------------
 * Owns device copies of the tables, the 1-D matrices and the grids, plus
 * pre-sized work buffers for blocks of up to max_nvec vectors: no device
 * allocation happens inside sg4_apply.                                       */
int sg4_plan_create(const sg4_tables *tables, int32_t max_nvec, sg4_plan **out);
/* A batch of ninst independent SG4 operators (different D, L, grids) in one
 * plan: the operator is block-diagonal over the concatenated packed rows
 * (instance i owns rows [row0_i, row0_i + nb_ba_i) in instance order), and
 * one apply runs every term of every instance in a single launch.  Operator
 * grids and kinetic coefficients are the instances' arrays concatenated in
 * the same order.  No counterpart in the reference (one operator per run). */
int sg4_plan_create_batch(int32_t ninst, const sg4_tables *tables, int32_t max_nvec,
                          sg4_plan **out);
/* hip_stream: a hipStream_t (NULL = default stream) */
int sg4_plan_set_stream(sg4_plan *p, void *hip_stream);
/* copy the operator grids to the device (host pointers) */
int sg4_plan_set_operator(sg4_plan *p, const sg4_operator *op);
/* General type-1 operator (Type_HamilOp = 1 with grid-dependent
 * coefficients: get_OpGrid_type1_OF_ONEDP_FOR_SG4 and case 1 of
 * sub_TabOpPsi_OF_ONEDP_FOR_SGtype4, SRC/sub_Operator/sub_OpPsi_SG4.f90
 * :1488-1544, :2818-3082): on the grid of every term
 *   H psi = sum_e F_e(q) D_e psi,   e = 0 .. nder-1,
 * with D_e given by der[2e], der[2e+1] (0-based modes, -1 = none; the
 * reference's derive_termQdyn(:,e) minus one): (-1,-1) the identity (the
 * potential term), (i,-1) dnRGG%d1 along mode i, (i,i) dnRGG%d2, (i,j) i != j
 * d1 along j and along i (DerivOp_TO_RDP_OF_SmolaykRep).  F: host array of
 * nder x sum_t nq_t grid values, F[e * sum_nq + point] (the reference's
 * GridOp(:,1,1,e) for one electronic state).  Single-instance plans, FP64
 * applies; sg4_plan_set_field and sg4_grad_V do not apply to it. */
int sg4_plan_set_operator_type1(sg4_plan *p, int32_t nder, const int32_t *der, const double *F);
/* evaluate the Henon-Heiles grid V on the device (first-call grid evaluation
 * of get_OpGrid_type1_OF_ONEDP_FOR_SG4, sub_OpPsi_SG4.f90:2978-3058, with
 * calcN_op of Working_tests/MPI_tests/6D_Davidson_S0/sub_system_HenonHeiles.f:14-43),
 * kinetic -1/2 sum d^2 (type 1).  Needs tables->nodes at plan creation. */
int sg4_plan_set_henon_heiles(sg4_plan *p, double lambda);
/* Dipole mu = sum_k q_k (the coupling of the field term, sg4_plan_set_field)
 * evaluated on the device at every grid point of every term of every
 * instance, in the layout of the operator grids (sum_t nq_t values, DEVICE
 * pointer).  The first-call grid evaluation of an operator that is the sum of
 * the coordinates; needs tables->nodes at plan creation. */
int sg4_plan_grid_dipole(sg4_plan *p, double *mu);
/* Type-10 operator of a triatomic in Jacobi coordinates (modes R, r, theta),
 * J = 0, evaluated on the device on every term grid (the first-call grid
 * evaluation of get_OpGrid_type10_OF_ONEDP_FOR_SG4,
 * SRC/sub_Operator/sub_OpPsi_SG4.f90:2536-2814, for this model):
 * prm = {mu_R, mu_r, D, a, Re, re, c}; V = D(1-e^{-a(R-Re)})^2
 * + D(1-e^{-a(r-re)})^2 + c cos^2(theta), G = diag(1/mu_R, 1/mu_r,
 * 1/(mu_R R^2) + 1/(mu_r r^2)), Jac = sin(theta), sqRhoOVERJac = 1. */
int sg4_plan_set_jacobi_triatomic(sg4_plan *p, const double *prm);

/* Block H.psi: sub_TabOpPsi_FOR_SGtype4(Psi, OpPsi, para_Op)
 * (SRC/sub_Operator/sub_OpPsi_SG4.f90:681-702).  psi and hpsi are DEVICE
 * pointers to nb_ba x nvec column-major blocks; hpsi is overwritten.  The two
 * blocks must not overlap (distinct objects, as Psi and OpPsi are in the
 * reference): overlapping blocks are rejected with SG4_ERR_ARG.  The same holds
 * for sg4_apply_rows and sg4_apply_graph.
 * Every DEVICE pointer of this interface (here and in sg4_apply_noalloc,
 * sg4_apply_rows, sg4_apply_graph, sg4_plan_set_field, sg4_grad_V and the
 * sg4_cheby_* calls) must be ordinary coarse-grained device memory of the GPU
 * the plan was created on (hipMalloc, hipMallocAsync, a torch CUDA tensor): the
 * kernels sum into their outputs with hardware floating-point atomics, which
 * are exact on that memory only.  Host, pinned, managed, fine-grained or
 * uncached memory, and memory of another GPU, is rejected with SG4_ERR_ARG
 * before anything is enqueued.
 * The block is staged in the rows layout padded with zero vectors to the next
 * power of two (nvec <= 16) or multiple of 16 within max_nvec: the kernels
 * carry 2^g vectors per work item (as many as the LDS budget allows), and
 * one wider item per term costs less than several narrow ones. */
int sg4_apply(sg4_plan *p, int32_t nvec, const double *psi, double *hpsi, int32_t prec);
/* Allocation-free variant of sg4_apply, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  It calls no hipMalloc,
 * hipFree or allocating library routine: the two rows-layout staging blocks
 * live in a file-scope device array of 256 MiB per GPU (held until the
 * process ends), so nb_ba x W <= 16777216, W being nvec padded as in
 * sg4_apply (larger blocks: SG4_ERR_ARG),
 * and it only uses work items, FP32 copies and matrix-core tables that an
 * earlier sg4_apply with the same nvec, precision and launch configuration
 * has prepared (otherwise SG4_ERR_STATE).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (several streams or threads)
 * race and must be serialised by the caller.  Nothing in the array needs to
 * start at zero (each call overwrites what it reads).  Kernel timing
 * (sg4_plan_timing) is not recorded by this entry point. */
int sg4_apply_noalloc(sg4_plan *p, int32_t nvec, const double *psi, double *hpsi, int32_t prec);
/* Same with HOST pointers (copies in and out over PCIe; synchronous). */
int sg4_apply_host(sg4_plan *p, int32_t nvec, const double *psi, double *hpsi, int32_t prec);
/* Same in the interleaved device layout used by the device-side drivers:
 * psi[row*nvec + v], row < nb_ba. */
int sg4_apply_rows(sg4_plan *p, int32_t nvec, const double *psi, double *hpsi, int32_t prec);

/* Field term of a piecewise-constant control: V = V0 + eps * mu on every grid
 * point, V0 being the grid held before the first call (mu: DEVICE pointer to
 * sum_t nq_t values, e.g. the dipole sum_i q_i; mu = NULL restores V0).  Used
 * per field segment by the propagators: H(t) = H0 + eps(t) mu is constant on a
 * segment.  The reference propagates fields with Taylor/RK only
 * (march_noD_field, SRC/sub_propagation/sub_module_propa_march.f90:1641). */
int sg4_plan_set_field(sg4_plan *p, const double *mu, double eps);

/* ---- Chebyshev propagation under a piecewise-constant field -------------
 * march_cheby (SRC/sub_propagation/sub_module_propa_march.f90:4145-4436,
 * recurrence :4297-4309) on the device: Hs = (H - E0)/Esc, w_0 = psi,
 * w_1 = -i Hs w_0, w_{k+1} = w_{k-1} - 2i Hs w_k, psi(t+dt) = exp(-i E0 dt)
 * sum_k c_k w_k with ncoef orders (the caller passes c_0 = J_0(Esc dt),
 * c_k = 2 J_k(Esc dt), as cof does, :5739-5760).  Segment s of nseg runs
 * steps_per_seg steps with H = H0 + eps[s] mu (sg4_plan_set_field); H0 is
 * restored at the end.  States: ns complex states as a DEVICE rows-layout
 * block of nb_ba x 2 ns values (row r: the real parts of the ns states, then
 * their imaginary parts).  The recurrence runs in element-wise kernels on the
 * plan's stream, one H.psi launch per order; no host synchronisation per
 * order.  The reference propagates fields with Taylor/RK only
 * (march_noD_field, :1641): Chebyshev per constant-field segment is new. */
typedef struct {
  double dt;              /* time step */
  double E0, Esc;         /* centre and half-width of the spectral range */
  int32_t ncoef;          /* Chebyshev orders per step */
  const double *coef;     /* host: the ncoef series coefficients */
  int32_t nseg;           /* field segments */
  int32_t steps_per_seg;  /* steps per segment */
  const double *eps;      /* host: nseg field amplitudes */
  int32_t ckpt_every;     /* gradient: segments between stored states */
  int32_t prec;           /* precision mode of the H.psi applications */
} sg4_cheby_params;
/* psi <- psi(T) (needs max_nvec >= 2 ns); mu: DEVICE grid of the dipole. */
int sg4_cheby_propagate(sg4_plan *H, const double *mu, int32_t ns, double *psi,
                        const sg4_cheby_params *prm);
/* J = sum_i |a_i|^2, a_i = sum over the rows [blk_rows[2i], blk_rows[2i+1])
 * and the ns states of conj(target) psi(T) (the instances of a batch), and
 * grad[s] = dJ/deps[s] for every segment (host array of nseg), exact for the
 * discrete propagator: the adjoint state is propagated backwards with
 * U^dagger and contracted with the forward tangent of each segment
 * (dHs/deps = mu/Esc, applied by Mu, a plan of the same rows holding the
 * potential mu and no kinetic part).  The states at the start of every
 * ckpt_every-th segment are stored and the backward sweep recomputes the
 * others, one checkpoint interval at a time.  psi0, target: DEVICE
 * rows-layout blocks as above; H needs max_nvec >= 4 ns, Mu >= 2 ns.  Mu runs
 * on H's stream during the call.  No counterpart in the reference (its only
 * derivative machinery is the field update of sub_control.f90:907-1011). */
int sg4_cheby_field_gradient(sg4_plan *H, sg4_plan *Mu, const double *mu, int32_t ns,
                             const double *psi0, const double *target, int32_t nblk,
                             const int64_t *blk_rows, const sg4_cheby_params *prm, double *J,
                             double *grad);

/* ---- Taylor propagation under a piecewise-constant field ----------------
 * march_noD_field (SRC/sub_propagation/sub_module_propa_march.f90:1641-1842)
 * with the field constant on each segment, so that the field's time
 * derivatives in its series vanish: H = H0 + eps[s] mu on segment s
 * (sg4_plan_set_field; the reference's field E enters as eps = -E, :1736), and
 * per step
 *   w_0 = psi,  w_{k+1} = (-i dt/(k+1)) (H - E0) w_k,  psi <- psi + w_{k+1}
 * (:1756-1787), cut after the first order whose squared norm is below
 * poly_tol for every state (:1791; the reference cuts each state on its own),
 * then psi <- exp(-i E0 dt) psi (:1812).  Errors (SG4_ERR_STATE), as the
 * reference's STOPs: a term's squared norm above 1e15 (time step too large,
 * :1792), no cut within npoly orders (npoly too small, march_noD :1959), a
 * state's squared norm above max_norm2 after a step when max_norm2 > 0
 * (:1817).  npoly from the spectral range is the caller's choice (cof_nOD,
 * :5771-5827).  States as for sg4_cheby_propagate (DEVICE rows layout,
 * nb_ba x 2 ns values, max_nvec >= 2 ns); orders: host array of
 * nseg x steps_per_seg entries receiving the number of terms each step used
 * (or NULL).  One H.psi launch and one host synchronisation per order (the new
 * term's norm decides the cut).  H0 is restored at the end, also on errors. */
typedef struct {
  double dt;              /* time step */
  double E0;              /* energy shift of the series */
  int32_t npoly;          /* maximal number of terms per step */
  double poly_tol;        /* cut: squared norm of a term (reference default 1e-20) */
  int32_t nseg;           /* field segments */
  int32_t steps_per_seg;  /* steps per segment */
  const double *eps;      /* host: nseg field amplitudes */
  double max_norm2;       /* > 0: bound on a state's squared norm (reference 1.3) */
  int32_t prec;           /* precision mode of the H.psi applications */
} sg4_taylor_params;
int sg4_taylor_propagate(sg4_plan *H, const double *mu, int32_t ns, double *psi,
                         const sg4_taylor_params *prm, int32_t *orders);

/* ---- dense products of the device block Davidson ------------------------
 * The projections, Gram matrices and rotations of sub_module_Davidson.f90
 * (Rayleigh-Ritz :59-763, double Gram-Schmidt :1529-1547), which the
 * reference computes with dot_product / matmul over the packed vectors, for
 * blocks of a few to ~100 rows of length n = nb_ba.  Rows are contiguous
 * (row i of A at A + i lda), all pointers DEVICE memory as for sg4_apply;
 * the kernels run on the plan's stream (the plan only provides stream and
 * device).  No counterpart in the reference as entry points.
 * sg4_block_gram: C (m x q, row-major) = A B^T, m, q <= 124 with
 *   ceil(m/4) ceil(q/4) <= 512; work: nwork doubles of scratch (at least
 *   m q; sg4_block_gram_work gives the size that lets every workgroup keep
 *   its partial, 2048 m q); partial sums are added in a fixed order, so the
 *   result is reproducible bit for bit.
 * sg4_block_combine: Y (k x n) = beta Y + U X, U (k x d, row-major, DEVICE),
 *   X d rows, k d <= 8192; Y must not overlap X; with beta = 0, Y is not
 *   read. */
int sg4_block_gram_work(int32_t m, int32_t q, int64_t *nwork);
int sg4_block_gram(sg4_plan *p, int32_t m, int32_t q, int64_t n, const double *A, int64_t lda,
                   const double *B, int64_t ldb, double *C, double *work, int64_t nwork);
int sg4_block_combine(sg4_plan *p, int32_t k, int32_t d, int64_t n, const double *U,
                      const double *X, int64_t ldx, double beta, double *Y, int64_t ldy);

/* Derivative of <phi_v|H|psi_v> with respect to every grid value of the
 * potential: grad[(p)*nvec + v] = W_t w(p) (B_t phi_v)(p) (B_t psi_v)(p) for the
 * grid point p of term t (w(p): product of the 1-D quadrature weights).  With
 * phi = psi = a normalised eigenvector this is the Hellmann-Feynman gradient
 * dE/dV(p); contracted with dV/dtheta it gives dE/dtheta.  No counterpart in
 * the reference (its only derivative machinery is the field update of
 * sub_control.f90:907-1011).  Rows layout, device pointers, grad has
 * sum_t nq_t x nvec entries; needs tables->qweights at plan creation. */
int sg4_grad_V(sg4_plan *p, int32_t nvec, const double *phi, const double *psi, double *grad);

/* Autotuner: times the H.psi kernel for blocks of nvec vectors over its
 * launch parameters (256/512 threads, 768 for the FP32 split path; LDS budget per workgroup, one workgroup
 * per work item or a persistent grid, XCD-grouped item order) and keeps the
 * fastest for later applies; best_ms receives its time.  Allocates two
 * temporary blocks.  The reference's only tuner scans OpenMP thread counts
 * (Tune_SG4threads_HPsi, SRC/sub_main/vib.f90:2013-2123). */
int sg4_plan_autotune(sg4_plan *p, int32_t nvec, int32_t prec, float *best_ms);
/* the launch parameters in use for (nvec, prec): threads, LDS budget, grid, xcd.
 * A block width with no parameters of its own (tuned or set) uses those of the
 * nearest configured width of the same precision (the narrowest wider one
 * first), else the defaults. */
int sg4_plan_config(const sg4_plan *p, int32_t nvec, int32_t prec, int32_t *cfg4);
/* set the launch parameters for (nvec, prec) (e.g. a configuration saved from
 * an earlier autotune): threads 256 or 512 (768 for the FP32 mode on the split
 * path), LDS budget in bytes (<= 160 KiB),
 * grid (0 = one workgroup per item, else the persistent grid size), xcd 0..8
 * (item order: 0 plain, 1 a term's items on one XCD, k >= 2 a class of terms
 * with equal levels in modes k..D-1 on one XCD).
 * No counterpart in the reference. */
int sg4_plan_set_config(sg4_plan *p, int32_t nvec, int32_t prec, const int32_t *cfg4);
/* sg4_apply_rows through a captured HIP graph (zeroing + H.psi kernel), one
 * graph per (nvec, prec, psi, hpsi); replays are ordered after the plan's
 * stream and the plan's stream waits for them. */
int sg4_apply_graph(sg4_plan *p, int32_t nvec, const double *psi, double *hpsi, int32_t prec);

/* Plan options.  SG4_OPT_GROUPING (default 1): terms whose active modes are
 * all square go through register-blocked passes over groups of up to three
 * modes, and a type-1 operator whose 1-D quadratures integrate products of
 * basis functions exactly (B^T W B = I) on a complete Smolyak term set takes
 * the split path: its kinetic part becomes one sparse matrix over the packed
 * rows, K[r][r'] = sum_k c_k T_k(L - Lambda_k(r))[i_k][i'_k] with
 * T_k = BtW D2 B (the combination weights telescope), and the terms keep
 * only the basis->grid, potential and grid->basis products.  0 forces one
 * pass per mode with the kinetic on the grid, as in the reference (all paths
 * give the same operator up to rounding). */
#define SG4_OPT_GROUPING 1
/* SG4_OPT_DETERMINISTIC (default 0), split type-1 path: the weighted
 * contributions of the terms are summed in a 128-bit fixed-point accumulator
 * per result element (scale 2^64, range |partial sums| < 2^63) with 64-bit
 * integer atomics, whose sum does not depend on the order of the additions:
 * repeated applies and every launch configuration give bitwise identical
 * results, and the only rounding of the sum is the truncation of each
 * contribution below 2^-64.  A contribution beyond the range makes the result
 * NaN.  The reference scatter-adds with OpenMP atomics (order-dependent,
 * sub_module_basis_BtoG_GtoB_SG4.f90:1288-1289); the default here is FP64
 * atomics.  Costs 16 bytes of device memory per result element. */
#define SG4_OPT_DETERMINISTIC 2
int sg4_plan_set_option(sg4_plan *p, int32_t option, int32_t value);
/* Whether FP64 and FP32 applies of this plan take the split path (split = 1)
 * and the number of stored entries of its kinetic matrix (nonzero whenever
 * the split algebra holds, which is all SG4_PREC_F32X2 needs; the FP64/FP32
 * split kernels also need a register-blocked shape for every term, e.g. no
 * mode of more than 24 points).  No counterpart in the reference. */
int sg4_plan_split_info(const sg4_plan *p, int32_t *split, int64_t *kin_nnz);

/* Statistics of the last apply for roofline reporting: its work items, and
 * the algorithmic flops and HBM bytes of its main H.psi launch (the one that
 * sg4_plan_timing brackets): basis->grid, grid->basis (and, off the split
 * path, kinetic) mode products of the active modes; bytes = potential grid,
 * map, one read of psi and one write of H.psi. */
int sg4_plan_stats(const sg4_plan *p, int32_t nvec, double *flops, double *bytes,
                   int64_t *items);
/* Split-precision mode (SG4_PREC_F32X2): flops the matrix cores executed in
 * the last such apply's main term launch (padded Kronecker tiles times the
 * six limb products) and the number of group matrices of the plan (-1 before
 * the first such apply).  No counterpart in the reference. */
int sg4_plan_mfma_stats(const sg4_plan *p, double *mfma_flops, int64_t *group_mats);
/* Kernel timing for roofline reporting: when enabled, every apply records a
 * pair of HIP events on the plan's stream around its main H.psi kernel launch
 * (on the split path the term kernel's main item list, not the kinetic rows);
 * sg4_plan_kernel_times waits for them, returns the elapsed milliseconds of
 * up to max recorded launches in order, and clears the record. */
int sg4_plan_timing(sg4_plan *p, int32_t enable);
int sg4_plan_kernel_times(sg4_plan *p, float *ms, int32_t max, int32_t *n);
int sg4_plan_destroy(sg4_plan *p);

#define SG4_MAX_D 32
#define SG4_MAX_L 40
#define SG4_MAX_N 64

#ifdef __cplusplus
}
#endif
#endif
