!===========================================================================
! mod_sg4_gpu -- iso_c_binding layer between an ElVibRot-style Fortran host
! and the MI355X SG4 H.psi library (include/sg4hpsi.h).
!
! It keeps the reference's entry points for the SG4 operator action:
!   sub_TabOpPsi_FOR_SGtype4(Psi,OpPsi,para_Op)   (SRC/sub_Operator/sub_OpPsi_SG4.f90:681)
!   sub_OpPsi_FOR_SGtype4(Psi,OpPsi,para_Op)      (SRC/sub_Operator/sub_OpPsi_SG4.f90:62)
! with the same argument meaning and the same error behaviour (STOP with the
! reference's messages on an empty block or a complex psi,
! sub_OpPsi_SG4.f90:741-753).
!
! Two layers:
!  * array-level routines a reference build calls with its own fields:
!      sg4_gpu_init_BasisnD  -- device plan from BasisnD's own tables
!                               (Tab_nDval, WeightSG, tab_iB_OF_SRep_TO_iB
!                               with 0 = absent, tab_basisPrimSG(l,k) matrices)
!      sg4_gpu_set_op_type1 / sg4_gpu_set_op_type10 -- operator grids
!      sg4_gpu_set_op_type1_nb0 -- grids of several electronic surfaces
!      sg4_gpu_TabOpPsi      -- block H.psi on host arrays PsiB(nb_ba,nvec)
!      sg4_gpu_TabOpPsi_dev  -- block H.psi on DEVICE blocks (c_ptr)
!  * sub_TabOpPsi_FOR_SGtype4 / sub_OpPsi_FOR_SGtype4 over small param_psi /
!    param_Op types holding only the fields the SG4 action reads (RvecB,
!    nb_tot, cplx; the device plan and nb_ba): the reference's own types
!    carry far more and live in modules this file cannot use.
!===========================================================================
MODULE mod_sg4_gpu
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE
  PRIVATE

  integer(c_int32_t), parameter, PUBLIC :: SG4_PREC_FP64 = 0, SG4_PREC_F32 = 1, SG4_PREC_F32X2 = 2
  integer(c_int32_t), parameter, PUBLIC :: SG4_BASIS_HO = 0, SG4_BASIS_LEGENDRE = 1
  integer(c_int), parameter, PUBLIC :: SG4_OK = 0, SG4_ERR_ARG = -1, SG4_ERR_HIP = -2, &
                                       SG4_ERR_OOM = -3, SG4_ERR_UNSUPPORTED = -4, SG4_ERR_STATE = -5

  TYPE, BIND(C), PUBLIC :: sg4_mode_desc
    integer(c_int32_t) :: kind = 0
    integer(c_int32_t) :: nq_A = 1, nq_B = 2, nb_A = 1, nb_B = 2
    real(c_double)     :: Q0 = 0.0_c_double, scaleQ = 1.0_c_double
  END TYPE sg4_mode_desc

  TYPE, BIND(C), PUBLIC :: sg4_tables
    integer(c_int32_t) :: D = 0, Lmax = 0
    integer(c_int64_t) :: nterms = 0, nb_ba = 0
    type(c_ptr) :: levels = c_null_ptr, weights = c_null_ptr, map = c_null_ptr
    type(c_ptr) :: nq = c_null_ptr, nb = c_null_ptr, mats = c_null_ptr, nodes = c_null_ptr
    type(c_ptr) :: qweights = c_null_ptr
  END TYPE sg4_tables

  TYPE, BIND(C) :: sg4_operator
    integer(c_int32_t) :: type_op = 1
    type(c_ptr) :: V = c_null_ptr, ck = c_null_ptr, G = c_null_ptr
    type(c_ptr) :: Jac = c_null_ptr, sqRJ = c_null_ptr
  END TYPE sg4_operator

  !> The fields of one primitive basis tab_basisPrimSG(l,k) (TYPE(basis),
  !! SRC/sub_Basis/sub_module_basis_set_alloc.f90:91-104) the SG4 action
  !! reads: a reference build fills it with
  !!   d0b = dnRGB%d0, d0bwrho = dnRBGwrho%d0, d1gg = dnRGG%d1(:,:,1),
  !!   d2gg = dnRGG%d2(:,:,1,1), x = x(1,:), wrho = wrho
  !! (x and wrho are optional: only the device-side grid evaluations and
  !! the grid-value gradients need them).
  TYPE, PUBLIC :: sg4_prim_basis
    real(c_double), allocatable :: d0b(:,:)      ! (nq, nb)
    real(c_double), allocatable :: d0bwrho(:,:)  ! (nb, nq)
    real(c_double), allocatable :: d1gg(:,:)     ! (nq, nq)
    real(c_double), allocatable :: d2gg(:,:)     ! (nq, nq)
    real(c_double), allocatable :: x(:), wrho(:) ! (nq)
  END TYPE sg4_prim_basis

  ! the parts of param_psi (SRC/sub_WP/sub_module_psi_set_alloc.f90:56-169) used here
  TYPE, PUBLIC :: param_psi
    logical                     :: cplx   = .FALSE.
    integer                     :: nb_tot = 0
    real(c_double), allocatable :: RvecB(:)
  END TYPE param_psi

  ! the parts of param_Op (SRC/sub_Operator/sub_module_SetOp.f90:59-154) used here
  TYPE, PUBLIC :: param_Op
    type(c_ptr)        :: setup = c_null_ptr
    type(c_ptr)        :: plan  = c_null_ptr
    ! several electronic surfaces (nb0 = nb_bie > 1): one plan per block
    ! (ib0, jb0) of GridOp(:,ib0,jb0,:), c_null_ptr for a zero block
    integer            :: nb0 = 1
    type(c_ptr), allocatable :: blocks(:,:)
    integer            :: nb_ba = 0
    integer            :: type_Op = 1
    integer            :: D = 0
    integer(c_int64_t) :: sum_nq = 0
    integer(c_int32_t) :: max_nvec = 0
    integer(c_int32_t) :: precision = SG4_PREC_FP64
  END TYPE param_Op

  INTERFACE
    FUNCTION sg4_setup_create(D, L, LB, modes, out) BIND(C, name='sg4_setup_create')
      IMPORT :: c_int, c_int32_t, c_ptr, sg4_mode_desc
      integer(c_int32_t), value :: D, L, LB
      type(sg4_mode_desc), intent(in) :: modes(*)
      type(c_ptr), intent(out) :: out
      integer(c_int) :: sg4_setup_create
    END FUNCTION
    FUNCTION sg4_setup_sizes(s, nterms, nb_ba, sum_nb, sum_nq) BIND(C, name='sg4_setup_sizes')
      IMPORT :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: s
      integer(c_int64_t), intent(out) :: nterms, nb_ba, sum_nb, sum_nq
      integer(c_int) :: sg4_setup_sizes
    END FUNCTION
    FUNCTION sg4_setup_tables(s, t) BIND(C, name='sg4_setup_tables')
      IMPORT :: c_int, c_ptr, sg4_tables
      type(c_ptr), value :: s
      type(sg4_tables), intent(out) :: t
      integer(c_int) :: sg4_setup_tables
    END FUNCTION
    FUNCTION sg4_setup_destroy(s) BIND(C, name='sg4_setup_destroy')
      IMPORT :: c_int, c_ptr
      type(c_ptr), value :: s
      integer(c_int) :: sg4_setup_destroy
    END FUNCTION
    FUNCTION sg4_plan_create(t, max_nvec, out) BIND(C, name='sg4_plan_create')
      IMPORT :: c_int, c_int32_t, c_ptr, sg4_tables
      type(sg4_tables), intent(in) :: t
      integer(c_int32_t), value :: max_nvec
      type(c_ptr), intent(out) :: out
      integer(c_int) :: sg4_plan_create
    END FUNCTION
    FUNCTION sg4_plan_set_henon_heiles(p, lambda) BIND(C, name='sg4_plan_set_henon_heiles')
      IMPORT :: c_int, c_ptr, c_double
      type(c_ptr), value :: p
      real(c_double), value :: lambda
      integer(c_int) :: sg4_plan_set_henon_heiles
    END FUNCTION
    FUNCTION sg4_apply_host(p, nvec, psi, hpsi, prec) BIND(C, name='sg4_apply_host')
      IMPORT :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: p
      integer(c_int32_t), value :: nvec, prec
      real(c_double), intent(in) :: psi(*)
      real(c_double), intent(inout) :: hpsi(*)
      integer(c_int) :: sg4_apply_host
    END FUNCTION
    FUNCTION sg4_apply(p, nvec, psi, hpsi, prec) BIND(C, name='sg4_apply')
      IMPORT :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: p, psi, hpsi
      integer(c_int32_t), value :: nvec, prec
      integer(c_int) :: sg4_apply
    END FUNCTION
    FUNCTION sg4_plan_set_operator(p, op) BIND(C, name='sg4_plan_set_operator')
      IMPORT :: c_int, c_ptr, sg4_operator
      type(c_ptr), value :: p
      type(sg4_operator), intent(in) :: op
      integer(c_int) :: sg4_plan_set_operator
    END FUNCTION
    FUNCTION sg4_plan_destroy(p) BIND(C, name='sg4_plan_destroy')
      IMPORT :: c_int, c_ptr
      type(c_ptr), value :: p
      integer(c_int) :: sg4_plan_destroy
    END FUNCTION
    FUNCTION sg4_last_error() BIND(C, name='sg4_last_error')
      IMPORT :: c_ptr
      type(c_ptr) :: sg4_last_error
    END FUNCTION
  END INTERFACE

  PUBLIC :: sub_TabOpPsi_FOR_SGtype4, sub_OpPsi_FOR_SGtype4
  PUBLIC :: init_para_Op_SG4_HenonHeiles, dealloc_para_Op_SG4, alloc_psi
  PUBLIC :: sg4_gpu_init_BasisnD, sg4_gpu_set_op_type1, sg4_gpu_set_op_type10, sg4_gpu_set_op_type1_nb0
  PUBLIC :: sg4_gpu_TabOpPsi, sg4_gpu_TabOpPsi_dev

CONTAINS

  SUBROUTINE sg4_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    IF (rc == 0) RETURN
    CALL c_f_pointer(sg4_last_error(), msg, [1024])
    n = 0
    DO WHILE (n < 1024)
      IF (msg(n+1) == c_null_char) EXIT
      n = n + 1
    END DO
    write(*,*) 'ERROR in ', where, ': sg4 status ', rc, ' ', msg(1:n)
    STOP 'ERROR in the SG4 GPU library'
  END SUBROUTINE sg4_check

  SUBROUTINE alloc_psi(psi, nb_tot)
    TYPE(param_psi), intent(inout) :: psi
    integer, intent(in) :: nb_tot
    IF (allocated(psi%RvecB)) deallocate(psi%RvecB)
    allocate(psi%RvecB(nb_tot))
    psi%nb_tot = nb_tot
    psi%RvecB(:) = 0.0_c_double
  END SUBROUTINE alloc_psi

  !> Henon-Heiles operator on the SG4 HO basis of the reference tests
  !! (Working_tests/MPI_tests/6D_Davidson_S0/shell_run:97-174): D modes,
  !! L = LB = LG, nq(l) = nb(l) = 1 + 2l, T = -1/2 sum d^2, V on the device.
  SUBROUTINE init_para_Op_SG4_HenonHeiles(para_Op, D, L, lambda, max_nvec)
    TYPE(param_Op), intent(inout) :: para_Op
    integer, intent(in) :: D, L, max_nvec
    real(c_double), intent(in) :: lambda
    TYPE(sg4_mode_desc) :: modes(D)
    TYPE(sg4_tables) :: tabs
    integer(c_int64_t) :: nterms, nb_ba, sum_nb, sum_nq
    CALL sg4_check(sg4_setup_create(int(D, c_int32_t), int(L, c_int32_t), int(L, c_int32_t), &
                                    modes, para_Op%setup), 'sg4_setup_create')
    CALL sg4_check(sg4_setup_sizes(para_Op%setup, nterms, nb_ba, sum_nb, sum_nq), 'sg4_setup_sizes')
    CALL sg4_check(sg4_setup_tables(para_Op%setup, tabs), 'sg4_setup_tables')
    CALL sg4_check(sg4_plan_create(tabs, int(max_nvec, c_int32_t), para_Op%plan), 'sg4_plan_create')
    CALL sg4_check(sg4_plan_set_henon_heiles(para_Op%plan, lambda), 'sg4_plan_set_henon_heiles')
    para_Op%nb_ba = int(nb_ba)
    para_Op%max_nvec = int(max_nvec, c_int32_t)
    para_Op%type_Op = 1
  END SUBROUTINE init_para_Op_SG4_HenonHeiles

  !> Device plan from the SG4 tables BasisnD holds after
  !! RecSparseGrid_ForDP_type4 (SRC/sub_Basis/sub_quadra_SparseBasis.f90:922),
  !! in the reference's own conventions:
  !!   Tab_nDval(D, nb_SG)   para_SGType2%nDind_SmolyakRep%Tab_nDval, 0-based
  !!                         levels (sub_quadra_SparseBasis.f90:778-790)
  !!   WeightSG(nb_SG)       BasisnD%WeightSG
  !!   tab_iB(:)             para_SGType2%tab_iB_OF_SRep_TO_iB: for the term
  !!                         blocks in term order (in-term index, mode 1
  !!                         fastest) the 1-based packed index, 0 = absent
  !!   prim(0:Lmax, D)       the tab_basisPrimSG(l,k) matrices
  !!   nb_ba                 the packed basis size (nDindB%Max_nDI)
  !! The map and the weights are read as SmolyakRepBasis_TO_tabPackedBasis
  !! reads them (sub_module_basis_BtoG_GtoB_SG4.f90:1005-1027): entries
  !! outside 1..nb_ba are absent and terms with |WeightSG| < 1e-6 are
  !! skipped (their weight is set to zero).
  SUBROUTINE sg4_gpu_init_BasisnD(para_Op, Tab_nDval, WeightSG, tab_iB, prim, nb_ba, max_nvec, nb0)
    TYPE(param_Op), intent(inout) :: para_Op
    integer, intent(in) :: Tab_nDval(:,:)
    real(c_double), intent(in) :: WeightSG(:)
    integer, intent(in) :: tab_iB(:)
    TYPE(sg4_prim_basis), intent(in) :: prim(0:,:)
    integer, intent(in) :: nb_ba, max_nvec
    integer, intent(in), optional :: nb0
    integer :: D, L1, Lmax, nterms, k, l, t, nq, nb, ib0, jb0
    integer(c_int64_t) :: sum_nb, sum_nq, nbt, nqt, off, noff
    integer(c_int32_t), allocatable, target :: lev(:,:), map0(:), nqv(:,:), nbv(:,:)
    real(c_double), allocatable, target :: w(:), mats(:), nodes(:), qw(:)
    logical :: have_x
    TYPE(sg4_tables) :: tabs

    D = size(Tab_nDval, 1)
    nterms = size(Tab_nDval, 2)
    L1 = size(prim, 1)
    Lmax = L1 - 1
    IF (size(prim, 2) /= D) STOP ' ERROR in sg4_gpu_init_BasisnD: prim(0:Lmax,:) and Tab_nDval(:,:) disagree on D'
    IF (size(WeightSG) /= nterms) STOP ' ERROR in sg4_gpu_init_BasisnD: size(WeightSG) /= nb_SG'
    allocate(lev(D, nterms), w(nterms), nqv(L1, D), nbv(L1, D))
    lev(:,:) = int(Tab_nDval, c_int32_t)
    w(:) = WeightSG
    WHERE (abs(w) < 1.0e-6_c_double) w = 0.0_c_double
    off = 0
    noff = 0
    have_x = .TRUE.
    DO k = 1, D
      DO l = 0, Lmax
        IF (.NOT. allocated(prim(l,k)%d0b)) STOP ' ERROR in sg4_gpu_init_BasisnD: prim(l,k)%d0b not set'
        nq = size(prim(l,k)%d0b, 1)
        nb = size(prim(l,k)%d0b, 2)
        IF (any(shape(prim(l,k)%d0bwrho) /= [nb, nq]) .OR. any(shape(prim(l,k)%d1gg) /= [nq, nq]) &
            .OR. any(shape(prim(l,k)%d2gg) /= [nq, nq])) &
          STOP ' ERROR in sg4_gpu_init_BasisnD: prim(l,k) matrices of inconsistent shapes'
        nqv(l+1, k) = nq
        nbv(l+1, k) = nb
        off = off + 2_c_int64_t * nq * nb + 2_c_int64_t * nq * nq
        noff = noff + nq
        have_x = have_x .AND. allocated(prim(l,k)%x) .AND. allocated(prim(l,k)%wrho)
      END DO
    END DO
    ! for k, for l: dnRGB%d0, dnRBGwrho%d0, dnRGG%d1, dnRGG%d2, column-major
    allocate(mats(off))
    off = 0
    DO k = 1, D
      DO l = 0, Lmax
        nq = nqv(l+1, k)
        nb = nbv(l+1, k)
        mats(off+1:off+nq*nb) = reshape(prim(l,k)%d0b, [nq*nb]);     off = off + nq*nb
        mats(off+1:off+nq*nb) = reshape(prim(l,k)%d0bwrho, [nq*nb]); off = off + nq*nb
        mats(off+1:off+nq*nq) = reshape(prim(l,k)%d1gg, [nq*nq]);    off = off + nq*nq
        mats(off+1:off+nq*nq) = reshape(prim(l,k)%d2gg, [nq*nq]);    off = off + nq*nq
      END DO
    END DO
    IF (have_x) THEN
      allocate(nodes(noff), qw(noff))
      noff = 0
      DO k = 1, D
        DO l = 0, Lmax
          nq = nqv(l+1, k)
          nodes(noff+1:noff+nq) = prim(l,k)%x
          qw(noff+1:noff+nq) = prim(l,k)%wrho
          noff = noff + nq
        END DO
      END DO
    END IF
    sum_nb = 0
    sum_nq = 0
    DO t = 1, nterms
      nbt = 1
      nqt = 1
      DO k = 1, D
        l = Tab_nDval(k, t)
        IF (l < 0 .OR. l > Lmax) STOP ' ERROR in sg4_gpu_init_BasisnD: Tab_nDval out of 0..Lmax'
        nbt = nbt * nbv(l+1, k)
        nqt = nqt * nqv(l+1, k)
      END DO
      sum_nb = sum_nb + nbt
      sum_nq = sum_nq + nqt
    END DO
    IF (size(tab_iB, kind=c_int64_t) < sum_nb) &
      STOP ' ERROR in sg4_gpu_init_BasisnD: tab_iB_OF_SRep_TO_iB shorter than the term blocks'
    ! 1-based, 0 (or anything outside 1..nb_ba) = absent  ->  0-based, -1 = absent
    allocate(map0(sum_nb))
    WHERE (tab_iB(1:sum_nb) > 0 .AND. tab_iB(1:sum_nb) <= nb_ba)
      map0 = int(tab_iB(1:sum_nb) - 1, c_int32_t)
    ELSEWHERE
      map0 = -1_c_int32_t
    END WHERE
    tabs%D = int(D, c_int32_t)
    tabs%Lmax = int(Lmax, c_int32_t)
    tabs%nterms = int(nterms, c_int64_t)
    tabs%nb_ba = int(nb_ba, c_int64_t)
    tabs%levels = c_loc(lev)
    tabs%weights = c_loc(w)
    tabs%map = c_loc(map0)
    tabs%nq = c_loc(nqv)
    tabs%nb = c_loc(nbv)
    tabs%mats = c_loc(mats)
    IF (have_x) THEN
      tabs%nodes = c_loc(nodes)
      tabs%qweights = c_loc(qw)
    END IF
    ! the plan copies every table: the local arrays may go afterwards
    CALL sg4_check(sg4_plan_create(tabs, int(max_nvec, c_int32_t), para_Op%plan), &
                   'sg4_gpu_init_BasisnD')
    para_Op%nb0 = 1
    IF (present(nb0)) para_Op%nb0 = nb0
    IF (para_Op%nb0 < 1) STOP ' ERROR in sg4_gpu_init_BasisnD: nb0 < 1'
    IF (para_Op%nb0 > 1) THEN
      ! one plan per surface block (ib0, jb0); block (1,1) is para_Op%plan
      allocate(para_Op%blocks(para_Op%nb0, para_Op%nb0))
      DO jb0 = 1, para_Op%nb0
        DO ib0 = 1, para_Op%nb0
          IF (ib0 == 1 .AND. jb0 == 1) THEN
            para_Op%blocks(1, 1) = para_Op%plan
          ELSE
            CALL sg4_check(sg4_plan_create(tabs, int(max_nvec, c_int32_t), para_Op%blocks(ib0, jb0)), &
                           'sg4_gpu_init_BasisnD')
          END IF
        END DO
      END DO
    END IF
    para_Op%nb_ba = nb_ba
    para_Op%D = D
    para_Op%sum_nq = sum_nq
    para_Op%max_nvec = int(max_nvec, c_int32_t)
  END SUBROUTINE sg4_gpu_init_BasisnD

  !> Type-1 operator grids (get_OpGrid_type1_OF_ONEDP_FOR_SG4,
  !! SRC/sub_Operator/sub_OpPsi_SG4.f90:2536): V on every grid point of every
  !! term (term order, sum_t nq_t values), ck(k) = -G_kk/2 constant.
  SUBROUTINE sg4_gpu_set_op_type1(para_Op, V, ck)
    TYPE(param_Op), intent(inout) :: para_Op
    real(c_double), intent(in), target :: V(:), ck(:)
    TYPE(sg4_operator) :: op
    IF (size(V, kind=c_int64_t) /= para_Op%sum_nq .OR. size(ck) /= para_Op%D) &
      STOP ' ERROR in sg4_gpu_set_op_type1: V(sum_nq) and ck(D) expected'
    op%type_op = 1
    op%V = c_loc(V)
    op%ck = c_loc(ck)
    CALL sg4_check(sg4_plan_set_operator(para_Op%plan, op), 'sg4_gpu_set_op_type1')
    para_Op%type_Op = 1
  END SUBROUTINE sg4_gpu_set_op_type1

  !> Type-1 operator on nb0 surfaces: GridV(:,ib0,jb0) = GridOp(:,ib0,jb0) of
  !! the potential term (sub_OpPsi_SG4.f90:1506-1522, every grid point in term
  !! order), the kinetic coefficients ck(k) = -G_kk/2 on the diagonal blocks
  !! only.  An off-diagonal block whose grid is zero everywhere is dropped
  !! (the reference skips grid_zero terms, :1508); block (ib0, jb0) then acts
  !! as the single-surface operator on its grid.
  SUBROUTINE sg4_gpu_set_op_type1_nb0(para_Op, GridV, ck)
    TYPE(param_Op), intent(inout) :: para_Op
    real(c_double), intent(in), target :: GridV(:,:,:), ck(:)
    real(c_double), allocatable, target :: Vb(:), c0(:)
    TYPE(sg4_operator) :: op
    integer :: ib0, jb0, n0
    integer(c_int) :: rc
    n0 = para_Op%nb0
    IF (n0 < 2) STOP ' ERROR in sg4_gpu_set_op_type1_nb0: the plan was built for one surface'
    IF (size(GridV, 1, kind=c_int64_t) /= para_Op%sum_nq .OR. size(GridV, 2) /= n0 .OR. &
        size(GridV, 3) /= n0 .OR. size(ck) /= para_Op%D) &
      STOP ' ERROR in sg4_gpu_set_op_type1_nb0: GridV(sum_nq,nb0,nb0) and ck(D) expected'
    allocate(Vb(para_Op%sum_nq), c0(para_Op%D))
    c0(:) = 0.0_c_double
    DO jb0 = 1, n0
      DO ib0 = 1, n0
        IF (.NOT. c_associated(para_Op%blocks(ib0, jb0))) CYCLE
        IF (ib0 /= jb0 .AND. all(GridV(:, ib0, jb0) == 0.0_c_double)) THEN
          rc = sg4_plan_destroy(para_Op%blocks(ib0, jb0))
          para_Op%blocks(ib0, jb0) = c_null_ptr
          CYCLE
        END IF
        Vb(:) = GridV(:, ib0, jb0)
        op%type_op = 1
        op%V = c_loc(Vb)
        IF (ib0 == jb0) THEN
          op%ck = c_loc(ck)
        ELSE
          op%ck = c_loc(c0)
        END IF
        CALL sg4_check(sg4_plan_set_operator(para_Op%blocks(ib0, jb0), op), 'sg4_gpu_set_op_type1_nb0')
      END DO
    END DO
    para_Op%type_Op = 1
  END SUBROUTINE sg4_gpu_set_op_type1_nb0

  !> Type-10 operator grids (get_OpGrid_type10_OF_ONEDP_FOR_SG4,
  !! SRC/sub_Operator/sub_OpPsi_SG4.f90:327): V, Jac, sqRhoOVERJac on every
  !! grid point and G(:,:,iq) = GGiq(iq,:,:), so that G(j,i,iq) multiplies
  !! d_j inside d_i as GGiq(:,j,i) does (:381-385).
  SUBROUTINE sg4_gpu_set_op_type10(para_Op, V, G, Jac, sqRhoOVERJac)
    TYPE(param_Op), intent(inout) :: para_Op
    real(c_double), intent(in), target :: V(:), G(:,:,:), Jac(:), sqRhoOVERJac(:)
    TYPE(sg4_operator) :: op
    IF (size(V, kind=c_int64_t) /= para_Op%sum_nq .OR. size(Jac, kind=c_int64_t) /= para_Op%sum_nq &
        .OR. size(sqRhoOVERJac, kind=c_int64_t) /= para_Op%sum_nq .OR. size(G, 1) /= para_Op%D &
        .OR. size(G, 2) /= para_Op%D .OR. size(G, 3, kind=c_int64_t) /= para_Op%sum_nq) &
      STOP ' ERROR in sg4_gpu_set_op_type10: V, Jac, sqRhoOVERJac(sum_nq) and G(D,D,sum_nq) expected'
    op%type_op = 10
    op%V = c_loc(V)
    op%G = c_loc(G)
    op%Jac = c_loc(Jac)
    op%sqRJ = c_loc(sqRhoOVERJac)
    CALL sg4_check(sg4_plan_set_operator(para_Op%plan, op), 'sg4_gpu_set_op_type10')
    para_Op%type_Op = 10
  END SUBROUTINE sg4_gpu_set_op_type10

  !> OpPsiB(:,v) = H PsiB(:,v) for host blocks PsiB(nb_ba, nvec) (the
  !! RvecB of a block of real param_psi side by side); the library is called
  !! on at most max_nvec columns at a time (copies over PCIe both ways).
  SUBROUTINE sg4_gpu_TabOpPsi(para_Op, PsiB, OpPsiB)
    TYPE(param_Op), intent(inout) :: para_Op
    real(c_double), intent(in)    :: PsiB(:,:)
    real(c_double), intent(inout) :: OpPsiB(:,:)
    integer :: nvec, v0, nv
    nvec = size(PsiB, 2)
    IF (nvec == 0) STOP ' ERROR in sub_TabOpPsi_FOR_SGtype4: size(Psi) = 0'
    IF (para_Op%nb0 > 1) THEN
      CALL tabOpPsi_surfaces(para_Op, PsiB, OpPsiB)
      RETURN
    END IF
    IF (size(PsiB, 1) /= para_Op%nb_ba .OR. any(shape(OpPsiB) /= shape(PsiB))) &
      STOP ' ERROR in sg4_gpu_TabOpPsi: PsiB and OpPsiB must be (nb_ba, nvec)'
    v0 = 0
    DO WHILE (v0 < nvec)
      nv = min(nvec - v0, int(para_Op%max_nvec))
      CALL sg4_check(sg4_apply_host(para_Op%plan, int(nv, c_int32_t), PsiB(:, v0+1:v0+nv), &
                                    OpPsiB(:, v0+1:v0+nv), para_Op%precision), 'sg4_gpu_TabOpPsi')
      v0 = v0 + nv
    END DO
  END SUBROUTINE sg4_gpu_TabOpPsi

  !> Several surfaces on host blocks PsiB(nb0 nb_ba, nvec): the nb0
  !! components one after the other (i_bhe = i_b + nb_ba (i_e - 1),
  !! SRC/sub_WP/sub_module_psi_io.f90:1525-1530), OpPsi_ib0 = sum_jb0
  !! H(ib0,jb0) Psi_jb0 with one library call per block and max_nvec columns.
  SUBROUTINE tabOpPsi_surfaces(para_Op, PsiB, OpPsiB)
    TYPE(param_Op), intent(inout) :: para_Op
    real(c_double), intent(in)    :: PsiB(:,:)
    real(c_double), intent(inout) :: OpPsiB(:,:)
    real(c_double), allocatable :: x(:,:), y(:,:)
    integer :: nb, nvec, v0, nv, ib0, jb0
    nb = para_Op%nb_ba
    nvec = size(PsiB, 2)
    IF (size(PsiB, 1) /= para_Op%nb0 * nb .OR. any(shape(OpPsiB) /= shape(PsiB))) &
      STOP ' ERROR in sg4_gpu_TabOpPsi: PsiB and OpPsiB must be (nb0*nb_ba, nvec)'
    allocate(x(nb, min(nvec, int(para_Op%max_nvec))), y(nb, min(nvec, int(para_Op%max_nvec))))
    OpPsiB(:,:) = 0.0_c_double
    v0 = 0
    DO WHILE (v0 < nvec)
      nv = min(nvec - v0, int(para_Op%max_nvec))
      DO jb0 = 1, para_Op%nb0
        x(:, 1:nv) = PsiB((jb0-1)*nb+1:jb0*nb, v0+1:v0+nv)
        DO ib0 = 1, para_Op%nb0
          IF (.NOT. c_associated(para_Op%blocks(ib0, jb0))) CYCLE
          CALL sg4_check(sg4_apply_host(para_Op%blocks(ib0, jb0), int(nv, c_int32_t), x, y, &
                                        para_Op%precision), 'sg4_gpu_TabOpPsi')
          OpPsiB((ib0-1)*nb+1:ib0*nb, v0+1:v0+nv) = OpPsiB((ib0-1)*nb+1:ib0*nb, v0+1:v0+nv) + y(:, 1:nv)
        END DO
      END DO
      v0 = v0 + nv
    END DO
    deallocate(x, y)
  END SUBROUTINE tabOpPsi_surfaces

  !> H.psi on DEVICE blocks: psi_d and hpsi_d point to nb_ba x nvec
  !! column-major FP64 blocks in device memory (hipMalloc, c_loc of a device
  !! array); nothing crosses PCIe and the call returns once the work is
  !! enqueued on the plan's stream.  Blocks wider than max_nvec are taken
  !! max_nvec columns at a time.  Host, managed or fine-grained memory is
  !! refused (status SG4_ERR_ARG).  With stat present the status is
  !! returned; without it a failure stops, as sg4_check does.
  SUBROUTINE sg4_gpu_TabOpPsi_dev(para_Op, nvec, psi_d, hpsi_d, stat)
    TYPE(param_Op), intent(inout) :: para_Op
    integer, intent(in) :: nvec
    type(c_ptr), intent(in) :: psi_d, hpsi_d
    integer, intent(out), optional :: stat
    integer(c_intptr_t) :: a, b, step
    integer(c_int) :: rc
    integer :: v0, nv
    IF (nvec < 1) STOP ' ERROR in sub_TabOpPsi_FOR_SGtype4: size(Psi) = 0'
    IF (para_Op%nb0 > 1) THEN
      ! the components of a device block are not contiguous per surface
      IF (present(stat)) THEN
        stat = int(SG4_ERR_UNSUPPORTED)
        RETURN
      END IF
      STOP ' ERROR in sg4_gpu_TabOpPsi_dev: several surfaces are applied on host arrays (sg4_gpu_TabOpPsi)'
    END IF
    a = transfer(psi_d, a)
    b = transfer(hpsi_d, b)
    rc = 0
    v0 = 0
    DO WHILE (v0 < nvec .AND. rc == 0)
      nv = min(nvec - v0, int(para_Op%max_nvec))
      step = int(v0, c_intptr_t) * int(para_Op%nb_ba, c_intptr_t) * 8_c_intptr_t
      rc = sg4_apply(para_Op%plan, int(nv, c_int32_t), transfer(a + step, psi_d), &
                     transfer(b + step, hpsi_d), para_Op%precision)
      v0 = v0 + nv
    END DO
    IF (present(stat)) THEN
      stat = int(rc)
    ELSE
      CALL sg4_check(rc, 'sg4_gpu_TabOpPsi_dev')
    END IF
  END SUBROUTINE sg4_gpu_TabOpPsi_dev

  SUBROUTINE dealloc_para_Op_SG4(para_Op)
    TYPE(param_Op), intent(inout) :: para_Op
    integer(c_int) :: rc
    integer :: ib, jb
    IF (allocated(para_Op%blocks)) THEN
      para_Op%blocks(1, 1) = c_null_ptr          ! = para_Op%plan, destroyed below
      DO jb = 1, size(para_Op%blocks, 2)
        DO ib = 1, size(para_Op%blocks, 1)
          IF (c_associated(para_Op%blocks(ib, jb))) rc = sg4_plan_destroy(para_Op%blocks(ib, jb))
        END DO
      END DO
      deallocate(para_Op%blocks)
    END IF
    para_Op%nb0 = 1
    IF (c_associated(para_Op%plan))  rc = sg4_plan_destroy(para_Op%plan)
    IF (c_associated(para_Op%setup)) rc = sg4_setup_destroy(para_Op%setup)
    para_Op%plan = c_null_ptr
    para_Op%setup = c_null_ptr
  END SUBROUTINE dealloc_para_Op_SG4

  !> OpPsi(:) = H Psi(:) for a block of real packed vectors.
  !! Same checks and zeroing semantics as the reference (sub_OpPsi_SG4.f90:741-770).
  SUBROUTINE sub_TabOpPsi_FOR_SGtype4(Psi, OpPsi, para_Op)
    TYPE(param_psi), intent(in)    :: Psi(:)
    TYPE(param_psi), intent(inout) :: OpPsi(:)
    TYPE(param_Op),  intent(inout) :: para_Op
    real(c_double), allocatable :: blk(:,:), hblk(:,:)
    integer :: itab, nvec, v0, nv
    IF (size(Psi) == 0) STOP ' ERROR in sub_TabOpPsi_FOR_SGtype4: size(Psi) = 0'
    IF (Psi(1)%cplx) STOP ' ERROR in sub_TabOpPsi_FOR_SGtype4: Psi(1) is complex'
    nvec = size(Psi)
    IF (para_Op%nb0 > 1) THEN
      ! nb_tot = nb0 nb_ba: the surface blocks on the whole block of vectors
      allocate(blk(para_Op%nb0 * para_Op%nb_ba, nvec), hblk(para_Op%nb0 * para_Op%nb_ba, nvec))
      DO itab = 1, nvec
        blk(:, itab) = Psi(itab)%RvecB(1:para_Op%nb0 * para_Op%nb_ba)
      END DO
      CALL tabOpPsi_surfaces(para_Op, blk, hblk)
      DO itab = 1, nvec
        IF (.NOT. allocated(OpPsi(itab)%RvecB)) CALL alloc_psi(OpPsi(itab), para_Op%nb0 * para_Op%nb_ba)
        OpPsi(itab)%cplx = .FALSE.
        OpPsi(itab)%RvecB(1:para_Op%nb0 * para_Op%nb_ba) = hblk(:, itab)
      END DO
      deallocate(blk, hblk)
      RETURN
    END IF
    allocate(blk(para_Op%nb_ba, min(nvec, int(para_Op%max_nvec))))
    allocate(hblk(para_Op%nb_ba, min(nvec, int(para_Op%max_nvec))))
    DO itab = 1, nvec
      IF (.NOT. allocated(OpPsi(itab)%RvecB)) CALL alloc_psi(OpPsi(itab), para_Op%nb_ba)
      OpPsi(itab)%cplx = .FALSE.
    END DO
    v0 = 0
    DO WHILE (v0 < nvec)   ! blocks of at most max_nvec vectors
      nv = min(nvec - v0, int(para_Op%max_nvec))
      DO itab = 1, nv
        blk(:, itab) = Psi(v0 + itab)%RvecB(1:para_Op%nb_ba)
      END DO
      CALL sg4_check(sg4_apply_host(para_Op%plan, int(nv, c_int32_t), blk, hblk, &
                                    para_Op%precision), 'sub_TabOpPsi_FOR_SGtype4')
      DO itab = 1, nv
        OpPsi(v0 + itab)%RvecB(1:para_Op%nb_ba) = hblk(:, itab)
      END DO
      v0 = v0 + nv
    END DO
    deallocate(blk, hblk)
  END SUBROUTINE sub_TabOpPsi_FOR_SGtype4

  !> single-vector form (sub_OpPsi_FOR_SGtype4, sub_OpPsi_SG4.f90:62)
  SUBROUTINE sub_OpPsi_FOR_SGtype4(Psi, OpPsi, para_Op)
    TYPE(param_psi), intent(in)    :: Psi
    TYPE(param_psi), intent(inout) :: OpPsi
    TYPE(param_Op),  intent(inout) :: para_Op
    TYPE(param_psi) :: tp(1), th(1)
    tp(1) = Psi
    CALL sub_TabOpPsi_FOR_SGtype4(tp, th, para_Op)
    OpPsi = th(1)
  END SUBROUTINE sub_OpPsi_FOR_SGtype4

END MODULE mod_sg4_gpu
