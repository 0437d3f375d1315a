!===========================================================================
! mod_sg4_gpu -- iso_c_binding layer between an ElVibRot-style Fortran host
! and the MI355X SG4 H.psi library (include/sg4hpsi.h).
!
! It keeps the reference's entry points for the SG4 operator action:
!   sub_TabOpPsi_FOR_SGtype4(Psi,OpPsi,para_Op)   (SRC/sub_Operator/sub_OpPsi_SG4.f90:681)
!   sub_OpPsi_FOR_SGtype4(Psi,OpPsi,para_Op)      (SRC/sub_Operator/sub_OpPsi_SG4.f90:62)
! with the same argument meaning and the same error behaviour (STOP with the
! reference's messages on an empty block or a complex psi,
! sub_OpPsi_SG4.f90:741-753).  The reference's param_psi / param_Op carry far
! more than this path needs; the types below hold only the fields the SG4
! action reads (RvecB, nb_tot, cplx; the device plan and nb_ba).
!===========================================================================
MODULE mod_sg4_gpu
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE
  PRIVATE

  integer(c_int32_t), parameter, PUBLIC :: SG4_PREC_FP64 = 0, SG4_PREC_F32 = 1
  integer(c_int32_t), parameter, PUBLIC :: SG4_BASIS_HO = 0, SG4_BASIS_LEGENDRE = 1

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
    integer            :: nb_ba = 0
    integer            :: type_Op = 1
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

  SUBROUTINE dealloc_para_Op_SG4(para_Op)
    TYPE(param_Op), intent(inout) :: para_Op
    integer(c_int) :: rc
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
