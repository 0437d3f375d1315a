! Fortran host given the SG4 tables in the reference's own layout (0-based
! Tab_nDval, WeightSG, 1-based tab_iB_OF_SRep_TO_iB with 0 = absent, the
! tab_basisPrimSG(l,k) matrices dnRGB%d0 / dnRBGwrho%d0 / dnRGG%d1 / dnRGG%d2)
! and the operator grids, read from a stream file:
!   int32  D, Lmax, nb_SG, nb_ba, sum_nb, sum_nq, nvec, type_op, max_nvec
!   int32  Tab_nDval(D, nb_SG);  real64 WeightSG(nb_SG);  int32 tab_iB(sum_nb)
!   for k = 1..D, l = 0..Lmax: int32 nq, nb; real64 d0b(nq,nb), d0bwrho(nb,nq),
!                              d1gg(nq,nq), d2gg(nq,nq), x(nq), wrho(nq)
!   type 1:  real64 V(sum_nq), ck(D)
!   type 10: real64 V(sum_nq), G(D,D,sum_nq), Jac(sum_nq), sqRhoOVERJac(sum_nq)
!   type 21 (type 1 on two electronic surfaces): real64 GridV(sum_nq,2,2), ck(D)
!   real64 psi(nb_ba, nvec)   (type 21: psi(2 nb_ba, nvec))
! It builds the device plan with sg4_gpu_init_BasisnD, applies H on host
! arrays (sg4_gpu_TabOpPsi) and on device blocks it allocates itself
! (sg4_gpu_TabOpPsi_dev), checks that a host array handed to the device
! entry is refused, and writes
!   real64 hpsi_host(nb_ba, nvec), hpsi_dev(nb_ba, nvec); int32 status_refused
! (type 21: hpsi_host from sg4_gpu_TabOpPsi and, in place of hpsi_dev, the
! same through sub_TabOpPsi_FOR_SGtype4 on param_psi; status_refused is the
! device entry's answer to a two-surface plan)
PROGRAM test_sg4_reflayout
  USE, INTRINSIC :: iso_c_binding
  USE mod_sg4_gpu
  IMPLICIT NONE
  INTERFACE
    FUNCTION hipMalloc(ptr, nbytes) BIND(C, name='hipMalloc')
      IMPORT :: c_ptr, c_size_t, c_int
      type(c_ptr), intent(out) :: ptr
      integer(c_size_t), value :: nbytes
      integer(c_int) :: hipMalloc
    END FUNCTION
    FUNCTION hipMemcpy(dst, src, nbytes, kind) BIND(C, name='hipMemcpy')
      IMPORT :: c_ptr, c_size_t, c_int
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: nbytes
      integer(c_int), value :: kind
      integer(c_int) :: hipMemcpy
    END FUNCTION
    FUNCTION hipFree(ptr) BIND(C, name='hipFree')
      IMPORT :: c_ptr, c_int
      type(c_ptr), value :: ptr
      integer(c_int) :: hipFree
    END FUNCTION
    FUNCTION hipDeviceSynchronize() BIND(C, name='hipDeviceSynchronize')
      IMPORT :: c_int
      integer(c_int) :: hipDeviceSynchronize
    END FUNCTION
  END INTERFACE
  integer(c_int), parameter :: H2D = 1, D2H = 2
  TYPE(param_Op) :: para_Op
  TYPE(sg4_prim_basis), allocatable :: prim(:,:)
  integer(c_int32_t) :: hdr(9), nq, nb
  integer(c_int32_t), allocatable :: tab_nDval(:,:), tab_iB(:)
  real(c_double), allocatable :: WeightSG(:), V(:), ck(:), G(:,:,:), Jac(:), sqRJ(:), GridV(:,:,:)
  TYPE(param_psi), allocatable :: tpsi(:), thpsi(:)
  integer :: nrow, it
  real(c_double), allocatable, target :: psi(:,:), hpsi_host(:,:), hpsi_dev(:,:), host_blk(:,:)
  integer :: D, Lmax, nterms, nb_ba, nvec, type_op, k, l, u, stat_refused
  integer(c_int64_t) :: sum_nb, sum_nq
  integer(c_size_t) :: nbytes
  type(c_ptr) :: psi_d, hpsi_d
  character(len=512) :: fin, fout

  CALL get_command_argument(1, fin)
  CALL get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  D = hdr(1); Lmax = hdr(2); nterms = hdr(3); nb_ba = hdr(4)
  sum_nb = hdr(5); sum_nq = hdr(6); nvec = hdr(7); type_op = hdr(8)
  allocate(tab_nDval(D, nterms), WeightSG(nterms), tab_iB(sum_nb), prim(0:Lmax, D))
  read(u) tab_nDval, WeightSG, tab_iB
  DO k = 1, D
    DO l = 0, Lmax
      read(u) nq, nb
      allocate(prim(l,k)%d0b(nq, nb), prim(l,k)%d0bwrho(nb, nq), prim(l,k)%d1gg(nq, nq), &
               prim(l,k)%d2gg(nq, nq), prim(l,k)%x(nq), prim(l,k)%wrho(nq))
      read(u) prim(l,k)%d0b, prim(l,k)%d0bwrho, prim(l,k)%d1gg, prim(l,k)%d2gg, &
              prim(l,k)%x, prim(l,k)%wrho
    END DO
  END DO
  allocate(V(sum_nq))
  nrow = nb_ba
  IF (type_op == 1) THEN
    allocate(ck(D))
    read(u) V, ck
  ELSE IF (type_op == 21) THEN
    allocate(GridV(sum_nq, 2, 2), ck(D))
    read(u) GridV, ck
    nrow = 2 * nb_ba
  ELSE
    allocate(G(D, D, sum_nq), Jac(sum_nq), sqRJ(sum_nq))
    read(u) V, G, Jac, sqRJ
  END IF
  allocate(psi(nrow, nvec), hpsi_host(nrow, nvec), hpsi_dev(nrow, nvec))
  read(u) psi
  close(u)

  IF (type_op == 21) THEN
    CALL sg4_gpu_init_BasisnD(para_Op, tab_nDval, WeightSG, tab_iB, prim, nb_ba, int(hdr(9)), nb0=2)
    CALL sg4_gpu_set_op_type1_nb0(para_Op, GridV, ck)
    CALL sg4_gpu_TabOpPsi(para_Op, psi, hpsi_host)
    allocate(tpsi(nvec), thpsi(nvec))
    DO it = 1, nvec
      CALL alloc_psi(tpsi(it), nrow)
      tpsi(it)%RvecB(:) = psi(:, it)
    END DO
    CALL sub_TabOpPsi_FOR_SGtype4(tpsi, thpsi, para_Op)
    DO it = 1, nvec
      hpsi_dev(:, it) = thpsi(it)%RvecB(1:nrow)
    END DO
    nbytes = int(8, c_size_t) * int(nrow, c_size_t) * int(nvec, c_size_t)
    IF (hipMalloc(psi_d, nbytes) /= 0 .OR. hipMalloc(hpsi_d, nbytes) /= 0) STOP 'hipMalloc failed'
    CALL sg4_gpu_TabOpPsi_dev(para_Op, nvec, psi_d, hpsi_d, stat=stat_refused)
    IF (hipFree(psi_d) /= 0 .OR. hipFree(hpsi_d) /= 0) STOP 'hipFree failed'
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) hpsi_host, hpsi_dev, int(stat_refused, c_int32_t)
    close(u)
    CALL dealloc_para_Op_SG4(para_Op)
    write(*,*) 'test_sg4_reflayout: two surfaces, nb_ba=', nb_ba, ' nvec=', nvec, &
               ' device entry status=', stat_refused
    STOP
  END IF

  CALL sg4_gpu_init_BasisnD(para_Op, tab_nDval, WeightSG, tab_iB, prim, nb_ba, int(hdr(9)))
  IF (type_op == 1) THEN
    CALL sg4_gpu_set_op_type1(para_Op, V, ck)
  ELSE
    CALL sg4_gpu_set_op_type10(para_Op, V, G, Jac, sqRJ)
  END IF

  ! host arrays, max_nvec columns per library call
  hpsi_host = 0.0_c_double
  CALL sg4_gpu_TabOpPsi(para_Op, psi, hpsi_host)

  ! device blocks
  nbytes = int(8, c_size_t) * int(nb_ba, c_size_t) * int(nvec, c_size_t)
  IF (hipMalloc(psi_d, nbytes) /= 0 .OR. hipMalloc(hpsi_d, nbytes) /= 0) STOP 'hipMalloc failed'
  IF (hipMemcpy(psi_d, c_loc(psi), nbytes, H2D) /= 0) STOP 'hipMemcpy failed'
  CALL sg4_gpu_TabOpPsi_dev(para_Op, nvec, psi_d, hpsi_d)
  IF (hipDeviceSynchronize() /= 0) STOP 'hipDeviceSynchronize failed'
  IF (hipMemcpy(c_loc(hpsi_dev), hpsi_d, nbytes, D2H) /= 0) STOP 'hipMemcpy failed'

  ! a host array given to the device entry is refused before anything runs
  allocate(host_blk(nb_ba, nvec))
  host_blk = 0.0_c_double
  CALL sg4_gpu_TabOpPsi_dev(para_Op, nvec, c_loc(host_blk), hpsi_d, stat=stat_refused)
  IF (hipFree(psi_d) /= 0 .OR. hipFree(hpsi_d) /= 0) STOP 'hipFree failed'

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) hpsi_host, hpsi_dev, int(stat_refused, c_int32_t)
  close(u)
  CALL dealloc_para_Op_SG4(para_Op)
  write(*,*) 'test_sg4_reflayout: done nb_ba=', nb_ba, ' nvec=', nvec, ' type=', type_op, &
             ' refused status=', stat_refused
END PROGRAM test_sg4_reflayout
