! Fortran host driving the GPU SG4 action through mod_sg4_gpu: builds the 6D
! Henon-Heiles operator (L=3), applies it to a block of deterministic vectors
! and writes psi and H psi (binary, stream access) for the GPU test to compare
! with the oracle.
PROGRAM test_sg4_fortran
  USE, INTRINSIC :: iso_c_binding
  USE mod_sg4_gpu
  IMPLICIT NONE
  TYPE(param_Op) :: para_Op
  TYPE(param_psi), allocatable :: Psi(:), HPsi(:)
  integer :: nvec, i, v, u
  character(len=512) :: fout
  nvec = 3
  CALL get_command_argument(1, fout)
  CALL init_para_Op_SG4_HenonHeiles(para_Op, 6, 3, 0.111803_c_double, 2)
  allocate(Psi(nvec), HPsi(nvec))
  DO v = 1, nvec
    CALL alloc_psi(Psi(v), para_Op%nb_ba)
    DO i = 1, para_Op%nb_ba
      Psi(v)%RvecB(i) = sin(0.37_c_double * i + 1.3_c_double * v) / (1.0_c_double + 0.01_c_double * i)
    END DO
  END DO
  ! 3 vectors with max_nvec = 2 exercises the block loop
  CALL sub_TabOpPsi_FOR_SGtype4(Psi, HPsi, para_Op)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) int(para_Op%nb_ba, c_int32_t), int(nvec, c_int32_t)
  DO v = 1, nvec
    write(u) Psi(v)%RvecB
  END DO
  DO v = 1, nvec
    write(u) HPsi(v)%RvecB
  END DO
  close(u)
  CALL dealloc_para_Op_SG4(para_Op)
  write(*,*) 'test_sg4_fortran: done nb_ba=', para_Op%nb_ba
END PROGRAM test_sg4_fortran
