! Reads wave-packet files with the Fortran runtime's own list-directed,
! namelist and unformatted sequential I/O, statement for statement as
! ElVibRot writes them, and dumps what it read (stream, little-endian) for
! the Python side to compare:
!   read_wp_files v1f  D cplx in out  -- version 1, formatted
!       read(u,nml=headerFile); nb_tot x read(u,*) ind(1:D+2);
!       nb_psi x read(u,*) T, vec(1:nb_tot)
!       (Write_header_saveFile_psi / Write_list_nDindBasis / Write_Psi_nDBasis
!        case 1, SRC/sub_WP/sub_module_psi_io.f90:1312-1477,1606-1617)
!   read_wp_files v1u  D cplx in out  -- version 1, unformatted sequential:
!       'X', nb_psi, nb_tot, version, 'X', nb_tot index records, nb_psi
!       records T, vec (Read_header_saveFile_psi :1241-1248)
!   read_wp_files rst  0 1    in out  -- propagation restart:
!       read(u,*) T, nb_WP, nb_tot; nb_WP x read(u,*) psi(1:nb_tot)
!       (ReadWP_restart, SRC/sub_propagation/sub_module_propagation.f90:429-495)
! out: int32 nb_psi, nb_tot, version; int32 ind(D+2, nb_tot); then per vector
!      real64 T and the nb_tot values (complex: re, im pairs).
! The reverse direction writes such files with the writer's own statements
! from the same dump layout (in: the dump, out: the file):
!   read_wp_files w1f  D cplx in out  -- write(u,nml=headerFile);
!       write(u,*) ind(:,i); write(u,*) T, vec(:)
!   read_wp_files wrst 0 1    in out  -- write(u,*) T, nb_WP, nb_tot;
!       write(u,*) psi(:)  (SaveWP_restart, sub_module_propagation.f90:379-427)
PROGRAM read_wp_files
  USE, INTRINSIC :: iso_fortran_env, only: real64, int32
  IMPLICIT NONE
  character(len=16) :: mode, arg
  character(len=1024) :: fin, fout
  character(len=1) :: ch
  integer :: D, icplx, u, o, i, j, ios
  integer :: Version_File, nb_psi, nb_tot
  integer, allocatable :: ind(:,:)
  real(real64) :: T
  real(real64), allocatable :: rv(:)
  complex(real64), allocatable :: cv(:)
  namelist / headerFile / Version_File, nb_psi, nb_tot

  CALL get_command_argument(1, mode)
  CALL get_command_argument(2, arg); read(arg, *) D
  CALL get_command_argument(3, arg); read(arg, *) icplx
  CALL get_command_argument(4, fin)
  CALL get_command_argument(5, fout)
  open(newunit=o, file=trim(fout), access='stream', form='unformatted', status='replace')

  SELECT CASE (trim(mode))
  CASE ('v1f')
    open(newunit=u, file=trim(fin), form='formatted', status='old')
    Version_File = -1; nb_psi = 0; nb_tot = 0
    read(u, nml=headerFile, iostat=ios)
    IF (ios /= 0) STOP 'namelist headerFile not read'
    allocate(ind(D+2, nb_tot))
    DO i = 1, nb_tot
      read(u, *) ind(:, i)
    END DO
    write(o) int(nb_psi, int32), int(nb_tot, int32), int(Version_File, int32), int(ind, int32)
    CALL read_vectors(.TRUE.)
  CASE ('v1u')
    open(newunit=u, file=trim(fin), form='unformatted', access='sequential', status='old')
    read(u) ch
    IF (ch /= 'X') STOP 'not a version-1 unformatted file'
    read(u) nb_psi
    read(u) nb_tot
    read(u) Version_File
    read(u) ch
    allocate(ind(D+2, nb_tot))
    DO i = 1, nb_tot
      read(u) ind(:, i)
    END DO
    write(o) int(nb_psi, int32), int(nb_tot, int32), int(Version_File, int32), int(ind, int32)
    CALL read_vectors(.FALSE.)
  CASE ('rst')
    open(newunit=u, file=trim(fin), form='formatted', status='old')
    read(u, *) T, nb_psi, nb_tot
    write(o) int(nb_psi, int32), int(nb_tot, int32), 0_int32
    allocate(cv(nb_tot))
    DO j = 1, nb_psi
      read(u, *) cv
      write(o) T, cv
    END DO
  CASE ('w1f', 'wrst')
    close(o)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) nb_psi, nb_tot, Version_File
    open(newunit=o, file=trim(fout), form='formatted', status='replace')
    IF (trim(mode) == 'w1f') THEN
      allocate(ind(D+2, nb_tot))
      read(u) ind
      write(o, nml=headerFile)
      DO i = 1, nb_tot
        write(o, *) ind(:, i)
      END DO
    END IF
    IF (icplx /= 0) THEN
      allocate(cv(nb_tot))
    ELSE
      allocate(rv(nb_tot))
    END IF
    DO j = 1, nb_psi
      IF (icplx /= 0) THEN
        read(u) T, cv
        IF (trim(mode) == 'w1f') THEN
          write(o, *) T, cv
        ELSE
          IF (j == 1) write(o, *) T, nb_psi, nb_tot
          write(o, *) cv
        END IF
      ELSE
        read(u) T, rv
        write(o, *) T, rv
      END IF
    END DO
  CASE DEFAULT
    STOP 'mode: v1f, v1u, rst, w1f or wrst'
  END SELECT
  close(u)
  close(o)

CONTAINS

  SUBROUTINE read_vectors(formatted)
    logical, intent(in) :: formatted
    IF (icplx /= 0) THEN
      allocate(cv(nb_tot))
    ELSE
      allocate(rv(nb_tot))
    END IF
    DO j = 1, nb_psi
      IF (icplx /= 0) THEN
        IF (formatted) THEN
          read(u, *) T, cv
        ELSE
          read(u) T, cv
        END IF
        write(o) T, cv
      ELSE
        IF (formatted) THEN
          read(u, *) T, rv
        ELSE
          read(u) T, rv
        END IF
        write(o) T, rv
      END IF
    END DO
  END SUBROUTINE read_vectors

END PROGRAM read_wp_files
