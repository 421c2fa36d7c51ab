!-----------------------------------------------------------------------
! subrec_check -- TEST INFRASTRUCTURE (host only, no device call): writes
! records through myqc_host's write_record_subrecords with a small subrecord
! limit, so a test can check the gfortran subrecord layout byte by byte.
!   subrec_check <file> <nbytes> <maxsub>
! The payload is bytes 0, 1, 2, ... (mod 256).
!-----------------------------------------------------------------------
PROGRAM subrec_check
  USE iso_c_binding
  USE myqc_host
  IMPLICIT NONE
  CHARACTER(LEN=512) :: fname, arg
  INTEGER(8) :: nbytes, maxsub, i, k
  INTEGER(1), ALLOCATABLE, TARGET :: payload(:)
  CALL get_command_argument(1, fname)
  CALL get_command_argument(2, arg)
  READ(arg,*) nbytes
  CALL get_command_argument(3, arg)
  READ(arg,*) maxsub
  ALLOCATE(payload(MAX(nbytes, 1_8)))
  DO i = 1, nbytes
    k = MOD(i - 1, 256_8)
    IF (k > 127) k = k - 256
    payload(i) = INT(k, 1)
  END DO
  CALL write_record_subrecords(TRIM(fname), C_LOC(payload), nbytes, maxsub)
END PROGRAM subrec_check
