!-----------------------------------------------------------------------
! int2e -- the dense two-electron integral file XX on the GPU.  Program
! contract of src/integrals/int2e.f90: reads envdat/nucpos/fmem, mybasis,
! Ftab; writes basinfo/setinfo and XX, one unformatted record of
! XX(i,j,g,h) = (ij|gh), n^4 column-major, 8-fold filled (int2e.f90:166,306);
! skips if XX exists (int2e.f90:58-67).  Every shell quartet is computed
! (the reference has no Schwarz screen).  The n^4 doubles are taken from the
! fmem ledger while held (int2e.f90:169-180,343-345).  A record above 2 GiB
! (n >= 128) is written as gfortran subrecords, the layout of the
! reference's own gfortran build (src/Makefile:11); MYQC_XX_RECORD=single
! writes it as one record instead (flang's runtime, whose 4-byte length
! markers then wrap modulo 2^32).
!-----------------------------------------------------------------------
PROGRAM int2e
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE, TARGET :: XX(:,:,:,:)
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem, t0, t1, mb
  INTEGER :: nnuc, nelcA, nelcB, n, rc, u, st
  LOGICAL :: ex
  CHARACTER(LEN=16) :: layout
  TYPE(c_ptr) :: ctx

  CALL CPU_TIME(t0)
  WRITE(*,*)
  WRITE(*,*) "int2e called"
  CALL read_env(nnuc, nelcA, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  INQUIRE(file='XX', EXIST=ex)
  IF (ex) THEN
    WRITE(*,*) "Reading two electron integrals from intermediate"
    STOP
  END IF
  WRITE(*,*) "Constructing two electron integrals"
  n = basinfo(1)
  mb = REAL(n, 8)**4 * 8.0D0 / 1.0D6
  CALL ledger_take(fmem, mb, "the two-electron integrals", 0)
  ALLOCATE(XX(0:n-1,0:n-1,0:n-1,0:n-1), STAT=st)
  IF (st /= 0) CALL host_fail("Could not allocate memory in int2e")
  CALL open_system(ctx, atoms, xyz, bas, basinfo, setv, setinfo)
  rc = myqc_eri_full(ctx, 0.0_c_double, XX)
  IF (rc /= 0) CALL host_fail("myqc_eri_full: " // TRIM(myqc_error_text(ctx)))
  rc = myqc_finalize(ctx)
  CALL get_environment_variable('MYQC_XX_RECORD', layout)
  IF (TRIM(layout) == 'single') THEN
    OPEN(newunit=u, file='XX', status='replace', access='sequential', form='unformatted')
    WRITE(u) XX
    CLOSE(u)
  ELSE
    CALL write_record_subrecords('XX', C_LOC(XX), 8_8 * INT(n, 8)**4, 2147483639_8)
  END IF
  DEALLOCATE(XX)
  CALL ledger_give(fmem, mb)
  CALL CPU_TIME(t1)
  WRITE(*,'(1x,A41,2x,F8.4)') "Two electron integrals constructed in (s)", t1 - t0
END PROGRAM int2e
