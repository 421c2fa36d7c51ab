!-----------------------------------------------------------------------
! ao2mo -- (ia|jb) for RHF-MP2 on the GPU.  Program contract of
! src/ao2mo/ao2mo.f90 (slow_ao2mo_MP2_RHF, :465-602): reads envdat/nucpos/
! fmem, basinfo and Cui; writes ijab_AA (one nvrt x nvrt record per i<j) and
! ijab_AB (one record per (i,j), i-major), a fastest (:566-582).  The AO
! integrals are recomputed on the device instead of read from XX.
!-----------------------------------------------------------------------
PROGRAM ao2mo
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:), C(:,:), rec(:,:,:,:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem
  INTEGER :: nnuc, nocc, nelcB, n, nv, rc, u, i, j, st
  INTEGER(c_long_long) :: nst
  TYPE(c_ptr) :: ctx

  WRITE(*,*)
  WRITE(*,*) "                STARTING AO TRANSFORM"
  CALL read_env(nnuc, nocc, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  IF (options(1) /= 1 .OR. options(3) /= 0) CALL host_fail("ao2mo: only the RHF-MP2 transform runs on this path")
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  n = basinfo(1)
  nv = n - nocc
  ALLOCATE(C(n,n), rec(0:nv-1,0:nv-1,0:nocc-1,0:nocc-1), STAT=st)
  IF (st /= 0) CALL host_fail("ao2mo: could not allocate memory")
  CALL read_text_matrix('Cui', C)
  CALL open_system(ctx, atoms, xyz, bas, basinfo, setv, setinfo)
  rc = myqc_eri_build(ctx, 0.0_c_double, nst)
  IF (rc /= 0) CALL host_fail("myqc_eri_build: " // TRIM(myqc_error_text(ctx)))
  ! rec(a,b,j,i) = (ia|jb): record (i,j) in i-major order, a fastest
  rc = myqc_ao2mo_ovov(ctx, C, INT(nocc, c_int), rec, MYQC_FP64)
  IF (rc /= 0) CALL host_fail("myqc_ao2mo_ovov: " // TRIM(myqc_error_text(ctx)))
  rc = myqc_finalize(ctx)
  WRITE(*,*) "Writing to ijab_AA"
  OPEN(newunit=u, file='ijab_AA', status='replace', form='unformatted')
  DO i = 0, nocc-2
    DO j = i+1, nocc-1
      WRITE(u) rec(:,:,j,i)
    END DO
  END DO
  CLOSE(u)
  WRITE(*,*) "Writing to ijab_AB"
  OPEN(newunit=u, file='ijab_AB', status='replace', form='unformatted')
  DO i = 0, nocc-1
    DO j = 0, nocc-1
      WRITE(u) rec(:,:,j,i)
    END DO
  END DO
  CLOSE(u)
END PROGRAM ao2mo
