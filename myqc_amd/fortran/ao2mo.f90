!-----------------------------------------------------------------------
! ao2mo -- (ia|jb) for MP2 on the GPU.  Program contract of
! src/ao2mo/ao2mo.f90: reads envdat/nucpos/fmem, basinfo and Cui.
!   RHF (slow_ao2mo_MP2_RHF, :465-602): writes ijab_AA (one nvrt x nvrt
!   record per i<j) and ijab_AB (one record per (i,j), i-major), a fastest
!   (:566-582).
!   UHF (slow_ao2mo_MP2_UHF, :614-904): Cui holds the alpha and beta
!   coefficients; writes ijab_AA and ijab_BB (records i<j of each spin) and
!   ijab_AB (records (i alpha, j beta), nvrtA x nvrtB), a fastest
!   (:718-724, :804-810, :890-896).
!   UHF-CIS (EXCITE=CIS with CALC=SCF; slow_ao2mo_CIS_UHF, :919-1227):
!   writes ajib_AA, ajbi_AA, ajib_AB, ajib_BB, ajbi_BB, one record per (j,b),
!   each over (i,a) with a fastest (dispatch :62-97).
! The AO integrals are recomputed on the device instead of read from XX.
!-----------------------------------------------------------------------
PROGRAM ao2mo
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem
  INTEGER :: nnuc, nelcA, nelcB, n, rc
  INTEGER(c_long_long) :: nst
  TYPE(c_ptr) :: ctx

  WRITE(*,*)
  WRITE(*,*) "                STARTING AO TRANSFORM"
  CALL read_env(nnuc, nelcA, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  IF (options(1) == 1) THEN
    IF (options(3) /= 0 .AND. options(3) /= 1) CALL host_fail("Sorry, that reference not coded yet")
  ELSE IF (options(13) == 1 .AND. options(1) == 0) THEN
    IF (options(3) /= 1) CALL host_fail("Sorry, only UHF CIS is coded")
  ELSE
    CALL host_fail("Sorry, that transform type has not been coded yet")
  END IF
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  n = basinfo(1)
  CALL open_system(ctx, atoms, xyz, bas, basinfo, setv, setinfo)
  rc = myqc_eri_build(ctx, 0.0_c_double, nst)
  IF (rc /= 0) CALL host_fail("myqc_eri_build: " // TRIM(myqc_error_text(ctx)))
  IF (options(1) == 1 .AND. options(3) == 0) THEN
    CALL transform_rhf(nelcA)
  ELSE IF (options(1) == 1) THEN
    CALL transform_uhf(nelcA, nelcB)
  ELSE
    CALL transform_cis(nelcA, nelcB)
  END IF

CONTAINS

  SUBROUTINE transform_rhf(nocc)
    INTEGER, INTENT(IN) :: nocc
    REAL(8), ALLOCATABLE :: C(:,:), rec(:,:,:,:)
    INTEGER :: nv, st, u, i, j
    nv = n - nocc
    ALLOCATE(C(n,n), rec(0:nv-1,0:nv-1,0:nocc-1,0:nocc-1), STAT=st)
    IF (st /= 0) CALL host_fail("ao2mo: could not allocate memory")
    CALL read_text_matrix('Cui', C)
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
  END SUBROUTINE transform_rhf

  SUBROUTINE transform_uhf(noccA, noccB)
    INTEGER, INTENT(IN) :: noccA, noccB
    REAL(8), ALLOCATABLE :: Ca(:,:), Cb(:,:), raa(:,:,:), rbb(:,:,:), rab(:,:,:)
    INTEGER :: nvA, nvB, st
    nvA = n - noccA
    nvB = n - noccB
    ! the record streams, one record per slice of the last index
    ALLOCATE(Ca(n,n), Cb(n,n), raa(0:nvA-1,0:nvA-1,noccA*(noccA-1)/2), rbb(0:nvB-1,0:nvB-1,noccB*(noccB-1)/2), &
             rab(0:nvA-1,0:nvB-1,noccA*noccB), STAT=st)
    IF (st /= 0) CALL host_fail("ao2mo: could not allocate memory")
    CALL read_text_matrix2('Cui', Ca, Cb)
    rc = myqc_ao2mo_ovov_uhf(ctx, Ca, Cb, INT(noccA, c_int), INT(noccB, c_int), raa, rbb, rab, MYQC_FP64)
    IF (rc /= 0) CALL host_fail("myqc_ao2mo_ovov_uhf: " // TRIM(myqc_error_text(ctx)))
    rc = myqc_finalize(ctx)
    CALL write_records('ijab_AA', raa)
    CALL write_records('ijab_BB', rbb)
    CALL write_records('ijab_AB', rab)
  END SUBROUTINE transform_uhf

  SUBROUTINE transform_cis(noccA, noccB)
    INTEGER, INTENT(IN) :: noccA, noccB
    REAL(8), ALLOCATABLE :: Ca(:,:), Cb(:,:), s1(:,:), s2(:,:), s3(:,:), s4(:,:), s5(:,:)
    INTEGER :: upA, upB, st
    upA = noccA * (n - noccA)
    upB = noccB * (n - noccB)
    ! each stream as (elements, records): record p is column p
    ALLOCATE(Ca(n,n), Cb(n,n), s1(upA,upA), s2(upA,upA), s3(upA,upB), s4(upB,upB), s5(upB,upB), STAT=st)
    IF (st /= 0) CALL host_fail("ao2mo: could not allocate memory")
    CALL read_text_matrix2('Cui', Ca, Cb)
    rc = myqc_ao2mo_cis_uhf(ctx, Ca, Cb, INT(noccA, c_int), INT(noccB, c_int), s1, s2, s3, s4, s5)
    IF (rc /= 0) CALL host_fail("myqc_ao2mo_cis_uhf: " // TRIM(myqc_error_text(ctx)))
    rc = myqc_finalize(ctx)
    CALL write_columns('ajib_AA', s1)
    CALL write_columns('ajbi_AA', s2)
    CALL write_columns('ajib_AB', s3)
    CALL write_columns('ajib_BB', s4)
    CALL write_columns('ajbi_BB', s5)
  END SUBROUTINE transform_cis

  SUBROUTINE write_columns(fname, r)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(8), INTENT(IN) :: r(:,:)
    INTEGER :: u, p
    WRITE(*,*) "Writing to " // fname
    OPEN(newunit=u, file=fname, status='replace', form='unformatted')
    DO p = 1, SIZE(r, 2)
      WRITE(u) r(:,p)
    END DO
    CLOSE(u)
  END SUBROUTINE write_columns

  SUBROUTINE write_records(fname, r)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(8), INTENT(IN) :: r(:,:,:)
    INTEGER :: u, p
    WRITE(*,*) "Writing to " // fname
    OPEN(newunit=u, file=fname, status='replace', form='unformatted')
    DO p = 1, SIZE(r, 3)
      WRITE(u) r(:,:,p)
    END DO
    CLOSE(u)
  END SUBROUTINE write_records
END PROGRAM ao2mo
