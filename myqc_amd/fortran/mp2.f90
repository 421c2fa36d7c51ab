!-----------------------------------------------------------------------
! mp2 -- MP2 energy on the GPU.  Program contract of src/mp2/mp2.f90:
!   RHF (mp2_rhf, :79-144): reads envdat, basinfo, eig and the ijab_AB
!   records; prints E(AA), E(AB) and E(MBPT2) = 2 E(AA) + E(AB).
!   UHF (mp2_uhf, :154-232): reads eig (alpha line, beta line) and the
!   ijab_AA, ijab_BB and ijab_AB records; prints E(AA), E(BB), E(AB) and
!   E(MBPT2) = E(AA) + E(BB) + E(AB).
!-----------------------------------------------------------------------
PROGRAM mp2
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem
  INTEGER :: nnuc, nelcA, nelcB, n, rc
  TYPE(c_ptr) :: ctx

  WRITE(*,*)
  WRITE(*,*) "mp2 called"
  CALL read_env(nnuc, nelcA, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  IF (options(3) /= 0 .AND. options(3) /= 1) CALL host_fail("Sorry, that spin case has not been coded for MP2 yet")
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  n = basinfo(1)
  IF (options(3) == 0) THEN
    CALL mp2_rhf(nelcA)
  ELSE
    CALL mp2_uhf(nelcA, nelcB)
  END IF

CONTAINS

  SUBROUTINE mp2_rhf(nocc)
    INTEGER, INTENT(IN) :: nocc
    REAL(8), ALLOCATABLE :: eig(:), rec(:,:,:,:)
    REAL(8) :: eaa, eab, etot
    INTEGER :: nv, u, i, j, ios
    nv = n - nocc
    ALLOCATE(eig(n), rec(0:nv-1,0:nv-1,0:nocc-1,0:nocc-1))
    OPEN(newunit=u, file='eig', status='old', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open eig")
    READ(u,*) eig
    CLOSE(u)
    OPEN(newunit=u, file='ijab_AB', status='old', form='unformatted', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open ijab_AB")
    DO i = 0, nocc-1
      DO j = 0, nocc-1
        READ(u) rec(:,:,j,i)
      END DO
    END DO
    CLOSE(u)
    rc = myqc_init(0_c_int, ctx)
    IF (rc /= 0) CALL host_fail("myqc_init: no usable gfx950 device")
    rc = myqc_mp2_rhf_records(ctx, rec, eig, INT(nocc, c_int), INT(nv, c_int), eaa, eab, etot)
    IF (rc /= 0) CALL host_fail("myqc_mp2_rhf_records: " // TRIM(myqc_error_text(ctx)))
    rc = myqc_finalize(ctx)
    WRITE(*,*) "E(AA) = ", eaa
    WRITE(*,*) "E(AB) = ", eab
    WRITE(*,*) "E(MBPT2) = ", etot
  END SUBROUTINE mp2_rhf

  SUBROUTINE mp2_uhf(noccA, noccB)
    INTEGER, INTENT(IN) :: noccA, noccB
    REAL(8), ALLOCATABLE :: eigA(:), eigB(:), raa(:,:,:), rbb(:,:,:), rab(:,:,:)
    REAL(8) :: eaa, ebb, eab, etot
    INTEGER :: nvA, nvB, u, ios
    nvA = n - noccA
    nvB = n - noccB
    ALLOCATE(eigA(n), eigB(n), raa(0:nvA-1,0:nvA-1,noccA*(noccA-1)/2), rbb(0:nvB-1,0:nvB-1,noccB*(noccB-1)/2), &
             rab(0:nvA-1,0:nvB-1,noccA*noccB))
    OPEN(newunit=u, file='eig', status='old', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open eig")
    READ(u,*) eigA
    READ(u,*) eigB
    CLOSE(u)
    CALL read_records('ijab_AA', raa)
    CALL read_records('ijab_BB', rbb)
    CALL read_records('ijab_AB', rab)
    rc = myqc_init(0_c_int, ctx)
    IF (rc /= 0) CALL host_fail("myqc_init: no usable gfx950 device")
    rc = myqc_mp2_uhf_records(ctx, raa, rbb, rab, eigA, eigB, INT(noccA, c_int), INT(noccB, c_int), &
                              INT(nvA, c_int), INT(nvB, c_int), eaa, ebb, eab, etot)
    IF (rc /= 0) CALL host_fail("myqc_mp2_uhf_records: " // TRIM(myqc_error_text(ctx)))
    rc = myqc_finalize(ctx)
    WRITE(*,*) "E(AA) = ", eaa
    WRITE(*,*) "E(BB) = ", ebb
    WRITE(*,*) "E(AB) = ", eab
    WRITE(*,*) "E(MBPT2) = ", etot
  END SUBROUTINE mp2_uhf

  SUBROUTINE read_records(fname, r)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(8), INTENT(OUT) :: r(:,:,:)
    INTEGER :: u, p, ios
    OPEN(newunit=u, file=fname, status='old', form='unformatted', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open " // fname)
    DO p = 1, SIZE(r, 3)
      READ(u) r(:,:,p)
    END DO
    CLOSE(u)
  END SUBROUTINE read_records
END PROGRAM mp2
