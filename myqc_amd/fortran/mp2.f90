!-----------------------------------------------------------------------
! mp2 -- RHF-MP2 energy on the GPU.  Program contract of src/mp2/mp2.f90
! (mp2_rhf, :79-144): reads envdat, basinfo, eig and the ijab_AB records;
! prints E(AA), E(AB) and E(MBPT2) = 2 E(AA) + E(AB).
!-----------------------------------------------------------------------
PROGRAM mp2
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:), eig(:), rec(:,:,:,:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem, eaa, eab, etot
  INTEGER :: nnuc, nocc, nelcB, n, nv, rc, u, i, j, ios
  TYPE(c_ptr) :: ctx

  WRITE(*,*)
  WRITE(*,*) "mp2 called"
  CALL read_env(nnuc, nocc, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  n = basinfo(1)
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
END PROGRAM mp2
