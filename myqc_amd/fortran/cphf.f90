!-----------------------------------------------------------------------
! cphf -- the SCF dipole moment on the GPU.  Program contract of
! src/props/cphf.f90 (cphf_dipole_rhf, :80-131): RHF only (:53-57); reads
! envdat/nucpos/fmem, mybasis and the total density Da; prints
! mu = -sum D(u,v) P(v,u) + sum_A Z_A R_A (atomic units) in the reference's
! table.  The dipole integrals are recomputed on the device rather than read
! from Pxuv.
!-----------------------------------------------------------------------
PROGRAM cphf
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:), D(:,:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem, mu(3)
  INTEGER :: nnuc, nelcA, nelcB, n, rc, u, ios
  TYPE(c_ptr) :: ctx

  WRITE(*,*)
  WRITE(*,*) "cphf called"
  WRITE(*,*)
  CALL read_env(nnuc, nelcA, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  IF (options(3) /= 0) THEN
    WRITE(*,*) "Sorry, only RHF CPHF equations are coded"
    STOP
  END IF
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  n = basinfo(1)
  ALLOCATE(D(n,n))
  OPEN(newunit=u, file='Da', status='old', form='unformatted', iostat=ios)
  IF (ios /= 0) CALL host_fail("Couldn't open Da")
  READ(u) D
  CLOSE(u)
  CALL open_system(ctx, atoms, xyz, bas, basinfo, setv, setinfo)
  rc = myqc_dipole_rhf(ctx, D, mu)
  IF (rc /= 0) CALL host_fail("myqc_dipole_rhf: " // TRIM(myqc_error_text(ctx)))
  rc = myqc_finalize(ctx)
  WRITE(*,*) "                Dipole Moments                "
  WRITE(*,*) "----------------------------------------------"
  WRITE(*,*) "       X              Y              Z        "
  WRITE(*,'(F15.6,F15.6,F15.6)') mu
  WRITE(*,*) "----------------------------------------------"
END PROGRAM cphf
