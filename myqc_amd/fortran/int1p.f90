!-----------------------------------------------------------------------
! int1p -- dipole integrals on the GPU.  Program contract of
! src/integrals/int1p.f90: reads envdat/nucpos/fmem, mybasis; writes Pxuv
! (the x, y and z matrices, one list-directed record each, :230-233); skips
! if Pxuv exists (:90-97).
!-----------------------------------------------------------------------
PROGRAM int1p
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:), P(:,:,:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem
  INTEGER :: nnuc, nelcA, nelcB, n, rc, u, c
  LOGICAL :: ex
  TYPE(c_ptr) :: ctx

  WRITE(*,*) "int1p called"
  CALL read_env(nnuc, nelcA, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  INQUIRE(file='Pxuv', EXIST=ex)
  IF (ex) THEN
    WRITE(*,*) "Reading Dipole integrals from Pxuv"
    STOP
  END IF
  n = basinfo(1)
  ALLOCATE(P(n,n,3))
  CALL open_system(ctx, atoms, xyz, bas, basinfo, setv, setinfo)
  rc = myqc_int1p(ctx, P)
  IF (rc /= 0) CALL host_fail("myqc_int1p: " // TRIM(myqc_error_text(ctx)))
  rc = myqc_finalize(ctx)
  OPEN(newunit=u, file='Pxuv', status='replace')
  DO c = 1, 3
    WRITE(u,*) P(:,:,c)
  END DO
  CLOSE(u)
  WRITE(*,*) "Dipole moment integrals written to Pxuv"
END PROGRAM int1p
