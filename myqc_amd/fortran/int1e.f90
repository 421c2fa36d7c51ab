!-----------------------------------------------------------------------
! int1e -- overlap Suv and core Hamiltonian Huv (text, column-major) on the
! GPU.  Program contract of src/integrals/int1e.f90: reads envdat/nucpos/
! fmem, mybasis, Ftab; writes basinfo/setinfo, Suv, Huv; skips if Suv and
! Huv exist (int1e.f90:98-111).  S and H are taken from the fmem ledger
! while held.
!-----------------------------------------------------------------------
PROGRAM int1e
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:), S(:,:), H(:,:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem
  INTEGER :: nnuc, nelcA, nelcB, n, rc
  LOGICAL :: ex1, ex2
  TYPE(c_ptr) :: ctx

  WRITE(*,*)
  WRITE(*,*) "int1e called"
  CALL read_env(nnuc, nelcA, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  n = basinfo(1)
  ! S and H from the fmem ledger (int1e.f90:81-96), before the skip check
  CALL ledger_take(fmem, 2.0D0 * n * n * 8.0D0 / 1.0D6, "int1e", 1)
  INQUIRE(file='Suv', EXIST=ex1)
  INQUIRE(file='Huv', EXIST=ex2)
  IF (ex1 .AND. ex2) THEN
    WRITE(*,*) "Reading one electron integrals from intermediate"
    CALL ledger_give(fmem, 2.0D0 * n * n * 8.0D0 / 1.0D6)
    STOP
  END IF
  ALLOCATE(S(n,n), H(n,n))
  CALL open_system(ctx, atoms, xyz, bas, basinfo, setv, setinfo)
  rc = myqc_int1e(ctx, S, H)
  IF (rc /= 0) CALL host_fail("myqc_int1e: " // TRIM(myqc_error_text(ctx)))
  rc = myqc_finalize(ctx)
  CALL write_text_matrix('Suv', S)
  CALL write_text_matrix('Huv', H)
  WRITE(*,*) "Overlap and core Hamiltonian written"
  CALL ledger_give(fmem, 2.0D0 * n * n * 8.0D0 / 1.0D6)
END PROGRAM int1e
