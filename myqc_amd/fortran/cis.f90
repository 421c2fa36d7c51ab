!-----------------------------------------------------------------------
! cis -- UHF-CIS on the GPU.  Program contract of src/CIS/cis.f90
! (cis_scf_uhf, :87-223): CALC=SCF with a UHF reference only (:57-69); reads
! envdat, basinfo, eig (alpha line, beta line) and the five streams of the
! transform (ajib_AA, ajbi_AA, ajib_AB, ajib_BB, ajbi_BB); writes ints_AA and
! ints_BB = <aj||ib> = ajib - ajbi (cis_scf_uhf_proc, :238-304); prints the
! lowest E_NUM (options(15)) eigenvalues and the largest components of their
! eigenvectors (cis_uhf_print / cis_uhf_vec, :319-441).  The matrix is built
! and diagonalised on the device.
!-----------------------------------------------------------------------
PROGRAM cis
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:), eigA(:), eigB(:), E(:), V(:,:)
  REAL(8), ALLOCATABLE :: s1(:,:), s2(:,:), s3(:,:), s4(:,:), s5(:,:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem, t0, t1
  INTEGER :: nnuc, noccA, noccB, n, nvA, nvB, upA, upB, neig, rc, u, ios, i
  TYPE(c_ptr) :: ctx

  CALL CPU_TIME(t0)
  WRITE(*,*)
  WRITE(*,*) "                      STARTING CIS"
  WRITE(*,*) "-----------------------------------------------------------"
  WRITE(*,*) "cis called"
  CALL read_env(nnuc, noccA, noccB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  IF (options(1) /= 0) CALL host_fail("Sorry, only SCF reference CIS has been coded.")
  IF (options(3) /= 1) CALL host_fail("Sorry, only UHF references have been coded.")
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  n = basinfo(1)
  nvA = n - noccA
  nvB = n - noccB
  upA = noccA * nvA
  upB = noccB * nvB
  ALLOCATE(s1(upA,upA), s2(upA,upA), s3(upA,upB), s4(upB,upB), s5(upB,upB), eigA(n), eigB(n))
  CALL read_columns('ajib_AA', s1)
  CALL read_columns('ajbi_AA', s2)
  CALL read_columns('ajib_AB', s3)
  CALL read_columns('ajib_BB', s4)
  CALL read_columns('ajbi_BB', s5)
  WRITE(*,*) "Pre-proccessing 2e- inegrals in MO basis"
  CALL write_columns('ints_AA', s1 - s2)
  CALL write_columns('ints_BB', s4 - s5)
  OPEN(newunit=u, file='eig', status='old', iostat=ios)
  IF (ios /= 0) CALL host_fail("Couldn't open eig")
  READ(u,*) eigA
  READ(u,*) eigB
  CLOSE(u)
  neig = MAX(MIN(options(15), upA + upB), 0)
  ALLOCATE(E(MAX(neig,1)), V(upA + upB, MAX(neig,1)))
  WRITE(*,*) "Constructing A matrix"
  WRITE(*,*) "Diagonalizing A matrix"
  rc = myqc_init(0_c_int, ctx)
  IF (rc /= 0) CALL host_fail("myqc_init: no usable gfx950 device")
  rc = myqc_cis_uhf(ctx, s1, s2, s3, s4, s5, eigA, eigB, INT(noccA, c_int), INT(noccB, c_int), INT(nvA, c_int), &
                    INT(nvB, c_int), INT(neig, c_int), E, V)
  IF (rc /= 0) CALL host_fail("myqc_cis_uhf: " // TRIM(myqc_error_text(ctx)))
  rc = myqc_finalize(ctx)
  WRITE(*,*) "========================================================================="
  WRITE(*,*) "                          CIS Results"
  WRITE(*,*) "========================================================================="
  DO i = 1, neig
    WRITE(*,*)
    WRITE(*,*) "CIS eigenpair #", i
    WRITE(*,*) "Eigenvalue (a.u.)", E(i)
    CALL print_vector(V(:,i))
    WRITE(*,*) "------------------------------------------------"
  END DO
  CALL CPU_TIME(t1)
  WRITE(*,'(1x,A22,2x,F8.4)') "CIS completed in (s): ", t1 - t0

CONTAINS

  SUBROUTINE read_columns(fname, r)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(8), INTENT(OUT) :: r(:,:)
    INTEGER :: uu, p, io
    OPEN(newunit=uu, file=fname, status='old', form='unformatted', iostat=io)
    IF (io /= 0) CALL host_fail("Couldn't open " // fname)
    DO p = 1, SIZE(r, 2)
      READ(uu) r(:,p)
    END DO
    CLOSE(uu)
  END SUBROUTINE read_columns

  SUBROUTINE write_columns(fname, r)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(8), INTENT(IN) :: r(:,:)
    INTEGER :: uu, p
    OPEN(newunit=uu, file=fname, status='replace', form='unformatted')
    DO p = 1, SIZE(r, 2)
      WRITE(uu) r(:,p)
    END DO
    CLOSE(uu)
  END SUBROUTINE write_columns

  ! the (up to) ten largest components of one eigenvector, largest first, as
  ! the table of cis_uhf_vec (cis.f90:354-441): i a j b (1-based orbital
  ! numbers, alpha pair then beta pair), the signed coefficient and the spin
  SUBROUTINE print_vector(x)
    REAL(8), INTENT(IN) :: x(:)
    LOGICAL, ALLOCATABLE :: used(:)
    INTEGER :: k, p, best, ii, aa
    ALLOCATE(used(SIZE(x)))
    used = .FALSE.
    WRITE(*,*) "     i   a   j   b      val             spin"
    WRITE(*,*) "------------------------------------------------"
    DO k = 1, MIN(SIZE(x), 10)
      best = 0
      DO p = 1, SIZE(x)
        IF (used(p)) CYCLE
        IF (best == 0) THEN
          best = p
        ELSE IF (ABS(x(p)) > ABS(x(best))) THEN
          best = p
        END IF
      END DO
      used(best) = .TRUE.
      IF (ABS(x(best)) <= 1.0D-10) EXIT
      IF (best <= upA) THEN
        ii = (best - 1) / nvA
        aa = MOD(best - 1, nvA)
        WRITE(*,'(4x,I3,1x,I3,1x,I3,1x,I3,4x,F15.10,4x,A2)') ii + 1, noccA + aa + 1, 0, 0, x(best), 'AA'
      ELSE
        ii = (best - upA - 1) / nvB
        aa = MOD(best - upA - 1, nvB)
        WRITE(*,'(4x,I3,1x,I3,1x,I3,1x,I3,4x,F15.10,4x,A2)') 0, 0, ii + 1, noccB + aa + 1, x(best), 'BB'
      END IF
    END DO
  END SUBROUTINE print_vector
END PROGRAM cis
