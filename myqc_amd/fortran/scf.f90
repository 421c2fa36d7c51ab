!-----------------------------------------------------------------------
! scf -- RHF or UHF on the GPU, in process.  Program contract of
! src/scf/scf.f90 (RHF :61-216, UHF :221-404): reads envdat/nucpos/fmem,
! basinfo, Huv, Suv and Cui when present (restart hook, :148-156 and
! :314-320; else the core-Hamiltonian guess, :410-512 and :518-622); writes
! Cui, Da and eig (UHF: two matrices / records / lines, alpha first).  The
! reference spawns dens and RHFI2G/UHFI2G every iteration (:877-878,
! :1091-1092) and re-reads XX from disk; here the screened unique integrals
! stay on the device and D, G, F never leave it.  The density criterion
! follows SCF_Conv (options(8), :1186-1258), with DIIS and |dE| < 1e-10 Eh.
! The working matrices are taken from the fmem ledger while held (:114-129,
! :210-211; UHF :272-292, :398-399), and the MOLDEN file is written as
! makeMOLDEN writes it (:194, :372, :1343-1466).
!-----------------------------------------------------------------------
PROGRAM scf
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:), S(:,:), H(:,:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem, enr, dtol, t0, t1, mb
  INTEGER :: nnuc, nelcA, nelcB, n, rc, k
  INTEGER(c_long_long) :: nst
  TYPE(c_ptr) :: ctx

  CALL CPU_TIME(t0)
  WRITE(*,*) ""
  WRITE(*,*) "                      STARTING SCF"
  WRITE(*,*) "------------------------------------------------------------"
  CALL read_env(nnuc, nelcA, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  IF (options(3) /= 0 .AND. options(3) /= 1) CALL host_fail("scf: only the RHF and UHF references run on this path")
  IF (options(3) == 0 .AND. nelcA /= nelcB) CALL host_fail("You cannot use RHF for open shell molecules!")
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  n = basinfo(1)
  IF (options(3) == 0) THEN
    mb = 6.0D0 * n * n * 8.0D0 / 1.0D6 + n * 8.0D0 / 1.0D6
    CALL ledger_take(fmem, mb, "rhf", 2)
  ELSE
    mb = 10.0D0 * n * n * 8.0D0 / 1.0D6 + 2.0D0 * n * 8.0D0 / 1.0D6
    CALL ledger_take(fmem, mb, "uhf", 2)
  END IF
  ALLOCATE(S(n,n), H(n,n))
  CALL read_text_matrix('Huv', H)
  CALL read_text_matrix('Suv', S)
  enr = nuclear_repulsion(xyz, atoms)
  WRITE(*,*) "Nuclear repulsion energy (hartrees) : ", enr
  CALL open_system(ctx, atoms, xyz, bas, basinfo, setv, setinfo)
  rc = myqc_eri_build(ctx, 1.0D-12, nst)
  IF (rc /= 0) CALL host_fail("myqc_eri_build: " // TRIM(myqc_error_text(ctx)))
  k = MIN(MAX(options(8), 0), 11)
  dtol = 10.0D0**(-k)
  IF (options(3) == 0) THEN
    CALL run_rhf()
  ELSE
    CALL run_uhf()
  END IF
  CALL ledger_give(fmem, mb)
  CALL CPU_TIME(t1)
  IF (options(3) == 0) THEN
    WRITE(*,'(1x,A16,F8.5)') "rhf ran in (s) :", t1 - t0
  ELSE
    WRITE(*,'(1x,A16,F8.5)') "uhf ran in (s) :", t1 - t0
  END IF

CONTAINS

  SUBROUTINE run_rhf()
    REAL(8), ALLOCATABLE :: C(:,:), D(:,:), eig(:)
    REAL(c_double), ALLOCATABLE, TARGET :: Cg(:,:)
    REAL(8) :: e
    INTEGER(c_int) :: iters
    INTEGER :: u
    LOGICAL :: ex
    TYPE(c_ptr) :: guess
    ALLOCATE(C(n,n), D(n,n), eig(n), Cg(n,n))
    guess = C_NULL_PTR
    INQUIRE(file='Cui', EXIST=ex)
    IF (ex) THEN
      WRITE(*,*) "Reading MO coef from Cui"
      CALL read_text_matrix('Cui', Cg)
      guess = C_LOC(Cg)
    END IF
    rc = myqc_rhf_scf(ctx, H, S, INT(nelcA, c_int), enr, guess, 1.0D-10, dtol, 500_c_int, MYQC_FP64, &
                      e, C, eig, iters)
    IF (rc /= 0) CALL host_fail("myqc_rhf_scf: " // TRIM(myqc_error_text(ctx)))
    rc = myqc_density_rhf(ctx, C, INT(nelcA, c_int), D)
    rc = myqc_finalize(ctx)
    WRITE(*,*) "SCF has converged!", iters, "iterations"
    WRITE(*,*)
    WRITE(*,*) "Total SCF energy (a.u)", e
    WRITE(*,*)
    CALL print_levels("Orbital eigenvalues (a.u.)", eig, nelcA)
    CALL write_molden(options(2), atoms, xyz, eig, C, .FALSE., eig, C)
    CALL write_text_matrix('Cui', C)
    OPEN(newunit=u, file='Da', status='replace', form='unformatted')
    WRITE(u) D
    CLOSE(u)
    OPEN(newunit=u, file='eig', status='replace')
    WRITE(u,*) eig
    CLOSE(u)
  END SUBROUTINE run_rhf

  SUBROUTINE run_uhf()
    REAL(8), ALLOCATABLE :: Ca(:,:), Cb(:,:), Da(:,:), Db(:,:), eigA(:), eigB(:)
    REAL(c_double), ALLOCATABLE, TARGET :: Cga(:,:), Cgb(:,:)
    REAL(8) :: e
    INTEGER(c_int) :: iters
    INTEGER :: u
    LOGICAL :: ex
    TYPE(c_ptr) :: ga, gb
    ALLOCATE(Ca(n,n), Cb(n,n), Da(n,n), Db(n,n), eigA(n), eigB(n), Cga(n,n), Cgb(n,n))
    ga = C_NULL_PTR
    gb = C_NULL_PTR
    INQUIRE(file='Cui', EXIST=ex)
    IF (ex) THEN
      WRITE(*,*) "Reading MO coef from Cui"
      CALL read_text_matrix2('Cui', Cga, Cgb)
      ga = C_LOC(Cga)
      gb = C_LOC(Cgb)
    END IF
    ! UHF gives up after 300 iterations (src/scf/scf.f90:336-339)
    rc = myqc_uhf_scf(ctx, H, S, INT(nelcA, c_int), INT(nelcB, c_int), enr, ga, gb, 1.0D-10, dtol, 300_c_int, &
                      e, Ca, Cb, eigA, eigB, iters)
    IF (rc /= 0) CALL host_fail("myqc_uhf_scf: " // TRIM(myqc_error_text(ctx)))
    rc = myqc_density_uhf(ctx, Ca, Cb, INT(nelcA, c_int), INT(nelcB, c_int), Da, Db)
    rc = myqc_finalize(ctx)
    WRITE(*,*) "SCF has converged!", iters, "iterations"
    WRITE(*,*)
    WRITE(*,*) "Total SCF energy (a.u)", e
    WRITE(*,*)
    CALL print_levels("Orbital eigenvalues of Alpha (a.u.)", eigA, nelcA)
    WRITE(*,*)
    CALL print_levels("Orbital eigenvalues of Beta (a.u.)", eigB, nelcB)
    CALL write_molden(options(2), atoms, xyz, eigA, Ca, .TRUE., eigB, Cb)
    CALL write_text_matrix2('Cui', Ca, Cb)
    OPEN(newunit=u, file='Da', status='replace', form='unformatted')
    WRITE(u) Da
    WRITE(u) Db
    CLOSE(u)
    OPEN(newunit=u, file='eig', status='replace')
    WRITE(u,*) eigA
    WRITE(u,*) eigB
    CLOSE(u)
  END SUBROUTINE run_uhf

  ! the level table of src/scf/scf.f90:178-190 and :345-381
  SUBROUTINE print_levels(title, eig, nocc)
    CHARACTER(LEN=*), INTENT(IN) :: title
    REAL(8), INTENT(IN) :: eig(:)
    INTEGER, INTENT(IN) :: nocc
    INTEGER :: i
    WRITE(*,*) "========================================================================="
    WRITE(*,*) title
    WRITE(*,*) "========================================================================="
    DO i = 1, SIZE(eig)
      IF (i == nocc + 1) WRITE(*,*) "-------------------------------------------------------------------------"
      WRITE(*,'(15x,I3,4x,F20.15)') i, eig(i)
    END DO
    WRITE(*,*) "========================================================================="
  END SUBROUTINE print_levels
END PROGRAM scf
