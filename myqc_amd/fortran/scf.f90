!-----------------------------------------------------------------------
! scf -- RHF on the GPU, in process.  Program contract of src/scf/scf.f90
! (RHF, :61-216): reads envdat/nucpos/fmem, basinfo, Huv, Suv and Cui when
! present (restart hook, :148-156; else the core-Hamiltonian guess,
! :410-512); writes Cui, Da and eig.  The reference spawns dens and RHFI2G
! every iteration (:877-878) and re-reads XX from disk; here the screened
! unique integrals stay on the device and D, G, F never leave it.  The
! density criterion follows SCF_Conv (options(8), :1186-1258), with DIIS and
! |dE| < 1e-10 Eh.
!-----------------------------------------------------------------------
PROGRAM scf
  USE iso_c_binding
  USE myqc_hip_api
  USE myqc_host
  IMPLICIT NONE
  REAL(8), ALLOCATABLE :: xyz(:,:), bas(:), setv(:), S(:,:), H(:,:), C(:,:), D(:,:), eig(:)
  REAL(c_double), ALLOCATABLE, TARGET :: Cg(:,:)
  INTEGER, ALLOCATABLE :: atoms(:), options(:), basinfo(:), setinfo(:)
  REAL(8) :: fmem, enr, e, dtol, t0, t1
  INTEGER :: nnuc, nelcA, nelcB, n, rc, u, i, k
  INTEGER(c_int) :: iters
  INTEGER(c_long_long) :: nst
  LOGICAL :: ex
  TYPE(c_ptr) :: ctx, guess

  CALL CPU_TIME(t0)
  WRITE(*,*) ""
  WRITE(*,*) "                      STARTING SCF"
  WRITE(*,*) "------------------------------------------------------------"
  CALL read_env(nnuc, nelcA, nelcB, xyz, atoms, fmem, options)
  CALL check_error_sentinel()
  IF (options(3) /= 0) CALL host_fail("scf: only the RHF reference runs on this path")
  IF (nelcA /= nelcB) CALL host_fail("You cannot use RHF for open shell molecules!")
  CALL build_basis(options(2), atoms, bas, basinfo, setv, setinfo)
  n = basinfo(1)
  ALLOCATE(S(n,n), H(n,n), C(n,n), D(n,n), eig(n), Cg(n,n))
  CALL read_text_matrix('Huv', H)
  CALL read_text_matrix('Suv', S)
  enr = nuclear_repulsion(xyz, atoms)
  WRITE(*,*) "Nuclear repulsion energy (hartrees) : ", enr
  guess = C_NULL_PTR
  INQUIRE(file='Cui', EXIST=ex)
  IF (ex) THEN
    WRITE(*,*) "Reading MO coef from Cui"
    CALL read_text_matrix('Cui', Cg)
    guess = C_LOC(Cg)
  END IF
  CALL open_system(ctx, atoms, xyz, bas, basinfo, setv, setinfo)
  rc = myqc_eri_build(ctx, 1.0D-12, nst)
  IF (rc /= 0) CALL host_fail("myqc_eri_build: " // TRIM(myqc_error_text(ctx)))
  k = MIN(MAX(options(8), 0), 11)
  dtol = 10.0D0**(-k)
  rc = myqc_rhf_scf(ctx, H, S, INT(nelcA, c_int), enr, guess, 1.0D-10, dtol, 500_c_int, MYQC_FP64, &
                    e, C, eig, iters)
  IF (rc /= 0) CALL host_fail("myqc_rhf_scf: " // TRIM(myqc_error_text(ctx)))
  rc = myqc_density_rhf(ctx, C, INT(nelcA, c_int), D)
  rc = myqc_finalize(ctx)
  WRITE(*,*) "SCF has converged!", iters, "iterations"
  WRITE(*,*)
  WRITE(*,*) "Total SCF energy (a.u)", e
  WRITE(*,*)
  WRITE(*,*) "Orbital eigenvalues (a.u.)"
  WRITE(*,*) "========================================================================="
  DO i = 1, n
    IF (i == nelcA + 1) WRITE(*,*) "-------------------------------------------------------------------------"
    WRITE(*,'(15x,I3,4x,F20.15)') i, eig(i)
  END DO
  WRITE(*,*) "========================================================================="
  CALL write_text_matrix('Cui', C)
  OPEN(newunit=u, file='Da', status='replace', form='unformatted')
  WRITE(u) D
  CLOSE(u)
  OPEN(newunit=u, file='eig', status='replace')
  WRITE(u,*) eig
  CLOSE(u)
  CALL CPU_TIME(t1)
  WRITE(*,'(1x,A16,F8.5)') "rhf ran in (s) :", t1 - t0
END PROGRAM scf
