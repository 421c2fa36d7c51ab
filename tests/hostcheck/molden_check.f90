!-----------------------------------------------------------------------
! molden_check -- TEST INFRASTRUCTURE (host only, no device call): calls
! myqc_host's write_molden on orbitals read from a text file, so a test can
! compare the MOLDEN text with the reference's own files.  Reads `molden_in`:
!   nnuc n uhf(0/1) / Z x y z per atom (bohr) / eigA / CA (n x n, column
!   major) / [eigB / CB]
! and writes MOLDEN in the working directory (mybasis must be there).
!-----------------------------------------------------------------------
PROGRAM molden_check
  USE myqc_host
  IMPLICIT NONE
  INTEGER :: nnuc, n, uhf, i, u
  INTEGER, ALLOCATABLE :: atoms(:)
  REAL(8), ALLOCATABLE :: xyz(:,:), eA(:), CA(:,:), eB(:), CB(:,:)
  OPEN(newunit=u, file='molden_in', status='old')
  READ(u,*) nnuc, n, uhf
  ALLOCATE(atoms(0:nnuc-1), xyz(0:nnuc-1,0:2), eA(n), CA(n,n), eB(n), CB(n,n))
  DO i = 0, nnuc-1
    READ(u,*) atoms(i), xyz(i,0), xyz(i,1), xyz(i,2)
  END DO
  READ(u,*) eA
  READ(u,*) CA
  IF (uhf == 1) THEN
    READ(u,*) eB
    READ(u,*) CB
  ELSE
    eB = eA
    CB = CA
  END IF
  CLOSE(u)
  CALL write_molden(0, atoms, xyz, eA, CA, uhf == 1, eB, CB)
END PROGRAM molden_check
