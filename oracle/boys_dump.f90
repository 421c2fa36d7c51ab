! Golden-vector harness: calls the reference's own Boys routine (module
! auxilary, compiled from src/integrals/auxilary.f90 into oracle/_ref/) on a
! grid of T for Q in {0,3,6,9,12} and prints F_0..F_Q with full precision.
! Reads the Boys table from ./Ftab like int2e does (int2e.f90:161-163).
PROGRAM boys_dump
  USE auxilary
  IMPLICIT NONE
  REAL(KIND=8), DIMENSION(0:120,0:22) :: Ft
  REAL(KIND=8), DIMENSION(0:16) :: Fj
  REAL(KIND=8) :: T
  INTEGER :: iq, Q, k, j
  OPEN(unit=1,file='Ftab',status='old',access='sequential',form='unformatted')
  READ(1) Ft
  CLOSE(unit=1)
  DO iq=0,4
    Q = 3*iq
    DO k=0,599
      ! T grid: 0..29.95 in steps of 0.05, plus off-grid points
      T = k*0.05D0 + 0.0123D0*MOD(k,3)
      Fj = 0.0D0
      CALL Boys(Fj(0:Q),Q,T,Ft)
      WRITE(*,'(I3,1X,ES25.17E3,17(1X,ES25.17E3))') Q, T, (Fj(j), j=0,Q)
    END DO
    ! Boys3 range, T >= 2Q+36
    DO k=0,7
      T = 2*Q + 36 + k*7.3D0
      Fj = 0.0D0
      CALL Boys(Fj(0:Q),Q,T,Ft)
      WRITE(*,'(I3,1X,ES25.17E3,17(1X,ES25.17E3))') Q, T, (Fj(j), j=0,Q)
    END DO
  END DO
END PROGRAM boys_dump
