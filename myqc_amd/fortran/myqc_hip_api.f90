!-----------------------------------------------------------------------
! myqc_hip_api -- iso_c_binding interfaces to libmyqc_hip (the C ABI of
! include/myqc_hip.h).  Arrays are passed as the Fortran host holds them
! (column-major, assumed size); the context is an opaque c_ptr.
!-----------------------------------------------------------------------
MODULE myqc_hip_api
  USE iso_c_binding
  IMPLICIT NONE

  INTEGER(c_int), PARAMETER :: MYQC_FP64 = 0, MYQC_FP32 = 1, MYQC_BF16X3 = 2
  INTEGER(c_int), PARAMETER :: MYQC_REF_COMPAT = 0

  INTERFACE
    INTEGER(c_int) FUNCTION myqc_init(device, ctx) BIND(C, name='myqc_init')
      IMPORT :: c_int, c_ptr
      INTEGER(c_int), VALUE :: device
      TYPE(c_ptr) :: ctx
    END FUNCTION myqc_init

    INTEGER(c_int) FUNCTION myqc_finalize(ctx) BIND(C, name='myqc_finalize')
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx
    END FUNCTION myqc_finalize

    TYPE(c_ptr) FUNCTION myqc_last_error(ctx) BIND(C, name='myqc_last_error')
      IMPORT :: c_ptr
      TYPE(c_ptr), VALUE :: ctx
    END FUNCTION myqc_last_error

    INTEGER(c_int) FUNCTION myqc_set_system(ctx, nnuc, Z, xyz, nset, alpha, setinfo, bas, norb, basinfo, &
                                            Ftab, flags) BIND(C, name='myqc_set_system')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int), VALUE :: nnuc, nset, norb, flags
      INTEGER(c_int) :: Z(*), setinfo(*), basinfo(*)
      REAL(c_double) :: xyz(*), alpha(*), bas(*), Ftab(*)
    END FUNCTION myqc_set_system

    INTEGER(c_int) FUNCTION myqc_eri_full(ctx, tol, XX) BIND(C, name='myqc_eri_full')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), VALUE :: tol
      REAL(c_double) :: XX(*)
    END FUNCTION myqc_eri_full

    INTEGER(c_int) FUNCTION myqc_eri_build(ctx, tol, nstored) BIND(C, name='myqc_eri_build')
      IMPORT :: c_int, c_ptr, c_double, c_long_long
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), VALUE :: tol
      INTEGER(c_long_long) :: nstored
    END FUNCTION myqc_eri_build

    INTEGER(c_int) FUNCTION myqc_fock_G(ctx, D, G, precision) BIND(C, name='myqc_fock_G')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: D(*), G(*)
      INTEGER(c_int), VALUE :: precision
    END FUNCTION myqc_fock_G

    INTEGER(c_int) FUNCTION myqc_fock_G_uhf(ctx, Da, Db, Ga, Gb) BIND(C, name='myqc_fock_G_uhf')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: Da(*), Db(*), Ga(*), Gb(*)
    END FUNCTION myqc_fock_G_uhf

    INTEGER(c_int) FUNCTION myqc_density_rhf(ctx, C, nocc, D) BIND(C, name='myqc_density_rhf')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: C(*), D(*)
      INTEGER(c_int), VALUE :: nocc
    END FUNCTION myqc_density_rhf

    INTEGER(c_int) FUNCTION myqc_ao2mo_ovov(ctx, C, nocc, iajb, precision) BIND(C, name='myqc_ao2mo_ovov')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: C(*), iajb(*)
      INTEGER(c_int), VALUE :: nocc, precision
    END FUNCTION myqc_ao2mo_ovov

    INTEGER(c_int) FUNCTION myqc_mp2_rhf_records(ctx, rec, eig, nocc, nvrt, e_aa, e_ab, e_tot) &
        BIND(C, name='myqc_mp2_rhf_records')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: rec(*), eig(*)
      INTEGER(c_int), VALUE :: nocc, nvrt
      REAL(c_double) :: e_aa, e_ab, e_tot
    END FUNCTION myqc_mp2_rhf_records

    INTEGER(c_int) FUNCTION myqc_int1e(ctx, S, H) BIND(C, name='myqc_int1e')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: S(*), H(*)
    END FUNCTION myqc_int1e

    INTEGER(c_int) FUNCTION myqc_rhf_scf(ctx, H, S, nocc, e_nuc, C_guess, e_tol, d_tol, max_iter, precision, &
                                         energy, C, eig, iters) BIND(C, name='myqc_rhf_scf')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: H(*), S(*)
      INTEGER(c_int), VALUE :: nocc, max_iter, precision
      REAL(c_double), VALUE :: e_nuc, e_tol, d_tol
      TYPE(c_ptr), VALUE :: C_guess
      REAL(c_double) :: energy, C(*), eig(*)
      INTEGER(c_int) :: iters
    END FUNCTION myqc_rhf_scf

    INTEGER(c_int) FUNCTION myqc_density_uhf(ctx, Ca, Cb, nalpha, nbeta, Da, Db) BIND(C, name='myqc_density_uhf')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: Ca(*), Cb(*), Da(*), Db(*)
      INTEGER(c_int), VALUE :: nalpha, nbeta
    END FUNCTION myqc_density_uhf

    INTEGER(c_int) FUNCTION myqc_uhf_scf(ctx, H, S, nalpha, nbeta, e_nuc, Ca_guess, Cb_guess, e_tol, d_tol, &
                                         max_iter, energy, Ca, Cb, eig_a, eig_b, iters) BIND(C, name='myqc_uhf_scf')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: H(*), S(*)
      INTEGER(c_int), VALUE :: nalpha, nbeta, max_iter
      REAL(c_double), VALUE :: e_nuc, e_tol, d_tol
      TYPE(c_ptr), VALUE :: Ca_guess, Cb_guess
      REAL(c_double) :: energy, Ca(*), Cb(*), eig_a(*), eig_b(*)
      INTEGER(c_int) :: iters
    END FUNCTION myqc_uhf_scf

    INTEGER(c_int) FUNCTION myqc_ao2mo_ovov_uhf(ctx, Ca, Cb, nalpha, nbeta, rec_aa, rec_bb, rec_ab, precision) &
        BIND(C, name='myqc_ao2mo_ovov_uhf')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: Ca(*), Cb(*), rec_aa(*), rec_bb(*), rec_ab(*)
      INTEGER(c_int), VALUE :: nalpha, nbeta, precision
    END FUNCTION myqc_ao2mo_ovov_uhf

    INTEGER(c_int) FUNCTION myqc_mp2_uhf_records(ctx, rec_aa, rec_bb, rec_ab, eig_a, eig_b, nalpha, nbeta, &
                                                 nvrt_a, nvrt_b, e_aa, e_bb, e_ab, e_tot) &
        BIND(C, name='myqc_mp2_uhf_records')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: rec_aa(*), rec_bb(*), rec_ab(*), eig_a(*), eig_b(*)
      INTEGER(c_int), VALUE :: nalpha, nbeta, nvrt_a, nvrt_b
      REAL(c_double) :: e_aa, e_bb, e_ab, e_tot
    END FUNCTION myqc_mp2_uhf_records

    INTEGER(c_int) FUNCTION myqc_ao2mo_cis_uhf(ctx, Ca, Cb, nalpha, nbeta, ajib_aa, ajbi_aa, ajib_ab, ajib_bb, &
                                               ajbi_bb) BIND(C, name='myqc_ao2mo_cis_uhf')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: Ca(*), Cb(*), ajib_aa(*), ajbi_aa(*), ajib_ab(*), ajib_bb(*), ajbi_bb(*)
      INTEGER(c_int), VALUE :: nalpha, nbeta
    END FUNCTION myqc_ao2mo_cis_uhf

    INTEGER(c_int) FUNCTION myqc_cis_uhf(ctx, ajib_aa, ajbi_aa, ajib_ab, ajib_bb, ajbi_bb, eig_a, eig_b, nalpha, &
                                         nbeta, nvrt_a, nvrt_b, nroots, energies, vectors) BIND(C, name='myqc_cis_uhf')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: ajib_aa(*), ajbi_aa(*), ajib_ab(*), ajib_bb(*), ajbi_bb(*), eig_a(*), eig_b(*)
      INTEGER(c_int), VALUE :: nalpha, nbeta, nvrt_a, nvrt_b, nroots
      REAL(c_double) :: energies(*), vectors(*)
    END FUNCTION myqc_cis_uhf

    INTEGER(c_int) FUNCTION myqc_int1p(ctx, P) BIND(C, name='myqc_int1p')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: P(*)
    END FUNCTION myqc_int1p

    INTEGER(c_int) FUNCTION myqc_dipole_rhf(ctx, D, mu) BIND(C, name='myqc_dipole_rhf')
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double) :: D(*), mu(3)
    END FUNCTION myqc_dipole_rhf
  END INTERFACE

CONTAINS

  ! text of the context's last error as a Fortran string
  FUNCTION myqc_error_text(ctx) RESULT(msg)
    TYPE(c_ptr), INTENT(IN) :: ctx
    CHARACTER(LEN=512) :: msg
    CHARACTER(KIND=c_char), POINTER :: p(:)
    TYPE(c_ptr) :: cp
    INTEGER :: i
    msg = ' '
    cp = myqc_last_error(ctx)
    IF (.NOT. C_ASSOCIATED(cp)) RETURN
    CALL C_F_POINTER(cp, p, [512])
    DO i = 1, 512
      IF (p(i) == C_NULL_CHAR) EXIT
      msg(i:i) = p(i)
    END DO
  END FUNCTION myqc_error_text

END MODULE myqc_hip_api
