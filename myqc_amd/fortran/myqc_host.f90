!-----------------------------------------------------------------------
! myqc_host -- host-side contracts of the myQC programs, kept so the
! programs here are drop-ins for int1e/int2e/scf/ao2mo/mp2:
!   * envdat / nucpos / fmem readers (layout of src/myQC/env.f90:16-73);
!   * the STO-3G basis arrays from `mybasis` with buildBasis's index layout
!     (src/myQC/basis.f90:23-226), also written to basinfo/setinfo;
!   * the Boys table `Ftab` (one unformatted record, int2e.f90:161-163);
!   * text matrices (Suv/Huv/Cui/eig) and unformatted records (XX, ijab_*);
!   * the `error` sentinel + STOP on failure (env.f90:36-41).
!-----------------------------------------------------------------------
MODULE myqc_host
  USE iso_c_binding
  USE myqc_hip_api
  IMPLICIT NONE

CONTAINS

  SUBROUTINE host_fail(msg)
    CHARACTER(LEN=*), INTENT(IN) :: msg
    INTEGER :: u
    WRITE(*,*) TRIM(msg)
    OPEN(newunit=u, file='error', status='replace')
    CLOSE(u)
    STOP 1
  END SUBROUTINE host_fail

  SUBROUTINE check_error_sentinel()
    LOGICAL :: ex
    INQUIRE(file='error', EXIST=ex)
    IF (ex) STOP
  END SUBROUTINE check_error_sentinel

  SUBROUTINE read_env(nnuc, nelcA, nelcB, xyz, atoms, fmem, options)
    INTEGER, INTENT(OUT) :: nnuc, nelcA, nelcB
    REAL(8), ALLOCATABLE, INTENT(OUT) :: xyz(:,:)          ! (0:nnuc-1, 0:2) bohr
    INTEGER, ALLOCATABLE, INTENT(OUT) :: atoms(:), options(:)
    REAL(8), INTENT(OUT) :: fmem
    INTEGER :: i, j, u1, u2, u3, ios
    OPEN(newunit=u1, file='envdat', status='old', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open envdat")
    OPEN(newunit=u2, file='nucpos', status='old', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open nucpos")
    OPEN(newunit=u3, file='fmem', status='old', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open fmem")
    READ(u3,*) fmem
    READ(u1,*) nnuc
    READ(u1,*) nelcA, nelcB
    READ(u1,*) j
    ALLOCATE(options(0:j-1))
    READ(u1,*) options
    ALLOCATE(atoms(0:nnuc-1), xyz(0:nnuc-1,0:2))
    DO i = 0, nnuc-1
      READ(u2,*) atoms(i), xyz(i,0), xyz(i,1), xyz(i,2)
    END DO
    CLOSE(u1); CLOSE(u2); CLOSE(u3)
  END SUBROUTINE read_env

  ! first blank-separated token of a line
  FUNCTION first_token(line) RESULT(tok)
    CHARACTER(LEN=*), INTENT(IN) :: line
    CHARACTER(LEN=16) :: tok
    INTEGER :: a, b
    tok = ' '
    a = VERIFY(line, ' ' // CHAR(9))
    IF (a == 0) RETURN
    b = SCAN(line(a:), ' ' // CHAR(9))
    IF (b == 0) THEN
      tok = line(a:)
    ELSE
      tok = line(a:a+b-2)
    END IF
  END FUNCTION first_token

  ! The basis arrays in buildBasis's 0-based layouts:
  !   bas(set*OpS+k), setv(set) = exponent,
  !   basinfo = [OpS, norb, {pri, ang, ori, nprim, centre} per orbital],
  !   setinfo = [nset, setl, {norb_in_set, maxL, centre, orb0..} per set].
  SUBROUTINE build_basis(bkey, atoms, bas, basinfo, setv, setinfo)
    INTEGER, INTENT(IN) :: bkey, atoms(0:)
    REAL(8), ALLOCATABLE, INTENT(OUT) :: bas(:), setv(:)
    INTEGER, ALLOCATABLE, INTENT(OUT) :: basinfo(:), setinfo(:)
    CHARACTER(LEN=8), PARAMETER :: bnames(0:3) = ['STO-3G  ', 'tester1 ', 'tester2 ', 'tester3 ']
    CHARACTER(LEN=2), PARAMETER :: anames(0:9) = ['H ', 'He', 'Li', 'Be', 'B ', 'C ', 'N ', 'O ', 'F ', 'Ne']
    CHARACTER(LEN=256), ALLOCATABLE :: lines(:)
    CHARACTER(LEN=256) :: buf
    REAL(8) :: vals(0:15), sid
    INTEGER :: u, ios, nl, i, k, p, start, Smax, Cmax, Omax, almax, OpS, maxN, maxL
    INTEGER :: anum, sec, orb, nset, func, ncoef, pri, ang, ori, setnum, orbnum, setl, s, so, m, j, f
    ! non-blank lines of mybasis
    OPEN(newunit=u, file='mybasis', status='old', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open mybasis")
    nl = 0
    DO
      READ(u, '(A)', iostat=ios) buf
      IF (ios /= 0) EXIT
      IF (LEN_TRIM(buf) > 0) nl = nl + 1
    END DO
    ALLOCATE(lines(nl))
    REWIND(u)
    i = 0
    DO
      READ(u, '(A)', iostat=ios) buf
      IF (ios /= 0) EXIT
      IF (LEN_TRIM(buf) > 0) THEN
        i = i + 1
        lines(i) = buf
      END IF
    END DO
    CLOSE(u)
    start = 0
    DO i = 1, nl
      IF (TRIM(first_token(lines(i))) == TRIM(bnames(bkey))) THEN
        start = i
        EXIT
      END IF
    END DO
    IF (start == 0) CALL host_fail("basis not found in mybasis")
    READ(lines(start+1),*) Smax, Cmax, Omax, almax, OpS
    READ(lines(start+2),*) maxN, maxL
    anum = SIZE(atoms)
    setl = 3 + OpS
    ALLOCATE(bas(0:anum*almax*OpS-1), setv(0:anum*almax-1))
    ALLOCATE(basinfo(0:2+5*Omax*anum-1), setinfo(0:2+anum*almax*setl-1))
    bas = 0.0D0; setv = 0.0D0; basinfo = 0; setinfo = 0
    setnum = 0
    orbnum = 0
    DO i = 0, anum-1
      k = 0
      DO p = start+3, nl
        IF (TRIM(first_token(lines(p))) == TRIM(anames(atoms(i)-1))) THEN
          k = p
          EXIT
        END IF
      END DO
      IF (k == 0) CALL host_fail("atom not found in mybasis")
      READ(lines(k+1),*) sec, orb, nset
      basinfo(0) = OpS
      basinfo(1) = basinfo(1) + orb
      setinfo(0) = setinfo(0) + nset
      setinfo(1) = setl
      p = k + 2
      DO j = 1, sec
        READ(lines(p),*) func, ncoef, pri, ang, ori
        p = p + 1
        DO f = 1, func
          READ(lines(p),*) vals(0:ncoef-1), sid
          p = p + 1
          s = setnum + NINT(sid)
          setv(s) = vals(ncoef-1)
          so = setinfo(2 + s*setl)
          setinfo(2 + s*setl + 2) = i
          IF (ori == -1) THEN
            setinfo(2 + s*setl + 3 + so) = orbnum
            bas(so + s*OpS) = vals(0)
            so = so + 1
          ELSE IF (ori == 2) THEN
            DO m = 0, 2
              setinfo(2 + s*setl + 3 + so) = orbnum + m
              bas(so + s*OpS) = vals(0)
              so = so + 1
            END DO
            IF (setinfo(2 + s*setl + 1) < 1) setinfo(2 + s*setl + 1) = 1
          ELSE
            CALL host_fail("only s and p sections are supported")
          END IF
          setinfo(2 + s*setl) = so
        END DO
        IF (ori == -1) THEN
          basinfo(2+5*orbnum:2+5*orbnum+4) = [pri, ang, ori, func, i]
          orbnum = orbnum + 1
        ELSE
          DO m = 0, 2
            basinfo(2+5*(orbnum+m):2+5*(orbnum+m)+4) = [pri, ang, m, func, i]
          END DO
          orbnum = orbnum + 3
        END IF
      END DO
      setnum = setinfo(0)
    END DO
    OPEN(newunit=u, file='basinfo', status='replace')
    WRITE(u,*) basinfo(:)
    WRITE(u,*)
    WRITE(u,*) "#orbitals per set, #orbitals, {principle qn., angular qn., orientation, #primatives, center number},... "
    CLOSE(u)
    OPEN(newunit=u, file='setinfo', status='replace')
    WRITE(u,*) setinfo(:)
    WRITE(u,*)
    WRITE(u,*) "#sets, length of each set,{#orbitals, max ang qn., center num, [orbital 0, orbital 1, ...]},..."
    CLOSE(u)
  END SUBROUTINE build_basis

  SUBROUTINE read_ftab(Ft)
    REAL(8), INTENT(OUT) :: Ft(0:120,0:22)
    INTEGER :: u, ios
    OPEN(newunit=u, file='Ftab', status='old', access='sequential', form='unformatted', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open Ftab")
    READ(u, iostat=ios) Ft
    IF (ios /= 0) CALL host_fail("Ftab: bad record")
    CLOSE(u)
  END SUBROUTINE read_ftab

  ! context on device 0 with the molecule and basis loaded
  SUBROUTINE open_system(ctx, atoms, xyz, bas, basinfo, setv, setinfo)
    TYPE(c_ptr), INTENT(OUT) :: ctx
    INTEGER, INTENT(IN) :: atoms(0:), basinfo(0:), setinfo(0:)
    REAL(8), INTENT(IN) :: xyz(0:,0:), bas(0:), setv(0:)
    REAL(8) :: Ft(0:120,0:22)
    INTEGER(c_int), ALLOCATABLE :: Zc(:), si(:), bi(:)
    REAL(c_double), ALLOCATABLE :: xc(:)
    INTEGER :: rc, nnuc
    CALL read_ftab(Ft)
    rc = myqc_init(0_c_int, ctx)
    IF (rc /= 0) CALL host_fail("myqc_init: no usable gfx950 device")
    nnuc = SIZE(atoms)
    ALLOCATE(Zc(nnuc), si(SIZE(setinfo)), bi(SIZE(basinfo)), xc(3*nnuc))
    Zc = atoms
    si = setinfo
    bi = basinfo
    xc = RESHAPE(xyz, [3*nnuc])
    rc = myqc_set_system(ctx, INT(nnuc, c_int), Zc, xc, INT(setinfo(0), c_int), setv, si, bas, &
                         INT(basinfo(1), c_int), bi, RESHAPE(Ft, [121*23]), MYQC_REF_COMPAT)
    IF (rc /= 0) CALL host_fail("myqc_set_system: " // TRIM(myqc_error_text(ctx)))
  END SUBROUTINE open_system

  SUBROUTINE read_text_matrix(fname, A)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(8), INTENT(OUT) :: A(:,:)
    INTEGER :: u, ios
    OPEN(newunit=u, file=fname, status='old', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open " // fname)
    READ(u,*) A(:,:)
    CLOSE(u)
  END SUBROUTINE read_text_matrix

  SUBROUTINE write_text_matrix(fname, A)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(8), INTENT(IN) :: A(:,:)
    INTEGER :: u
    OPEN(newunit=u, file=fname, status='replace')
    WRITE(u,*) A(:,:)
    CLOSE(u)
  END SUBROUTINE write_text_matrix

  ! two matrices, one list-directed record each (the UHF Cui file,
  ! src/scf/scf.f90:316-320 read, :1081-1084 write)
  SUBROUTINE read_text_matrix2(fname, A, B)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(8), INTENT(OUT) :: A(:,:), B(:,:)
    INTEGER :: u, ios
    OPEN(newunit=u, file=fname, status='old', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open " // fname)
    READ(u,*) A(:,:)
    READ(u,*) B(:,:)
    CLOSE(u)
  END SUBROUTINE read_text_matrix2

  SUBROUTINE write_text_matrix2(fname, A, B)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(8), INTENT(IN) :: A(:,:), B(:,:)
    INTEGER :: u
    OPEN(newunit=u, file=fname, status='replace')
    WRITE(u,*) A(:,:)
    WRITE(u,*) B(:,:)
    CLOSE(u)
  END SUBROUTINE write_text_matrix2

  REAL(8) FUNCTION nuclear_repulsion(xyz, atoms) RESULT(e)
    REAL(8), INTENT(IN) :: xyz(0:,0:)
    INTEGER, INTENT(IN) :: atoms(0:)
    INTEGER :: a, b
    e = 0.0D0
    DO a = 0, SIZE(atoms)-1
      DO b = 0, a-1
        e = e + atoms(a)*atoms(b) / SQRT((xyz(a,0)-xyz(b,0))**2 + (xyz(a,1)-xyz(b,1))**2 + (xyz(a,2)-xyz(b,2))**2)
      END DO
    END DO
  END FUNCTION nuclear_repulsion

  ! The fmem ledger (src/myQC/env.f90:77-112, nmem): a program takes the
  ! host memory it allocates from the free-memory figure (MB) before it
  ! allocates, fails when the figure goes negative, and gives it back when it
  ! is done; the file is rewritten at each step, as nmem does.
  SUBROUTINE write_fmem(fmem)
    REAL(8), INTENT(IN) :: fmem
    INTEGER :: u, ios
    OPEN(newunit=u, file='fmem', status='replace', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open fmem")
    WRITE(u,*) fmem
    CLOSE(u)
  END SUBROUTINE write_fmem

  ! check: 0 none (int2e only records the figure, int2e.f90:179-180), 1 fail
  ! below zero (int1e.f90:81-85), 2 fail at or below zero (scf.f90:114-133)
  SUBROUTINE ledger_take(fmem, mb, what, check)
    REAL(8), INTENT(INOUT) :: fmem
    REAL(8), INTENT(IN) :: mb
    CHARACTER(LEN=*), INTENT(IN) :: what
    INTEGER, INTENT(IN) :: check
    WRITE(*,'(1x,A,2x,F15.4)') "Allocating memory for " // what // " (MB)", mb
    fmem = fmem - mb
    IF ((check == 1 .AND. fmem < 0.0D0) .OR. (check == 2 .AND. fmem <= 0.0D0)) THEN
      WRITE(*,*) "You need more memory for ", what, ": ", -fmem, " MB short"
      CALL host_fail("not enough memory (MEMORY= / fmem) for " // what)
    END IF
    CALL write_fmem(fmem)
  END SUBROUTINE ledger_take

  SUBROUTINE ledger_give(fmem, mb)
    REAL(8), INTENT(INOUT) :: fmem
    REAL(8), INTENT(IN) :: mb
    fmem = fmem + mb
    CALL write_fmem(fmem)
  END SUBROUTINE ledger_give

  ! One unformatted sequential record of nbytes bytes at `buf`, in gfortran's
  ! layout (the compiler of the reference's own build, src/Makefile:11): a
  ! record longer than maxsub bytes (gfortran: 2147483639 = 2^31 - 9) is
  ! split into subrecords of at most that many bytes, each framed by 4-byte
  ! length markers; a head marker is negative when another subrecord follows,
  ! a tail marker when one precedes.  A record that fits one subrecord is the
  ! plain single record.
  SUBROUTINE write_record_subrecords(fname, buf, nbytes, maxsub)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    TYPE(c_ptr), INTENT(IN) :: buf
    INTEGER(8), INTENT(IN) :: nbytes, maxsub
    INTEGER(1), POINTER :: b(:)
    INTEGER(8) :: pos, ln
    INTEGER(4) :: head, tail
    INTEGER :: u, ios
    CALL c_f_pointer(buf, b, [nbytes])
    OPEN(newunit=u, file=fname, status='replace', access='stream', form='unformatted', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open " // fname)
    pos = 0
    DO
      ln = MIN(nbytes - pos, maxsub)
      head = INT(ln, 4)
      tail = INT(ln, 4)
      IF (pos + ln < nbytes) head = -head
      IF (pos > 0) tail = -tail
      WRITE(u) head
      IF (ln > 0) WRITE(u) b(pos+1:pos+ln)
      WRITE(u) tail
      pos = pos + ln
      IF (pos >= nbytes) EXIT
    END DO
    CLOSE(u)
  END SUBROUTINE write_record_subrecords

  ! The MOLDEN file of makeMOLDEN (src/scf/scf.f90:1343-1466): the atoms in
  ! bohr, the contracted GTOs of each atom as mybasis lists them (exponent,
  ! coefficient per primitive), and the orbitals (energy, spin, occupation
  ! 1.0 below zero energy else 0.0, coefficients); the beta set follows for
  ! UHF (uhf = .TRUE.).  The fixed-format lines use the reference's formats.
  SUBROUTINE write_molden(bkey, atoms, xyz, eigA, CA, uhf, eigB, CB)
    INTEGER, INTENT(IN) :: bkey, atoms(0:)
    REAL(8), INTENT(IN) :: xyz(0:,0:), eigA(:), CA(:,:)
    LOGICAL, INTENT(IN) :: uhf
    REAL(8), INTENT(IN) :: eigB(:), CB(:,:)
    CHARACTER(LEN=2), PARAMETER :: anames(1:10) = ['H ', 'He', 'Li', 'Be', 'B ', 'C ', 'N ', 'O ', 'F ', 'Ne']
    CHARACTER(LEN=8), PARAMETER :: bnames(0:3) = ['STO-3G  ', 'tester1 ', 'tester2 ', 'tester3 ']
    CHARACTER(LEN=1), PARAMETER :: lnames(0:5) = ['s', 'p', 'd', 'f', 'g', 'h']
    CHARACTER(LEN=256), ALLOCATABLE :: lines(:)
    CHARACTER(LEN=256) :: buf
    REAL(8) :: vals(0:15)
    INTEGER :: u, m, ios, nl, i, k, p, start, j, f, sec, orb, nset, func, ncoef, pri, ang, ori, norb
    CHARACTER(LEN=*), PARAMETER :: fatom = '(A2,4x,I3,4x,I3,4x,F15.8,4x,F15.8,4x,F15.8)'
    CHARACTER(LEN=*), PARAMETER :: fene = '(1x,A4,1x,F15.8)', fcoef = '(1x,I4,4x,F15.10)'
    OPEN(newunit=m, file='mybasis', status='old', iostat=ios)
    IF (ios /= 0) CALL host_fail("Couldn't open mybasis")
    nl = 0
    DO
      READ(m, '(A)', iostat=ios) buf
      IF (ios /= 0) EXIT
      IF (LEN_TRIM(buf) > 0) nl = nl + 1
    END DO
    ALLOCATE(lines(nl))
    REWIND(m)
    i = 0
    DO
      READ(m, '(A)', iostat=ios) buf
      IF (ios /= 0) EXIT
      IF (LEN_TRIM(buf) > 0) THEN
        i = i + 1
        lines(i) = buf
      END IF
    END DO
    CLOSE(m)
    start = 0
    DO i = 1, nl
      IF (TRIM(first_token(lines(i))) == TRIM(bnames(bkey))) THEN
        start = i
        EXIT
      END IF
    END DO
    IF (start == 0) CALL host_fail("basis not found in mybasis")
    OPEN(newunit=u, file='MOLDEN', status='replace')
    WRITE(u,*) "[Molden Format]"
    WRITE(u,*) "[ATOMS] AU"
    DO i = 0, SIZE(atoms)-1
      WRITE(u,fatom) anames(atoms(i)), i+1, atoms(i), xyz(i,0:2)
    END DO
    WRITE(u,*) "[Molden Format]"
    WRITE(u,*) "[GTO]"
    DO i = 0, SIZE(atoms)-1
      WRITE(u,*) i+1, 0
      k = 0
      DO p = start+3, nl
        IF (TRIM(first_token(lines(p))) == TRIM(anames(atoms(i)))) THEN
          k = p
          EXIT
        END IF
      END DO
      IF (k == 0) CALL host_fail("atom not found in mybasis")
      READ(lines(k+1),*) sec, orb, nset
      p = k + 2
      DO j = 1, sec
        READ(lines(p),*) func, ncoef, pri, ang, ori
        p = p + 1
        WRITE(u,*) lnames(ang), func, "1.00"
        DO f = 1, func
          READ(lines(p),*) vals(0:ncoef-1)
          p = p + 1
          WRITE(u,*) vals(1:ncoef-1), vals(0)
        END DO
      END DO
      WRITE(u,*)
    END DO
    WRITE(u,*)
    WRITE(u,*) "[MO]"
    norb = SIZE(eigA)
    CALL molden_set(eigA, CA, "Alpha")
    IF (uhf) CALL molden_set(eigB, CB, "Beta")
    CLOSE(u)
  CONTAINS
    SUBROUTINE molden_set(eig, C, spin)
      REAL(8), INTENT(IN) :: eig(:), C(:,:)
      CHARACTER(LEN=*), INTENT(IN) :: spin
      INTEGER :: a, b
      DO a = 1, norb
        WRITE(u,*) "Sym= A"
        WRITE(u,fene) "Ene=", eig(a)
        WRITE(u,*) " Spin= " // spin
        IF (eig(a) < 0.0D0) THEN
          WRITE(u,*) "Occup= 1.0"
        ELSE
          WRITE(u,*) "Occup= 0.0"
        END IF
        DO b = 1, norb
          WRITE(u,fcoef) b, C(b,a)
        END DO
      END DO
    END SUBROUTINE molden_set
  END SUBROUTINE write_molden

END MODULE myqc_host
