#!/usr/bin/env bash
# Builds the reference's own two-electron-path programs from their Fortran
# sources, where they lie under /root/reference/src, with AMD flang.  Outputs
# (objects, module files, executables) go only to oracle/_ref/.  Nothing of the
# reference is copied into the repository.
#
# Built: int1e, int2e, RHFI2G, UHFI2G, dens, ao2mo, mp2, int1p, cphf (they need nothing beyond the
# Fortran runtime).  Not built: scf and cis -- they call LAPACK DSYGV/DSYEV and the image
# has no LAPACK library, so they are unbuildable here; the SCF driver loops and
# the CIS matrix are restated in oracle/ref_driver.py on top of the reference
# programs built here.
set -euo pipefail
SRC=${MYQC_REF_SRC:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
if [ ! -d "$SRC" ]; then echo "reference sources not found at $SRC; skipping"; exit 0; fi
mkdir -p "$OUT/mod"
FC=${FC:-flang}
FFLAGS="-O3 -w -module-dir $OUT/mod -I$OUT/mod"
$FC $FFLAGS -c "$SRC/myQC/env.f90" -o "$OUT/env.o"
$FC $FFLAGS -c "$SRC/myQC/basis.f90" -o "$OUT/basis.o"
$FC $FFLAGS -c "$SRC/integrals/auxilary.f90" -o "$OUT/auxilary.o"
$FC $FFLAGS -o "$OUT/int1e" "$SRC/integrals/int1e.f90" "$OUT/auxilary.o" "$OUT/env.o" "$OUT/basis.o"
$FC $FFLAGS -o "$OUT/int2e" "$SRC/integrals/int2e.f90" "$OUT/auxilary.o" "$OUT/env.o" "$OUT/basis.o"
$FC $FFLAGS -o "$OUT/RHFI2G" "$SRC/I2G/RHFI2G.f90" "$OUT/env.o"
$FC $FFLAGS -o "$OUT/UHFI2G" "$SRC/I2G/UHFI2G.f90" "$OUT/env.o"
$FC $FFLAGS -o "$OUT/dens" "$SRC/dens/dens.f90" "$OUT/env.o" "$OUT/basis.o"
$FC $FFLAGS -o "$OUT/ao2mo" "$SRC/ao2mo/ao2mo.f90" "$OUT/env.o"
$FC $FFLAGS -o "$OUT/mp2" "$SRC/mp2/mp2.f90" "$OUT/env.o"
# int1p's own file faults when built at -O3 here (the cause was not traced; it
# also reads stat2 unset, int1p.f90:81).  That one file is built at -O0; its
# Pxuv then equals that of a build with every file at -O0.
$FC ${FFLAGS/-O3/-O0} -o "$OUT/int1p" "$SRC/integrals/int1p.f90" "$OUT/auxilary.o" "$OUT/env.o" "$OUT/basis.o"
$FC $FFLAGS -o "$OUT/cphf" "$SRC/props/cphf.f90" "$OUT/env.o"
$FC --version | head -1 > "$OUT/FLANG_VERSION"
echo "reference programs built in $OUT"
# Golden-vector harness over the reference's Boys routine
$FC $FFLAGS -o "$OUT/boys_dump" "$HERE/boys_dump.f90" "$OUT/auxilary.o"
