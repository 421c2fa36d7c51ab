"""myqc_amd: the two-electron path of myQC (STO-3G RHF/MP2) on MI355X.

Host side: the reference's program contracts (int2e, scf, ao2mo, mp2) as
Fortran programs under myqc_amd/fortran, and this Python mirror of the same
C ABI (myqc_amd.hip).  Device side: hand-written gfx950 kernels in
myqc_amd/csrc, built into myqc_amd/libmyqc_hip.so.
"""
from .basis import Basis, build_basis, read_ftab  # noqa: F401
from .molecules import co2, to_bohr, water_cluster  # noqa: F401
