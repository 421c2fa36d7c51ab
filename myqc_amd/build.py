"""Build libmyqc_hip.so (gfx950) in-tree with hipcc, and the Fortran host
programs with flang.

    python -m myqc_amd.build            # library + Fortran programs
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmyqc_hip.so")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    # the shell-quartet recursions are nested constant-trip loops that must
    # unroll completely, or their arrays fall back to scratch memory
    "-mllvm", "-pragma-unroll-threshold=100000",
    f"-I{os.path.join(ROOT, 'include')}",
]
SOURCES = ["capi.hip", "eri.hip", "jk.hip", "scf.hip", "sp2.hip", "batch.hip", "mlp.hip", "one_e.hip", "ao2mo.hip", "mfma_gemm.hip", "mfma_f64.hip", "cis.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_lib(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []

    def comp(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 2)) as ex:
        objs = list(ex.map(comp, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fortran(verbose: bool = True):
    """Fortran host programs (int1e, int2e, scf, ao2mo, mp2, cis, int1p, cphf) over the C ABI."""
    fc = shutil.which("flang") or shutil.which("amdflang")
    if fc is None or not os.path.isdir(FSRC):
        if verbose:
            print("flang or Fortran sources missing; skipping host programs")
        return
    os.makedirs(BIN, exist_ok=True)
    mod = os.path.join(BUILD, "fmod")
    os.makedirs(mod, exist_ok=True)
    flags = ["-O2", "-module-dir", mod, "-I", mod]
    mods = ["myqc_hip_api.f90", "myqc_host.f90"]
    objs = []
    for m in mods:
        o = os.path.join(BUILD, m.replace(".f90", ".o"))
        _run([fc, *flags, "-c", os.path.join(FSRC, m), "-o", o])
        objs.append(o)
    for prog in ["int1e", "int2e", "scf", "ao2mo", "mp2", "cis", "int1p", "cphf"]:
        src = os.path.join(FSRC, prog + ".f90")
        if not os.path.exists(src):
            continue
        _run([fc, *flags, "-o", os.path.join(BIN, prog), src, *objs, f"-L{PKG}", "-lmyqc_hip",
              f"-Wl,-rpath,{PKG}", "-Wl,-rpath,/opt/rocm/lib"])
    if verbose:
        print(f"built Fortran host programs in {BIN}")


def build_all(verbose: bool = True):
    build_lib(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build_all()
    sys.exit(0)
