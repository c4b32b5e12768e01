"""Build recipe for oracle/_ref/libref_strucf.so -- TEST INFRASTRUCTURE ONLY.

Compiles DISCUS's per-type structure-factor kernel and its corner-phase
construction from the reference sources where they lie
(discus/prog/four_strucf_OMP_mod.f90, discus/prog/fourier_lmn_mod.f90 and the
modules they use, 51 files of discus/prog and lib_f90) with the image's flang,
plus the C entries of oracle/ref_strucf_bind.f90, into oracle/_ref/.  Nothing of the reference is
copied: the sources are read from REF (default /root/reference) and only
objects, module files and the shared library are written, under oracle/_ref/.

Flags follow the reference's GNU build (CMakeLists.txt:129 and
discus/prog/CMakeLists.txt:127): -fPIC -cpp -O3 -ffast-math -fopenmp.
lib_f90's compiler-specific module sys_compiler comes from sys_gfortran.f90,
the file lib_f90/CMakeLists.txt:71-78 picks for GNU-compatible compilers.

The rest of the Fourier path (fourier_sup.f90: four_run, four_stltab,
four_cexpt) is not built: its module closure includes lib_f90_fftw3.f90 and
the hdf5 module, which need the FFTW3 and HDF5 Fortran interfaces that this
image lacks.

Four external symbols stay unresolved in the library: fstat_, stat_ and
ctime_ (GNU Fortran extensions called by the file-information routines of
sys_gfortran.f90 and support_mod.f90, which flang does not provide) and
calc_intr_ (the expression evaluator's intrinsic dispatcher, another part of
lib_f90).  None of them is reachable from four_strucf; nothing stands in for
them, and the library is opened with lazy binding (oracle/oracle.py).

    python oracle/build_ref.py          (no-op with a message when REF is absent)
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
LIB = os.path.join(OUT, "libref_strucf.so")
REF = os.environ.get("DC_REFERENCE", "/root/reference")
ROOT_SOURCE = "discus/prog/four_strucf_OMP_mod.f90"
ROOT_SOURCES = [ROOT_SOURCE, "discus/prog/fourier_lmn_mod.f90"]
FLAGS = ["-fPIC", "-cpp", "-O3", "-ffast-math", "-fopenmp"]
INTRINSIC = {"omp_lib", "iso_c_binding", "iso_fortran_env", "ieee_arithmetic", "mpi"}
# modules that several reference files define: the file the reference's own
# build picks for a GNU compiler on Linux without CUDA (lib_f90/CMakeLists.txt:
# 71-78 for sys_compiler, 99-104 for parallel_mod; the kernel module is the
# OpenMP one, ROOT_SOURCE)
PREFER = {"sys_compiler": "lib_f90/sys_gfortran.f90", "parallel_mod": "lib_f90/parallel_mod.f90",
          "four_strucf_mod": ROOT_SOURCE, "operating_mod": "lib_f90/op_linux.f90"}


def _flang():
    for c in (os.environ.get("FC"), shutil.which("flang"), "/opt/rocm/llvm/bin/flang"):
        if c and os.path.exists(c):
            return c
    return None


def _module_map(ref):
    mods = {}
    for root, _, files in os.walk(ref):
        for f in sorted(files):
            if not f.endswith(".f90"):
                continue
            p = os.path.join(root, f)
            for line in open(p, errors="ignore"):
                m = re.match(r"\s*module\s+(\w+)\s*(!.*)?$", line, re.I)
                if not m or m.group(1).lower() == "procedure":
                    continue
                mods.setdefault(m.group(1).lower(), p)
    for name, rel in PREFER.items():
        mods[name] = os.path.join(ref, rel)
    return mods


def _uses(path):
    out = []
    for line in open(path, errors="ignore"):
        m = re.match(r"\s*use\s*(?:,\s*\w+\s*::)?\s*(\w+)", line, re.I)
        if m and m.group(1).lower() not in INTRINSIC:
            out.append(m.group(1).lower())
    return out


def compile_order(ref=REF):
    """Source files of the kernel's module closure, dependencies first."""
    mods = _module_map(ref)
    order, state = [], {}

    def visit(p):
        if state.get(p):
            return
        state[p] = 1
        for m in _uses(p):
            if m not in mods:
                raise RuntimeError(f"module {m} (used by {p}) not found under {ref}")
            if mods[m] != p:
                visit(mods[m])
        order.append(p)

    for root in ROOT_SOURCES:
        visit(os.path.join(ref, root))
    return order


def build(force=False, verbose=False):
    if not os.path.isdir(REF):
        print(f"build_ref: {REF} absent, nothing built")
        return None
    fc = _flang()
    if fc is None:
        print("build_ref: no flang, nothing built")
        return None
    if os.path.exists(LIB) and not force:
        return LIB
    obj = os.path.join(OUT, "obj")
    os.makedirs(obj, exist_ok=True)
    objs = []
    for src in compile_order() + [os.path.join(HERE, "ref_strucf_bind.f90")]:
        o = os.path.join(obj, os.path.basename(src) + ".o")
        cmd = [fc, "-c", *FLAGS, "-module-dir", obj, "-I", obj, src, "-o", o]
        if verbose:
            print(" ".join(cmd))
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            sys.stderr.write(p.stdout + p.stderr)
            raise RuntimeError("build_ref: compile failed: " + src)
        objs.append(o)
    libdir = os.path.join(os.path.dirname(os.path.dirname(fc)), "lib")
    tmp = LIB + ".tmp"
    # fourier_lmn computes in REAL(16): flang calls libquadmath for it
    cmd = [fc, "-shared", "-fopenmp", "-o", tmp, *objs, "-lquadmath", "-Wl,-rpath," + libdir]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("build_ref: link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose="-v" in sys.argv))
