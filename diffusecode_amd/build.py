"""In-tree build of libdiffuse_turbo.so (hipcc, gfx950) and of the Fortran
ISO_C_BINDING module (flang).  Used by __graft_entry__.build() and by
``python -m diffusecode_amd.build``."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdiffuse_turbo.so")
SOURCES = ["dc_gemm.hip", "dc_lot.hip", "dc_api.hip", "dc_adjoint.hip", "dc_refine.hip", "dc_batch.hip", "dc_post.hip", "dc_nufft.hip",
           "dc_peak.hip", "dc_noalloc.hip"]
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
ARCH = os.environ.get("DC_OFFLOAD_ARCH", "gfx950")
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            # packed FP32 VALU ops stay off (wrong results seen on this hardware pool)
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include")]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return p


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError("build failed: " + " ".join(cmd))
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          "-L", ROCM_LIB, "-lhipfft", "-Wl,-rpath," + ROCM_LIB])
    os.replace(tmp, LIB)
    return LIB


FORTRAN_SRC = os.path.join(HERE, "fortran", "diffuse_turbo_mod.f90")


FORTRAN_EXAMPLE = os.path.join(HERE, "fortran", "example_strucf.f90")
FORTRAN_EXE = os.path.join(OBJ, "fortran", "example_strucf")
FORTRAN_EXAMPLE_POST = os.path.join(HERE, "fortran", "example_post.f90")
FORTRAN_EXE_POST = os.path.join(OBJ, "fortran", "example_post")
FORTRAN_EXAMPLE_REFINE = os.path.join(HERE, "fortran", "example_refine.f90")
FORTRAN_EXE_REFINE = os.path.join(OBJ, "fortran", "example_refine")
FORTRAN_EXAMPLE_DIFFEV = os.path.join(HERE, "fortran", "example_diffev.f90")
FORTRAN_EXE_DIFFEV = os.path.join(OBJ, "fortran", "example_diffev")


def build_fortran() -> str | None:
    """Compile the ISO_C_BINDING interface module and link the example host
    program against libdiffuse_turbo.so (rpath to the package directory)."""
    fc = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    if not os.path.exists(FORTRAN_SRC) or not os.path.exists(fc):
        return None
    out = os.path.join(OBJ, "fortran")
    os.makedirs(out, exist_ok=True)
    obj = os.path.join(out, "diffuse_turbo_mod.o")
    _run([fc, "-c", "-O2", "-fPIC", "-module-dir", out, FORTRAN_SRC, "-o", obj])
    for src, exe in ((FORTRAN_EXAMPLE, FORTRAN_EXE), (FORTRAN_EXAMPLE_POST, FORTRAN_EXE_POST),
                     (FORTRAN_EXAMPLE_REFINE, FORTRAN_EXE_REFINE), (FORTRAN_EXAMPLE_DIFFEV, FORTRAN_EXE_DIFFEV)):
        _run([fc, "-O2", "-module-dir", out, "-I", out, src, obj, "-o", exe,
              "-L", HERE, "-ldiffuse_turbo", "-Wl,-rpath,$ORIGIN/../..", "-Wl,-rpath," + HERE])
    return FORTRAN_EXE


YELL_SRC = os.path.join(CSRC, "dc_yell.c")
YELL_LIB = os.path.join(HERE, "libdiffuse_yell.so")
FORTRAN_EXAMPLE_YELL = os.path.join(HERE, "fortran", "example_yell.f90")
FORTRAN_EXE_YELL = os.path.join(OBJ, "fortran", "example_yell")


def hdf5_root() -> str | None:
    """Prefix holding include/hdf5.h and lib/libhdf5.so: $HDF5_ROOT, else the
    image's conda prefix, else /usr."""
    for r in (os.environ.get("HDF5_ROOT"), "/opt/conda", "/usr"):
        if r and os.path.exists(os.path.join(r, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(r, "lib", "libhdf5.so")):
            return r
    return None


def build_yell() -> str | None:
    """libdiffuse_yell.so, the Yell 1.0 HDF5 writer (gcc, linked to HDF5 with
    an rpath).  Without an HDF5 installation there is no writer, as DISCUS
    links hdf_write_no.f90 then; returns None."""
    root = hdf5_root()
    if root is None:
        return None
    inc, lib = os.path.join(root, "include"), os.path.join(root, "lib")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", inc,
          YELL_SRC, "-o", YELL_LIB, "-L", lib, "-lhdf5", "-Wl,-rpath," + lib])
    # the Fortran host example, when the interface module has been built (the
    # module's own procedures reference libdiffuse_turbo, so both are linked)
    fc = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    out = os.path.join(OBJ, "fortran")
    mod = os.path.join(out, "diffuse_turbo_mod.o")
    if os.path.exists(fc) and os.path.exists(mod):
        _run([fc, "-O2", "-module-dir", out, "-I", out, FORTRAN_EXAMPLE_YELL, mod, "-o", FORTRAN_EXE_YELL,
              "-L", HERE, "-ldiffuse_yell", "-ldiffuse_turbo", "-Wl,-rpath," + HERE])
    return YELL_LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran())
    print(build_yell())
