"""C-ABI checks that need no GPU: the library loads, exports every entry
point declared in include/diffuse_turbo.h, and its argument checks return
the documented status codes without touching a device."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from diffusecode_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "diffuse_turbo.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) <= set(names)


def test_version_and_tiles():
    L = _lib.lib()
    L.dc_version.restype = C.c_char_p
    assert b"gfx950" in L.dc_version()
    tiles = _lib.Context.gemm_tiles()
    assert len(tiles) >= 4
    for rt, ct, nt in tiles:
        assert rt % 16 == 0 and ct % 16 == 0 and nt % 64 == 0
    r = C.c_int()
    assert L.dc_gemm_tile_shape(len(tiles), C.byref(r), None, None) == _lib.DC_ERR_ARG


def test_null_context_arguments():
    L = _lib.lib()
    assert L.dc_set_precision(None, 0) == _lib.DC_ERR_ARG
    assert L.dc_set_gemm_tile(None, 0) == _lib.DC_ERR_ARG
    assert L.dc_set_slice_atoms(None, 0) == _lib.DC_ERR_ARG
    assert L.dc_set_phase_dither(None, 1) == _lib.DC_ERR_ARG
    assert L.dc_set_adjoint_walk(None, 0) == _lib.DC_ERR_ARG
    assert L.dc_crystal_set(None, None) == _lib.DC_ERR_ARG
    assert L.dc_four_run(None, None, None, None, None, None, None) == _lib.DC_ERR_ARG
    assert L.dc_strucf(None, None, 0, 0, None, None, None, None, None, None, None, None, 0, None, None, None, 0,
                       None) == _lib.DC_ERR_ARG
    assert L.dc_create(None, 0) == _lib.DC_ERR_ARG
    assert L.dc_four_post(None, None, None, None, None) == _lib.DC_ERR_ARG
    assert L.dc_four_post_dev(None, None, None, None, None) == _lib.DC_ERR_ARG
    assert L.dc_four_fill(None, None, None, None) == _lib.DC_ERR_ARG
    assert L.dc_four_accumulate_dev(None, 0, None, None, None, None, None) == _lib.DC_ERR_ARG
    assert L.dc_qval(None, None, None, None, None, None) == _lib.DC_ERR_ARG
    assert L.dc_batch_chi2(None, None, None, None, None, None, None) == _lib.DC_ERR_ARG
    assert L.dc_batch_chi2_dev(None, None, None, None, None, None, None) == _lib.DC_ERR_ARG
    assert L.dc_measure_peaks(None, 0, None) == _lib.DC_ERR_ARG


def _gpu_present():
    return os.path.exists("/dev/kfd") and bool(shutil.which("rocminfo")) and \
        subprocess.run(["rocminfo"], capture_output=True).returncode == 0


@pytest.mark.skipif(_gpu_present(), reason="checks the no-device path")
def test_create_without_device_reports_nodev():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.dc_create(C.byref(h), 0) == _lib.DC_ERR_NODEV


def test_fortran_interface_module_compiles(tmp_path):
    fc = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    src = os.path.join(ROOT, "diffusecode_amd", "fortran", "diffuse_turbo_mod.f90")
    if not os.path.exists(fc) or not os.path.exists(src):
        pytest.skip("flang or the Fortran module not present")
    p = subprocess.run([fc, "-c", "-module-dir", str(tmp_path), src, "-o", str(tmp_path / "m.o")],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


# descriptor fields, in declaration order, as both languages name them
_LAYOUTS = {
    "dc_model_desc": ["pos", "type", "occ", "bdw", "natoms", "nscat", "ncfact", "cfact", "eck", "vi", "inc", "rten"],
    "dc_batch_desc": ["nmembers", "atom_beg", "pos", "type", "occ", "bdw", "nscat", "ncfact", "cfact", "rten", "eck",
                      "vi", "inc"],
    "dc_run_desc": ["eck", "vi", "inc", "rten", "cfact", "ncfact", "ilots", "ls", "nlots", "lbeg", "lperiod", "fave",
                    "sel_cell", "four_friedel", "ano", "diff_scale"],
    "dc_crystal_desc": ["cr_pos", "cr_iscat", "natoms", "icc", "ncatoms", "nscat", "dim0", "dim_lo"],
}


def test_fortran_descriptor_layouts_match_c(tmp_path):
    """Every bind(C) descriptor of diffuse_turbo_mod has the size and field
    offsets of its C struct (a Fortran host passes them by reference)."""
    fc = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    cc = shutil.which("gcc")
    src = os.path.join(ROOT, "diffusecode_amd", "fortran", "diffuse_turbo_mod.f90")
    if not os.path.exists(fc) or not cc:
        pytest.skip("flang or gcc not present")
    fl = ["program layouts", "use iso_c_binding", "use diffuse_turbo_mod", "implicit none"]
    for t in _LAYOUTS:
        fl.append(f"type({t}), target :: v_{t}")
    for t, fields in _LAYOUTS.items():
        fl.append(f"write(*, '(a,1x,i0)') '{t} size', c_sizeof(v_{t})")
        for f in fields:
            fl.append(f"write(*, '(a,1x,i0)') '{t} {f}', "
                      f"transfer(c_loc(v_{t}%{f}), 0_c_intptr_t) - transfer(c_loc(v_{t}), 0_c_intptr_t)")
    fl.append("end program layouts")
    (tmp_path / "layouts.f90").write_text("\n".join(fl) + "\n")
    cl = ["#include <stdio.h>", "#include <stddef.h>", '#include "diffuse_turbo.h"', "int main(void) {"]
    for t, fields in _LAYOUTS.items():
        cl.append(f'printf("{t} size %zu\\n", sizeof({t}));')
        for f in fields:
            cl.append(f'printf("{t} {f} %zu\\n", offsetof({t}, {f}));')
    cl.append("return 0; }")
    (tmp_path / "layouts.c").write_text("\n".join(cl) + "\n")
    p = subprocess.run([fc, "-c", "-module-dir", str(tmp_path), src, "-o", str(tmp_path / "m.o")],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    libdir = os.path.dirname(_lib.lib()._name)
    p = subprocess.run([fc, "-module-dir", str(tmp_path), "-I", str(tmp_path), str(tmp_path / "layouts.f90"),
                        str(tmp_path / "m.o"), "-o", str(tmp_path / "lf"), "-L", libdir, "-ldiffuse_turbo",
                        "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    p = subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(tmp_path / "layouts.c"), "-o",
                        str(tmp_path / "lc")], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    f = subprocess.run([str(tmp_path / "lf")], capture_output=True, text=True).stdout.split()
    c = subprocess.run([str(tmp_path / "lc")], capture_output=True, text=True).stdout.split()
    assert len(f) > 40 and f == c, (f, c)
