"""CPU checks of the C-ABI boundary (no GPU work): the library loads, exports
every entry point include/icar_adv.h declares, and rejects bad arguments in
host code before touching the device."""
import ctypes
import os
import re

import pytest

from hicar_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "icar_adv.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(icar_adv_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 17
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.EXPORTS)


def test_create_rejects_small_grid_without_device():
    lib = _lib.load()
    g = _lib.Grid(1, 2, 1, 5, 1, 5, 1, 2, 1, 5, 1000.0)
    o = _lib.Opts(2, 1, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    rc = lib.icar_adv_create(ctypes.byref(g), ctypes.byref(o), ctypes.byref(h))
    assert rc == -2 and not h.value
    assert b"at least 3" in lib.icar_adv_last_error(None)


@pytest.mark.parametrize("tile", [(0, 6, 3, 4), (3, 9, 3, 4), (3, 6, 0, 4), (3, 6, 3, 7),
                                  (9, 6, 3, 4), (3, 6, 3, 0)])
def test_create_rejects_a_tile_outside_memory_without_device(tile):
    """its, ite, jts and jte must lie inside ims..ime x jms..jme (here 1..8 x
    1..6): the halo strips address the tile.  Checked before any device
    call."""
    lib = _lib.load()
    g = _lib.Grid(1, 8, 1, 5, 1, 6, *tile, 1000.0)
    o = _lib.Opts(2, 1, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    rc = lib.icar_adv_create(ctypes.byref(g), ctypes.byref(o), ctypes.byref(h))
    assert rc == -2 and not h.value
    assert b"inside the memory bounds" in lib.icar_adv_last_error(None)


def test_create_accepts_an_empty_tile_inside_memory_without_device():
    """A 3-cell-wide domain with a halo of 2 has its = 3 > ite = 1 (the
    reference's test_mpdata domain): allowed; without a device create then
    fails only at the device."""
    lib = _lib.load()
    g = _lib.Grid(1, 3, 1, 5, 1, 100, 3, 1, 3, 98, 1000.0)
    o = _lib.Opts(2, 1, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    rc = lib.icar_adv_create(ctypes.byref(g), ctypes.byref(o), ctypes.byref(h))
    assert b"inside the memory bounds" not in (lib.icar_adv_last_error(None) or b"")
    if rc == 0:
        lib.icar_adv_destroy(h)


@pytest.mark.parametrize("field,val", [("mpdata_order", -1), ("kernel_variant", 7)])
def test_create_rejects_bad_options(field, val):
    lib = _lib.load()
    g = _lib.Grid(1, 8, 1, 5, 1, 6, 3, 6, 3, 4, 1000.0)
    o = _lib.Opts(2, 1, 0, 0, 0, 0)
    setattr(o, field, val)
    h = ctypes.c_void_p()
    assert lib.icar_adv_create(ctypes.byref(g), ctypes.byref(o), ctypes.byref(h)) == -1


def test_null_arguments():
    lib = _lib.load()
    assert lib.icar_adv_create(None, None, None) == -1
    assert lib.icar_adv_advect(None, ctypes.c_float(1.0)) == -1
    assert lib.icar_adv_destroy(None) == 0


def test_kernel_names():
    lib = _lib.load()
    n = lib.icar_adv_kernel_count()
    names = [lib.icar_adv_kernel_name(i).decode() for i in range(n)]
    assert "upwind" in names and "fct" in names and len(set(names)) == n
    assert lib.icar_adv_kernel_name(n) == b""


def test_fortran_interface_binds_declared_symbols(tmp_path):
    """Every bind(C) name of the iso_c_binding module is a declared entry
    point, and the module compiles with the image's flang."""
    import shutil
    import subprocess
    src = os.path.join(ROOT, "fortran", "icar_adv_c.f90")
    names = set(re.findall(r'bind\(C,\s*name="(icar_adv_\w+)"\)', open(src).read()))
    assert names and names <= set(declared_symbols())
    fc = shutil.which("flang") or shutil.which("amdflang")
    if fc is None:
        pytest.skip("no Fortran compiler in this environment")
    subprocess.run([fc, "-c", src, "-module-dir", str(tmp_path), "-o", str(tmp_path / "m.o")],
                   check=True, timeout=120)


def test_extension_entry_points_check_arguments_without_device():
    """The division check, the allocation-free budget, the divergence result
    and the compensation switch refuse bad arguments in host code, before any
    device call."""
    import numpy as np
    lib = _lib.load()
    bad, first = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    sig = 1 << 23
    for d0, n in ((0, 0), (sig, 1), (sig - 4, 5)):
        assert lib.icar_adv_verify_division(d0, n, ctypes.byref(bad), ctypes.byref(first)) \
            == _lib.EINVAL, (d0, n)
        assert b"verify_division" in lib.icar_adv_last_error(None)
    assert lib.icar_adv_verify_division(0, 1, None, ctypes.byref(first)) == _lib.EINVAL

    fp = ctypes.POINTER(ctypes.c_float)
    a = np.zeros(8, np.float32).ctypes.data_as(fp)
    out = (ctypes.c_double * 1)()
    assert lib.icar_adv_mass_budget_noalloc(2, 8, 8, a, a, a, a, out, None) == _lib.EINVAL
    assert lib.icar_adv_mass_budget_noalloc(8, 8, 8, a, None, a, a, out, None) == _lib.EINVAL
    assert b"noalloc" in lib.icar_adv_last_error(None)

    res = (ctypes.c_double * 2)()
    assert lib.icar_adv_divergence(None, res) == _lib.EINVAL
    assert lib.icar_adv_set_compensation(None, 1) == _lib.EINVAL
