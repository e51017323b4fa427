"""The C-ABI library loads without a GPU and exports every function that
include/lbm_amd.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9]*\s*\**\s*(lbm_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "lbm_create" in names and "lbm_step" in names and len(names) >= 15


def test_library_exports_every_declared_symbol(native_lib):
    from cudalbm_amd import native

    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert hasattr(native_lib, n)
    # the Python binding lists exactly the declared set
    assert sorted(native.EXPORTED) == declared_functions()


def test_create_rejects_bad_config_without_gpu_work(native_lib):
    """Argument validation happens before any device call."""
    from cudalbm_amd import native

    cfg = native.lbm_config()
    cfg.velocity_set = 21
    cfg.nx = cfg.ny = cfg.nz = 16
    cfg.Re, cfg.u_inf = 100.0, 0.05
    h = ctypes.c_void_p()
    assert native_lib.lbm_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"velocity_set" in native_lib.lbm_create_error()
    cfg.velocity_set = 19
    cfg.nx = 20  # not a multiple of any footprint with ny=16
    cfg.tile_x, cfg.tile_y = 16, 16
    assert native_lib.lbm_create(ctypes.byref(cfg), ctypes.byref(h)) == -49
    # moment arrays are addressed with 32-bit byte offsets: at most 2^30 nodes
    cfg.nx = cfg.ny = cfg.nz = 1056
    cfg.tile_x = cfg.tile_y = cfg.tile_z = 0
    assert native_lib.lbm_create(ctypes.byref(cfg), ctypes.byref(h)) == -49
    assert b"2^30" in native_lib.lbm_create_error()
