"""CPU checks of the C ABI: the library builds for gfx950, loads, and exports
every function include/cac_hip.h declares.  No device calls."""

import ctypes
import os
import subprocess

import cac_amd


def test_library_exports_every_declared_symbol():
    assert os.path.exists(cac_amd.LIBPATH), "build libcachip.so first (__graft_entry__.build())"
    names = cac_amd.header_functions()
    assert len(names) >= 30
    L = ctypes.CDLL(cac_amd.LIBPATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", cac_amd.LIBPATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_binding_covers_header():
    L = cac_amd.lib()
    for n in cac_amd.header_functions():
        fn = getattr(L, n)
        assert fn.restype in (ctypes.c_int, ctypes.c_char_p), n


def test_library_holds_gfx950_code():
    blob = open(cac_amd.LIBPATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
