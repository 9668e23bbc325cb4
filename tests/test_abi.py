"""The C-ABI library loads and exports every entry point the header declares
(no compute call: these run without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gasoline_amd.h")
LIB = os.path.join(ROOT, "gasoline_amd", "libgasoline_amd.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "gg_grav_all" in names and "gg_smooth" in names and "gg_resmooth" in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libgasoline_amd.so not built"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_the_header():
    from gasoline_amd import pkd
    assert sorted(pkd.ABI) == declared()


def test_last_error_callable_without_device():
    lib = ctypes.CDLL(LIB)
    lib.gg_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.gg_last_error(), bytes)


def test_host_sequences_library_loads():
    # the C host sequences (gasoline_amd/host/libggmsr.so) over the ABI
    from gasoline_amd import pkd
    L = pkd.msr_lib()
    for name in ("ggMsrTopStepKDK", "ggMsrForces", "ggMsrSphRung", "ggMsrGravity", "ggMsrSph"):
        assert hasattr(L, name), name
