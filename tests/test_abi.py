"""The C-ABI library loads on a machine without a GPU and exports every entry point that
include/fr.h declares (no compute call that needs a device)."""
import ctypes
import os
import re

import frhip
from frhip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(fr\w*)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "fr2d_euler_rhs" in names and "fr2d_euler_rhs_dev" in names and "fr2d_step" in names
    assert len(names) > 30


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_declared_symbols():
    # every symbol the header declares has a ctypes signature in the host package or is 1D/ensemble
    sig = set(_lib.SIGNATURES)
    names = [n for n in declared_symbols() if n.startswith(("fr_", "fr2d_"))]
    assert set(names) <= sig, sorted(set(names) - sig)


def test_version_and_errors_without_gpu():
    L = frhip.lib()
    assert b"gfx950" in L.fr_version()
    assert L.fr_device_count() >= 0
    # argument validation happens before any device call
    rc = L.fr_legendre_point(-1, None, None)
    assert rc == 1 and b"fr_legendre_point" in L.fr_last_error()
