"""The C-ABI library loads on a host without a GPU and exports every entry
point that include/*.h declares (no compute calls)."""
import ctypes
import glob
import os
import re

import pytest

from tandav_amd import engine

pytestmark = pytest.mark.cpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(tdv_\w+)\s*\(", src,
                             flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "tdv_create" in names and "tdv_compute_acceleration" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(engine.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_loader_fails_loudly_without_library(tmp_path):
    with pytest.raises(engine.TDVError, match="no CPU fallback"):
        engine._lib = None
        try:
            engine.load_library(str(tmp_path / "missing.so"))
        finally:
            engine._lib = None
