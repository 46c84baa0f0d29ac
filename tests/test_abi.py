"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU call is made)."""
import glob
import os
import re

from lbcuda_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b(lbm_[A-Za-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"lbm_create", "lbm_step", "lbm_destroy", "lbm_get_pops"} <= names


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    handle = L.lib()
    missing = [n for n in sorted(names) if not hasattr(handle, n)]
    assert not missing, missing
    assert set(L.EXPORTED) >= names - {"lbm_ctx"}


def test_version_string():
    assert "gfx950" in L.version()


def test_config_struct_matches_c_layout():
    import ctypes
    from lbcuda_amd.abi import LbmConfig, default_config
    c = LbmConfig()
    L.lib().lbm_config_default(ctypes.byref(c), 64)
    py = default_config(64)
    for name, _ in LbmConfig._fields_:
        a, b = getattr(c, name), getattr(py, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert abs(a - b) <= 1e-7 * max(1.0, abs(a)), (name, a, b)
