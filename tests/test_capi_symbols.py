"""The C-ABI library loads and exports every symbol include/enblend_amd.h
declares (no GPU work)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "enblend_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(enbl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header(enbl):
    lib = ctypes.CDLL(enbl.library_path())
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_auto_levels_host_only(enbl, oracle):
    for (w, h) in [(1024, 768), (7168, 3072), (6000, 4000), (5, 5), (3000, 17)]:
        assert enbl.auto_levels(w, h) == oracle.auto_levels(w, h)
    # --levels handling of roiBounds (src/bounds.h:136-156)
    assert enbl.auto_levels(7168, 3072, 1, 4) == 4
    assert enbl.auto_levels(7168, 3072, 1, 40) == 10
    assert enbl.auto_levels(7168, 3072, 1, -3) == 7
    assert enbl.auto_levels(7168, 3072, 1, -30) == 1


def test_context_without_device_fails_cleanly(enbl):
    import torch
    if torch.cuda.is_available():
        return
    import ctypes as c
    h = c.c_void_p()
    rc = enbl.lib().enbl_context_create(0, c.byref(h))
    assert rc in (enbl.ENBL_E_ARG, enbl.ENBL_E_HIP)
