"""The C-ABI library loads and exports every function include/*.h declares (no device calls)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_functions():
    names = []
    for h in ("icflow_hip.h", "icflow_host.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?(?:int|double|char\s*\*|const\s+char\s*\*)\s*\*?\s*([A-Za-z_]\w*)\s*\(",
                            txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 25, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from icflow_amd import _lib
    assert sorted(_lib.EXPORTED) == declared_functions()


def test_consts_struct_layout_matches_c(lib):
    """icf_consts is 36+144+7+16+162+9 doubles; the ctypes mirror must agree."""
    from icflow_amd import _lib
    assert ctypes.sizeof(_lib.IcfConsts) == 8 * (36 + 144 + 7 + 16 + 162 + 9)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from icflow_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    try:
        _lib.load()
    except OSError as e:
        assert "missing" in str(e)
    else:
        raise AssertionError("load() must raise when the library is absent")
