"""The C ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "thecodecombo_7_amd", "libttc.so")


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(ttc_\w+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from thecodecombo_7_amd import build
        build.build()
    return C.CDLL(LIB)


def test_header_declares_entry_points():
    d = _declared()
    assert "ttc_create" in d and "ttc_cnonbondnl" in d and len(d) >= 20


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ttc_\w+)$", out, flags=re.M))
    assert set(_declared()) <= exported


def test_library_carries_gfx950_code_object():
    # the embedded offload bundle names its target triple
    with open(LIB, "rb") as fh:
        assert b"amdgcn-amd-amdhsa--gfx950" in fh.read()


def test_python_mirror_fails_loudly_without_library(tmp_path, monkeypatch):
    from thecodecombo_7_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.lib()
