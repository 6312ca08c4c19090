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


REFDIR = "/root/reference"
SHIM = os.path.join(ROOT, "thecodecombo_7_amd", "host", "ttc_refshim.c")
# entry points the shim stands in for, with the reference file that defines them
SHIM_SYMBOLS = {
    "nblist": "nblist.c", "nl_check": "nblist.c", "nblist_pivot": "nblist_pivot.c",
    "cnonbondnl": "cnonbond_nblist.c", "pme_calc": "pme_calc.c", "pme_force": "pme_calc.c",
    "charge_grid": "pme_qgrid.c", "pme_init": "pme_setup.c", "pme_clean": "pme_setup.c",
    "ewald_setup": "ewald.c", "enbnd_mc": "enbnd_mc.c", "erfc": "ewald.c",
}


@pytest.mark.skipif(not os.path.isdir(REFDIR), reason="reference tree absent")
def test_shim_compiles_against_reference_headers(tmp_path):
    """The reference-side binding compiles against the reference's own
    defines.h (flags NLIST DLIST STYPE S14 EWALD PRESSURE) and defines the
    replaced entry points, each of which the reference defines by that name."""
    obj = str(tmp_path / "shim.o")
    r = subprocess.run(["gcc", "-std=c99", "-O2", "-c", "-DNLIST", "-DDLIST", "-DSTYPE", "-DS14", "-DEWALD",
                        "-DPRESSURE", f"-I{REFDIR}", f"-I{ROOT}/include", SHIM, "-o", obj],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run(["nm", "--defined-only", obj], capture_output=True, text=True).stdout
    defined = set(re.findall(r" T (\w+)$", out, flags=re.M))
    assert set(SHIM_SYMBOLS) <= defined
    # each one is really a reference entry point of that name
    for name, src in SHIM_SYMBOLS.items():
        text = open(os.path.join(REFDIR, src), errors="replace").read()
        assert re.search(rf"^\s*(void|double)\s+{name}\s*\(", text, flags=re.M), (name, src)


@pytest.mark.skipif(not os.path.isdir(REFDIR), reason="reference tree absent")
def test_shim_fftw_allocation_pair(tmp_path):
    """With -DTTC_SHIM_FFTW_ALLOC the binding defines FFTW's allocation pair
    (fftw_malloc / fftw_free, the only FFTW calls outside the replaced
    pme_*.c files: allocate.c:500), so an SPME program links without
    -lfftw3.  (The SPME flag itself is not compiled here: defines.h:41
    includes <fftw3.h>, which this image does not have.)"""
    obj = str(tmp_path / "shim3.o")
    r = subprocess.run(["gcc", "-std=c99", "-O2", "-c", "-DNLIST", "-DDLIST", "-DSTYPE", "-DS14", "-DEWALD",
                        "-DPRESSURE", "-DTTC_SHIM_FFTW_ALLOC", f"-I{REFDIR}", f"-I{ROOT}/include", SHIM, "-o", obj],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run(["nm", "--defined-only", obj], capture_output=True, text=True).stdout
    defined = set(re.findall(r" T (\w+)$", out, flags=re.M))
    assert {"fftw_malloc", "fftw_free"} <= defined
    # and no other FFTW symbol is used by the reference outside pme_*.c
    for f in os.listdir(REFDIR):
        if f.endswith(".c") and not f.startswith("pme_"):
            calls = set(re.findall(r"\bfftw_\w+(?=\s*\()", open(os.path.join(REFDIR, f), errors="replace").read()))
            assert calls <= {"fftw_malloc", "fftw_free"}, (f, calls)


@pytest.mark.skipif(not os.path.isdir(REFDIR), reason="reference tree absent")
def test_shim_configt_entry_point(tmp_path):
    """In a CONFIGT build the binding also defines cnonbondnl2
    (cnonbond_nblist2.c:47) over the reference's config[k] struct."""
    obj = str(tmp_path / "shim2.o")
    r = subprocess.run(["gcc", "-std=c99", "-O2", "-c", "-DNLIST", "-DDLIST", "-DSTYPE", "-DS14", "-DEWALD",
                        "-DPRESSURE", "-DCONFIGT", f"-I{REFDIR}", f"-I{ROOT}/include", SHIM, "-o", obj],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run(["nm", "--defined-only", obj], capture_output=True, text=True).stdout
    assert re.search(r" T cnonbondnl2$", out, flags=re.M)
    text = open(os.path.join(REFDIR, "cnonbond_nblist2.c"), errors="replace").read()
    assert re.search(r"^\s*void\s+cnonbondnl2\s*\(", text, flags=re.M)
