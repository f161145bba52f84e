"""The C-ABI library loads and exports every symbol the public headers
declare (CPU only: no compute call is made)."""

import ctypes
import glob
import os
import re

import flursys_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for header in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(header).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(fl_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_functions():
    names = declared_functions()
    assert "fl_create" in names and "fl_cavity3d_step" in names


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_signatures_cover_the_header():
    assert declared_functions() == set(L.SIGNATURES)


def test_version_string():
    v = L.load().fl_version()
    assert v.startswith(b"flursys-hip")


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "absent.so"))
    try:
        L.load()
    except L.FlursysError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("load() must raise when the library is absent")
    finally:
        monkeypatch.undo()
        L._lib = None


def test_library_links_no_oracle():
    """The product library must not depend on the CPU oracle."""
    data = open(L.LIB_PATH, "rb").read()
    assert b"liboracle" not in data and b"or3d_" not in data and b"or_cg" not in data
    ctypes.CDLL(L.LIB_PATH)
