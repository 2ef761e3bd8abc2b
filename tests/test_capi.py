"""The C-ABI library loads without a GPU and exports every entry point that
include/ksgf.h declares (no compute calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "ksgf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(ksgf_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "ksgf_extrap_vel_to_faces" in names and "ksgf_macproj_project" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol(ksgf):
    L = ksgf.load()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(ksgf.EXPORTS) == set(_declared())
    assert L.ksgf_version().startswith(b"ksgf-gfx950")


def test_library_is_gfx950_code_object():
    path = os.path.join(ROOT, "kynema-sgf_amd", "libksgf.so")
    data = open(path, "rb").read()
    assert b"gfx950" in data
