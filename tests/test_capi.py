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


def test_boxarray_rejects_overlapping_boxes(ksgf):
    """Two boxes that overlap (with a gap of the same size elsewhere, so the
    cell count still equals the domain's) are refused before any device
    call; so are overlapping boxes on a level that does not cover its
    domain."""
    import ctypes as C

    L = ksgf.load()
    g = ksgf.KsgfGeometry()
    for d in range(3):
        g.domain_lo[d], g.domain_hi[d] = 0, 7
        g.prob_lo[d], g.prob_hi[d] = 0.0, 1.0
        g.is_periodic[d] = 1
    boxes = (C.c_int * 12)(0, 0, 0, 7, 7, 3, 0, 0, 3, 7, 7, 6)  # z 0..3 and 3..6: overlap at z 3, gap at z 7
    out = C.c_void_p()
    assert L.ksgf_boxarray_create(C.byref(g), boxes, 2, C.byref(out)) == ksgf.KSGF_ERR_INVALID
    assert b"overlap" in L.ksgf_last_error()
    assert L.ksgf_boxarray_create_level(C.byref(g), boxes, 2, 0, C.byref(out)) == ksgf.KSGF_ERR_INVALID
