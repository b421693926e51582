"""The C-ABI library exports every entry point include/dgx.h declares (no GPU needed)."""
import ctypes
import os
import re

import dg_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dgx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^\s*typedef[^;]*;", "", text, flags=re.M)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "sizeof", "defined")))


def test_header_declares_reference_launchers():
    names = declared_functions()
    for ref in ("LmapToFace", "LmapToElement", "Lredistribute_sf", "Lredistribute_q", "LaddSFQ", "Levaluate_sf",
                "Levaluate_q", "LrflctiveBoundary", "Lsolver", "LfindUPA", "Lhrl1D", "Lhri1D", "Lcalc_sensors",
                "Lcopy_detected"):
        assert ref in names, ref


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(dg_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_functions()) >= 35
