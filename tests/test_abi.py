"""CPU checks of the C-ABI boundary: the library loads without a GPU and
exports every function include/pixie_hip.h declares."""
import ctypes
import os
import re

import pytest

import pixie_amd as px

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"\b(pxh_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(px.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_without_device():
    lib = px.lib()
    assert b"gfx950" in lib.pxh_version()
    assert isinstance(lib.pxh_last_error(), bytes)


def test_enums_match_pixie_ordinals():
    # BlendMode, common.nim:6-29
    assert px.BlendMode.Normal == 0 and px.BlendMode.Luminosity == 15
    assert px.BlendMode.Mask == 16 and px.BlendMode.ExcludeMask == 19
    # WindingRule paths.nim:5-8, LineCap / LineJoin paths.nim:10-16
    assert (px.NonZero, px.EvenOdd) == (0, 1)
    assert (px.ButtCap, px.RoundCap, px.SquareCap) == (0, 1, 2)
    assert (px.MiterJoin, px.RoundJoin, px.BevelJoin) == (0, 1, 2)
