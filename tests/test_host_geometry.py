"""CPU tests of libpixie_hip's host geometry against the oracle.

The product flattens, strokes, builds segments and y-bands on the host
(paths.nim:654-1262, 1922-2082) and hands the bands to the raster kernel.
These tests compare that geometry with the oracle's restatement field by
field: fill locals (paths.nim:1602-1619), band layout and flags, entry
(at.y, to.y, m, b) and windings, all bit for bit.  No GPU is needed.
"""
import math

import numpy as np
import pytest

import pixie_amd as px
from pixie_amd import workloads as WL
from pixie_amd.paths import Path, translate, scale


def _compare(orc, cmds, w, h, mat=None, stroke=False, width=1.0, cap=0, join=0, miter=4.0):
    info, parts, ents, wind = px.debug_fill_geometry(cmds, w, h, transform=mat, stroke=stroke,
                                                     width=width, cap=cap, join=join, miter=miter)
    if stroke:
        shapes = orc.stroke_shapes(cmds, mat=mat, width=width, cap=cap, join=join,
                                   miter_limit=miter)
    else:
        shapes = orc.flatten(cmds, mat=mat, close=True)
    P, E, bounds = orc.partitions(shapes, w, h)
    assert tuple(info[:4]) == tuple(bounds), (info[:4], bounds)
    if bounds[2] <= 0:
        return 0
    assert len(parts) == len(P)
    # band tops/bottoms from the layout
    startY, nP, bandH, pathH = info[1], info[4], info[5], info[3]
    tops = [startY + i * bandH for i in range(nP)]
    bots = [startY + (i + 1) * bandH for i in range(nP)]
    bots[-1] = pathH
    assert np.array_equal(P[:, 0], tops)
    assert np.array_equal(P[:, 1], bots)
    assert np.array_equal(parts[:, 1], P[:, 2]), "entry counts differ"
    assert np.array_equal(parts[:, 2], P[:, 3]), "band flags differ"
    assert len(ents) == len(E)
    # bit-exact float comparison
    assert np.array_equal(ents.view(np.uint32), E[:, :4].astype(np.float32).view(np.uint32))
    assert np.array_equal(wind.astype(np.int32), E[:, 6].astype(np.int32))
    return len(E)


def test_parse_matches_oracle(orc):
    for s in ["M 10 10 H 90 V 90 H 10 L 10 10",
              "M 10,30 A 20,20 0,0,1 50,30 A 20,20 0,0,1 90,30 Q 90,60 50,90 Q 10,60 10,30 z",
              "m 1 2 3 4 5 6", "M 0.1E-10 0.1e10 L2+2 L3-3 L0.1E+10-1",
              "M1 0.5C1 0.776142 0.776142 1 0.5 1C0.223858 1 0 0.776142 0 0.5Z",
              "M 3 3 L 20 3 L 20 20 L 3 20 Z", "a 7 7 7 7 7 7 7 z", ""]:
        a = px.parse_path(s)
        b = orc.parse_path(s)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), s
    with pytest.raises(px.PixieError):
        px.parse_path("M 1 2 3")


@pytest.mark.parametrize("path", [
    "M 10 10 H 90 V 90 H 10 L 10 10",
    "M 10.1 10.1 H 60.1 V 60.1 H 10.1 z",
    "M30 60 A 20 20 0 0 0 90 60 L 30 60",
    "M 20 50 A 20 10 45 1 1 80 50 L 20 50",
    "M 40 40 L 40 80 L 80 80 L 80 40 C 80 -20 40 100 40 40",
    "M99 99 L999 99 L999 100 L99 100 Z",
    "M0 0 L0 0 L60 0 L60 60 L0 60",
    "M 10,30 A 20,20 0,0,1 50,30 A 20,20 0,0,1 90,30 Q 90,60 50,90 Q 10,60 10,30 z",
])
def test_fill_geometry_reference_paths(orc, path):
    _compare(orc, px.parse_path(path), 100, 100)


def test_fill_geometry_transformed(orc):
    p = px.parse_path("M1 0.5C1 0.776142 0.776142 1 0.5 1C0.223858 1 0 0.776142 0 0.5"
                      "C0 0.223858 0.223858 0 0.5 0C0.776142 0 1 0.223858 1 0.5Z")
    _compare(orc, p, 200, 200, mat=scale(200, 200))
    _compare(orc, p, 200, 200, mat=scale(200, 200), stroke=True, width=0.01)


@pytest.mark.parametrize("cap,join", [(0, 0), (1, 1), (2, 2), (0, 2), (1, 0)])
def test_stroke_geometry(orc, cap, join):
    _compare(orc, px.parse_path("M 15 10 L 30 90 60 30 90 90"), 100, 100, stroke=True,
             width=10, cap=cap, join=join)
    _compare(orc, px.parse_path("M 3 3 L 20 3 L 20 20 L 3 20 Z"), 60, 60, mat=translate(10, 10),
             stroke=True, width=10, cap=cap, join=join)


def test_star_and_mixed_geometry(orc):
    cmd, w, h = WL.star512()
    _compare(orc, cmd.cmds, w, h)
    cmds, _ = WL.mixed_scene(n_paths=40, W=1024, H=1024)
    n = 0
    for c in cmds:
        if c.stroke:
            n += _compare(orc, c.cmds, 1024, 1024, stroke=True, width=c.width, cap=c.cap,
                          join=c.join, miter=c.miter)
        else:
            n += _compare(orc, c.cmds, 1024, 1024)
    assert n > 1000


def test_glyph_geometry(orc):
    rng = np.random.default_rng(4)
    for _ in range(20):
        p = WL.glyph_path(rng, int(rng.integers(8, 128)), int(rng.integers(8, 128)))
        _compare(orc, p.array(), 256, 256)


def test_empty_path_draws_nothing(orc):
    # an all-horizontal or empty path has no segment; computeBounds keeps its
    # infinite initial extremes.  The reference's tests fill such paths and
    # expect no error and nothing drawn (test_svg.nim's dashes.svg and tiger,
    # test_fonts.nim's spans6), so both the product and the oracle draw
    # nothing; finite bounds beyond the int range still raise.
    for path in ("M 0 0 L 10 0 Z", "M 5 5", ""):
        info, parts, entries, _ = px.debug_fill_geometry(path, 100, 100)
        assert info[2] == 0 and len(parts) == 0 and len(entries) == 0, (path, info)
        img = orc.new_image(100, 100)
        orc.fill(img, (1, 2, 3, 4))
        for mode in (0, 16, 17):
            orc.fill_path(img, path, (255, 0, 0, 255), mode=mode)
        assert (img.reshape(-1, 4) == (1, 2, 3, 4)).all()
    with pytest.raises(px.PixieError, match="overflow"):
        px.debug_fill_geometry("M 1e30 0 L -1e30 0 L 0 10 Z", 100, 100)
    with pytest.raises(orc.OracleError, match="overflow"):
        orc.fill_path(orc.new_image(100, 100), "M 1e30 0 L -1e30 0 L 0 10 Z", (255, 0, 0, 255))
