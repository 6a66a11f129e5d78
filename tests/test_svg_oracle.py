"""SVG ingestion on the host (parseSvg) and its oracle.

The product's parser (pixie_amd/csrc/pxh_svg.cpp through the C-ABI) is
compared with the oracle (oracle/svg.py: svg.nim restated over Python's
expat XML reader and Pillow's CSS colour table) element by element, bit for
bit.  The oracle's drawing of every document of the reference's SVG tests
(test_svg.nim) is scored against the reference's masters; those masters were
drawn by an older pixie whose coverage has 16 levels, so they pin geometry,
colours and order, not bits.
"""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SVG_DIR = os.path.join(ROOT, "tests", "golden", "svg")
FILES = ["line01", "polyline01", "polygon01", "rect01", "rect02", "circle01", "ellipse01",
         "triangle01", "quad01", "Ghostscript_Tiger", "scale", "miterlimit", "dashes", "dragon2",
         "accessibility-outline"]
SIZE = {"accessibility-outline": (512, 512)}  # test_svg.nim:24-31


@pytest.fixture(scope="module")
def px():
    import pixie_amd
    return pixie_amd


@pytest.fixture(scope="module")
def osvg():
    from oracle import svg
    return svg


def data(name):
    with open(os.path.join(SVG_DIR, name + ".svg"), "rb") as fh:
        return fh.read()


def same(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def rgbx_tuple(v):
    return tuple((v >> (8 * k)) & 255 for k in range(4))


def compare_docs(orc, osvg, px, doc, width=0, height=0):
    s = px.parseSvg(doc, width, height)
    o = osvg.parse_svg(orc, doc, width, height)
    assert (s.width, s.height) == (o["width"], o["height"])
    els = s.elements
    assert len(els) == len(o["elements"])
    for i, ((pc, pp), (oc, op)) in enumerate(zip(els, o["elements"])):
        assert same(pc, oc), i
        assert same(pp["transform"], op["transform"]), i
        assert pp["fill"] == op["fill"] and pp["fillRule"] == op["fillRule"], i
        assert rgbx_tuple(pp["stroke"]) == tuple(op["stroke"]), i
        for k in ("strokeWidth", "lineCap", "lineJoin", "miterLimit", "opacity", "fillOpacity",
                  "strokeOpacity", "display"):
            assert pp[k] == op[k], (i, k)
        assert same(pp["dashes"], op["dashes"]), i
    return s, o


@pytest.mark.parametrize("name", FILES)
def test_parse_matches_oracle(orc, osvg, px, name):
    compare_docs(orc, osvg, px, data(name), *SIZE.get(name, (0, 0)))
    compare_docs(orc, osvg, px, data(name), 333, 250)  # scaled to another size


SYNTH = [
    # nested groups, transforms of every kind, style overriding, inherited
    # dashes appended, opacities, display none, evenodd, rounded rects
    """<svg viewBox="-10 -20 300 200" xmlns="http://www.w3.org/2000/svg" fill="#0a0">
      <g transform="translate(5,6) rotate(30 40 50) scale(1.5, .5)" stroke="navy"
         stroke-dasharray="4 2" style="stroke-width:3px; fill-rule: evenodd">
        <g transform="matrix(1 0.25 -0.5 1 3 4)" stroke-dasharray="1" opacity="0.6">
          <path d="M 10 10 h 50 v 40 h -50 z m 10 10 h 20 v 10 h -20 z" fill-opacity="0.5"/>
          <rect x="1.5" y="2.25" width="80" height="30" rx="7" stroke-linejoin="round"/>
          <rect x="3" y="4" width="20" height="10" ry="3" stroke-linecap="square"/>
        </g>
        <ellipse cx="100" cy="50" rx="30.5" ry="10.25" stroke="currentColor" display="none"/>
        <circle cx="150" cy="80" r="12.5" fill="rgb(200, 100, 50)" stroke-opacity="0.3"/>
        <polygon points="1 1 40 1 20 30" fill="#fa0" stroke="none"/>
        <polyline points="0,0  10,10
                          20,0" fill="none" style="stroke:#123456;stroke-miterlimit:2"/>
        <line x1="0" y1="5" x2="100" y2="5" stroke-width="0"/>
      </g>
      <rect width="0" height="10"/>
      <!-- a comment -->
      <title>t</title><desc>d</desc><defs><rect width="5" height="5"/></defs>
    </svg>""",
    """<?xml version="1.0"?>
    <!DOCTYPE svg PUBLIC "-//W3C//DTD SVG 1.1//EN" "x.dtd">
    <svg viewBox="0 0 100 100" xmlns="http://www.w3.org/2000/svg">
      <linearGradient id="g1" gradientUnits="userSpaceOnUse" x1="0" y1="0" x2="100" y2="0">
        <stop offset="0" stop-color="#f00"/><stop offset="1" style="stop-color:blue"/>
      </linearGradient>
      <rect x="10" y="10" width="80" height="80" fill="url(#g1)" transform="rotate(10)"/>
      <circle cx="50" cy="50" r="20" fill="WHITE" fill-opacity="0.75" opacity="0.5"/>
    </svg>""",
]


@pytest.mark.parametrize("i", range(len(SYNTH)))
def test_synthetic_documents(orc, osvg, px, i):
    s, o = compare_docs(orc, osvg, px, SYNTH[i].encode())
    assert len(s) > 0


def test_named_colours(orc, osvg, px):
    from PIL import ImageColor
    names = sorted(n for n, v in ImageColor.colormap.items() if len(v) == 7)
    doc = '<svg viewBox="0 0 10 10">%s</svg>' % "".join(
        '<line x2="1" stroke="%s"/>' % n for n in names)
    s = px.parseSvg(doc)
    for n, (_, p) in zip(names, s.elements):
        assert rgbx_tuple(p["stroke"]) == osvg.color_rgbx(osvg.parse_color(n)), n
    for c in ("#abc", "#A1B2C3", "rgb(1, 2, 3)", "rgba(10, 20, 30, 0.5)", "#11223344", " red "):
        p = px.parseSvg('<svg viewBox="0 0 1 1"><line stroke="%s"/></svg>' % c).element(0)[1]
        assert rgbx_tuple(p["stroke"]) == osvg.color_rgbx(osvg.parse_color(c)), c


@pytest.mark.parametrize("doc,msg", [
    ('<svg viewBox="0 0 10 10"><foo/></svg>', "Unsupported SVG tag: foo"),
    ('<svg viewBox="0 0 10 10"><path d="M0 0" fill-rule="odd"/></svg>', "Invalid fill-rule"),
    ('<svg viewBox="0 0 10 10"><path d="M0 0" transform="skewX(3)"/></svg>', "transform"),
    ('<svg viewBox="0 0 10 10"><path d="M0 0" stroke-linecap="bad"/></svg>', "stroke-linecap"),
    ('<svg viewBox="0 0 10 10"><polygon points="1,2 3"/></svg>', "Invalid SVG data"),
    ('<svg><path d="M0 0"/></svg>', None),
    ('<svg viewBox="0 0 10.5 10"></svg>', "invalid integer"),
    ('<g viewBox="0 0 10 10"></g>', "Invalid SVG data"),
    ('<svg viewBox="0 0 10 10"><path d="M0 0"></svg>', "XML"),
    ('<svg viewBox="0 0 10 10"><circle r="x1"/></svg>', "invalid float"),
])
def test_errors(px, doc, msg):
    with pytest.raises(px.PixieError, match=msg):
        px.parseSvg(doc)


@pytest.mark.parametrize("name", FILES)
def test_oracle_svg_masters(masters, osvg, name):
    """The oracle's drawing, scored as the reference scores masters
    (images.nim:138-166), is within 2 and at least 1.5x closer than the same
    drawing moved by one pixel."""
    import oracle as orc
    img = orc.to_straight_alpha(osvg.render(orc, osvg.parse_svg(orc, data(name),
                                                                 *SIZE.get(name, (0, 0)))))
    m = masters("tests/fileformats/svg/masters/%s.png" % name, premultiply=False)
    score = orc.diff_score(m, img)
    moved = orc.diff_score(m, np.ascontiguousarray(np.roll(img, 1, axis=1)))
    assert score <= 2.0 and score * 1.5 <= moved, (name, score, moved)
