"""newImage(svg) on the device against the oracle, bit for bit: every
document of the reference's SVG tests (test_svg.nim) and synthetic documents
with gradients, opacities and hidden elements."""
import os

import numpy as np
import pytest

from test_svg_oracle import FILES, SIZE, SYNTH, data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def px():
    import pixie_amd
    pixie_amd.init()
    return pixie_amd


def _cmp(name, got, want):
    d = np.abs(got.astype(int) - want.astype(int))
    assert d.max() == 0, "%s: %d pixels differ (max %d), first %s" % (
        name, (d.max(-1) > 0).sum(), d.max(), np.argwhere(d.max(-1) > 0)[0])


def _oracle(doc, w=0, h=0):
    import oracle as orc
    from oracle import svg as osvg
    return osvg.render(orc, osvg.parse_svg(orc, doc, w, h))


@pytest.mark.parametrize("name", FILES)
def test_svg_documents(px, name):
    print("svg", name, flush=True)
    doc = data(name)
    s = px.parseSvg(doc, *SIZE.get(name, (0, 0)))
    img = px.svg.newImage(s)
    want = _oracle(doc, *SIZE.get(name, (0, 0)))
    _cmp(name, img.download(), want)
    # a second render of the same document (the scene replays) is identical
    from pixie_amd import lib, _check
    _check(lib().pxh_svg_render(s._h, img._h))
    _cmp(name + " again", img.download(), want)


@pytest.mark.parametrize("i", range(len(SYNTH)))
def test_synthetic_documents(px, i):
    doc = SYNTH[i].encode()
    for w, h in ((0, 0), (211, 157)):
        img = px.svg.newImage(px.parseSvg(doc, w, h))
        _cmp("synthetic %d %dx%d" % (i, w, h), img.download(), _oracle(doc, w, h))


def test_bad_fill_colour_raises(px):
    s = px.parseSvg('<svg viewBox="0 0 10 10"><rect width="5" height="5" fill="nocolour"/></svg>')
    with pytest.raises(px.PixieError, match="Invalid color"):
        px.svg.newImage(s)
