"""fillText / strokeText on the device against the oracle, bit for bit.

Every text test of the reference (test_fonts.nim, as tests/text_cases.py
states them) is drawn by the product -- typeset and span paths on the host,
fills and strokes on the device, through the C-ABI -- and by the oracle
(oracle/fonts.py paths drawn by liboracle), on the same canvas.
"""
import numpy as np
import pytest

import text_cases as tc

pytestmark = pytest.mark.gpu
pytest.importorskip("fontTools")


@pytest.fixture(scope="module")
def px():
    import pixie_amd
    pixie_amd.init()
    return pixie_amd


def _cmp(name, got, want):
    d = np.abs(got.astype(int) - want.astype(int))
    assert d.max() == 0, "%s: %d pixels differ (max %d), first %s" % (
        name, (d.max(-1) > 0).sum(), d.max(), np.argwhere(d.max(-1) > 0)[0])


@pytest.mark.parametrize("name", list(tc.CASES))
def test_text_cases(px, name):
    import oracle as orc
    print("text case", name, flush=True)
    _cmp(name, tc.render_product(name), tc.render_oracle(orc, name))


def test_empty_span_path_draws_nothing(px):
    """A span whose path is empty (one space) has no segment: nothing is
    drawn and no error raised, as the reference's spans6 test expects."""
    img = px.Image(100, 60)
    img.fill((9, 8, 7, 255))
    f = tc._p_font("Ubuntu-Regular_1.ttf", 20, {})
    a = px.typeset([px.Span(" ", f)])
    img.fillText(a)
    assert (img.download().reshape(-1, 4) == (9, 8, 7, 255)).all()


def test_transformed_paragraph(px):
    """A paragraph (device geometry: its path holds more than 4096 floats)
    rotated and scaled, filled then stroked with dashes."""
    import oracle as orc
    from oracle import fonts as of
    from pixie_amd.paths import sinf, cosf
    th = np.float32(0.3)
    c, s = cosf(th), sinf(th)
    m = np.array([c * np.float32(1.5), s * np.float32(1.5), 0, -s, c, 0, 40, -20, 1], np.float32)
    text = tc.PARAS[2]
    f = tc._p_font("NotoSans-Regular_4.ttf", 14, {})
    a = px.typeset(f, text, bounds=(700, 0))
    img = px.Image(1200, 700)
    img.fill((255, 255, 255, 255))
    img.fillText(a, transform=m)
    img.strokeText(a, transform=m, strokeWidth=0.75, dashes=(3.0, 1.5))
    o = of.OFont(tc.o_typeface("NotoSans-Regular_4.ttf"), 14)
    b = of.typeset([(text, o)], (700, 0))
    want = orc.new_image(1200, 700)
    orc.fill(want, (255, 255, 255, 255))
    (p,) = of.compute_paths(b)
    assert len(p) > 4096
    orc.fill_path(want, p, (0, 0, 0, 255), mat=m)
    orc.stroke_path(want, p, (0, 0, 0, 255), width=0.75, dashes=(3.0, 1.5), mat=m)
    _cmp("paragraph", img.download(), want)
