"""Text ingestion on the host (typeface parsing, glyph paths, typesetting,
span paths) and its oracle.

The product's host code (pixie_amd/csrc/pxh_font.cpp through the C-ABI) is
compared with the oracle (oracle/fonts.py: pixie's rules restated over
fontTools' independent table decoding) bit for bit.  The oracle is pinned by
the reference's own asserts (test_fonts.nim:22-34, 1043-1045), by closed forms
of the TrueType rules on a synthetic font, and by the reference's font
masters within anti-aliasing (the masters were drawn by an older pixie whose
coverage has 16 levels, so they pin glyph outlines and layout, not bits).
"""
import os
import random

import numpy as np
import pytest

import synthetic_font as sf
import text_cases as tc

pytest.importorskip("fontTools")

FONT_FILES = ["Roboto-Regular_1.ttf", "Ubuntu-Regular_1.ttf", "Aclonica-Regular_1.ttf",
              "IBMPlexSans-Regular_2.ttf", "NotoSans-Regular_4.ttf", "Pacifico-Regular_4.ttf",
              "PinyonScript.ttf", "Inter-Regular.ttf", "Inter-Bold.ttf", "NotoEmoji-subset.otf"]


@pytest.fixture(scope="module")
def px():
    import pixie_amd
    return pixie_amd


@pytest.fixture(scope="module")
def of():
    from oracle import fonts
    return fonts


def bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def same(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(bits(a), bits(b))


# ---- the reference's asserts ----

def test_layout_bounds_asserts(px, of):
    # test_fonts.nim:22-34
    f = px.readFont(os.path.join(tc.FONTS, "Roboto-Regular_1.ttf"))
    f.size = 24
    assert f.layoutBounds("Word") == (56, 28)
    assert f.layoutBounds("Word\n") == (56, 56)
    o = of.OFont(tc.o_typeface("Roboto-Regular_1.ttf"), 24)
    assert of.layout_bounds(of.typeset([("Word", o)])) == (56, 28)
    assert of.layout_bounds(of.typeset([("Word\n", o)])) == (56, 56)


def test_kerning_assert(px, of):
    # test_fonts.nim:1043-1045
    assert tc.p_typeface("Roboto-Regular_1.ttf").getKerningAdjustment("T", "e") == -99.0
    assert tc.o_typeface("Roboto-Regular_1.ttf").kerning(ord("T"), ord("e")) == -99.0


# ---- product vs oracle on the reference's fonts ----

@pytest.mark.parametrize("name", FONT_FILES)
def test_glyph_paths_cmap_advances(px, of, name):
    p, o = tc.p_typeface(name), tc.o_typeface(name)
    runes = sorted(o.cmap)
    for r in range(0x10000):
        assert p.hasGlyph(r) == o.has_glyph(r), r
    for r in runes + [0x10FFFF, 0xFFFF]:
        assert same(p.getGlyphPath(r), o.glyph_path(r)), (name, r)
        assert p.getAdvance(r) == o.advance(r), (name, r)
    m = [p.scale, p.ascent, p.descent, p.lineGap, p.capHeight, p.underlinePosition,
         p.underlineThickness, p.strikeoutPosition, p.strikeoutThickness]
    assert m == [o.unitsPerEm, o.ascent, o.descent, o.lineGap, o.capHeight, o.underlinePosition,
                 o.underlineThickness, o.strikeoutPosition, o.strikeoutSize]


@pytest.mark.parametrize("name", ["Inter-Regular.ttf", "Inter-Bold.ttf", "NotoEmoji-subset.otf"])
def test_cff_outlines_match_fonttools(of, name):
    """The oracle's CFF charstring interpreter (a restatement of
    opentype.nim:1028-1297) against fontTools' own: on every mapped glyph the
    same points in the same order.  Where a contour is closed differs by
    design (fontTools closes each one; pixie only at endchar), so only the
    points are compared."""
    from fontTools.pens.recordingPen import RecordingPen
    o = tc.o_typeface(name)
    gs = o.font.getGlyphSet()
    count = {0: 0, 1: 2, 2: 2, 5: 6, 7: 4}
    for r in sorted(o.cmap):
        pen = RecordingPen()
        gs[o.order[o.glyph_id(r)]].draw(pen)
        want = [c for _, args in pen.value for pt in args for c in (float(pt[0]), float(pt[1]))]
        got, cmds, k = [], list(o.glyph_path(r)), 0
        while k < len(cmds):
            n = count[int(cmds[k])]
            for j in range(k + 1, k + 1 + n, 2):
                got += [float(cmds[j]), -float(cmds[j + 1])]  # back to y up
            k += 1 + n
        assert got == want, (name, r)


@pytest.mark.parametrize("name", FONT_FILES)
def test_kerning_pairs(px, of, name):
    p, o = tc.p_typeface(name), tc.o_typeface(name)
    rng = random.Random(hash(name) & 0xFFFF)
    latin = [r for r in sorted(o.cmap) if r < 0x250] or sorted(o.cmap)  # the emoji font has none
    nonzero = 0
    for _ in range(6000):
        a, b = rng.choice(latin), rng.choice(latin)
        k = p.getKerningAdjustment(a, b)
        assert k == o.kerning(a, b), (name, a, b)
        nonzero += k != 0
    if name not in ("PinyonScript.ttf", "NotoEmoji-subset.otf"):  # fonts without pair adjustments
        assert nonzero > 0


def _random_spans(rng, px, of, n):
    words = ("She had come to the conclusion that you could tell a lot about a person by their "
             "ears The way they stuck out\nand the size of the earlobes, Supercalifragilistic! "
             "V( V) AV To Te Wa àéîõü ÀÉÎÕÜ ÿ µ").split(" ")
    ps, os_ = [], []
    for _ in range(n):
        name = rng.choice(FONT_FILES)
        size = rng.choice([8, 12, 16, 18.5, 24, 36, 50, 72, 100])
        opts = dict(lineHeight=rng.choice([-1, -1, 20, 30, 60.5, 100]), underline=rng.random() < 0.2,
                    strikethrough=rng.random() < 0.2, noKerningAdjustments=rng.random() < 0.3)
        case = rng.choice([0, 0, 0, 1, 2, 3])
        text = " ".join(rng.choice(words) for _ in range(rng.randint(1, 12)))
        if rng.random() < 0.2:
            text += "\n"
        f = tc._p_font(name, size, opts)
        f.textCase = case
        o = of.OFont(tc.o_typeface(name), size, opts["lineHeight"], opts["underline"],
                     opts["strikethrough"], opts["noKerningAdjustments"], case)
        ps.append(px.Span(text, f))
        os_.append((text, o))
    return ps, os_


def test_typeset_and_span_paths_random(px, of):
    rng = random.Random(7)
    for it in range(150):
        ps, os_ = _random_spans(rng, px, of, rng.randint(1, 4))
        bounds = (rng.choice([0, 75, 200, 360, 1000]), rng.choice([0, 60, 100, 400]))
        h, v, wrap = rng.randint(0, 2), rng.randint(0, 2), rng.random() < 0.9
        a = px.typeset(ps, bounds=bounds, hAlign=h, vAlign=v, wrap=wrap)
        b = of.typeset(os_, bounds, h, v, wrap)
        assert a.runes == b.runes, it
        assert a.lines == list(b.lines), it
        assert a.spans == b.spans, it
        assert same(a.positions, np.array(b.positions, np.float32).reshape(-1, 2)), it
        assert same(a.selectionRects, np.array(b.rects, np.float32).reshape(-1, 4)), it
        assert a.layoutBounds() == of.layout_bounds(b), it
        pa, pb = a.computePaths(), of.compute_paths(b)
        assert len(pa) == len(pb)
        for x, y in zip(pa, pb):
            assert same(x, y), it


def test_compute_bounds(px, of):
    import oracle as orc
    rng = random.Random(3)
    for it in range(20):
        ps, os_ = _random_spans(rng, px, of, rng.randint(1, 3))
        a = px.typeset(ps, bounds=(300, 0))
        b = of.typeset(os_, (300, 0))
        mat = None if it % 2 else [1, 0.25, 0, -0.5, 1, 0, 7.5, -3, 1]
        got = a.computeBounds(None if mat is None else np.array(mat, np.float32))
        assert same(got, of.compute_bounds(orc, b, mat)), it


def test_empty_arrangement(px, of):
    f = tc._p_font("Roboto-Regular_1.ttf", 24, {})
    a = px.typeset(f, "\t\r\x01")  # control runes other than LF are dropped
    assert a.runes == [] and a.lines == [] and a.layoutBounds() == (0, 0)
    assert a.computePaths() == []


def test_text_case(px, of):
    f = tc._p_font("Roboto-Regular_1.ttf", 24, {})
    text = "hello wörld ÿ µ Àb\nnew LINE"
    want = {1: "HELLO WÖRLD Ÿ Μ ÀB\nNEW LINE", 2: "hello wörld ÿ µ àb\nnew line",
            3: "Hello Wörld Ÿ Μ Àb\nNew LINE"}
    for case, w in want.items():
        f.textCase = case
        assert "".join(chr(r) for r in px.typeset(f, text).runes) == w


def test_utf8_decoding(px):
    f = tc._p_font("Roboto-Regular_1.ttf", 24, {})
    text = "aé€\U0001F600z"
    assert px.typeset(f, text).runes == [ord(c) for c in text]


# ---- the synthetic font: closed forms of the TrueType rules ----

@pytest.fixture(scope="module")
def synth(px, of):
    data = sf.build()
    return px.parseTtf(data), of.OTypeface(data), data


def _flip(cmds):
    out, i = [], 0
    while i < len(cmds):
        k = int(cmds[i])
        n = {0: 0, 1: 2, 2: 2, 7: 4}[k]
        out.append(k)
        for j in range(n):
            v = cmds[i + 1 + j]
            out.append(-v if j % 2 else v)
        i += 1 + n
    return out


def _alloff_path(xf=lambda x, y: (x, y)):
    # every point off-curve: on-curve points implied at each midpoint, the
    # contour starting at the midpoint of its last and first points
    p = [xf(x, y) for x, y in sf.ALLOFF]
    mid = lambda a, b: ((a[0] + b[0]) / 2, (a[1] + b[1]) / 2)  # noqa: E731
    cmds = [1, *mid(p[3], p[0])]
    for i in range(4):
        cmds += [7, *p[i], *mid(p[i], p[(i + 1) % 4])]
    return cmds + [0]


def test_synthetic_closed_forms(synth):
    p, o, _ = synth
    assert np.array_equal(p.getGlyphPath(0x43), np.array(_flip(_alloff_path()), np.float32))
    (a, b), (c, d), (dx, dy) = sf.TWO_BY_TWO  # x' = a x + c y + dx, y' = b x + d y + dy
    want = _flip(_alloff_path(lambda x, y: (a * x + c * y + dx, b * x + d * y + dy)))
    assert np.array_equal(p.getGlyphPath(0x46), np.array(want, np.float32))
    # a contour starting off-curve starts at its first on-curve point
    got = p.getGlyphPath(0x42)
    assert list(got[:3]) == [1, 300, -720] and got[-1] == 0
    # matched-point arguments place the component at the origin
    assert np.array_equal(p.getGlyphPath(0x48)[-len(_alloff_path()):],
                          np.array(_flip(_alloff_path()), np.float32))
    assert len(p.getGlyphPath(0x20)) == 0  # empty glyph
    assert len(p.getGlyphPath(0x1F600)) > 0  # cmap format 12
    # the kern table's accumulation (opentype.nim:838-855): after each
    # subtable every subtable so far is applied again, so (sq, offstart) is
    # -40 | -80 - 7 = -87 | -87 - 40 - 7 = -134, (offstart, sq) 25 | 50 | 75,
    # and the override subtable sets (c_xy, sq) to 3
    assert p.getKerningAdjustment(0x41, 0x42) == -134
    assert p.getKerningAdjustment(0x42, 0x41) == 75
    assert p.getKerningAdjustment(0x45, 0x41) == 3
    assert p.getKerningAdjustment(0x43, 0x43) == 0


def test_synthetic_vs_oracle(synth):
    p, o, _ = synth
    for r in list(sf.CMAP) + [0x99]:
        assert same(p.getGlyphPath(r), o.glyph_path(r)), hex(r)
        assert p.getAdvance(r) == o.advance(r)
        for r2 in sf.CMAP:
            assert p.getKerningAdjustment(r, r2) == o.kerning(r, r2)


def test_fallbacks(px, of, synth):
    _, _, data = synth
    p, o = px.parseTtf(data), of.OTypeface(data)  # unitsPerEm 1000
    rp, ro = tc.p_typeface("Roboto-Regular_1.ttf"), tc.o_typeface("Roboto-Regular_1.ttf")
    p.fallbacks.append(rp)
    o.fallbacks.append(ro)
    for r in [0x41, 0x61, 0x62, 0x20, 0x2603]:
        assert same(p.getGlyphPath(r), o.glyph_path(r)), r
        assert p.getAdvance(r) == o.advance(r)
    for a, b in [(0x41, 0x42), (0x54, 0x65), (0x41, 0x65)]:
        assert p.getKerningAdjustment(a, b) == o.kerning(a, b)
    assert p.getKerningAdjustment(0x54, 0x65) != 0  # both from the fallback
    assert p.getKerningAdjustment(0x41, 0x65) == 0  # different typefaces
    # a glyph taken from the 2048-unit fallback is scaled by 1000 / 2048
    ratio = np.float32(1000) / np.float32(2048)
    g = rp.getGlyphPath(0x61)
    assert p.getGlyphPath(0x61)[1] == np.float32(g[1] * ratio)


def test_parse_errors(px, synth):
    _, _, data = synth
    with pytest.raises(px.PixieError, match="EOF"):
        px.parseTtf(data[:200])
    with pytest.raises(px.PixieError, match="Missing required font table"):
        px.parseTtf(sf.rename_table(data, b"hhea", b"xxxx"))
    with pytest.raises(px.PixieError, match="glyph outlines"):
        px.parseTtf(sf.rename_table(data, b"glyf", b"xxxx"))
    with pytest.raises(px.PixieError, match="EOF|CFF"):  # glyf bytes are no CFF table
        px.parseTtf(sf.rename_table(data, b"glyf", b"CFF "))
    with pytest.raises(px.PixieError):
        px.readTypeface("font.woff")


def test_collection(px, synth, tmp_path):
    from fontTools.ttLib import TTCollection, TTFont
    import io
    _, _, data = synth
    col = TTCollection()
    col.fonts = [TTFont(io.BytesIO(data)), TTFont(io.BytesIO(tc.p_typeface("Roboto-Regular_1.ttf")._data))]
    path = str(tmp_path / "two.ttc")
    col.save(path)
    faces = px.readTypefaces(path)
    assert len(faces) == 2
    assert faces[0].scale == 1000 and faces[1].scale == 2048
    assert faces[1].name == "Roboto-Regular"
    assert same(faces[1].getGlyphPath(0x61), tc.p_typeface("Roboto-Regular_1.ttf").getGlyphPath(0x61))


# ---- the oracle against the reference's masters ----

SPARSE = {"tofu_advance"}
PINS = [n for n in tc.CASES if n not in tc.NOT_PINS and os.path.exists(
    os.path.join(tc.ROOT, "tests", "golden", "masters", tc.master_name(n).replace("/", "_")))]


@pytest.mark.parametrize("name", PINS)
def test_oracle_text_masters(masters, name):
    """The oracle's drawing of the reference's text test, scored as the
    reference scores masters (images.nim:138-166), must be closer to the
    master than the same drawing moved by one pixel is, by at least 1.5x,
    and within a score of 2 (percent of the maximum difference)."""
    import oracle as orc
    img = orc.to_straight_alpha(tc.render_oracle(orc, name))
    m = masters(tc.master_name(name), premultiply=False)
    score = orc.diff_score(m, img)
    moved = orc.diff_score(m, np.ascontiguousarray(np.roll(img, 1, axis=1)))
    if name in SPARSE:
        # little ink on a white canvas: the master's coarser anti-aliasing
        # weighs as much as a shift, so the drawing must beat every shift of
        # 1-4 px by 10 %
        for k in (1, 2, 3, 4):
            mk = orc.diff_score(m, np.ascontiguousarray(np.roll(img, k, axis=1)))
            assert score <= 2.0 and score * 1.1 <= mk, (name, k, score, mk)
        return
    assert score <= 2.0 and score * 1.5 <= moved, (name, score, moved)
