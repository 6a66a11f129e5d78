"""The reference's text tests (tests/test_fonts.nim) as data, rendered by the
oracle (CPU) or by the product (device).

A case is (master name, canvas w, h, ops).  Ops run in order on a canvas
filled white:
  ("text", spans, typeset kwargs, transform, stroke) -- spans are
      (text, font file, size, font options) with font options lineHeight,
      underline, strikethrough, noKerningAdjustments and paints (a list of
      solid (r, g, b, a) float colours or ("linear", handles, stops));
  ("rects", color) -- fillPath of each selection rect of the previous text op
      (Context.fillRect, contexts.nim:392-398);
  ("rect", x, y, w, h, color) -- a filled rectangle;
  ("bounded", spans, kwargs, dx, dy) -- spans7 (test_fonts.nim:981-1016): the
      arrangement rendered into an image of its snapped computeBounds size,
      then drawn at the bounds' corner plus (dx, dy).
"""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FONTS = os.path.join(ROOT, "tests", "golden", "fonts")
f32 = np.float32

R, U, A, I, N = ("Roboto-Regular_1.ttf", "Ubuntu-Regular_1.ttf", "Aclonica-Regular_1.ttf",
                 "IBMPlexSans-Regular_2.ttf", "NotoSans-Regular_4.ttf")
P = "PinyonScript.ttf"
IN, IB, EMO = "Inter-Regular.ttf", "Inter-Bold.ttf", "NotoEmoji-subset.otf"  # CFF outlines
# the emoji text of test_fonts.nim:7-20
EMOJI_TEXT = ('🚑🐑👭🔉🚷🦣💆🔁💺🚵🕦🔦🗓🦟😶🦄⌛🍙😄🇽\n'
              '🐠💓🦗🎭🏛🔴🫕🧶🍖🦁🏋🌗🛬🕐💡👉🎯🕔🚏🚲\n'
              '🐵🎍💳🥬🟦🪘📠📊🎧🎦🎁🌌🪲🦩🤢☎🚺🚾👺🚃\n'
              '🐨🌆🥉💭🗳🦵🟪📆🥮⏯🩴💷🦲➗🌶🧜🖖⏰🛗🔻\n'
              '📁🧞😃🌴🚶\U000fe82b🦙🔎⏲🔵🖐☦😪🌯🙆🇺😂🍅🇿🚟🤜\n'
              '📼👰🍁📽☪🔄🤝🔧🦸🏰🏳🔜🎥🚋🇫🦨🏜🆖🏤🪖⏏')
BLACK = (0.0, 0.0, 0.0, 1.0)


def rgba(r, g, b, a):
    """pixie's ColorRGBA -> Color: each byte / 255 in float32."""
    return tuple(float(f32(v) / f32(255)) for v in (r, g, b, a))


def hexc(s):
    return rgba(int(s[1:3], 16), int(s[3:5], 16), int(s[5:7], 16), 255)


def T(text, font, size, bounds=None, stroke=False, mat=None, **kw):
    """One text op over one span."""
    tk = {} if bounds is None else {"bounds": bounds}
    for k in ("hAlign", "vAlign"):
        if k in kw:
            tk[k] = kw.pop(k)
    return ("text", [(text, font, size, kw)], tk, mat, stroke)


def S(spans, bounds, mat=None):
    return ("text", spans, {"bounds": bounds}, mat, False)


W8 = "Wrapping text to new line"
PARAS = [
    "ShehadcometotheconclusionthatyoucouldtellalotaboutapersonbytheirearsThewaytheystuckout"
    "andthesizeoftheearlobescouldgiveyou",
    "She had come to the conclusion that you could tell a lot about a person by their ears The "
    "way they stuck out and the size of the earlobes could give you wonderful insights into the "
    "person Of course she couldnt scientifically prove any of this but that didnt matter to her "
    "Before anything else she would size up the ears of the person she was talking to Shes asked "
    "the question so many times that she barely listened to the answers anymore The answers were "
    "always the same Well not exactly the same but the same in a general sense A more accurate "
    "description was the answers never surprised her",
    "She had come to the conclusion that you could tell a lot about a person by their ears The "
    "way they stuck out and the size of the earlobes could give you wonderful insights into the "
    "person. Of course, she couldn't scientifically prove any of this, but that didn't matter to "
    "her. Before anything else, she would size up the ears of the person she was talking to. "
    "She's asked the question so many times that she barely listened to the answers anymore. The "
    "answers were always the same. Well, not exactly the same, but the same in a general sense. "
    "A more accurate description was the answers never surprised her.",
]
ROBO_SPANS = [("One span ", R, 80, {}), ("Two span", A, 100, {}), (" Three span", U, 48, {})]
GREY = rgba(127, 127, 127, 127)


def _cases():
    c = {}
    c["basic1"] = (200, 100, [T("asdf", R, 72)])
    c["basic2"] = (200, 100, [T("A cow", R, 72)])
    c["basic3"] = (200, 100, [T("A bit of text HERE", R, 24)])
    c["basic4"] = (200, 100, [T("Line height", R, 24, lineHeight=100)])
    c["basic5"] = (200, 100, [T("Another font", U, 24)])
    c["basic6"] = (200, 100, [T("Different font", A, 24)])
    c["basic7"] = (200, 100, [T("First line", R, 24),
                              ("text", [("Second line", R, 24, {})], {}, "defaultLineHeight", False)])
    c["basic8"] = (200, 100, [T(W8, R, 24, (200, 0))])
    c["basic8b"] = (200, 100, [T("Long words: Supercalifragilisticexpialidocious\n"
                                 "Antidisestablishmentarianism", R, 24, (100, 0))])
    c["basic9"] = (200, 100, [T("Supercalifragilisticexpialidocious", R, 24, (200, 0))])
    c["basic10"] = (200, 100, [T("a b c d e f g h i j k l m n o p", R, 24, (200, 0), hAlign=2)])
    c["image_fill"] = (200, 100, [T("fill", R, 64)])
    c["image_stroke"] = (200, 100, [T("stroke", R, 64, stroke=True)])
    c["image_paint_fill"] = (100, 100, [T("Text", I, 48, paints=[
        ("linear", [(0, 50), (100, 50)], [((1, 0, 0, 1), 0.0), ((1, 0, 0, 0.5), 1.0)])])])
    for k, (deco, stroke, paint) in {
            "underline1": ("underline", False, BLACK), "underline2": ("underline", False, rgba(0, 0, 0, 127)),
            "underline3": ("underline", True, BLACK), "strikethrough1": ("strikethrough", False, BLACK),
            "strikethrough2": ("strikethrough", False, rgba(0, 0, 0, 127)),
            "strikethrough3": ("strikethrough", True, BLACK)}.items():
        c[k] = (200, 100, [T(W8, R, 24, (200, 0), stroke=stroke, paints=[paint], **{deco: True})])
    # CFF outlines (test_fonts.nim:7-20, 1051-1089, 1130-1138)
    grumpy = "Grumpy wizards make toxic brew for the evil Queen and Jack."
    c["emoji"] = (800, 300, [T(EMOJI_TEXT, EMO, 26)])
    c["cff"] = (800, 100, [T(grumpy, IN, 26)])
    c["cff_underline"] = (800, 100, [T(grumpy, IN, 26, underline=True)])
    c["cff_strikethrough"] = (800, 100, [T(grumpy, IN, 26, strikethrough=True)])
    c["tofu_advance"] = (800, 100, [T("This[]Advance!", IN, 26)])
    c["paints1"] = (200, 100, [T("Multiple fills", R, 36, (200, 0),
                                 paints=[rgba(0, 0, 255, 127), rgba(255, 0, 0, 127)])])
    names = [["TopLeft", "TopCenter", "TopRight"], ["MiddleLeft", "MiddleCenter", "MiddleRight"],
             ["BottomLeft", "BottomCenter", "BottomRight"]]
    c["alignments"] = (800, 800, [T(names[v][h], R, 36, (800, 800), hAlign=h, vAlign=v)
                                  for v in range(3) for h in range(3)])
    c["spans1"] = (700, 250, [S(ROBO_SPANS, (700, 250))])
    c["spans2"] = (475, 400, [S(ROBO_SPANS, (475, 400))])
    c["selection_rects1"] = (700, 250, [S(ROBO_SPANS, (700, 250)), ("rects", GREY)])
    c["selection_rects2"] = (475, 400, [S(ROBO_SPANS, (475, 400)), ("rects", GREY)])
    c["selection_rects3"] = (75, 75, [T("Wrapping text to the next line", R, 16, (75, 75)),
                                      ("rects", GREY)])
    c["spans4"] = (600, 600, [S([
        ("Using spans, ", R, 64, {}), ("Pixie ", A, 80, {}), ("can arrange and rasterize ", I, 40, {}),
        ("very complex text layouts. ", U, 56, {}), ("Spans", N, 72, {}),
        (" can have different ", R, 48, {}), ("font sizes,", N, 64, {}),
        (" colors", U, 54, {"paints": [(1.0, 0.0, 0.0, 1.0)]}), (" and ", R, 48, {}),
        ("line heights.", A, 48, {"lineHeight": 120})], (600, 600))])
    c["spans5"] = (400, 400, [S([
        ("verb [with object] ", U, 15, {"paints": [hexc("#CACACA")]}), ("strallow\n", U, 84, {}),
        ("\nstral·low\n", U, 18, {"paints": [hexc("#007FF4")]}),
        ("\n1. free (something) from restrictive restrictions \"the regulations are intended to "
         "strallow changes in public policy\" ", U, 20, {"paints": [hexc("#4F4F4F")]})],
        (360, 360), mat=[1, 0, 0, 0, 1, 0, 20, 20, 1])])
    c["spans6"] = (400, 400, [S([
        ("verb [with object] ", U, 15, {"paints": [hexc("#CACACA")]}), ("strallow\n", U, 84, {}),
        ("\nstral·low\n", U, 18, {"paints": [hexc("#007FF4")]}),
        ("\n1. free (something) from ", U, 20, {"paints": [hexc("#4F4F4F")]}),
        ("restrictive restrictions", U, 20, {"paints": [hexc("#4F4F4F")], "underline": True}),
        (" ", U, 20, {"paints": [hexc("#4F4F4F")]}),
        ("\"the regulations are intended to strallow changes in public policy\" ", U, 20,
         {"paints": [hexc("#4F4F4F")], "strikethrough": True})],
        (360, 360), mat=[1, 0, 0, 0, 1, 0, 20, 20, 1])])
    c["spans7"] = (400, 400, [("rect", 40, 170, 320, 60, hexc("#FFD6D6")),
                              ("bounded", [("Fancy text", P, 82, {"lineHeight": 60})],
                               {"bounds": (320, 60)}, 40, 170)])
    c["lines1"] = (200, 150, [T("First line\nSecond line\nThird line\nFourth line\nFifth line\n"
                                "Sixth line\nSeventh line", R, 18, (200, 150))])
    c["lines2"] = (200, 150, [T("First line\nSecond line\nThird line\nFourth line\nFifth line",
                                R, 18, (200, 150), lineHeight=30)])
    for font, base in ((R, "paragraph1"), (U, "paragraph2"), (N, "paragraph4")):
        for nk in (False, True):
            for j, text in enumerate(PARAS):
                name = base + ("_nokern" if nk else "") + ("" if j == 0 else "_%d" % (j + 1))
                c[name] = (1000, 150, [T(text, font, 16, (1000, 150), noKerningAdjustments=nk)])
    for font, base in ((R, "huge1"), (U, "huge2")):
        for nk in (False, True):
            c[base + ("_nokern" if nk else "")] = (1400, 200, [T(
                "Shehadcometotheconclusion", font, 100, (1400, 200), noKerningAdjustments=nk)])
    c["pairs1"] = (1400, 100, [T("HA HT HX HY IA IT IX IY MA MT MX MY NA NT NX NY", R, 50,
                                 (1400, 100))])
    c["huge3"] = (1400, 200, [T("Wrapping text to the next line", R, 100, (1400, 200))])
    for font, base in ((I, "paragraph3"), ("Pacifico-Regular_4.ttf", "paragraph5")):
        for j, text in enumerate(PARAS):
            c[base + ("" if j == 0 else "_%d" % (j + 1))] = (1000, 150, [T(text, font, 16, (1000, 150))])
    c["pairs2"] = (1400, 100, [T("V( V) V- V/ V: v; v? v@ VT VV VW VX VY V] Vu Vz V{", U, 50,
                                 (1400, 100))])
    c["pairs3"] = (1400, 100, [T("B, B. BA BJ BT BW BY Bf Bg Bt bw By", I, 50, (1400, 100))])
    return c


CASES = _cases()

# Masters made by an older pixie whose layout differs from v6's typeset (text
# 11 px lower; a Roboto / Pacifico kerning that drifts by ~0.04 px a glyph;
# IBM Plex baselines 1 px lower at 16 px; selection rects as tall as the
# whole line rather than each span's font): not pins.  Their text is still
# rendered by the GPU tests against the oracle.
NOT_PINS = {"pairs1", "paragraph1_2", "paragraph1_3", "paragraph1_nokern_2",
            "paragraph1_nokern_3", "huge3", "paragraph3", "paragraph3_2", "paragraph3_3",
            "paragraph5", "paragraph5_2", "paragraph5_3", "selection_rects1",
            "selection_rects2"}


def master_name(name):
    return "tests/fonts/masters/%s.png" % name


def r_rect(x, y, w, h):
    x, y, w, h = f32(x), f32(y), f32(w), f32(h)
    return np.array([1, x, y, 2, f32(x + w), y, 2, f32(x + w), f32(y + h), 2, x, f32(y + h), 0],
                    np.float32)


def snap(b):
    """snapToPixels (common.nim:92-101) in float32."""
    x, y, w, h = [f32(v) for v in b]
    xmin, xmax, ymin, ymax = x, f32(x + w), y, f32(y + h)
    rx, ry = f32(np.floor(xmin)), f32(np.floor(ymin))
    return rx, ry, f32(f32(np.ceil(xmax)) - rx), f32(f32(np.ceil(ymax)) - ry)


# ---- the oracle's rendering ----

_otf = {}


def o_typeface(name):
    from oracle import fonts as of
    if name not in _otf:
        _otf[name] = of.load(os.path.join(FONTS, name))
    return _otf[name]


def _o_font(font, size, opts):
    from oracle import fonts as of
    return of.OFont(o_typeface(font), size, opts.get("lineHeight", -1.0), opts.get("underline", False),
                    opts.get("strikethrough", False), opts.get("noKerningAdjustments", False))


def _o_paint(orc, img, p, paint, mat, stroke):
    if isinstance(paint, tuple) and paint and paint[0] == "linear":
        _, handles, stops = paint
        orc.fill_path_paint(img, p, orc.make_paint(orc.LINEAR, handles, stops), mat=mat)
        return
    if not paint[3] > 0:  # fillPath / strokePath draw nothing (paths.nim:2106, 2161)
        return
    c = orc.color_to_rgbx(*paint)
    if stroke:
        orc.stroke_path(img, p, c, mat=mat)
    else:
        orc.fill_path(img, p, c, mat=mat)


def o_text(orc, img, spans, tk, mat, stroke):
    from oracle import fonts as of
    fonts = [_o_font(f, size, opts) for _, f, size, opts in spans]
    a = of.typeset([(t, f) for (t, _, _, _), f in zip(spans, fonts)], **tk)
    if isinstance(mat, str):  # translate(0, font.defaultLineHeight)
        mat = [1, 0, 0, 0, 1, 0, 0, float(fonts[0].defaultLineHeight), 1]
    for s, p in enumerate(of.compute_paths(a)):
        for paint in spans[s][3].get("paints", [BLACK]):
            _o_paint(orc, img, p, paint, mat, stroke)
    return a


def render_oracle(orc, name):
    w, h, ops = CASES[name]
    img = orc.new_image(w, h)
    orc.fill(img, (255, 255, 255, 255))
    last = None
    for op in ops:
        if op[0] == "text":
            last = o_text(orc, img, *op[1:])
        elif op[0] == "rects":
            for r in last.rects:
                orc.fill_path(img, r_rect(*r), orc.color_to_rgbx(*op[1]))
        elif op[0] == "rect":
            orc.fill_path(img, r_rect(*op[1:5]), orc.color_to_rgbx(*op[5]))
        elif op[0] == "bounded":
            from oracle import fonts as of
            _, spans, tk, dx, dy = op
            fonts = [_o_font(f, size, opts) for _, f, size, opts in spans]
            a = of.typeset([(t, f) for (t, _, _, _), f in zip(spans, fonts)], **tk)
            bx, by, bw, bh = snap(of.compute_bounds(orc, a))
            text = orc.new_image(int(bw), int(bh))
            for s, p in enumerate(of.compute_paths(a)):
                for paint in spans[s][3].get("paints", [BLACK]):
                    _o_paint(orc, text, p, paint, [1, 0, 0, 0, 1, 0, -bx, -by, 1], False)
            orc.draw_transform(img, text, [1, 0, 0, 0, 1, 0, f32(bx + f32(dx)), f32(by + f32(dy)), 1])
    return img


# ---- the product's rendering (device) ----

_ptf = {}


def p_typeface(name):
    import pixie_amd as px
    if name not in _ptf:
        _ptf[name] = px.readTypeface(os.path.join(FONTS, name))
    return _ptf[name]


def _p_font(font, size, opts):
    import pixie_amd as px
    f = px.Font(p_typeface(font))
    f.size = size
    f.lineHeight = opts.get("lineHeight", -1.0)
    f.underline = opts.get("underline", False)
    f.strikethrough = opts.get("strikethrough", False)
    f.noKerningAdjustments = opts.get("noKerningAdjustments", False)
    paints = []
    for p in opts.get("paints", [BLACK]):
        if isinstance(p, tuple) and p and p[0] == "linear":
            _, handles, stops = p
            paints.append(px.Paint(px.PaintKind.LinearGradientPaint, gradientHandlePositions=handles,
                                   gradientStops=[px.ColorStop(c, pos) for c, pos in stops]))
        else:
            paints.append(px.Paint(px.PaintKind.SolidPaint, color=p))
    f.paints = paints
    return f


def p_arrangement(spans, tk):
    import pixie_amd as px
    fonts = [_p_font(f, size, opts) for _, f, size, opts in spans]
    return px.typeset([px.Span(t, f) for (t, _, _, _), f in zip(spans, fonts)], **tk), fonts


def render_product(name):
    import pixie_amd as px
    w, h, ops = CASES[name]
    img = px.Image(w, h)
    img.fill((255, 255, 255, 255))
    last = None
    for op in ops:
        if op[0] == "text":
            _, spans, tk, mat, stroke = op
            a, fonts = p_arrangement(spans, tk)
            if isinstance(mat, str):
                mat = px.translate(0, float(fonts[0].defaultLineHeight))
            m = None if mat is None else np.asarray(mat, np.float32)
            if stroke:
                img.strokeText(a, transform=m)
            else:
                img.fillText(a, transform=m)
            last = a
        elif op[0] == "rects":
            paint = px.Paint(px.PaintKind.SolidPaint, color=op[1])
            for r in last.selectionRects:
                img.fillPath(r_rect(*r), paint)
        elif op[0] == "rect":
            img.fillPath(r_rect(*op[1:5]), px.Paint(px.PaintKind.SolidPaint, color=op[5]))
        elif op[0] == "bounded":
            _, spans, tk, dx, dy = op
            a, fonts = p_arrangement(spans, tk)
            bx, by, bw, bh = snap(a.computeBounds())
            text = px.Image(int(bw), int(bh))
            text.fillText(a, transform=px.translate(-bx, -by))
            img.draw(text, transform=px.translate(f32(bx + f32(dx)), f32(by + f32(dy))))
    return img.download()
