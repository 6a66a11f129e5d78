"""A small TrueType font built with fontTools for the text tests: simple
glyphs whose contours start off-curve or are all off-curve, composites with a
scale, an x/y scale, a 2x2 transform, nesting and matched-point arguments,
an empty glyph, a supplementary-plane rune (cmap format 12) and a kern table
of three subtables (accumulating and overriding) with no GPOS table."""
import array
import io

from fontTools.fontBuilder import FontBuilder
from fontTools.ttLib import newTable
from fontTools.ttLib.tables import ttProgram
from fontTools.ttLib.tables._g_l_y_f import Glyph, GlyphComponent, GlyphCoordinates
from fontTools.ttLib.tables._k_e_r_n import KernTable_format_0

# rune -> glyph name
CMAP = {0x41: "sq", 0x42: "offstart", 0x43: "alloff", 0x44: "c_scale", 0x45: "c_xy",
        0x46: "c_2x2", 0x47: "c_nested", 0x48: "c_match", 0x20: "empty", 0x1F600: "emoji"}
ORDER = [".notdef", "sq", "offstart", "alloff", "c_scale", "c_xy", "c_2x2", "c_nested",
         "c_match", "empty", "emoji"]
ALLOFF = [(0, 0), (400, 0), (400, 400), (0, 400)]   # every point off-curve
TWO_BY_TWO = ((0.75, 0.25), (-0.5, 1.5), (5, 7))     # c_2x2: fontTools transform, (dx, dy)


def raw_glyph(contours):
    """A simple glyph from contours of (x, y, on-curve) points."""
    g = Glyph()
    g.numberOfContours = len(contours)
    pts = [p for c in contours for p in c]
    g.coordinates = GlyphCoordinates([(x, y) for x, y, _ in pts])
    g.flags = array.array("B", [1 if on else 0 for _, _, on in pts])
    ends, n = [], 0
    for c in contours:
        n += len(c)
        ends.append(n - 1)
    g.endPtsOfContours = ends
    g.program = ttProgram.Program()
    g.program.fromBytecode(b"")
    return g


def comp(name, x=0, y=0, transform=None, xy=True):
    c = GlyphComponent()
    c.glyphName = name
    c.flags = 0
    if xy:
        c.x, c.y = x, y
    else:  # matched points: parent point x, child point y
        c.firstPt, c.secondPt = x, y
    if transform is not None:
        c.transform = [list(r) for r in transform]
    return c


def composite(components):
    g = Glyph()
    g.numberOfContours = -1
    g.components = components
    return g


def build(kern=True):
    fb = FontBuilder(1000, isTTF=True)
    fb.setupGlyphOrder(ORDER)
    fb.setupCharacterMap(CMAP)
    empty = Glyph()
    empty.numberOfContours = 0
    glyphs = {
        ".notdef": raw_glyph([[(0, 0, 1), (500, 0, 1), (500, 700, 1), (0, 700, 1)]]),
        "sq": raw_glyph([[(100, 0, 1), (600, 0, 1), (600, 500, 1), (100, 500, 1)],
                         [(200, 100, 1), (200, 400, 1), (500, 400, 1), (500, 100, 1)]]),
        "offstart": raw_glyph([[(50, 350, 0), (50, 0, 1), (300, -20, 0), (550, 0, 1),
                                (550, 700, 0), (300, 720, 1)]]),
        "alloff": raw_glyph([[(x, y, 0) for x, y in ALLOFF]]),
        "empty": empty,
        "emoji": raw_glyph([[(10, 10, 1), (90, 10, 0), (90, 90, 1), (10, 90, 0)]]),
        "c_scale": composite([comp("sq", 10, -20, [[0.5, 0], [0, 0.5]])]),
        "c_xy": composite([comp("offstart", -30, 40, [[1.25, 0], [0, -0.75]]), comp("sq", 300, 0)]),
        "c_2x2": composite([comp("alloff", TWO_BY_TWO[2][0], TWO_BY_TWO[2][1], TWO_BY_TWO[:2])]),
        "c_nested": composite([comp("c_2x2", 100, 100, [[0.5, 0], [0, 0.5]]), comp("c_scale", -100, 0)]),
        "c_match": composite([comp("sq", 30, 40), comp("alloff", 2, 1, None, xy=False)]),
    }
    fb.setupGlyf(glyphs)
    fb.setupHorizontalMetrics({n: (600 + 10 * i, 0) for i, n in enumerate(ORDER)})
    fb.setupHorizontalHeader(ascent=800, descent=-200, lineGap=90)
    fb.setupNameTable({"familyName": "Synth", "styleName": "Regular"})
    fb.setupOS2(sTypoAscender=800, usWinAscent=800, usWinDescent=200, yStrikeoutSize=50,
                yStrikeoutPosition=300, sCapHeight=700)
    fb.setupPost(underlinePosition=-100, underlineThickness=40)
    font = fb.font
    if kern:
        k = newTable("kern")
        k.version = 0
        subs = []
        for coverage, pairs in ((1, {("sq", "offstart"): -40, ("offstart", "sq"): 25,
                                     ("alloff", "alloff"): 0}),
                                (1, {("sq", "offstart"): -7, ("c_xy", "sq"): 11}),
                                (1 | 8, {("c_xy", "sq"): 3})):
            st = KernTable_format_0()
            st.version, st.coverage, st.format = 0, coverage, 0
            st.kernTable = pairs
            subs.append(st)
        k.kernTables = subs
        font["kern"] = k
    buf = io.BytesIO()
    font.save(buf)
    return buf.getvalue()


def rename_table(data, old, new):
    """The font with one table tag renamed in its table directory."""
    b = bytearray(data)
    n = int.from_bytes(b[4:6], "big")
    for i in range(n):
        at = 12 + 16 * i
        if bytes(b[at:at + 4]) == old:
            b[at:at + 4] = new
    return bytes(b)
