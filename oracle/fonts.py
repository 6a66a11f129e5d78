"""Oracle for text ingestion: pixie's typeface queries, glyph paths,
typesetting and span paths restated in Python.

TEST INFRASTRUCTURE ONLY.  Only tests/ may import this module, and only as
the checker.  The font tables are decoded by fontTools (an independent
OpenType decoder); on top of the decoded tables this module restates the
reference's own rules, each function citing the file:line it follows:

* which cmap subtables count and in which order they overwrite
  (opentype.nim:433-565: Windows platform only, formats 4 and 12);
* the contour -> path rule with implied on-curve midpoints
  (opentype.nim:2282-2319) and composite placement (2321-2414);
* the GPOS pair-adjustment lookup order and the kern-table accumulation
  quirk (opentype.nim:800-860, 1952-2164, 2481-2515);
* fallbacks, typeset, layoutBounds and computePaths (fonts.nim:124-564).

Arithmetic is float32 in the reference's operation order.
"""
import io
import math

import numpy as np
from fontTools.ttLib import TTFont

f32 = np.float32
LF, SP = 10, 32
MOVE, LINE, CUBIC, QUAD, CLOSE = 1, 2, 5, 7, 0


def nim_round(x):
    """Nim's round on float32 (roundf): half away from zero."""
    x = float(f32(x))
    return f32(math.copysign(math.floor(abs(x) + 0.5), x))


def nim_max(x, y):
    return x if (y <= x or y != y) else y


def m3_apply(m, x, y):
    """vmath Mat3 * Vec2, m[3 * col + row]."""
    x, y = f32(x), f32(y)
    return (f32(f32(f32(m[0] * x) + f32(m[3] * y)) + m[6]),
            f32(f32(f32(m[1] * x) + f32(m[4] * y)) + m[7]))


def m3_mul(a, b):
    r = [f32(0)] * 9
    for i in range(3):
        for j in range(3):
            r[3 * i + j] = f32(f32(f32(b[3 * i] * a[j]) + f32(b[3 * i + 1] * a[3 + j]))
                               + f32(b[3 * i + 2] * a[6 + j]))
    return r


IDENTITY = [f32(v) for v in (1, 0, 0, 0, 1, 0, 0, 0, 1)]


def path_transform(cmds, m):
    """Path.transform (paths.nim:264-333) for the commands glyph paths hold
    (moves, lines, quadratic and cubic curves, closes)."""
    m = [f32(v) for v in m]
    if m == IDENTITY:
        return list(cmds)
    out = list(cmds)
    i = 0
    while i < len(out):
        k = int(out[i])
        i += 1
        if k == CLOSE:
            continue
        n = {MOVE: 1, LINE: 1, QUAD: 2, CUBIC: 3}[k]
        for p in range(n):
            out[i + 2 * p], out[i + 2 * p + 1] = m3_apply(m, out[i + 2 * p], out[i + 2 * p + 1])
        i += 2 * n
    return out


class OTypeface:
    """One OpenType font with glyf or CFF outlines, decoded by fontTools."""

    def __init__(self, data):
        font = TTFont(io.BytesIO(data), lazy=False)
        self.font = font
        self.order = font.getGlyphOrder()
        self.gid = {n: i for i, n in enumerate(self.order)}
        self.unitsPerEm = f32(font["head"].unitsPerEm)
        hhea = font["hhea"]
        self.ascent, self.descent, self.lineGap = f32(hhea.ascent), f32(hhea.descent), f32(hhea.lineGap)
        os2 = font["OS/2"]
        self.strikeoutSize = f32(os2.yStrikeoutSize)
        self.strikeoutPosition = f32(os2.yStrikeoutPosition)
        self.capHeight = f32(os2.sCapHeight if os2.version >= 2 else 0)
        post = font["post"]
        self.underlinePosition = f32(post.underlinePosition)
        self.underlineThickness = f32(post.underlineThickness)
        # opentype.nim:433-565: Windows (3, *) subtables of format 4 or 12, later
        # ones overwriting earlier mappings; format 4 never maps U+FFFF.
        self.cmap = {}
        for t in font["cmap"].tables:
            if t.platformID != 3 or t.format not in (4, 12):
                continue
            for code, name in t.cmap.items():
                if t.format == 4 and code == 0xFFFF:
                    continue
                self.cmap[code] = self.gid[name]
        hm = font["hmtx"].metrics
        n = hhea.numberOfHMetrics
        self.advances = [hm[self.order[g]][0] for g in range(min(n, len(self.order)))]
        self.glyf = font["glyf"] if "glyf" in font else None
        self.cff = None
        if self.glyf is None:  # CFF outlines (opentype.nim:1299-1678)
            cff = font["CFF "].cff
            top = cff.topDictIndex[0]
            self.cff = top.CharStrings
            self.cff_gsubrs = cff.GlobalSubrs
            self.cff_cid = hasattr(top, "ROS")
            if self.cff_cid:
                self.cff_fdselect = top.FDSelect
                self.cff_fdarray = top.FDArray
            else:
                priv = top.Private
                self.cff_subrs = priv.Subrs if hasattr(priv, "Subrs") else []
        self.paths = {}
        self.fallbacks = []
        self._kerning_tables()

    def _kerning_tables(self):
        font = self.font
        self.gpos = None
        self.kern = None
        if "GPOS" in font:
            # opentype.nim:2085-2132: LookupType 2 subtables in lookup order,
            # kept when they hold any non-zero xAdvance of the first value
            tables = []
            for lookup in font["GPOS"].table.LookupList.Lookup:
                if lookup.LookupType != 2:
                    continue
                for st in lookup.SubTable:
                    t = self._pair_pos(st)
                    if t["pairs"] or t["classPairs"]:
                        tables.append(t)
            self.gpos = tables
        elif "kern" in font:
            # opentype.nim:800-860: after each subtable, the pairs of every
            # subtable read so far are applied again (accumulating unless the
            # override bit is set)
            k = font["kern"]
            if k.version == 0:
                pairs = {}
                seen = []
                for st in k.kernTables:
                    seen.append(st)
                    for t in seen:
                        if (t.coverage & 1) == 0:
                            continue
                        for (a, b), v in t.kernTable.items():
                            if v == 0:
                                continue
                            key = (self.gid[a], self.gid[b])
                            value = f32(v)
                            if (t.coverage & 0x8) == 0:
                                value = f32(value + pairs.get(key, f32(0)))
                            pairs[key] = value
                self.kern = pairs

    def _pair_pos(self, st):
        """parsePairPos (opentype.nim:1952-2083)."""
        t = {"format": st.Format, "covered": set(), "pairs": {}, "class1": {}, "class2": {},
             "classPairs": {}}
        cov = [self.gid[n] for n in st.Coverage.glyphs]
        t["covered"] = set(cov)
        has_xadv = (st.ValueFormat1 & 4) != 0
        if st.Format == 1:
            if has_xadv:
                for ci, g in enumerate(cov):
                    for rec in st.PairSet[ci].PairValueRecord:
                        xa = getattr(rec.Value1, "XAdvance", 0) if rec.Value1 is not None else 0
                        if xa:
                            t["pairs"][(g, self.gid[rec.SecondGlyph])] = xa
        elif st.Format == 2:
            t["class1"] = {self.gid[n]: c for n, c in st.ClassDef1.classDefs.items() if c != 0}
            t["class2"] = {self.gid[n]: c for n, c in st.ClassDef2.classDefs.items() if c != 0}
            if has_xadv:
                for c1, r1 in enumerate(st.Class1Record):
                    for c2, r2 in enumerate(r1.Class2Record):
                        xa = getattr(r2.Value1, "XAdvance", 0) if r2.Value1 is not None else 0
                        if xa:
                            t["classPairs"][(c1, c2)] = xa
        return t

    # ---- opentype.nim queries ----

    def glyph_id(self, rune):
        return self.cmap.get(rune, 0)

    def has_glyph(self, rune):
        return rune in self.cmap

    def _glyf_path(self, gid):
        """parseGlyfGlyph (opentype.nim:2416-2439) as a command list in font
        units (y up)."""
        g = self.glyf[self.order[gid]]
        if g.numberOfContours == 0:
            return []
        if g.isComposite():
            out = []
            for c in g.components:  # opentype.nim:2321-2414
                xs, s10, s01, ys = f32(1), f32(0), f32(0), f32(1)
                if hasattr(c, "transform"):
                    (xs, s10), (s01, ys) = [[f32(v) for v in row] for row in c.transform]
                dx = f32(c.x) if hasattr(c, "x") else f32(0)
                dy = f32(c.y) if hasattr(c, "y") else f32(0)
                sub = self._glyf_path(self.gid[c.glyphName])
                out += path_transform(sub, [xs, s10, 0, s01, ys, 0, dx, dy, 1])
            return out
        coords = g.coordinates
        flags = g.flags
        out = []
        start = 0
        for end in g.endPtsOfContours:  # opentype.nim:2282-2319
            pts = [(f32(coords[i][0]), f32(coords[i][1]), bool(flags[i] & 1))
                   for i in range(start, end + 1)]
            start = end + 1
            n = len(pts)
            curr, nxt = pts[-1], pts[0]
            if curr[2]:
                out += [MOVE, curr[0], curr[1]]
            elif nxt[2]:
                out += [MOVE, nxt[0], nxt[1]]
            else:
                out += [MOVE, f32((curr[0] + nxt[0]) / f32(2)), f32((curr[1] + nxt[1]) / f32(2))]
            for p in range(n):
                curr, nxt = nxt, pts[(p + 1) % n]
                if curr[2]:
                    out += [LINE, curr[0], curr[1]]
                else:
                    n2 = nxt
                    if not nxt[2]:
                        n2 = (f32((curr[0] + nxt[0]) / f32(2)), f32((curr[1] + nxt[1]) / f32(2)), True)
                    out += [QUAD, curr[0], curr[1], n2[0], n2[1]]
            out += [CLOSE]
        return out

    def _cff_path(self, gid):
        """parseCFFCharstring (opentype.nim:1028-1297) over the charstring
        bytes fontTools decoded: float32 stack and pen, operands shifted off
        the stack's front, moveTo without closing the contour before it,
        endchar closing the path without ending the charstring."""
        cs = self.cff[self.order[gid]]
        if self.cff_cid:
            priv = self.cff_fdarray[self.cff_fdselect[gid]].Private
            subrs = priv.Subrs if hasattr(priv, "Subrs") else []
        else:
            subrs = self.cff_subrs

        def bias(s):  # calcCFFSubroutineBias 1020-1026
            return 107 if len(s) < 1240 else 1131 if len(s) < 33900 else 32768

        out = []
        st = {"stack": [], "nStems": 0, "haveWidth": False, "x": f32(0), "y": f32(0)}
        stack = st["stack"]

        def shift():
            return stack.pop(0)

        def stems():
            if len(stack) % 2 != 0 and not st["haveWidth"]:
                shift()
            st["nStems"] += len(stack) >> 1
            del stack[:]
            st["haveWidth"] = True

        def width_if(more):
            if len(stack) > more and not st["haveWidth"]:
                shift()
                st["haveWidth"] = True

        def curve(c1x, c1y, c2x, c2y):
            out.extend([CUBIC, c1x, c1y, c2x, c2y, st["x"], st["y"]])

        def run(code, depth):
            assert depth <= 64
            i = 0
            while i < len(code):
                v = code[i]
                i += 1
                x, y = st["x"], st["y"]
                if v in (1, 3, 18, 23):
                    stems()
                elif v == 4:
                    width_if(1)
                    st["y"] = f32(y + stack.pop())
                    out.extend([MOVE, st["x"], st["y"]])
                elif v == 5:
                    while stack:
                        st["x"] = f32(st["x"] + shift())
                        st["y"] = f32(st["y"] + shift())
                        out.extend([LINE, st["x"], st["y"]])
                elif v in (6, 7):
                    horiz = v == 6
                    while stack:
                        k = "x" if horiz else "y"
                        st[k] = f32(st[k] + shift())
                        out.extend([LINE, st["x"], st["y"]])
                        horiz = not horiz
                elif v == 8 or v == 24:
                    while len(stack) > (2 if v == 24 else 0):
                        c1x = f32(st["x"] + shift()); c1y = f32(st["y"] + shift())
                        c2x = f32(c1x + shift()); c2y = f32(c1y + shift())
                        st["x"] = f32(c2x + shift()); st["y"] = f32(c2y + shift())
                        curve(c1x, c1y, c2x, c2y)
                    if v == 24:
                        st["x"] = f32(st["x"] + shift()); st["y"] = f32(st["y"] + shift())
                        out.extend([LINE, st["x"], st["y"]])
                elif v in (10, 29):
                    s = subrs if v == 10 else self.cff_gsubrs
                    k = int(stack.pop()) + bias(s)
                    sub = s[k]
                    if not hasattr(sub, "bytecode") or sub.bytecode is None:
                        sub.compile()
                    if len(sub.bytecode):
                        run(sub.bytecode, depth + 1)
                elif v == 11:
                    return
                elif v == 14:
                    width_if(2)
                    out.append(CLOSE)
                elif v in (19, 20):
                    stems()
                    i += (st["nStems"] + 7) >> 3
                elif v == 21:
                    width_if(2)
                    st["y"] = f32(y + stack.pop())
                    st["x"] = f32(x + stack.pop())
                    out.extend([MOVE, st["x"], st["y"]])
                elif v == 22:
                    width_if(1)
                    st["x"] = f32(x + stack.pop())
                    out.extend([MOVE, st["x"], st["y"]])
                elif v == 25:
                    while len(stack) > 6:
                        st["x"] = f32(st["x"] + shift()); st["y"] = f32(st["y"] + shift())
                        out.extend([LINE, st["x"], st["y"]])
                    c1x = f32(st["x"] + shift()); c1y = f32(st["y"] + shift())
                    c2x = f32(c1x + shift()); c2y = f32(c1y + shift())
                    st["x"] = f32(c2x + shift()); st["y"] = f32(c2y + shift())
                    curve(c1x, c1y, c2x, c2y)
                elif v == 26:
                    if len(stack) % 2:
                        st["x"] = f32(st["x"] + shift())
                    while stack:
                        c1x = st["x"]; c1y = f32(st["y"] + shift())
                        c2x = f32(c1x + shift()); c2y = f32(c1y + shift())
                        st["x"] = c2x; st["y"] = f32(c2y + shift())
                        curve(c1x, c1y, c2x, c2y)
                elif v == 27:
                    if len(stack) % 2:
                        st["y"] = f32(st["y"] + shift())
                    while stack:
                        c1x = f32(st["x"] + shift()); c1y = st["y"]
                        c2x = f32(c1x + shift()); c2y = f32(c1y + shift())
                        st["x"] = f32(c2x + shift()); st["y"] = c2y
                        curve(c1x, c1y, c2x, c2y)
                elif v == 28:
                    stack.append(f32(int.from_bytes(bytes(code[i:i + 2]), "big", signed=True)))
                    i += 2
                elif v in (30, 31):
                    vert = v == 30
                    while stack:
                        if vert:
                            c1x = st["x"]; c1y = f32(st["y"] + shift())
                            c2x = f32(c1x + shift()); c2y = f32(c1y + shift())
                            st["x"] = f32(c2x + shift())
                            st["y"] = f32(c2y + (shift() if len(stack) == 1 else f32(0)))
                        else:
                            c1x = f32(st["x"] + shift()); c1y = st["y"]
                            c2x = f32(c1x + shift()); c2y = f32(c1y + shift())
                            st["y"] = f32(c2y + shift())
                            st["x"] = f32(c2x + (shift() if len(stack) == 1 else f32(0)))
                        curve(c1x, c1y, c2x, c2y)
                        vert = not vert
                elif v < 32:
                    raise ValueError("CFF unknown operator: %d" % v)
                elif v < 247:
                    stack.append(f32(v - 139))
                elif v < 251:
                    stack.append(f32((v - 247) * 256 + code[i] + 108)); i += 1
                elif v < 255:
                    stack.append(f32(-(v - 251) * 256 - code[i] - 108)); i += 1
                else:
                    u = int.from_bytes(bytes(code[i:i + 4]), "big")
                    stack.append(f32(f32(u) / f32(65536))); i += 4

        if cs.bytecode is None:
            cs.compile()
        run(cs.bytecode, 0)
        return out

    def ot_glyph_path(self, rune):
        """getGlyphPath (opentype.nim:2456-2463): flipped to y down, cached."""
        if rune not in self.paths:
            gid = self.glyph_id(rune)
            raw = self._glyf_path(gid) if self.glyf is not None else self._cff_path(gid)
            self.paths[rune] = path_transform(raw, [1, 0, 0, 0, -1, 0, 0, 0, 1])
        return self.paths[rune]

    def ot_advance(self, rune):  # opentype.nim:2474-2479
        g = self.glyph_id(rune)
        return f32(self.advances[g] if g < len(self.advances) else self.advances[-1])

    def ot_kerning(self, left, right):  # opentype.nim:2481-2515
        if left not in self.cmap or right not in self.cmap:
            return f32(0)
        lg, rg = self.cmap[left], self.cmap[right]
        if self.gpos is not None:
            for t in self.gpos:
                if lg not in t["covered"]:
                    continue
                if t["format"] == 1:
                    if (lg, rg) in t["pairs"]:
                        return f32(t["pairs"][(lg, rg)])
                elif t["format"] == 2:
                    key = (t["class1"].get(lg, 0), t["class2"].get(rg, 0))
                    if key in t["classPairs"]:
                        return f32(t["classPairs"][key])
            return f32(0)
        if self.kern is not None:
            return f32(self.kern.get((lg, rg), f32(0)))
        return f32(0)

    # ---- fonts.nim typeface level (with fallbacks) ----

    @property
    def scale(self):
        return self.unitsPerEm

    def fallback(self, rune):  # fonts.nim:124-131
        if self.has_glyph(rune):
            return self
        for f in self.fallbacks:
            t = f.fallback(rune)
            if t is not None:
                return t
        return None

    def glyph_path(self, rune):  # fonts.nim:133-150
        t2 = self.fallback(rune)
        if t2 is None:
            return []
        p = list(t2.ot_glyph_path(rune))
        ratio = f32(self.scale / t2.scale)
        if ratio != f32(1):
            p = path_transform(p, [ratio, 0, 0, 0, ratio, 0, 0, 0, 1])
        return p

    def advance(self, rune):  # fonts.nim:152-165
        t2 = self.fallback(rune) or self
        return f32(t2.ot_advance(rune) * f32(self.scale / t2.scale))

    def kerning(self, left, right):  # fonts.nim:167-184
        tr, tl = self.fallback(right), self.fallback(left)
        if tr is None or tl is None or tr is not tl:
            return f32(0)
        return f32(tr.ot_kerning(left, right) * f32(self.scale / tr.scale))


def load(path):
    with open(path, "rb") as fh:
        return OTypeface(fh.read())


class OFont:
    """fonts.nim:17-26 with the parameters typeset and computePaths read."""

    def __init__(self, typeface, size=12.0, lineHeight=-1.0, underline=False,
                 strikethrough=False, noKerningAdjustments=False, textCase=0):
        self.typeface = typeface
        self.size = f32(size)
        self.lineHeight = f32(lineHeight)
        self.underline, self.strikethrough = underline, strikethrough
        self.noKerningAdjustments = noKerningAdjustments
        self.textCase = textCase

    @property
    def scale(self):  # fonts.nim:186-188
        return f32(self.size / self.typeface.scale)

    @property
    def defaultLineHeight(self):  # fonts.nim:190-194
        t = self.typeface
        return nim_round(f32(f32(f32(t.ascent - t.descent) + t.lineGap) * self.scale))

    def line_height(self):
        return self.lineHeight if self.lineHeight >= 0 else self.defaultLineHeight

    def line_gap(self):  # fonts.nim:196-205
        lh = self.line_height()
        if lh == self.defaultLineHeight:
            return self.typeface.lineGap
        t = self.typeface
        return f32(f32(f32(lh / self.scale) - t.ascent) + t.descent)


def _upper(r):
    if ord("a") <= r <= ord("z") or (0xE0 <= r <= 0xFE and r != 0xF7):
        return r - 32
    return {0xFF: 0x178, 0xB5: 0x39C}.get(r, r)


def _lower(r):
    if ord("A") <= r <= ord("Z") or (0xC0 <= r <= 0xDE and r != 0xD7):
        return r + 32
    return r


def _white(r):
    return (9 <= r <= 13 or r in (32, 0x85, 0xA0, 0x1680, 0x2028, 0x2029, 0x202F, 0x205F, 0x3000)
            or 0x2000 <= r <= 0x200A)


def _case(runes, tc):  # fonts.nim:235-251
    if tc == 1:
        return [_upper(r) for r in runes]
    if tc == 2:
        return [_lower(r) for r in runes]
    if tc == 3:
        out, prev = [], SP
        for r in runes:
            if _white(prev):
                r = _upper(r)
            out.append(r)
            prev = r
        return out
    return list(runes)


class OArrangement:
    pass


def typeset(spans, bounds=(0, 0), hAlign=0, vAlign=0, wrap=True):
    """typeset (fonts.nim:256-466); spans are (text, OFont) pairs."""
    bx, by = f32(bounds[0]), f32(bounds[1])
    a = OArrangement()
    a.runes, a.spans, a.fonts, a.lines = [], [], [], []
    start = 0
    for text, font in spans:
        rs = [ord(c) for c in text if ord(c) >= SP or ord(c) == LF]
        if rs:
            rs = _case(rs, font.textCase)
            a.runes += rs
            a.spans.append((start, start + len(rs) - 1))
            a.fonts.append(font)
            start += len(rs)
    n = len(a.runes)
    a.positions = [[f32(0), f32(0)] for _ in range(n)]
    a.rects = [[f32(0)] * 4 for _ in range(n)]
    if n == 0:
        return a
    lines = [[0, 0]]
    P, R, runes = a.positions, a.rects, a.runes

    def advance(font, i):
        r = f32(0)
        if not font.noKerningAdjustments and i + 1 < n:
            r = f32(r + font.typeface.kerning(runes[i], runes[i + 1]))
        r = f32(r + font.typeface.advance(runes[i]))
        return f32(r * font.scale)

    atx, aty = f32(0), f32(0)
    prevCanWrap = 0
    for si, (s0, s1) in enumerate(a.spans):
        font = a.fonts[si]
        for ri in range(s0, s1 + 1):
            rune = runes[ri]
            if rune == LF:
                adv = f32(font.typeface.advance(SP) * font.scale)
                P[ri] = [atx, aty]
                R[ri] = [atx, aty, adv, f32(0)]
                atx, aty = f32(0), f32(aty + f32(1))
                prevCanWrap = 0
                lines[-1][1] = ri
                if ri + 1 < n:
                    lines.append([ri + 1, 0])
            else:
                adv = advance(font, ri)
                if wrap and rune != SP and bx > 0 and f32(atx + adv) > bx:
                    atx, aty = f32(0), f32(aty + f32(1))
                    lineStart = ri
                    if prevCanWrap > 0 and prevCanWrap != ri:
                        for i in range(prevCanWrap + 1, ri):
                            P[i] = [atx, aty]
                            R[i][0], R[i][1] = atx, aty
                            atx = f32(atx + advance(font, i))
                            lineStart -= 1
                        prevCanWrap = 0
                    lines[-1][1] = lineStart - 1
                    lines.append([lineStart, 0])
                if rune == SP or _white(rune):
                    prevCanWrap = ri
                P[ri] = [atx, aty]
                R[ri] = [atx, aty, adv, f32(0)]
                atx = f32(atx + adv)
    lines[-1][1] = n - 1

    if hAlign != 0:
        for (l0, l1) in lines:
            furthest = f32(0)
            for i in range(l1, l0 - 1, -1):
                if runes[i] != SP and runes[i] != LF:
                    furthest = f32(R[i][0] + R[i][2])
                    break
            adj = f32(0)
            if hAlign == 1:
                adj = f32(f32(bx - furthest) / f32(2))
            elif hAlign == 2:
                adj = f32(bx - furthest)
            if adj != 0:
                for i in range(l0, l1 + 1):
                    P[i][0] = f32(P[i][0] + adj)
                    R[i][0] = f32(R[i][0] + adj)

    L = len(lines)
    initialY = f32(0)
    for si, (s0, s1) in enumerate(a.spans):
        font = a.fonts[si]
        units = f32(font.typeface.ascent + f32(font.line_gap() / f32(2)))
        initialY = nim_max(initialY, nim_round(f32(units * font.scale)))
        if s1 >= lines[0][1]:
            break
    heights = [f32(0)] * L
    line = 0
    for si, (s0, s1) in enumerate(a.spans):
        flh = a.fonts[si].line_height()
        heights[line] = nim_max(heights[line], flh)
        for ri in range(s0, s1 + 1):
            if line + 1 < L and ri == lines[line + 1][0]:
                line += 1
                heights[line] = nim_max(heights[line], flh)
        if line + 1 < L and s1 == lines[line][1]:
            line += 1
    line, baseline = 0, initialY
    for si, (s0, s1) in enumerate(a.spans):
        font = a.fonts[si]
        lh = font.line_height()
        top = nim_round(f32(f32(font.typeface.ascent + f32(font.line_gap() / f32(2))) * font.scale))
        for ri in range(s0, s1 + 1):
            if line + 1 < L and ri == lines[line + 1][0]:
                line += 1
                baseline = f32(baseline + heights[line])
            P[ri][1] = baseline
            R[ri][1] = f32(baseline - top)
            R[ri][3] = lh
    if vAlign != 0:
        furthestY = f32(R[-1][1] + R[-1][3])
        adj = f32(0)
        if vAlign == 1:
            adj = nim_round(f32(f32(by - furthestY) / f32(2)))
        elif vAlign == 2:
            adj = f32(by - furthestY)
        if adj != 0:
            for i in range(n):
                P[i][1] = f32(P[i][1] + adj)
                R[i][1] = f32(R[i][1] + adj)
    for r in R:
        minX, maxX = nim_round(r[0]), nim_round(f32(r[0] + r[2]))
        minY, maxY = nim_round(r[1]), nim_round(f32(r[1] + r[3]))
        r[0], r[2], r[1], r[3] = minX, f32(maxX - minX), minY, f32(maxY - minY)
    a.lines = [tuple(v) for v in lines]
    return a


def layout_bounds(a):  # fonts.nim:478-492
    x = y = f32(0)
    if a.runes:
        for i, r in enumerate(a.runes):
            if r != LF:
                x = nim_max(x, f32(a.rects[i][0] + a.rects[i][2]))
        f = a.rects[-1]
        y = f32(f[1] + f[3])
        if a.runes[-1] == LF:
            y = f32(y + f[3])
    return (x, y)


def _rect(x, y, w, h, clockwise=True):  # paths.nim:506-521
    x, y, w, h = f32(x), f32(y), f32(w), f32(h)
    if clockwise:
        return [MOVE, x, y, LINE, f32(x + w), y, LINE, f32(x + w), f32(y + h), LINE, x, f32(y + h),
                CLOSE]
    return [MOVE, x, y, LINE, x, f32(y + h), LINE, f32(x + w), f32(y + h), LINE, f32(x + w), y,
            CLOSE]


def compute_paths(a):
    """computePaths (fonts.nim:517-564): one command list per span."""
    out = []
    line = 0
    for si, (s0, s1) in enumerate(a.spans):
        font = a.fonts[si]
        t = font.typeface
        sc = font.scale
        ut, up = f32(t.underlineThickness * sc), f32(t.underlinePosition * sc)
        st, sp = f32(t.strikeoutSize * sc), f32(t.strikeoutPosition * sc)
        path = []
        for ri in range(s0, s1 + 1):
            px, py = a.positions[ri]
            g = t.glyph_path(a.runes[ri])
            g = path_transform(g, m3_mul([1, 0, 0, 0, 1, 0, px, py, 1],
                                         [sc, 0, 0, 0, sc, 0, 0, 0, 1]))
            decorate = True
            if line < len(a.lines) and ri == a.lines[line][1]:
                line += 1
                if a.runes[ri] == SP:
                    decorate = False
            if decorate:
                r = a.rects[ri]
                cw = t.glyf is not None  # isCCW (opentype.nim:2517-2521)
                if font.underline:
                    g += _rect(r[0], f32(f32(py - up) + f32(ut / f32(2))), r[2], ut, cw)
                if font.strikethrough:
                    g += _rect(r[0], f32(py - sp), r[2], st, cw)
            path += g
        out.append(np.array(path, np.float32))
    return out


def compute_bounds(orc, a, mat=None):
    """computeBounds (fonts.nim:590-597 -> paths.nim:1098-1125): the span
    paths flattened with closed subpaths, every segment kept (horizontal
    ones too), y quantized to 1/256; (x, y, w, h)."""
    full = np.concatenate(compute_paths(a) + [np.zeros(0, np.float32)])
    if mat is not None:
        full = np.array(path_transform(list(full), mat), np.float32)
    shapes = orc.flatten(full, None, True) if len(full) else []
    big = f32(3.4028235e38)
    xmin, xmax, ymin, ymax = big, -big, big, -big

    def q(v):
        v = f32(v)
        return f32(math.copysign(math.floor(abs(float(v)) * 256.0) / 256.0, float(v)))
    for s in shapes:
        s = np.asarray(s, np.float32).reshape(-1, 2)
        if len(s) < 2:
            continue
        v1 = (s[-1][0], q(s[-1][1]))
        for i in range(len(s)):
            v2 = (s[i][0], q(s[i][1]))
            if i == 0 and v1 == v2:
                continue
            at, to = v1, v2
            v1 = v2
            if at[1] > to[1]:
                at, to = to, at
            xmin = min(xmin, min(at[0], to[0]))
            xmax = max(xmax, max(at[0], to[0]))
            ymin = min(ymin, at[1])
            ymax = max(ymax, to[1])
    return (f32(xmin), f32(ymin), f32(xmax - xmin), f32(ymax - ymin))
