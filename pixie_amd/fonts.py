"""Text through the native library: pixie's Typeface / Font / Span / typeset /
fillText / strokeText (fonts.nim) over the C-ABI of include/pixie_hip.h.

Fonts are OpenType files with TrueType (glyf) outlines, parsed by the native
library (opentype.nim:2531-2588); typesetting (fonts.nim:256-466) and the span
paths (computePaths, fonts.nim:517-564) are computed there too, and the paths
are filled or stroked on the device with each of the font's paints.
"""
import ctypes

import numpy as np

from . import (lib, _check, _fp, _mat, Paint, PaintKind, PixieError, _PxhPaint)

autoLineHeight = -1.0  # fonts.nim:6


class TextCase:
    """fonts.nim:50-55 (upper / lower / title case convert ASCII and
    Latin-1 letters)."""
    NormalCase, UpperCase, LowerCase, TitleCase = range(4)


class HorizontalAlignment:
    LeftAlign, CenterAlign, RightAlign = range(3)


class VerticalAlignment:
    TopAlign, MiddleAlign, BottomAlign = range(3)


LeftAlign, CenterAlign, RightAlign = range(3)
TopAlign, MiddleAlign, BottomAlign = range(3)


class _PxhFont(ctypes.Structure):
    _fields_ = [("typeface", ctypes.c_void_p), ("size", ctypes.c_float),
                ("line_height", ctypes.c_float), ("text_case", ctypes.c_int32),
                ("underline", ctypes.c_int32), ("strikethrough", ctypes.c_int32),
                ("no_kerning_adjustments", ctypes.c_int32), ("npaints", ctypes.c_int32),
                ("paints", ctypes.POINTER(_PxhPaint))]


class _PxhSpan(ctypes.Structure):
    _fields_ = [("text", ctypes.c_char_p), ("len", ctypes.c_int64),
                ("font", ctypes.POINTER(_PxhFont))]


def _declare(L):
    P = ctypes.POINTER
    vp, f32p, i32p, u32p = ctypes.c_void_p, P(ctypes.c_float), P(ctypes.c_int32), P(ctypes.c_uint32)
    i32, f, u32, i64 = ctypes.c_int32, ctypes.c_float, ctypes.c_uint32, ctypes.c_int64
    L.pxh_typeface_parse.restype = vp
    L.pxh_typeface_parse.argtypes = [ctypes.c_char_p, i64]
    L.pxh_typeface_parse_collection.restype = vp
    L.pxh_typeface_parse_collection.argtypes = [ctypes.c_char_p, i64, i32, i32p]
    L.pxh_typeface_free.argtypes = [vp]
    L.pxh_typeface_set_fallbacks.argtypes = [vp, P(vp), i32]
    L.pxh_typeface_metrics.argtypes = [vp, f32p]
    L.pxh_typeface_name.restype = i32
    L.pxh_typeface_name.argtypes = [vp, ctypes.c_char_p, i32]
    L.pxh_typeface_has_glyph.restype = i32
    L.pxh_typeface_has_glyph.argtypes = [vp, u32]
    L.pxh_typeface_advance.restype = f
    L.pxh_typeface_advance.argtypes = [vp, u32]
    L.pxh_typeface_kerning.restype = f
    L.pxh_typeface_kerning.argtypes = [vp, u32, u32]
    L.pxh_typeface_glyph_path.argtypes = [vp, u32, P(f32p), i32p]
    L.pxh_typeset.restype = vp
    L.pxh_typeset.argtypes = [P(_PxhSpan), i32, f, f, i32, i32, i32]
    L.pxh_arrangement_free.argtypes = [vp]
    L.pxh_arrangement_count.restype = i32
    L.pxh_arrangement_count.argtypes = [vp, i32]
    L.pxh_arrangement_read.argtypes = [vp, u32p, f32p, f32p, i32p, i32p]
    L.pxh_arrangement_layout_bounds.argtypes = [vp, f32p]
    L.pxh_arrangement_span_path.argtypes = [vp, i32, P(f32p), i32p]
    L.pxh_arrangement_compute_bounds.argtypes = [vp, f32p, f32p]
    L.pxh_fill_text.argtypes = [vp, vp, f32p]
    L.pxh_stroke_text.argtypes = [vp, vp, f32p, f, i32, i32, f, f32p, i32]


def _rune(r):
    return ord(r) if isinstance(r, str) else int(r)


def _take_cmds(out, n):
    if n.value == 0:
        return np.zeros(0, np.float32)
    arr = np.ctypeslib.as_array(out, shape=(n.value,)).copy()
    lib().pxh_free(out)
    return arr


def _f32(x):
    return np.float32(x)


def _round(x):
    """Nim's round on float32: half away from zero."""
    x = np.float32(x)
    return np.float32(np.sign(x) * np.floor(np.abs(np.float64(x)) + 0.5))


class Typeface:
    """A parsed font file (fonts.nim:11-15)."""

    def __init__(self, handle, data, filePath=""):
        self._h = ctypes.c_void_p(handle)
        self._data = data
        self.filePath = filePath
        self._fallbacks = []
        m = (ctypes.c_float * 9)()
        _check(lib().pxh_typeface_metrics(self._h, m))
        (self.scale, self.ascent, self.descent, self.lineGap, self.capHeight,
         self.underlinePosition, self.underlineThickness, self.strikeoutPosition,
         self.strikeoutThickness) = [np.float32(v) for v in m]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().pxh_typeface_free(h)
            except Exception:
                pass
            self._h = None

    @property
    def fallbacks(self):
        return _FallbackList(self)

    @fallbacks.setter
    def fallbacks(self, tfs):
        self._fallbacks = list(tfs)
        self._sync_fallbacks()

    def _sync_fallbacks(self):
        arr = (ctypes.c_void_p * max(1, len(self._fallbacks)))(*[t._h.value for t in self._fallbacks])
        _check(lib().pxh_typeface_set_fallbacks(self._h, arr, len(self._fallbacks)))

    @property
    def lineHeight(self):
        """fonts.nim:83-86, in font units."""
        return self.ascent - self.descent + self.lineGap

    @property
    def name(self):
        n = lib().pxh_typeface_name(self._h, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib().pxh_typeface_name(self._h, buf, n + 1)
        return buf.value.decode("utf-8", "replace")

    def hasGlyph(self, rune):
        return bool(lib().pxh_typeface_has_glyph(self._h, _rune(rune)))

    def getAdvance(self, rune):
        return np.float32(lib().pxh_typeface_advance(self._h, _rune(rune)))

    def getKerningAdjustment(self, left, right):
        return np.float32(lib().pxh_typeface_kerning(self._h, _rune(left), _rune(right)))

    def getGlyphPath(self, rune):
        """The glyph outline as a pixie command array (font units, y down)."""
        out = ctypes.POINTER(ctypes.c_float)()
        n = ctypes.c_int32()
        _check(lib().pxh_typeface_glyph_path(self._h, _rune(rune), ctypes.byref(out),
                                             ctypes.byref(n)))
        return _take_cmds(out, n)


class _FallbackList(list):
    """Typeface.fallbacks as a list whose changes reach the native typeface."""

    def __init__(self, owner):
        super().__init__(owner._fallbacks)
        self._owner = owner

    def _push(self):
        self._owner._fallbacks = list(self)
        self._owner._sync_fallbacks()

    def append(self, t):
        super().append(t)
        self._push()

    add = append

    def extend(self, ts):
        super().extend(ts)
        self._push()


def _load(data, filePath=""):
    data = bytes(data)
    h = lib().pxh_typeface_parse(data, len(data))
    if not h:
        raise PixieError(5, lib().pxh_last_error().decode())
    return Typeface(h, data, filePath)


def parseOtf(data):
    """parseOtf (fonts.nim:500-502)."""
    return _load(data)


parseTtf = parseOtf


def readTypeface(filePath):
    """readTypeface (fonts.nim:644-661) for .ttf / .otf files."""
    if not filePath.lower().endswith((".ttf", ".otf")):
        raise PixieError(6, "Unsupported font format")
    with open(filePath, "rb") as fh:
        return _load(fh.read(), filePath)


def readTypefaces(filePath):
    """readTypefaces (fonts.nim:663-672): every face of a collection."""
    with open(filePath, "rb") as fh:
        data = fh.read()
    count = ctypes.c_int32(0)
    out = []
    i = 0
    while True:
        h = lib().pxh_typeface_parse_collection(data, len(data), i, ctypes.byref(count))
        if not h:
            if i < count.value or count.value == 0:
                raise PixieError(5, lib().pxh_last_error().decode())
            break
        out.append(Typeface(h, data, filePath))
        i += 1
        if i >= count.value:
            break
    return out


def _solid(color):
    """A paint from a Paint, an "#rrggbb" string, (r, g, b, a) bytes (pixie's
    ColorRGBA converter: each byte / 255 in float32) or float channels."""
    if isinstance(color, Paint):
        return color
    if isinstance(color, str):
        h = color.lstrip("#")
        if len(h) != 6:
            raise PixieError(6, "unsupported colour string " + color)
        color = tuple(int(h[k:k + 2], 16) for k in (0, 2, 4)) + (255,)
    c = tuple(color)
    if all(isinstance(v, (int, np.integer)) for v in c):
        c = tuple(float(np.float32(v) / np.float32(255)) for v in c)
    return Paint(PaintKind.SolidPaint, color=c)


class Font:
    """A Font (fonts.nim:17-26, newFont 213-219): size 12, automatic line
    height, one black solid paint."""

    def __init__(self, typeface):
        self.typeface = typeface
        self.size = 12.0
        self.lineHeight = autoLineHeight
        self.paints = [Paint(PaintKind.SolidPaint, color=(0.0, 0.0, 0.0, 1.0))]
        self.textCase = TextCase.NormalCase
        self.underline = False
        self.strikethrough = False
        self.noKerningAdjustments = False

    def copy(self):
        f = Font(self.typeface)
        f.size, f.lineHeight, f.paints = self.size, self.lineHeight, list(self.paints)
        f.textCase, f.underline = self.textCase, self.underline
        f.strikethrough, f.noKerningAdjustments = self.strikethrough, self.noKerningAdjustments
        return f

    @property
    def paint(self):
        return self.paints[0]

    @paint.setter
    def paint(self, p):
        self.paints = [_solid(p)]

    @property
    def scale(self):
        """fonts.nim:186-188."""
        return np.float32(np.float32(self.size) / self.typeface.scale)

    @property
    def defaultLineHeight(self):
        """fonts.nim:190-194."""
        t = self.typeface
        units = np.float32(np.float32(t.ascent - t.descent) + t.lineGap)
        return _round(np.float32(units * self.scale))

    def layoutBounds(self, text):
        return typeset(self, text).layoutBounds()

    def _fill(self, st):
        """Writes the font's current values into its native struct."""
        paints = [_solid(p) for p in self.paints]
        arr = (_PxhPaint * max(1, len(paints)))()
        keep = []
        for i, p in enumerate(paints):
            n, k = p._native()
            arr[i] = n
            keep.append(k)
        st.typeface = self.typeface._h.value
        st.size = float(self.size)
        st.line_height = float(self.lineHeight)
        st.text_case = int(self.textCase)
        st.underline = int(bool(self.underline))
        st.strikethrough = int(bool(self.strikethrough))
        st.no_kerning_adjustments = int(bool(self.noKerningAdjustments))
        st.npaints = len(paints)
        st.paints = ctypes.cast(arr, ctypes.POINTER(_PxhPaint))
        return (arr, keep, paints)


def newFont(typeface):
    return Font(typeface)


def readFont(filePath):
    """readFont (fonts.nim:680-682)."""
    return Font(readTypeface(filePath))


class Span:
    """fonts.nim:28-30, 232-234."""

    def __init__(self, text, font):
        self.text = text
        self.font = font


newSpan = Span


class Arrangement:
    """The typeset runes (fonts.nim:32-39).  Holds the fonts it was typeset
    with; their current values are used when it is filled or stroked, as
    pixie's Arrangement holds Font refs."""

    def __init__(self, handle, fonts, structs):
        self._h = ctypes.c_void_p(handle)
        self.fonts = fonts
        self._structs = structs
        self._keep = []
        L = lib()
        n = L.pxh_arrangement_count(self._h, 0)
        nl = L.pxh_arrangement_count(self._h, 1)
        ns = L.pxh_arrangement_count(self._h, 2)
        runes = np.zeros(max(n, 1), np.uint32)
        pos = np.zeros(max(n, 1) * 2, np.float32)
        rects = np.zeros(max(n, 1) * 4, np.float32)
        lines = np.zeros(max(nl, 1) * 2, np.int32)
        spans = np.zeros(max(ns, 1) * 2, np.int32)
        P = ctypes.POINTER
        _check(L.pxh_arrangement_read(self._h, runes.ctypes.data_as(P(ctypes.c_uint32)),
                                      _fp(pos), _fp(rects),
                                      lines.ctypes.data_as(P(ctypes.c_int32)),
                                      spans.ctypes.data_as(P(ctypes.c_int32))))
        self.runes = [int(r) for r in runes[:n]]
        self.positions = pos[:2 * n].reshape(-1, 2)
        self.selectionRects = rects[:4 * n].reshape(-1, 4)
        self.lines = [tuple(v) for v in lines[:2 * nl].reshape(-1, 2).tolist()]
        self.spans = [tuple(v) for v in spans[:2 * ns].reshape(-1, 2).tolist()]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().pxh_arrangement_free(h)
            except Exception:
                pass
            self._h = None

    def _sync(self):
        self._keep = [f._fill(st) for f, st in zip(self.fonts, self._structs)]

    def layoutBounds(self):
        out = (ctypes.c_float * 2)()
        _check(lib().pxh_arrangement_layout_bounds(self._h, out))
        return (np.float32(out[0]), np.float32(out[1]))

    def computePaths(self):
        """One command array per span (fonts.nim:517-564)."""
        self._sync()
        paths = []
        for s in range(len(self.spans)):
            out = ctypes.POINTER(ctypes.c_float)()
            n = ctypes.c_int32()
            _check(lib().pxh_arrangement_span_path(self._h, s, ctypes.byref(out),
                                                   ctypes.byref(n)))
            paths.append(_take_cmds(out, n))
        return paths

    def computeBounds(self, transform=None):
        """fonts.nim:590-597: (x, y, w, h)."""
        self._sync()
        m = _mat(transform)
        out = (ctypes.c_float * 4)()
        _check(lib().pxh_arrangement_compute_bounds(self._h, _fp(m), out))
        return tuple(np.float32(v) for v in out)


def typeset(spans, text=None, bounds=(0.0, 0.0), hAlign=LeftAlign, vAlign=TopAlign, wrap=True):
    """typeset (fonts.nim:256-476): `spans` is a list of Span, or a Font with
    `text`."""
    if isinstance(spans, Font):
        spans = [Span(text, spans)]
    fonts = [s.font for s in spans]
    structs = [_PxhFont() for _ in spans]
    keep = [f._fill(st) for f, st in zip(fonts, structs)]
    texts = [s.text.encode("utf-8") for s in spans]
    arr = (_PxhSpan * max(1, len(spans)))()
    for i, (t, st) in enumerate(zip(texts, structs)):
        arr[i].text = t
        arr[i].len = len(t)
        arr[i].font = ctypes.pointer(st)
    h = lib().pxh_typeset(arr, len(spans), float(bounds[0]), float(bounds[1]), int(hAlign),
                          int(vAlign), int(bool(wrap)))
    if not h:
        raise PixieError(6, lib().pxh_last_error().decode())
    a = Arrangement(h, fonts, structs)
    a._keep = keep
    return a


def _arrangement(target, text, bounds, hAlign, vAlign):
    if isinstance(target, Arrangement):
        return target
    return typeset(target, text, bounds, hAlign, vAlign)


def fill_text(image, arrangement, transform=None):
    """fillText (fonts.nim:599-604)."""
    arrangement._sync()
    m = _mat(transform)
    _check(lib().pxh_fill_text(image._h, arrangement._h, _fp(m)))


def stroke_text(image, arrangement, transform=None, strokeWidth=1.0, lineCap=0, lineJoin=0,
                miterLimit=4.0, dashes=()):
    """strokeText (fonts.nim:622-641)."""
    arrangement._sync()
    m = _mat(transform)
    d = np.ascontiguousarray(dashes, dtype=np.float32) if len(dashes) else None
    _check(lib().pxh_stroke_text(image._h, arrangement._h, _fp(m), strokeWidth, lineCap,
                                 lineJoin, miterLimit, _fp(d), 0 if d is None else len(d)))
