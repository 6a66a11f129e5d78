"""SVG through the native library: pixie's parseSvg / newImage(svg)
(fileformats/svg.nim) over the C-ABI of include/pixie_hip.h.

parseSvg reads the document on the host into path commands and resolved
properties; newImage draws it on the device, solid fills and strokes in one
scene pass.
"""
import ctypes

import numpy as np

from . import lib, _check, Image, PixieError


class _PxhSvgProps(ctypes.Structure):
    _fields_ = [("display", ctypes.c_int32), ("fill_rule", ctypes.c_int32),
                ("fill", ctypes.c_char * 64), ("stroke_rgbx", ctypes.c_uint32),
                ("stroke_width", ctypes.c_float), ("line_cap", ctypes.c_int32),
                ("line_join", ctypes.c_int32), ("miter_limit", ctypes.c_float),
                ("ndashes", ctypes.c_int32), ("dashes", ctypes.POINTER(ctypes.c_float)),
                ("transform", ctypes.c_float * 9), ("opacity", ctypes.c_float),
                ("fill_opacity", ctypes.c_float), ("stroke_opacity", ctypes.c_float)]


def _declare(L):
    P = ctypes.POINTER
    vp, f32p, i32p, i32 = ctypes.c_void_p, P(ctypes.c_float), P(ctypes.c_int32), ctypes.c_int32
    L.pxh_svg_parse.restype = vp
    L.pxh_svg_parse.argtypes = [ctypes.c_char_p, ctypes.c_int64, i32, i32]
    L.pxh_svg_free.argtypes = [vp]
    L.pxh_svg_size.argtypes = [vp, i32p, i32p]
    L.pxh_svg_count.restype = i32
    L.pxh_svg_count.argtypes = [vp]
    L.pxh_svg_element.argtypes = [vp, i32, P(f32p), i32p, P(_PxhSvgProps)]
    L.pxh_svg_render.argtypes = [vp, vp]


class Svg:
    """A parsed SVG document (svg.nim:13-17)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        w, h = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().pxh_svg_size(self._h, ctypes.byref(w), ctypes.byref(h)))
        self.width, self.height = w.value, h.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().pxh_svg_free(h)
            except Exception:
                pass
            self._h = None

    def __len__(self):
        return lib().pxh_svg_count(self._h)

    def element(self, i):
        """(commands, properties) of element i."""
        out = ctypes.POINTER(ctypes.c_float)()
        n = ctypes.c_int32()
        p = _PxhSvgProps()
        _check(lib().pxh_svg_element(self._h, i, ctypes.byref(out), ctypes.byref(n),
                                     ctypes.byref(p)))
        cmds = np.zeros(0, np.float32)
        if n.value:
            cmds = np.ctypeslib.as_array(out, shape=(n.value,)).copy()
            lib().pxh_free(out)
        props = {
            "display": bool(p.display), "fillRule": p.fill_rule, "fill": p.fill.decode(),
            "stroke": p.stroke_rgbx, "strokeWidth": np.float32(p.stroke_width),
            "lineCap": p.line_cap, "lineJoin": p.line_join,
            "miterLimit": np.float32(p.miter_limit),
            "dashes": [np.float32(p.dashes[k]) for k in range(p.ndashes)],
            "transform": np.array(p.transform[:], np.float32),
            "opacity": np.float32(p.opacity), "fillOpacity": np.float32(p.fill_opacity),
            "strokeOpacity": np.float32(p.stroke_opacity)}
        return cmds, props

    @property
    def elements(self):
        return [self.element(i) for i in range(len(self))]


def parseSvg(data, width=0, height=0):
    """parseSvg (svg.nim:547-555)."""
    if isinstance(data, str):
        data = data.encode("utf-8")
    h = lib().pxh_svg_parse(data, len(data), int(width), int(height))
    if not h:
        raise PixieError(5, lib().pxh_last_error().decode())
    return Svg(h)


def newImage(svg):
    """newImage(svg) (svg.nim:557-602): the drawn document."""
    img = Image(svg.width, svg.height)
    _check(lib().pxh_svg_render(svg._h, img._h))
    return img


def readSvg(path, width=0, height=0):
    with open(path, "rb") as fh:
        return parseSvg(fh.read(), width, height)
