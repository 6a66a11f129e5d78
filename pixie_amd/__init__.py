"""pixie_amd -- pixie's raster hot path on MI355X (gfx950) through libpixie_hip.so.

Python mirror of pixie's Image / Path / Paint API for the hot path
(fillPath, strokePath, draw, blur, spread, shadow, applyOpacity, fill),
calling the C-ABI declared in include/pixie_hip.h.  Images live in HBM; pixels
are premultiplied RGBA8 exactly as pixie's Image.data (common.nim:34-37).

There is no CPU fallback: importing this package fails if the native
library has not been built, and every call reports device errors.
"""
import ctypes
import math
import os

import numpy as np

from .paths import Path, mat3, translate, scale, f32  # noqa: F401

__all__ = ["Image", "Path", "PixieError", "parse_path", "lib", "blend_pixels",
           "BlendMode", "NonZero", "EvenOdd", "mat3", "translate", "scale", "Paint",
           "PaintKind", "ColorStop", "Typeface", "Font", "Span", "Arrangement", "typeset",
           "readFont", "readTypeface", "readTypefaces", "parseOtf", "parseTtf", "Svg", "parseSvg",
           "readSvg"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpixie_hip.so")

NonZero, EvenOdd = 0, 1
ButtCap, RoundCap, SquareCap = 0, 1, 2
MiterJoin, RoundJoin, BevelJoin = 0, 1, 2


class BlendMode:
    """pixie's BlendMode ordinals (common.nim:6-29)."""
    Normal, Darken, Multiply, ColorBurn, Lighten, Screen, ColorDodge, Overlay, SoftLight, \
        HardLight, Difference, Exclusion, Hue, Saturation, Color, Luminosity, Mask, Overwrite, \
        SubtractMask, ExcludeMask = range(20)


class PaintKind:
    """pixie's PaintKind ordinals (paints.nim:6-12)."""
    SolidPaint, ImagePaint, TiledImagePaint, LinearGradientPaint, RadialGradientPaint, \
        AngularGradientPaint = range(6)


class ColorStop:
    """A gradient stop (paints.nim:24-27): straight float colour (r, g, b, a)
    in 0..1 and its position."""
    __slots__ = ("color", "position")

    def __init__(self, color, position):
        self.color = tuple(float(v) for v in color)
        self.position = float(position)


class Paint:
    """pixie's Paint (paints.nim:12-24) for the non-solid kinds: blend mode,
    opacity; `image` (an Image) and `imageMat` for ImagePaint /
    TiledImagePaint; gradient handle positions (image space; 2 for linear, 3
    for radial and angular) and colour stops for the gradients."""

    def __init__(self, kind, blendMode=0, opacity=1.0, gradientHandlePositions=(),
                 gradientStops=(), image=None, imageMat=None, color=(0.0, 0.0, 0.0, 0.0)):
        self.kind = kind
        self.color = tuple(float(v) for v in color)  # SolidPaint: straight float rgba
        self.blendMode = blendMode
        self.opacity = float(opacity)
        self.gradientHandlePositions = [tuple(p) for p in gradientHandlePositions]
        self.gradientStops = list(gradientStops)
        self.image = image
        self.imageMat = mat3() if imageMat is None else np.asarray(imageMat, np.float32)

    def solid_rgbx(self):
        """A SolidPaint's colour as fillPath hands it to fillShapes
        (paths.nim:2101-2112): opacity clamped and multiplied into alpha, then
        converted to premultiplied bytes (round to 8 bits, then (c * a + 127)
        div 255)."""
        o = min(max(self.opacity, 0.0), 1.0)
        r, g, b, a = self.color
        a = float(np.float32(a) * np.float32(o))
        # roundf of the float32 product: half away from zero
        byte = [int(math.floor(float(np.float32(min(max(v, 0.0), 1.0)) * np.float32(255)) + 0.5))
                for v in (r, g, b, a)]
        r, g, b, a = byte
        if a != 255:
            r, g, b = [(c * a + 127) // 255 for c in (r, g, b)]
        return (r, g, b, a)

    def _native(self):
        hp = [float(v) for xy in self.gradientHandlePositions[:3] for v in xy]
        hp += [0.0] * (6 - len(hp))
        stops = np.array([list(s.color) + [s.position] for s in self.gradientStops],
                         np.float32).reshape(-1, 5)
        stops = np.ascontiguousarray(stops)
        m = (ctypes.c_float * 9)(*[float(v) for v in np.asarray(self.imageMat).ravel()])
        p = _PxhPaint(self.kind, self.blendMode, self.opacity, len(self.gradientHandlePositions),
                      (ctypes.c_float * 6)(*hp), len(stops), _fp(stops) if len(stops) else None,
                      self.image._h if self.image is not None else None, m,
                      (ctypes.c_float * 4)(*self.color))
        return p, stops


class _PxhPaint(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("blend_mode", ctypes.c_int32),
                ("opacity", ctypes.c_float), ("nhandles", ctypes.c_int32),
                ("handles", ctypes.c_float * 6), ("nstops", ctypes.c_int32),
                ("stops", ctypes.POINTER(ctypes.c_float)), ("image", ctypes.c_void_p),
                ("image_mat", ctypes.c_float * 9), ("color", ctypes.c_float * 4)]


class PixieError(RuntimeError):
    """Raised when a call fails (pixie raises PixieError, common.nim:4)."""

    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


_lib = None


def lib():
    """The loaded native library.  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libpixie_hip.so is not built (run make -C pixie_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp, f32p, i32p = ctypes.c_void_p, P(ctypes.c_float), P(ctypes.c_int32)
        u32, i32, f = ctypes.c_uint32, ctypes.c_int32, ctypes.c_float
        L.pxh_init.restype = ctypes.c_int
        L.pxh_last_error.restype = ctypes.c_char_p
        L.pxh_version.restype = ctypes.c_char_p
        L.pxh_image_new.restype = vp
        L.pxh_image_new.argtypes = [i32, i32]
        L.pxh_image_new_a8.restype = vp
        L.pxh_image_new_a8.argtypes = [i32, i32]
        L.pxh_image_free.argtypes = [vp]
        L.pxh_image_width.argtypes = [vp]
        L.pxh_image_height.argtypes = [vp]
        L.pxh_image_format.argtypes = [vp]
        L.pxh_image_device_ptr.restype = vp
        L.pxh_image_device_ptr.argtypes = [vp]
        L.pxh_image_upload.argtypes = [vp, vp]
        L.pxh_image_download.argtypes = [vp, vp]
        L.pxh_image_copy.argtypes = [vp, vp]
        L.pxh_image_fill.argtypes = [vp, u32]
        L.pxh_parse_path.argtypes = [ctypes.c_char_p, P(f32p), i32p]
        L.pxh_free.argtypes = [vp]
        L.pxh_fill_path.argtypes = [vp, f32p, i32, f32p, u32, i32, i32]
        L.pxh_fill_path_paint.argtypes = [vp, f32p, i32, f32p, P(_PxhPaint), i32]
        L.pxh_stroke_path_paint.argtypes = [vp, f32p, i32, f32p, P(_PxhPaint), f, i32, i32, f,
                                            f32p, i32]
        L.pxh_draw_transform.argtypes = [vp, vp, f32p, i32]
        L.pxh_draw_tiled.argtypes = [vp, vp, f32p, i32]
        L.pxh_minify_by2.restype = vp
        L.pxh_minify_by2.argtypes = [vp, i32]
        L.pxh_magnify_by2.restype = vp
        L.pxh_magnify_by2.argtypes = [vp, i32]
        L.pxh_stroke_path.argtypes = [vp, f32p, i32, f32p, u32, f, i32, i32, f, f32p, i32, i32]
        L.pxh_fill_shapes.argtypes = [vp, f32p, i32p, i32, u32, i32, i32]
        L.pxh_draw.argtypes = [vp, vp, i32, i32, i32]
        L.pxh_blur.argtypes = [vp, f, u32]
        L.pxh_spread.argtypes = [vp, f]
        L.pxh_shadow.argtypes = [vp, vp, f, f, f, f, u32]
        L.pxh_apply_opacity.argtypes = [vp, f]
        L.pxh_blend_pixels.argtypes = [i32, vp, vp, vp, ctypes.c_int64]
        L.pxh_touched_pixels.argtypes = [P(ctypes.c_int64), i32]
        L.pxh_event_record.argtypes = [i32]
        L.pxh_profile_end.argtypes = [P(ctypes.c_double), P(ctypes.c_int64)]
        L.pxh_profile_kernel.argtypes = [ctypes.c_char_p, P(ctypes.c_double), P(ctypes.c_int64)]
        L.pxh_event_elapsed.argtypes = [i32, i32, P(ctypes.c_float)]
        L.pxh_debug_fill_geometry.argtypes = [f32p, i32, f32p, i32, i32, i32, f, i32, i32, f,
                                              i32p, P(i32p), i32p, P(f32p),
                                              P(P(ctypes.c_int8)), i32p]
        L.pxh_lanes.restype = i32
        L.pxh_lane_synchronize.argtypes = [i32]
        L.pxh_scene_render_on.argtypes = [vp, vp, i32]
        L.pxh_scene_autotune.argtypes = [vp, vp, i32, i32p, P(ctypes.c_float)]
        L.pxh_scene_graph_active.argtypes = [vp]
        L.pxh_scene_set_shape.argtypes = [vp, i32]
        L.pxh_blur_on.argtypes = [vp, f, u32, i32]
        L.pxh_image_fill_on.argtypes = [vp, u32, i32]
        L.pxh_image_view.restype = vp
        L.pxh_image_view.argtypes = [vp, i32, i32, i32]
        L.pxh_scene_new.restype = vp
        L.pxh_scene_free.argtypes = [vp]
        L.pxh_scene_clear.argtypes = [vp]
        L.pxh_scene_count.argtypes = [vp]
        L.pxh_scene_fill_path.argtypes = [vp, f32p, i32, f32p, u32, i32, i32]
        L.pxh_scene_stroke_path.argtypes = [vp, f32p, i32, f32p, u32, f, i32, i32, f, f32p, i32,
                                            i32]
        L.pxh_scene_render.argtypes = [vp, vp]
        L.pxh_scene_status.argtypes = [vp, i32p, i32]
        L.pxh_scene_debug_geometry.argtypes = [vp, i32, i32p, P(i32p), i32p, P(f32p),
                                               P(P(ctypes.c_int8)), i32p]
        L.pxh_scene_debug_limit_runs.argtypes = [vp, i32]
        L.pxh_server_new.restype = vp
        L.pxh_server_new.argtypes = [i32, ctypes.c_int64]
        L.pxh_server_free.argtypes = [vp]
        L.pxh_server_enqueue.restype = ctypes.c_int64
        L.pxh_server_enqueue.argtypes = [vp, vp, i32, i32, f]
        L.pxh_server_poll.argtypes = [vp, ctypes.c_int64, i32p, i32p]
        L.pxh_server_wait.argtypes = [vp, ctypes.c_int64, i32p]
        L.pxh_server_image.restype = vp
        L.pxh_server_image.argtypes = [vp, ctypes.c_int64]
        L.pxh_server_message.restype = ctypes.c_char_p
        L.pxh_server_message.argtypes = [vp, ctypes.c_int64]
        d = ctypes.POINTER(ctypes.c_double)
        L.pxh_server_times.argtypes = [vp, ctypes.c_int64, d, d, d]
        L.pxh_server_now.restype = ctypes.c_double
        L.pxh_server_now.argtypes = [vp]
        L.pxh_server_release.argtypes = [vp, ctypes.c_int64]
        from . import fonts as _fonts, svg as _svg
        _fonts._declare(L)
        _svg._declare(L)
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise PixieError(rc, lib().pxh_last_error().decode())


def init():
    _check(lib().pxh_init())


def synchronize():
    _check(lib().pxh_synchronize())


def pack(c):
    """(r, g, b, a) premultiplied bytes -> packed uint32 (r in the low byte)."""
    r, g, b, a = [int(v) & 255 for v in c]
    return r | (g << 8) | (b << 16) | (a << 24)


def _cmds(path):
    if isinstance(path, str):
        path = parse_path(path)
    elif isinstance(path, Path):
        path = path.array()
    return np.ascontiguousarray(path, dtype=np.float32)


def _paint_skipped(color, mode):
    """fillPath / strokePath draw nothing for a transparent solid paint unless
    it overwrites (paths.nim:2106-2107); colours here are bytes, so the float
    test color.a > 0 is alpha byte > 0."""
    return not (int(color[3]) > 0 or mode == BlendMode.Overwrite)


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def _mat(m):
    return None if m is None else np.ascontiguousarray(m, dtype=np.float32)


def parse_path(s):
    """parsePath (paths.nim:119-262) by the native parser."""
    out = ctypes.POINTER(ctypes.c_float)()
    n = ctypes.c_int32()
    _check(lib().pxh_parse_path(s.encode(), ctypes.byref(out), ctypes.byref(n)))
    if n.value == 0:
        return np.zeros(0, np.float32)
    arr = np.ctypeslib.as_array(out, shape=(n.value,)).copy()
    lib().pxh_free(out)
    return arr


class Image:
    """A device-resident image (pixie's Image, common.nim:34-37)."""

    def __init__(self, width, height, a8=False):
        L = lib()
        h = L.pxh_image_new_a8(width, height) if a8 else L.pxh_image_new(width, height)
        if not h:
            raise PixieError(6, L.pxh_last_error().decode())
        self._h = ctypes.c_void_p(h)
        self.width, self.height, self.a8 = width, height, a8

    @classmethod
    def view(cls, device_ptr, width, height, a8=False):
        """A non-owning image over device memory the caller keeps alive."""
        L = lib()
        h = L.pxh_image_view(ctypes.c_void_p(device_ptr), width, height, 1 if a8 else 0)
        if not h:
            raise PixieError(6, L.pxh_last_error().decode())
        img = cls.__new__(cls)
        img._h = ctypes.c_void_p(h)
        img.width, img.height, img.a8 = width, height, a8
        return img

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr, dtype=np.uint8)
        a8 = arr.ndim == 2
        img = cls(arr.shape[1], arr.shape[0], a8=a8)
        img.upload(arr)
        return img

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.pxh_image_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def device_ptr(self):
        return lib().pxh_image_device_ptr(self._h)

    def upload(self, arr):
        shape = (self.height, self.width) if self.a8 else (self.height, self.width, 4)
        arr = np.ascontiguousarray(arr, dtype=np.uint8)
        assert arr.shape == shape, (arr.shape, shape)
        _check(lib().pxh_image_upload(self._h, arr.ctypes.data))

    def download(self):
        shape = (self.height, self.width) if self.a8 else (self.height, self.width, 4)
        out = np.empty(shape, np.uint8)
        _check(lib().pxh_image_download(self._h, out.ctypes.data))
        return out

    def copy(self):
        o = Image(self.width, self.height, self.a8)
        _check(lib().pxh_image_copy(o._h, self._h))
        return o

    def fill(self, color, lane=0):
        if lane:
            _check(lib().pxh_image_fill_on(self._h, pack(color), lane))
        else:
            _check(lib().pxh_image_fill(self._h, pack(color)))

    def fillPath(self, path, color, transform=None, windingRule=NonZero, blendMode=0):
        """`color` is premultiplied (r, g, b, a) bytes, or a gradient Paint
        (whose own blend mode applies, as in pixie)."""
        if isinstance(color, Paint) and color.kind == PaintKind.SolidPaint:
            if color.opacity <= 0:
                return
            color, blendMode = color.solid_rgbx(), color.blendMode
        if isinstance(color, Paint):
            c = _cmds(path)
            m = _mat(transform)
            p, keep = color._native()
            _check(lib().pxh_fill_path_paint(self._h, _fp(c), len(c), _fp(m), ctypes.byref(p),
                                             windingRule))
            del keep
            return
        if _paint_skipped(color, blendMode):
            return
        c = _cmds(path)
        m = _mat(transform)
        _check(lib().pxh_fill_path(self._h, _fp(c), len(c), _fp(m), pack(color), windingRule,
                                   blendMode))

    def strokePath(self, path, color, transform=None, strokeWidth=1.0, lineCap=ButtCap,
                   lineJoin=MiterJoin, miterLimit=4.0, dashes=(), blendMode=0):
        """`color` is premultiplied (r, g, b, a) bytes, or a non-solid Paint
        (whose own blend mode applies, as in pixie)."""
        c = _cmds(path)
        m = _mat(transform)
        d = np.ascontiguousarray(dashes, dtype=np.float32) if len(dashes) else None
        nd = 0 if d is None else len(d)
        if isinstance(color, Paint) and color.kind == PaintKind.SolidPaint:
            if color.opacity <= 0:
                return
            color, blendMode = color.solid_rgbx(), color.blendMode
        if isinstance(color, Paint):
            p, keep = color._native()
            _check(lib().pxh_stroke_path_paint(self._h, _fp(c), len(c), _fp(m), ctypes.byref(p),
                                               strokeWidth, lineCap, lineJoin, miterLimit, _fp(d),
                                               nd))
            del keep
            return
        if _paint_skipped(color, blendMode):
            return
        _check(lib().pxh_stroke_path(self._h, _fp(c), len(c), _fp(m), pack(color), strokeWidth,
                                     lineCap, lineJoin, miterLimit, _fp(d), nd, blendMode))

    def fillText(self, font, text=None, transform=None, bounds=(0.0, 0.0), hAlign=0, vAlign=0):
        """fillText (fonts.nim:599-620): a Font and text (typeset here) or an
        Arrangement."""
        from . import fonts
        fonts.fill_text(self, fonts._arrangement(font, text, bounds, hAlign, vAlign), transform)

    def strokeText(self, font, text=None, transform=None, strokeWidth=1.0, bounds=(0.0, 0.0),
                   hAlign=0, vAlign=0, lineCap=ButtCap, lineJoin=MiterJoin, miterLimit=4.0,
                   dashes=()):
        """strokeText (fonts.nim:622-666)."""
        from . import fonts
        fonts.stroke_text(self, fonts._arrangement(font, text, bounds, hAlign, vAlign), transform,
                          strokeWidth, lineCap, lineJoin, miterLimit, dashes)

    def fillShapes(self, shapes, color, windingRule=NonZero, blendMode=0):
        pts = np.ascontiguousarray(np.concatenate([np.asarray(s, np.float32).reshape(-1, 2)
                                                   for s in shapes]) if shapes else
                                   np.zeros((0, 2), np.float32), dtype=np.float32)
        ofs = np.zeros(len(shapes) + 1, np.int32)
        for i, s in enumerate(shapes):
            ofs[i + 1] = ofs[i] + len(s)
        _check(lib().pxh_fill_shapes(self._h, _fp(pts), ofs.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int32)), len(shapes), pack(color), windingRule, blendMode))

    def draw(self, other, dx=0, dy=0, blendMode=0, transform=None):
        """Integer-offset composite (blendRect), or with `transform` pixie's
        draw(a, b, transform, blendMode) (images.nim:636-678)."""
        if transform is not None:
            m = _mat(transform)
            _check(lib().pxh_draw_transform(self._h, other._h, _fp(m), blendMode))
        else:
            _check(lib().pxh_draw(self._h, other._h, int(dx), int(dy), blendMode))

    def drawTiled(self, other, mat, blendMode=0):
        """drawTiled (images.nim:680-683)."""
        m = _mat(mat)
        _check(lib().pxh_draw_tiled(self._h, other._h, _fp(m), blendMode))

    def _wrap(self, h):
        if not h:
            raise PixieError(6, lib().pxh_last_error().decode())
        L = lib()
        img = Image.__new__(Image)
        img._h = ctypes.c_void_p(h)
        img.width, img.height, img.a8 = L.pxh_image_width(h), L.pxh_image_height(h), False
        return img

    def minifyBy2(self, power=1):
        """minifyBy2 (images.nim:168-228): a new image."""
        return self._wrap(lib().pxh_minify_by2(self._h, int(power)))

    def magnifyBy2(self, power=1):
        """magnifyBy2 (images.nim:230-259): a new image."""
        return self._wrap(lib().pxh_magnify_by2(self._h, int(power)))

    def blur(self, radius, outOfBounds=(0, 0, 0, 0), lane=0):
        _check(lib().pxh_blur_on(self._h, float(radius), pack(outOfBounds), lane))

    def spread(self, s):
        _check(lib().pxh_spread(self._h, float(s)))

    def shadow(self, offset, spread, blur, color):
        out = Image(self.width, self.height)
        _check(lib().pxh_shadow(out._h, self._h, float(offset[0]), float(offset[1]),
                                float(spread), float(blur), pack(color)))
        return out

    def applyOpacity(self, opacity):
        _check(lib().pxh_apply_opacity(self._h, float(opacity)))


class Scene:
    """An ordered recording of fillPath / strokePath calls rendered in one
    device pass (the batched form of paths.nim:2093-2176).  Rendering a scene
    onto an image leaves the pixels the same calls leave when issued one by
    one on that image."""

    def __init__(self):
        h = lib().pxh_scene_new()
        if not h:
            raise PixieError(7, "scene allocation failed")
        self._h = ctypes.c_void_p(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.pxh_scene_free(h)
            self._h = None

    def __len__(self):
        return lib().pxh_scene_count(self._h)

    def clear(self):
        _check(lib().pxh_scene_clear(self._h))

    def fillPath(self, path, color, transform=None, windingRule=NonZero, blendMode=0):
        if _paint_skipped(color, blendMode):
            return
        c = _cmds(path)
        m = _mat(transform)
        _check(lib().pxh_scene_fill_path(self._h, _fp(c), len(c), _fp(m), pack(color),
                                         windingRule, blendMode))

    def strokePath(self, path, color, transform=None, strokeWidth=1.0, lineCap=ButtCap,
                   lineJoin=MiterJoin, miterLimit=4.0, dashes=(), blendMode=0):
        if _paint_skipped(color, blendMode):
            return
        c = _cmds(path)
        m = _mat(transform)
        d = np.ascontiguousarray(dashes, dtype=np.float32) if len(dashes) else None
        _check(lib().pxh_scene_stroke_path(self._h, _fp(c), len(c), _fp(m), pack(color),
                                           strokeWidth, lineCap, lineJoin, miterLimit, _fp(d),
                                           0 if d is None else len(d), blendMode))

    def render(self, image, lane=0):
        """Renders onto `image`, queued on execution lane `lane`."""
        _check(lib().pxh_scene_render_on(self._h, image.handle, lane))

    def autotune(self, image, reps=3):
        """Picks the fastest phase-A launch shape for renders onto images of
        this size (on a scratch image).  Returns (shape index, ms per render)."""
        k, ms = ctypes.c_int32(), ctypes.c_float()
        _check(lib().pxh_scene_autotune(self._h, image.handle, reps, ctypes.byref(k),
                                        ctypes.byref(ms)))
        return k.value, ms.value

    def set_shape(self, shape):
        """Fixes the phase-A launch shape (what autotune() chooses)."""
        _check(lib().pxh_scene_set_shape(self._h, shape))
        return shape

    def graph_active(self):
        """Whether renders replay a captured HIP graph."""
        return bool(lib().pxh_scene_graph_active(self._h))

    def status(self):
        n = len(self)
        out = np.zeros(max(n, 1), np.int32)
        _check(lib().pxh_scene_status(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                      n))
        return out[:n]

    def debug_limit_runs(self, cap):
        """Test access: lowers the run-pool capacity later renders see."""
        _check(lib().pxh_scene_debug_limit_runs(self._h, cap))

    def debug_geometry(self, i):
        """Device geometry of call i: (info[16], parts[n,4], entries[m,4], windings[m])."""
        info = (ctypes.c_int32 * 16)()
        P = ctypes.POINTER
        parts, ents, wind = P(ctypes.c_int32)(), P(ctypes.c_float)(), P(ctypes.c_int8)()
        np_, ne = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().pxh_scene_debug_geometry(self._h, i, info, ctypes.byref(parts),
                                              ctypes.byref(np_), ctypes.byref(ents),
                                              ctypes.byref(wind), ctypes.byref(ne)))
        pa = np.ctypeslib.as_array(parts, shape=(max(np_.value, 1) * 4,)).copy()[:np_.value * 4]
        en = np.ctypeslib.as_array(ents, shape=(max(ne.value, 1) * 4,)).copy()[:ne.value * 4]
        wd = np.ctypeslib.as_array(wind, shape=(max(ne.value, 1),)).copy()[:ne.value]
        for p in (parts, ents, wind):
            lib().pxh_free(p)
        return np.array(info[:], np.int32), pa.reshape(-1, 4), en.reshape(-1, 4), wd


def lanes():
    """Number of execution lanes (independent device streams)."""
    return lib().pxh_lanes()


def lane_synchronize(lane):
    _check(lib().pxh_lane_synchronize(lane))


def touched_pixels(reset=True):
    """Pixels written by fills since the last reset (N_touch)."""
    v = ctypes.c_int64()
    _check(lib().pxh_touched_pixels(ctypes.byref(v), int(reset)))
    return v.value


def profile_begin():
    _check(lib().pxh_profile_begin())


def profile_end():
    """(summed kernel ms, launches) since profile_begin()."""
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _check(lib().pxh_profile_end(ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def profile_kernel(name):
    """(ms, launches) of one kernel or scene stage in the last profile."""
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _check(lib().pxh_profile_kernel(name.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def event_record(slot):
    _check(lib().pxh_event_record(slot))


def event_elapsed_ms(a, b):
    ms = ctypes.c_float()
    _check(lib().pxh_event_elapsed(a, b, ctypes.byref(ms)))
    return ms.value


def blend_pixels(mode, backdrop, source):
    """blender(mode) over pixel arrays on the device (uint8 [..., 4])."""
    b = np.ascontiguousarray(backdrop, dtype=np.uint8)
    s = np.ascontiguousarray(source, dtype=np.uint8)
    out = np.empty_like(b)
    _check(lib().pxh_blend_pixels(mode, b.ctypes.data, s.ctypes.data, out.ctypes.data,
                                  b.size // 4))
    return out


def debug_fill_geometry(path, w, h, transform=None, stroke=False, width=1.0, cap=0, join=0,
                        miter=4.0):
    """Host geometry of one fill: (info[16], parts[n,4], entries[m,4], windings[m])."""
    c = _cmds(path)
    m = _mat(transform)
    info = (ctypes.c_int32 * 16)()
    P = ctypes.POINTER
    parts, ents, wind = P(ctypes.c_int32)(), P(ctypes.c_float)(), P(ctypes.c_int8)()
    np_, ne = ctypes.c_int32(), ctypes.c_int32()
    rc = lib().pxh_debug_fill_geometry(_fp(c), len(c), _fp(m), w, h, int(stroke), width, cap,
                                       join, miter, info, ctypes.byref(parts), ctypes.byref(np_),
                                       ctypes.byref(ents), ctypes.byref(wind), ctypes.byref(ne))
    pa = np.ctypeslib.as_array(parts, shape=(max(np_.value, 1) * 4,)).copy()[:np_.value * 4]
    en = np.ctypeslib.as_array(ents, shape=(max(ne.value, 1) * 4,)).copy()[:ne.value * 4]
    wd = np.ctypeslib.as_array(wind, shape=(max(ne.value, 1),)).copy()[:ne.value]
    for p in (parts, ents, wind):
        lib().pxh_free(p)
    _check(rc)
    return np.array(info[:], np.int32), pa.reshape(-1, 4), en.reshape(-1, 4), wd


from .fonts import (Typeface, Font, Span, Arrangement, typeset, readFont, readTypeface,  # noqa: E402
                    readTypefaces, parseOtf, parseTtf)
from .svg import Svg, parseSvg, readSvg  # noqa: E402
