"""ctypes front-end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle restates pixie v6.0.0's scalar raster
path in C (see pixie_oracle.h).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module, and only as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATHS = {"scalar": os.path.join(_HERE, "liboracle.so"),
              # pixie's SIMD coverage path, -O3 -mavx2: the CPU baseline's
              # second form, not a checker
              "avx2": os.path.join(_HERE, "liboracle_avx2.so")}
_libs = {}
_variant = "scalar"

NONZERO, EVENODD = 0, 1


class OracleError(RuntimeError):
    pass


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "touched", "rows_trap", "rows_aa", "rows_nonaa", "rows_span2",
        "dropped", "segments", "partitions", "max_part", "max_sel")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Stop(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("r", "g", "b", "a", "pos")]


class Paint(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("mode", ctypes.c_int), ("opacity", ctypes.c_float),
                ("nhandles", ctypes.c_int), ("h", ctypes.c_float * 6), ("nstops", ctypes.c_int),
                ("stops", ctypes.POINTER(Stop)), ("image", ctypes.POINTER(ctypes.c_uint8)),
                ("iw", ctypes.c_int), ("ih", ctypes.c_int), ("image_mat", ctypes.c_float * 9)]


IMAGE, TILED, LINEAR, RADIAL, ANGULAR = 1, 2, 3, 4, 5


def make_paint(kind, handles, stops, opacity=1.0, mode=0):
    """kind: LINEAR / RADIAL / ANGULAR; handles: [(x, y), ...]; stops:
    [((r, g, b, a) floats, position), ...]."""
    arr = (Stop * max(1, len(stops)))(*[Stop(c[0], c[1], c[2], c[3], p) for c, p in stops])
    hv = [float(v) for xy in handles[:3] for v in xy]
    h = (ctypes.c_float * 6)(*(hv + [0.0] * (6 - len(hv))))
    p = Paint(kind, mode, opacity, len(handles), h, len(stops), arr)
    p._keep = arr
    return p


def make_image_paint(image, mat, tiled=False, opacity=1.0, mode=0):
    """ImagePaint / TiledImagePaint over a premultiplied RGBA8 array, with
    imageMat `mat` (9 floats, m[3 * col + row])."""
    img = np.ascontiguousarray(image, dtype=np.uint8)
    ih, iw = img.shape[:2]
    m = (ctypes.c_float * 9)(*[float(v) for v in np.asarray(mat, np.float32).ravel()])
    p = Paint(TILED if tiled else IMAGE, mode, opacity, 0, (ctypes.c_float * 6)(), 0, None,
              _u8(img), iw, ih, m)
    p._keep = img
    return p


class RGBX(ctypes.Structure):
    _fields_ = [("r", ctypes.c_uint8), ("g", ctypes.c_uint8),
                ("b", ctypes.c_uint8), ("a", ctypes.c_uint8)]


def build():
    """Compile liboracle.so from the C restatement (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class simd_mode:
    """Context manager: the functions below run on liboracle_avx2.so (the
    model of pixie's AVX2 build) inside the block."""

    def __enter__(self):
        global _variant
        self._prev, _variant = _variant, "avx2"
        return self

    def __exit__(self, *exc):
        global _variant
        _variant = self._prev
        return False


def lib():
    if _variant not in _libs:
        path = _LIB_PATHS[_variant]
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        f32p, i32p, u8p = P(ctypes.c_float), P(ctypes.c_int), P(ctypes.c_uint8)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_blend.restype = RGBX
        L.orc_blend.argtypes = [ctypes.c_int, RGBX, RGBX]
        L.orc_blend_pixels.argtypes = [ctypes.c_int, u8p, u8p, u8p, ctypes.c_int64]
        L.orc_mul_cov.restype = RGBX
        L.orc_mul_cov.argtypes = [RGBX, ctypes.c_uint8]
        L.orc_mul_opacity.restype = RGBX
        L.orc_mul_opacity.argtypes = [RGBX, ctypes.c_float]
        L.orc_rgbx_from_color.restype = RGBX
        L.orc_rgbx_from_color.argtypes = [ctypes.c_float] * 4
        L.orc_rgbx_from_rgba.restype = RGBX
        L.orc_rgbx_from_rgba.argtypes = [RGBX]
        L.orc_rgba_from_rgbx.restype = RGBX
        L.orc_rgba_from_rgbx.argtypes = [RGBX]
        L.orc_to_straight_alpha.argtypes = [u8p, ctypes.c_int64]
        L.orc_to_premultiplied_alpha.argtypes = [u8p, ctypes.c_int64]
        L.orc_gaussian_kernel.argtypes = [ctypes.c_int, P(ctypes.c_uint16)]
        L.orc_fill.argtypes = [u8p, ctypes.c_int, ctypes.c_int, RGBX]
        L.orc_apply_opacity.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float]
        L.orc_blur.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float, RGBX]
        L.orc_spread.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float]
        L.orc_shadow.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_float, ctypes.c_float, RGBX]
        L.orc_draw_rect.argtypes = [u8p, ctypes.c_int, ctypes.c_int, u8p, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_diff_score.restype = ctypes.c_float
        L.orc_diff_score.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int]
        L.orc_parse_path.argtypes = [ctypes.c_char_p, P(f32p), i32p]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_flatten.argtypes = [f32p, ctypes.c_int, f32p, ctypes.c_int, P(f32p), P(i32p), i32p]
        L.orc_stroke_shapes.argtypes = [f32p, ctypes.c_int, f32p, ctypes.c_float, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_float, f32p, ctypes.c_int,
                                        P(f32p), P(i32p), i32p]
        L.orc_fill_shapes.argtypes = [u8p, ctypes.c_int, ctypes.c_int, f32p, i32p, ctypes.c_int,
                                      RGBX, ctypes.c_int, ctypes.c_int, P(Stats)]
        L.orc_fill_path.argtypes = [u8p, ctypes.c_int, ctypes.c_int, f32p, ctypes.c_int, f32p,
                                    RGBX, ctypes.c_int, ctypes.c_int, P(Stats)]
        L.orc_stroke_path.argtypes = [u8p, ctypes.c_int, ctypes.c_int, f32p, ctypes.c_int, f32p,
                                      RGBX, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, f32p, ctypes.c_int, ctypes.c_int, P(Stats)]
        L.orc_segments.argtypes = [f32p, i32p, ctypes.c_int, P(f32p), P(P(ctypes.c_int16)), i32p]
        L.orc_partitions.argtypes = [f32p, i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     P(i32p), i32p, P(f32p), i32p, i32p]
        L.orc_fill_overlaps.argtypes = [f32p, ctypes.c_int, f32p, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_int, i32p]
        L.orc_fill_gradient.argtypes = [u8p, ctypes.c_int, ctypes.c_int, P(Paint)]
        L.orc_fill_path_paint.argtypes = [u8p, ctypes.c_int, ctypes.c_int, f32p, ctypes.c_int, f32p,
                                          P(Paint), ctypes.c_int, P(Stats)]
        L.orc_radial_matrix.argtypes = [P(Paint), f32p]
        L.orc_minify_size.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p]
        L.orc_minify_by2.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, u8p]
        L.orc_magnify_by2.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, u8p]
        L.orc_mix.restype = RGBX
        L.orc_mix.argtypes = [RGBX, RGBX, ctypes.c_float]
        L.orc_get_rgba_smooth.restype = RGBX
        L.orc_get_rgba_smooth.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                          ctypes.c_float, ctypes.c_int]
        L.orc_draw.argtypes = [u8p, ctypes.c_int, ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int,
                               f32p, ctypes.c_int]
        L.orc_draw_tiled.argtypes = [u8p, ctypes.c_int, ctypes.c_int, u8p, ctypes.c_int,
                                     ctypes.c_int, f32p, ctypes.c_int]
        _libs[_variant] = L
    return _libs[_variant]


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


def _u8(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _f32(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def rgbx(c):
    return RGBX(*[int(v) for v in c])


def blend(mode, backdrop, source):
    r = lib().orc_blend(mode, rgbx(backdrop), rgbx(source))
    return (r.r, r.g, r.b, r.a)


def blend_pixels(mode, backdrop, source):
    b = np.ascontiguousarray(backdrop, dtype=np.uint8)
    s = np.ascontiguousarray(source, dtype=np.uint8)
    out = np.empty_like(b)
    lib().orc_blend_pixels(mode, _u8(b), _u8(s), _u8(out), b.size // 4)
    return out


def mul_cov(c, cov):
    r = lib().orc_mul_cov(rgbx(c), cov)
    return (r.r, r.g, r.b, r.a)


def mul_opacity(c, o):
    r = lib().orc_mul_opacity(rgbx(c), o)
    return (r.r, r.g, r.b, r.a)


def color_to_rgbx(r, g, b, a):
    c = lib().orc_rgbx_from_color(r, g, b, a)
    return (c.r, c.g, c.b, c.a)


def rgba_to_rgbx(c):
    o = lib().orc_rgbx_from_rgba(rgbx(c))
    return (o.r, o.g, o.b, o.a)


def rgbx_to_rgba(c):
    o = lib().orc_rgba_from_rgbx(rgbx(c))
    return (o.r, o.g, o.b, o.a)


def to_straight_alpha(img):
    img = np.ascontiguousarray(img, dtype=np.uint8).copy()
    lib().orc_to_straight_alpha(_u8(img), img.size // 4)
    return img


def to_premultiplied_alpha(img):
    img = np.ascontiguousarray(img, dtype=np.uint8).copy()
    lib().orc_to_premultiplied_alpha(_u8(img), img.size // 4)
    return img


def gaussian_kernel(radius):
    out = np.zeros(2 * radius + 1, dtype=np.uint16)
    lib().orc_gaussian_kernel(radius, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    return out


def new_image(w, h):
    return np.zeros((h, w, 4), dtype=np.uint8)


def fill(img, c):
    h, w = img.shape[:2]
    lib().orc_fill(_u8(img), w, h, rgbx(c))


def apply_opacity(img, o):
    h, w = img.shape[:2]
    lib().orc_apply_opacity(_u8(img), w, h, o)


def blur(img, radius, oob=(0, 0, 0, 0)):
    h, w = img.shape[:2]
    _check(lib().orc_blur(_u8(img), w, h, radius, rgbx(oob)))


def spread(img, s):
    h, w = img.shape[:2]
    _check(lib().orc_spread(_u8(img), w, h, s))


def shadow(img, offset, spread_, blur_, color):
    h, w = img.shape[:2]
    out = np.empty_like(img)
    src = np.ascontiguousarray(img)
    _check(lib().orc_shadow(_u8(out), _u8(src), w, h, offset[0], offset[1], spread_, blur_,
                            rgbx(color)))
    return out


def draw(a, b, dx=0, dy=0, mode=0):
    ah, aw = a.shape[:2]
    bh, bw = b.shape[:2]
    b = np.ascontiguousarray(b)
    lib().orc_draw_rect(_u8(a), aw, ah, _u8(b), bw, bh, int(dx), int(dy), mode)


def minify_by2(img, power=1):
    """minifyBy2 (images.nim:168-228): a new array."""
    h, w = img.shape[:2]
    ow, oh = ctypes.c_int(), ctypes.c_int()
    lib().orc_minify_size(w, h, max(power, 0), ctypes.byref(ow), ctypes.byref(oh))
    out = np.zeros((oh.value, ow.value, 4), np.uint8)
    src = np.ascontiguousarray(img, dtype=np.uint8)
    _check(lib().orc_minify_by2(_u8(src), w, h, power, _u8(out)))
    return out


def magnify_by2(img, power=1):
    """magnifyBy2 (images.nim:230-259): a new array."""
    h, w = img.shape[:2]
    s = 1 << max(power, 0)
    out = np.zeros((h * s, w * s, 4), np.uint8)
    src = np.ascontiguousarray(img, dtype=np.uint8)
    _check(lib().orc_magnify_by2(_u8(src), w, h, power, _u8(out)))
    return out


def mix(a, b, t):
    r = lib().orc_mix(rgbx(a), rgbx(b), t)
    return (r.r, r.g, r.b, r.a)


def get_rgba_smooth(img, x, y, wrapped=False):
    h, w = img.shape[:2]
    src = np.ascontiguousarray(img, dtype=np.uint8)
    r = lib().orc_get_rgba_smooth(_u8(src), w, h, x, y, int(wrapped))
    return (r.r, r.g, r.b, r.a)


def draw_transform(a, b, mat, mode=0):
    """draw(a, b, transform, blendMode) (images.nim:636-678)."""
    ah, aw = a.shape[:2]
    bh, bw = b.shape[:2]
    b = np.ascontiguousarray(b, dtype=np.uint8)
    m = _f32(np.asarray(mat, np.float32).ravel())
    _check(lib().orc_draw(_u8(a), aw, ah, _u8(b), bw, bh, m[1], mode))


def draw_tiled(a, b, mat, mode=0):
    """drawTiled (images.nim:680-683)."""
    ah, aw = a.shape[:2]
    bh, bw = b.shape[:2]
    b = np.ascontiguousarray(b, dtype=np.uint8)
    m = _f32(np.asarray(mat, np.float32).ravel())
    _check(lib().orc_draw_tiled(_u8(a), aw, ah, _u8(b), bw, bh, m[1], mode))


def diff_score(master, img):
    h, w = master.shape[:2]
    return lib().orc_diff_score(_u8(np.ascontiguousarray(master)),
                                _u8(np.ascontiguousarray(img)), w, h)


def parse_path(s):
    out = ctypes.POINTER(ctypes.c_float)()
    n = ctypes.c_int()
    _check(lib().orc_parse_path(s.encode(), ctypes.byref(out), ctypes.byref(n)))
    arr = np.ctypeslib.as_array(out, shape=(n.value,)).copy() if n.value else np.zeros(0, np.float32)
    if n.value:
        lib().orc_free(out)
    return arr.astype(np.float32)


def _shapes_out(xy, ofs, n):
    nn = n.value
    offs = np.ctypeslib.as_array(ofs, shape=(nn + 1,)).copy()
    npts = int(offs[-1])
    pts = np.ctypeslib.as_array(xy, shape=(max(npts, 1) * 2,)).copy()[:npts * 2].reshape(-1, 2)
    lib().orc_free(xy)
    lib().orc_free(ofs)
    return [pts[offs[i]:offs[i + 1]] for i in range(nn)]


def flatten(cmds, mat=None, close=True):
    c, cp = _f32(cmds)
    m = _f32(mat)
    xy, ofs, n = ctypes.POINTER(ctypes.c_float)(), ctypes.POINTER(ctypes.c_int)(), ctypes.c_int()
    _check(lib().orc_flatten(cp, len(c), m[1] if m else None, int(close), ctypes.byref(xy),
                             ctypes.byref(ofs), ctypes.byref(n)))
    return _shapes_out(xy, ofs, n)


def stroke_shapes(cmds, mat=None, width=1.0, cap=0, join=0, miter_limit=4.0, dashes=()):
    c, cp = _f32(cmds)
    m = _f32(mat)
    d = _f32(np.asarray(dashes, np.float32)) if len(dashes) else None
    xy, ofs, n = ctypes.POINTER(ctypes.c_float)(), ctypes.POINTER(ctypes.c_int)(), ctypes.c_int()
    _check(lib().orc_stroke_shapes(cp, len(c), m[1] if m else None, width, cap, join,
                                   miter_limit, d[1] if d else None, len(dashes),
                                   ctypes.byref(xy), ctypes.byref(ofs), ctypes.byref(n)))
    return _shapes_out(xy, ofs, n)


def _pack_shapes(shapes):
    pts = np.concatenate([np.asarray(s, np.float32).reshape(-1, 2) for s in shapes]) \
        if shapes else np.zeros((0, 2), np.float32)
    ofs = np.zeros(len(shapes) + 1, np.int32)
    for i, s in enumerate(shapes):
        ofs[i + 1] = ofs[i] + len(s)
    return np.ascontiguousarray(pts, np.float32), ofs


def fill_shapes(img, shapes, color, winding=NONZERO, mode=0):
    h, w = img.shape[:2]
    pts, ofs = _pack_shapes(shapes)
    st = Stats()
    _check(lib().orc_fill_shapes(_u8(img), w, h, pts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 ofs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(shapes),
                                 rgbx(color), winding, mode, ctypes.byref(st)))
    return st.as_dict()


def fill_path(img, cmds, color, winding=NONZERO, mode=0, mat=None):
    """fillPath with a solid premultiplied colour; returns the fill stats."""
    h, w = img.shape[:2]
    if isinstance(cmds, str):
        cmds = parse_path(cmds)
    c, cp = _f32(cmds)
    m = _f32(mat)
    st = Stats()
    _check(lib().orc_fill_path(_u8(img), w, h, cp, len(c), m[1] if m else None, rgbx(color),
                               winding, mode, ctypes.byref(st)))
    return st.as_dict()


def stroke_path(img, cmds, color, width=1.0, cap=0, join=0, miter_limit=4.0, dashes=(),
                mode=0, mat=None):
    h, w = img.shape[:2]
    if isinstance(cmds, str):
        cmds = parse_path(cmds)
    c, cp = _f32(cmds)
    m = _f32(mat)
    d = _f32(np.asarray(dashes, np.float32)) if len(dashes) else None
    st = Stats()
    _check(lib().orc_stroke_path(_u8(img), w, h, cp, len(c), m[1] if m else None, rgbx(color),
                                 width, cap, join, miter_limit, d[1] if d else None,
                                 len(dashes), mode, ctypes.byref(st)))
    return st.as_dict()


def fill_gradient(img, paint):
    h, w = img.shape[:2]
    _check(lib().orc_fill_gradient(_u8(img), w, h, ctypes.byref(paint)))


def fill_path_paint(img, cmds, paint, winding=NONZERO, mat=None):
    """fillPath with a gradient paint (paths.nim:2115-2142)."""
    h, w = img.shape[:2]
    if isinstance(cmds, str):
        cmds = parse_path(cmds)
    c, cp = _f32(cmds)
    m = _f32(mat)
    st = Stats()
    _check(lib().orc_fill_path_paint(_u8(img), w, h, cp, len(c), m[1] if m else None,
                                     ctypes.byref(paint), winding, ctypes.byref(st)))
    return st.as_dict()


def segments(shapes):
    pts, ofs = _pack_shapes(shapes)
    segs, wind, n = ctypes.POINTER(ctypes.c_float)(), ctypes.POINTER(ctypes.c_int16)(), ctypes.c_int()
    _check(lib().orc_segments(pts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                              ofs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(shapes),
                              ctypes.byref(segs), ctypes.byref(wind), ctypes.byref(n)))
    nn = n.value
    s = np.ctypeslib.as_array(segs, shape=(max(nn, 1) * 4,)).copy()[:nn * 4].reshape(-1, 4)
    wd = np.ctypeslib.as_array(wind, shape=(max(nn, 1),)).copy()[:nn]
    lib().orc_free(segs)
    lib().orc_free(wind)
    return s, wd


def partitions(shapes, w, h):
    pts, ofs = _pack_shapes(shapes)
    parts, np_ = ctypes.POINTER(ctypes.c_int)(), ctypes.c_int()
    ents, ne = ctypes.POINTER(ctypes.c_float)(), ctypes.c_int()
    bounds = (ctypes.c_int * 4)()
    _check(lib().orc_partitions(pts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                ofs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(shapes),
                                w, h, ctypes.byref(parts), ctypes.byref(np_), ctypes.byref(ents),
                                ctypes.byref(ne), bounds))
    P = np.ctypeslib.as_array(parts, shape=(np_.value * 4,)).copy().reshape(-1, 4) \
        if np_.value else np.zeros((0, 4), np.int32)
    E = np.ctypeslib.as_array(ents, shape=(ne.value * 7,)).copy().reshape(-1, 7) \
        if ne.value else np.zeros((0, 7), np.float32)
    if np_.value:
        lib().orc_free(parts)
        lib().orc_free(ents)
    return P, E, tuple(bounds)


def fill_overlaps(cmds, x, y, winding=NONZERO, mat=None):
    if isinstance(cmds, str):
        cmds = parse_path(cmds)
    c, cp = _f32(cmds)
    m = _f32(mat)
    r = ctypes.c_int()
    _check(lib().orc_fill_overlaps(cp, len(c), m[1] if m else None, x, y, winding,
                                   ctypes.byref(r)))
    return bool(r.value)
