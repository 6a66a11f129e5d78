"""Transformed image draws and image paints on the GPU against the oracle
(images.nim:168-259, 367-445, 531-683; paths.nim:2115-2142, 2176-2214).

Every step is integer arithmetic or float32 in pixie's operation order (the
transform inverse and vectors on the host, the running sample positions,
floors and mix weights on the device), so the bar is bit-exact.
"""
import math

import numpy as np
import pytest

import pixie_amd as px
from test_oracle_draw import SMOOTH, T, S, R, mul, rad, _input, _canvas, WHITE
from test_oracle_paints import HEART

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    px.init()
    return px


def _rand(seed, w, h, transparent=0.2):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    a[..., :3] = np.minimum(a[..., :3], a[..., 3:4])
    a[rng.random((h, w)) < transparent] = 0
    return a


def _exact(got, want, what):
    if not np.array_equal(got, want):
        d = np.argwhere(np.any(got != want, axis=-1))
        raise AssertionError("%s: %d pixels differ, first %s got %s want %s" % (
            what, len(d), d[0], got[tuple(d[0])], want[tuple(d[0])]))


@pytest.mark.parametrize("w,h,power", [(64, 48, 1), (101, 77, 1), (257, 129, 3), (1, 1, 2),
                                       (3, 640, 2)])
def test_minify(gpu, orc, w, h, power):
    src = _rand(w * h + power, w, h)
    got = px.Image.from_array(src).minifyBy2(power).download()
    _exact(got, orc.minify_by2(src, power), "minifyBy2 %dx%d power %d" % (w, h, power))


def test_magnify(gpu, orc):
    src = _rand(9, 37, 21)
    for power in (0, 1, 3):
        got = px.Image.from_array(src).magnifyBy2(power).download()
        _exact(got, orc.magnify_by2(src, power), "magnifyBy2 power %d" % power)


def _draw_both(orc, base, b, mat, mode, tiled=False):
    img = px.Image.from_array(base)
    src = px.Image.from_array(b)
    if tiled:
        img.drawTiled(src, mat, mode)
    else:
        img.draw(src, transform=mat, blendMode=mode)
    want = base.copy()
    if tiled:
        orc.draw_tiled(want, b, mat, mode)
    else:
        orc.draw_transform(want, b, mat, mode)
    return img.download(), want


@pytest.mark.parametrize("name,src,color,mat", SMOOTH, ids=[s[0] for s in SMOOTH])
def test_reference_draws(gpu, orc, name, src, color, mat):
    # the transforms of test_images_draw.nim:131-231 on the same inputs
    base = _canvas(orc, 100, 100, WHITE)
    b = _input(orc, src) if isinstance(src, str) else _canvas(orc, src[0], src[1], color)
    got, want = _draw_both(orc, base, b, mat(), 0)
    _exact(got, want, name)


TRANSFORMS = {
    "rot30": lambda: mul(T(200, 40), R(rad(30))),
    "rot-73_scale1.3": lambda: mul(mul(T(150, 260), R(-rad(73))), S(1.3, 1.3)),
    "scale0.3": lambda: mul(T(12.5, 7.25), S(0.3, 0.3)),        # two minify steps
    "scale0.5_exact": lambda: mul(T(40, 30), S(0.5, 0.5)),      # minify then composite
    "scale2_int": lambda: mul(T(-30, 11), S(2, 2)),             # magnify then composite
    "scale3.7_skew": lambda: mul(T(20.3, -50.6), np.array([3.7, 0.4, 0, -0.9, 3.1, 0, 0, 0, 1],
                                                          np.float32)),
    "frac_translate": lambda: T(33.37, -12.81),
    "int_translate": lambda: T(-17, 23),
    "mirror": lambda: mul(T(300, 10), S(-1, 1)),
}


@pytest.mark.parametrize("tname", sorted(TRANSFORMS))
@pytest.mark.parametrize("mode", [0, 2, 10, 16, 17, 18])
def test_draw_transforms(gpu, orc, tname, mode):
    base = _rand(11, 320, 280, transparent=0.1)
    b = _rand(12, 150, 110)
    got, want = _draw_both(orc, base, b, TRANSFORMS[tname](), mode)
    _exact(got, want, "%s mode %d" % (tname, mode))


@pytest.mark.parametrize("mode", range(20))
def test_draw_every_mode(gpu, orc, mode):
    base = _rand(20 + mode, 211, 173, transparent=0.1)
    b = _rand(40 + mode, 97, 131)
    got, want = _draw_both(orc, base, b, mul(mul(T(90.5, 20.25), R(rad(21))), S(1.15, 0.9)), mode)
    _exact(got, want, "mode %d" % mode)


@pytest.mark.parametrize("mode", [0, 16, 11])
@pytest.mark.parametrize("mat", ["scale0.3", "rot30", "mirror", "scale3.7_skew"])
def test_draw_tiled(gpu, orc, mode, mat):
    base = _rand(30, 260, 190, transparent=0.1)
    b = _rand(31, 45, 37)
    got, want = _draw_both(orc, base, b, TRANSFORMS[mat](), mode, tiled=True)
    _exact(got, want, "tiled %s mode %d" % (mat, mode))


def test_large_rotated_draw(gpu, orc):
    # a 2048x1536 canvas, a 1500x1100 source under a rotation and scale 1.1
    base = _rand(50, 2048, 1536, transparent=0.05)
    b = _rand(51, 1500, 1100)
    got, want = _draw_both(orc, base, b, mul(mul(T(900, -200), R(rad(37))), S(1.1, 1.1)), 0)
    _exact(got, want, "2048x1536 rotated")


@pytest.mark.parametrize("name,tiled,mat,opacity", [
    ("paintImage", False, (0.2, 0.2), 1.0),
    ("paintImageOpacity", False, (0.2, 0.2), 0.5),
    ("paintImageTiled", True, (0.02, 0.02), 1.0),
    ("paintImageTiledOpacity", True, (0.02, 0.02), 0.5),
])
def test_image_paint_masters(gpu, masters, orc, name, tiled, mat, opacity):
    # test_paints.nim:20-55, compared in straight alpha as pixie writes them
    mandrill = px.Image.from_array(_input(orc, "mandrill.png"))
    paint = px.Paint(px.PaintKind.TiledImagePaint if tiled else px.PaintKind.ImagePaint,
                     opacity=opacity, image=mandrill, imageMat=S(*mat))
    img = px.Image(100, 100)
    img.fillPath(HEART, paint)
    got = orc.to_straight_alpha(img.download())
    _exact(got, masters("tests/paths/%s.png" % name, premultiply=False), name)


@pytest.mark.parametrize("mode", [0, 3, 11, 16, 17, 19])
@pytest.mark.parametrize("tiled", [False, True])
def test_image_paint_fill_and_stroke(gpu, orc, mode, tiled):
    W, H = 300, 220
    base = _rand(60 + mode, W, H, transparent=0.3)
    pic = _rand(61, 70, 50)
    mat = mul(mul(T(40.5, 17.25), R(rad(12))), S(1.7, 1.7))
    kind = px.PaintKind.TiledImagePaint if tiled else px.PaintKind.ImagePaint
    paint = px.Paint(kind, blendMode=mode, opacity=0.8, image=px.Image.from_array(pic),
                     imageMat=mat)
    path = "M 20 30 C 120 -40 260 60 280 200 L 60 190 Q 0 120 20 30 Z"
    for stroke in (False, True):
        img = px.Image.from_array(base)
        want = base.copy()
        op = orc.make_image_paint(pic, mat, tiled=tiled, opacity=0.8, mode=mode)
        if stroke:
            img.strokePath(path, paint, strokeWidth=9.5, lineJoin=px.RoundJoin)
            _stroke_paint_oracle(orc, want, path, pic, mat, tiled, 0.8, mode, 9.5)
        else:
            img.fillPath(path, paint, windingRule=px.EvenOdd)
            orc.fill_path_paint(want, path, op, winding=orc.EVENODD)
        _exact(img.download(), want, "image paint mode %d tiled %s stroke %s" % (mode, tiled, stroke))


def _stroke_paint_oracle(orc, want, path, pic, mat, tiled, opacity, mode, width):
    # strokePath's route (paths.nim:2176-2214) from the oracle's pieces: the
    # white stroke as the mask, the image drawn onto transparent, the mask
    # at the paint's opacity, MaskBlend, then the paint's blend mode
    h, w = want.shape[:2]
    mask = orc.new_image(w, h)
    orc.stroke_path(mask, path, (255, 255, 255, 255), width=width, join=1)
    fill = orc.new_image(w, h)
    if tiled:
        orc.draw_tiled(fill, pic, mat, 0)
    else:
        orc.draw_transform(fill, pic, mat, 0)
    if opacity != 1:
        orc.apply_opacity(mask, opacity)
    orc.draw(fill, mask, 0, 0, 16)
    orc.draw(want, fill, 0, 0, mode)


def test_gradient_stroke(gpu, orc):
    # strokePath with a gradient paint: the stroke is the mask
    W, H = 200, 150
    base = _rand(70, W, H, transparent=0.5)
    stops = [((1, 0, 0, 1), 0.0), ((0, 0.5, 1, 0.6), 1.0)]
    paint = px.Paint(px.PaintKind.LinearGradientPaint, blendMode=2, opacity=0.9,
                     gradientHandlePositions=[(10, 10), (190, 140)],
                     gradientStops=[px.ColorStop(c, p) for c, p in stops])
    img = px.Image.from_array(base)
    img.strokePath("M 20 20 L 180 40 L 100 130 Z", paint, strokeWidth=7)
    want = base.copy()
    mask = orc.new_image(W, H)
    orc.stroke_path(mask, "M 20 20 L 180 40 L 100 130 Z", (255, 255, 255, 255), width=7)
    fill = orc.new_image(W, H)
    orc.fill_gradient(fill, orc.make_paint(3, [(10, 10), (190, 140)], stops))
    orc.apply_opacity(mask, 0.9)
    orc.draw(fill, mask, 0, 0, 16)
    orc.draw(want, fill, 0, 0, 2)
    _exact(img.download(), want, "gradient stroke")


def test_draw_errors(gpu):
    a = px.Image(50, 50)
    b = px.Image(10, 10)
    with pytest.raises(px.PixieError, match="degenerate"):
        a.draw(b, transform=S(0, 0))
    with pytest.raises(px.PixieError, match="differ"):
        a.draw(a, transform=T(1.5, 0))
    with pytest.raises(px.PixieError, match="negative power"):
        b.minifyBy2(-1)
    bad = px.Paint(px.PaintKind.LinearGradientPaint, gradientHandlePositions=[(0, 0)],
                   gradientStops=[px.ColorStop((1, 0, 0, 1), 0)])
    with pytest.raises(px.PixieError, match="requires 2 handles"):
        a.fillPath("M 0 0 L 50 0 L 50 50 Z", bad)


@pytest.mark.parametrize("offset,spread,blur", [((3.25, -1.5), 2, 4), ((0.5, 0.5), 0, 6),
                                                 ((-7.75, 12.1), -3, 0), ((12, 12), 4, 16)])
def test_shadow_offsets(gpu, orc, offset, spread, blur):
    src = _rand(80, 400, 300, transparent=0.4)
    color = (20, 40, 60, 200)
    got = px.Image.from_array(src).shadow(offset, spread, blur, color).download()
    _exact(got, orc.shadow(src, offset, spread, blur, color), "shadow %s" % (offset,))
