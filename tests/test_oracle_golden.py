"""Pins the CPU oracle against the reference's own fixtures.

Every case below restates a block of pixie's test-suite (file:line cited) and
compares the oracle with the master PNG that block writes, or with the value
its doAssert checks.  Masters are straight alpha; they are premultiplied the
way pixie's readImage does before comparing.

Two classes of master:
  * EXACT: the oracle's premultiplied output equals the master bit for bit.
  * ALPHA_EXACT: coverage (alpha) is exact; colour channels of partially
    covered pixels may differ by 1 LSB.  These masters were written by a SIMD
    build of pixie, whose coverage multiply floors where the scalar path
    (the oracle's semantics) rounds.
"""
import math

import numpy as np
import pytest

from pixie_amd.paths import Path, translate, scale, f32

RED = (255, 0, 0, 255)
WHITE = (255, 255, 255, 255)
BLACK = (0, 0, 0, 255)


def _cmp(orc, masters, name, img, exact=True, alpha_exact=True):
    m = masters(name)
    assert m.shape == img.shape
    d = np.abs(m.astype(int) - img.astype(int))
    if exact:
        assert d.max() == 0, "%s: %d pixels differ (max %d)" % (name, (d.max(-1) > 0).sum(), d.max())
    else:
        if alpha_exact:
            assert d[..., 3].max() == 0, "%s: alpha differs" % name
        assert d.max() <= 1, "%s: colour differs by %d" % (name, d.max())


def _img(orc, w, h, fill=None):
    a = orc.new_image(w, h)
    if fill is not None:
        orc.fill(a, fill)
    return a


def _pink(orc):
    return orc.rgba_to_rgbx((0xFC, 0x42, 0x7B, 255))


# test_paths.nim:51-110 rectangles and strokes
@pytest.mark.parametrize("name,path,color,stroke", [
    ("tests/paths/pathStroke1.png", "M 10 10 L 90 90", RED, True),
    ("tests/paths/pathStroke2.png", "M 10 10 L 50 60 90 90", RED, True),
    ("tests/paths/pathStroke3.png", "M 15 10 L 30 90 60 30 90 90", (255, 255, 0, 255), True),
    ("tests/paths/pathBlackRectangle.png", "M 10 10 H 90 V 90 H 10 L 10 10", BLACK, False),
    ("tests/paths/pathBlackRectangleZ.png", "M 10 10 H 90 V 90 H 10 Z", BLACK, False),
    ("tests/paths/pathYellowRectangle.png", "M 10 10 H 90 V 90 H 10 L 10 10", (255, 255, 0, 255), False),
    ("tests/paths/pathStroke1Big.png", "M0 0 L200 200", RED, True),       # :603-609
    ("tests/paths/path1pxCover.png", "M99 99 L999 99 L999 100 L99 100 Z", RED, False),  # :611-617
    ("tests/paths/path0pxCover.png", "M100 100 L999 100 L999 101 L100 101 Z", RED, False),
])
def test_paths_basic(orc, masters, name, path, color, stroke):
    a = _img(orc, 100, 100)
    if stroke:
        orc.stroke_path(a, path, color, width=10)
    else:
        orc.fill_path(a, path, color)
    _cmp(orc, masters, name, a)


def test_red_rectangle_builder(orc, masters):
    # test_paths.nim:100-110 (moveTo/lineTo builder, no close)
    p = Path()
    p.move_to(10, 10); p.line_to(10, 90); p.line_to(90, 90); p.line_to(90, 10); p.line_to(10, 10)
    a = _img(orc, 100, 100)
    orc.fill_path(a, p.array(), RED)
    _cmp(orc, masters, "tests/paths/pathRedRectangle.png", a)


@pytest.mark.parametrize("name,path", [
    ("tests/paths/pathBottomArc.png", "M30 60 A 20 20 0 0 0 90 60 L 30 60"),           # :112-118
    ("tests/paths/pathHeart.png", "M 10,30 A 20,20 0,0,1 50,30 A 20,20 0,0,1 90,30 "
                                  "Q 90,60 50,90 Q 10,60 10,30 z"),                     # :120-132
    ("tests/paths/pathRotatedArc.png", "M 20 50 A 20 10 45 1 1 80 50 L 20 50"),        # :134-140
    ("tests/paths/pathInvertedCornerArc.png", "M 0 50 A 50 50 0 0 0 50 0 L 50 50 L 0 50"),
    ("tests/paths/pathCornerArc.png", "M 0 50 A 50 50 0 0 1 50 0 L 50 50 L 0 50"),
])
def test_paths_arcs(orc, masters, name, path):
    a = _img(orc, 100, 100)
    orc.fill_path(a, path, _pink(orc))
    _cmp(orc, masters, name, a, exact=False)


def test_pixel_scale(orc, masters):
    # test_paths.nim:175-185: tolerance scaled by the transform
    a = _img(orc, 200, 200, WHITE)
    p = orc.parse_path("M1 0.5C1 0.776142 0.776142 1 0.5 1C0.223858 1 0 0.776142 0 0.5"
                       "C0 0.223858 0.223858 0 0.5 0C0.776142 0 1 0.223858 1 0.5Z")
    orc.fill_path(a, p, RED, mat=scale(200, 200))
    orc.stroke_path(a, p, (0, 255, 0, 255), width=0.01, mat=scale(200, 200))
    _cmp(orc, masters, "tests/paths/pixelScale.png", a)


@pytest.mark.parametrize("name,cap,join,closed", [
    ("tests/paths/boxRound.png", 1, 1, True), ("tests/paths/boxBevel.png", 1, 2, True),
    ("tests/paths/boxMiter.png", 1, 0, True), ("tests/paths/ButtCap.png", 0, 2, False),
    ("tests/paths/RoundCap.png", 1, 2, False), ("tests/paths/SquareCap.png", 2, 2, False),
])
def test_stroke_caps_joins(orc, masters, name, cap, join, closed):
    # test_paths.nim:187-251
    path = "M 3 3 L 20 3 L 20 20 L 3 20" + (" Z" if closed else "")
    a = _img(orc, 60, 60, WHITE)
    orc.stroke_path(a, path, BLACK, width=10, cap=cap, join=join, mat=translate(10, 10))
    _cmp(orc, masters, name, a)


def test_dashes(orc, masters):
    # test_paths.nim:253-288
    a = _img(orc, 60, 120, WHITE)
    for ty, d in [(5, ()), (25, (2, 2)), (45, (4, 4)), (65, (2, 4, 6, 2)), (85, (1,)),
                  (105, (1, 2, 3, 4, 5, 6, 7, 8, 9))]:
        orc.stroke_path(a, "M 0 0 L 50 0", BLACK, width=10, cap=0, join=2, dashes=d,
                        mat=translate(5, ty))
    _cmp(orc, masters, "tests/paths/dashes.png", a)


@pytest.mark.parametrize("angle,limit", [(10, 2), (145, 2), (155, 2), (165, 2), (165, 10),
                                         (145, 3.32), (145, 3.33)])
def test_miter_limit(orc, masters, angle, limit):
    # test_paths.nim:290-313.  degToRad returns float32; `+ PI/2` and sin/cos
    # are float64, narrowed at lineTo.
    a = _img(orc, 60, 60, WHITE)
    p = Path()
    p.move_to(-20, 0)
    p.line_to(0, 0)
    th = float(f32(float(f32(angle)) * math.pi / 180)) + math.pi / 2
    p.line_to(f32(math.sin(th) * 20), f32(math.cos(th) * 20))
    orc.stroke_path(a, p.array(), BLACK, width=8, cap=0, join=0, miter_limit=f32(limit),
                    mat=translate(30, 30))
    _cmp(orc, masters, "tests/paths/miterLimit_%ddeg_%.2fnum.png" % (angle, limit), a)


def test_selfclosing(orc, masters):
    # test_paths.nim:315-322
    a = _img(orc, 60, 60, WHITE)
    orc.fill_path(a, "M0 0 L0 0 L60 0 L60 60 L0 60", (127, 127, 127, 255))
    _cmp(orc, masters, "tests/paths/selfclosing.png", a)


@pytest.mark.parametrize("name,p1,p2,mode,stroke", [
    ("tests/paths/rectExcludeMask.png", "M 10 10 H 60 V 60 H 10 z", "M 30 30 H 80 V 80 H 30 z", 19, False),
    ("tests/paths/rectExcludeMaskAA.png", "M 10.1 10.1 H 60.1 V 60.1 H 10.1 z",
     "M 30.1 30.1 H 80.1 V 80.1 H 30.1 z", 19, False),
    ("tests/paths/rectMask.png", "M 10 10 H 60 V 60 H 10 z", "M 30 30 H 80 V 80 H 30 z", 16, False),
    ("tests/paths/rectMaskAA.png", "M 10.1 10.1 H 60.1 V 60.1 H 10.1 z",
     "M 30.1 30.1 H 80.1 V 80.1 H 30.1 z", 16, False),
    ("tests/paths/rectMaskStroke.png", "M 10 10 H 60 V 60 H 10 z", "M 30 30 H 50 V 50 H 30 z", 16, True),
])
def test_mask_modes(orc, masters, name, p1, p2, mode, stroke):
    # test_paths.nim:362-445
    a = _img(orc, 100, 100)
    orc.fill_path(a, p1, RED)
    if stroke:
        orc.stroke_path(a, p2, (0, 255, 0, 255), width=10, mode=mode)
    else:
        orc.fill_path(a, p2, (0, 255, 0, 255), mode=mode)
    _cmp(orc, masters, name, a)


@pytest.mark.parametrize("sides", range(3, 9))
def test_polygons(orc, masters, sides):
    # test_paths.nim:639-646
    p = Path()
    p.polygon(50, 50, 30, sides)
    a = _img(orc, 100, 100)
    orc.fill_path(a, p.array(), WHITE)
    _cmp(orc, masters, "tests/paths/polygon%d.png" % sides, a)


def test_swish(orc, masters):
    # test_paths.nim:689-698: cubic that folds back over itself
    a = _img(orc, 100, 100, WHITE)
    orc.fill_path(a, "M 40 40 L 40 80 L 80 80 L 80 40 C 80 -20 40 100 40 40", BLACK)
    _cmp(orc, masters, "tests/paths/pathSwish.png", a)


@pytest.mark.parametrize("stroke", [False, True])
def test_opacity(orc, masters, stroke):
    # test_paths.nim:582-601: color(1,0,1,1) with opacity 0.5 -> alpha 0.5
    p = Path()
    p.circle(50, 50, 30)
    a = _img(orc, 100, 100)
    c = orc.color_to_rgbx(1, 0, 1, 0.5)
    assert c == (128, 0, 128, 128)
    if stroke:
        orc.stroke_path(a, p.array(), c, width=10)
    else:
        orc.fill_path(a, p.array(), c)
    # With a translucent paint even alpha goes through the floor-vs-round
    # coverage multiply, so only the 1 LSB bound holds.
    _cmp(orc, masters, "tests/paths/opacity%s.png" % ("Stroke" if stroke else "Fill"), a,
         exact=False, alpha_exact=False)


def test_zero_width_fill_does_not_raise(orc):
    # test_paths.nim:648-655
    a = _img(orc, 100, 100)
    orc.fill_path(a, "M0 0 L0 1 L0 0 Z", RED)
    assert a.max() == 0


def test_fill_overlaps(orc):
    # test_paths.nim:544-560
    p = Path(); p.rect(0, 0, 10, 10)
    c = p.array()
    assert orc.fill_overlaps(c, 5, 5)
    assert orc.fill_overlaps(c, 0, 0)
    assert orc.fill_overlaps(c, 9, 0)
    assert orc.fill_overlaps(c, 0, 9)
    assert not orc.fill_overlaps(c, 10, 10)
    p = Path(); p.ellipse(20, 20, 20, 10)
    c = p.array()
    assert not orc.fill_overlaps(c, 0, 0)
    assert orc.fill_overlaps(c, 20, 20)
    assert orc.fill_overlaps(c, 10, 20)
    assert orc.fill_overlaps(c, 30, 20)


def test_parse_errors_raise(orc):
    with pytest.raises(orc.OracleError):
        orc.parse_path("M 1 2 3")


# ---- images.nim ----

@pytest.mark.parametrize("oob,name", [((0, 0, 0, 0), "imageblur20"), ((0, 0, 0, 255), "imageblur20oob")])
def test_blur_masters(orc, masters, oob, name):
    # test_images.nim:126-140
    a = _img(orc, 100, 100, BLACK)
    p = Path(); p.rect(25, 25, 50, 50)
    orc.fill_path(a, p.array(), WHITE)
    orc.blur(a, 20, oob)
    _cmp(orc, masters, "tests/images/%s.png" % name, a)


def test_mask_clears_on_draw(orc, masters):
    # test_images_draw.nim:307-360
    p = Path(); p.rect(10, 10, 80, 80)
    mask = _img(orc, 100, 100)
    orc.fill_path(mask, p.array(), WHITE)
    for i, (tx, ty) in enumerate([(0, 0), (50, -50), (50, 50), (-50, 50), (-50, -50)]):
        a = _img(orc, 100, 100, (0, 0, 255, 255))
        orc.draw(a, mask, tx, ty, 16)
        _cmp(orc, masters, "tests/images/maskClearsOnDraw%d.png" % i, a)
    for tx, ty in [(100, 100), (100, -100), (-100, 100), (-100, -100), (0, 1000), (0, -1000),
                   (1000, 0), (-1000, 0), (1000, 1000), (1000, -1000), (-1000, 1000), (-1000, -1000)]:
        a = _img(orc, 100, 100, (0, 0, 255, 255))
        orc.draw(a, mask, tx, ty, 16)
        assert a.max() == 0


def test_fill_optimization_known_answer(orc, masters):
    # test_images_draw.nim:266-282: fill + stroke composite, doAssert pixel
    img = _img(orc, 20, 20)
    s = _img(orc, 20, 20)
    ps = "M 0 0 L 20 0 L 20 20 L 0 20 z"
    orc.fill_path(img, ps, orc.color_to_rgbx(1.0, 0.5, 0.25, 1.0))
    orc.stroke_path(s, ps, WHITE, width=4)
    orc.draw(img, s)
    assert tuple(img[10, 10]) == (255, 128, 64, 255)
    _cmp(orc, masters, "tests/images/fillOptimization.png", img)


def test_shadow_example(orc, masters):
    # examples/shadow.nim
    image = _img(orc, 200, 200, WHITE)
    p = Path(); p.polygon(100, 100, 70, 8)
    poly = _img(orc, 200, 200)
    orc.fill_path(poly, p.array(), WHITE)
    sh = orc.shadow(poly, (2, 2), 2, 10, orc.rgba_to_rgbx((0, 0, 0, 200)))
    orc.draw(image, sh)
    orc.draw(image, poly)
    _cmp(orc, masters, "examples/shadow.png", image)


def test_blur_mask_example(orc, masters):
    # examples/blur.nim: blur(20) of a photo, masked by a hexagon
    trees = masters("examples/data/trees.png")
    bl = trees.copy()
    image = _img(orc, 200, 200, WHITE)
    p = Path(); p.polygon(100, 100, 70, 6)
    mk = _img(orc, 200, 200)
    orc.fill_path(mk, p.array(), WHITE)
    orc.blur(bl, 20)
    orc.draw(bl, mk, 0, 0, 16)
    orc.draw(image, trees)
    orc.draw(image, bl)
    _cmp(orc, masters, "examples/blur.png", image)


def test_known_answers_images(orc):
    # test_images.nim:19-23 premultiply; :25-29 straight alpha; :31-36 applyOpacity
    a = np.zeros((10, 10, 4), np.uint8); a[:] = (255, 0, 0, 128)
    assert tuple(orc.to_premultiplied_alpha(a)[9, 9]) == (128, 0, 0, 128)
    d = np.zeros((1, 100, 4), np.uint8); d[:] = (100, 0, 0, 128)
    assert tuple(orc.to_straight_alpha(d)[0, 10]) == (199, 0, 0, 128)
    img = _img(orc, 100, 100, (200, 200, 200, 200))
    orc.apply_opacity(img, 0.5)
    assert tuple(img[0, 0]) == (100, 100, 100, 100)
    assert tuple(img[88, 88]) == (100, 100, 100, 100)


def test_premultiply_identity_exhaustive(orc):
    # test_images.nim:204-228: all 65,536 (r, a) pairs; float round path ==
    # integer (c*a+127) div 255 == chroma rgbx(ColorRGBA)
    r, a = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    px = np.zeros((256, 256, 4), np.uint8)
    px[..., 0] = r
    px[..., 3] = a
    conv = orc.to_premultiplied_alpha(px)
    fl = np.round((r / 255.0) * (a / 255.0) * 255).astype(np.uint8)  # float path of the test
    assert np.array_equal(conv[..., 0], fl)
    for rr, aa in [(0, 0), (1, 1), (255, 128), (77, 200), (254, 254), (3, 129)]:
        assert orc.rgba_to_rgbx((rr, 0, 0, aa))[0] == conv[rr, aa, 0]
