"""Pins the oracle's transformed image draws (images.nim:168-259, 367-445,
531-683) and image paints (paths.nim:2115-2142 with ImagePaint /
TiledImagePaint).

* Exact (bit for bit, in the straight alpha pixie writes): the draw masters
  rotate0/90/270 and scaleHalf (tests/test_images_draw.nim:3-51, 121-129),
  smooth8/9 (magnification, :195-211) and minify_odd (:245-252); the
  minifyBy2 / magnifyBy2 masters of tests/test_images.nim:90-118 (odd
  edges, two powers, mandrill); and the image-paint masters
  (tests/test_paints.nim:20-55: mandrill.png under scale(0.2) and, tiled,
  scale(0.02), at opacity 1 and 0.5), which run minifyBy2 and bilinear
  sampling at fractional positions.
* rotate180 / rotate360: the float32 sine of -pi and -2 pi is not 0, which
  puts one extra row inside drawSmooth's y-range; those masters differ from
  this restatement in that row only (its samples fall a pixel outside the
  source, which reads transparent here and the source edge in the master),
  and are exact elsewhere.
* tests/images/masters/smooth1-7, 10-12 are not pixie's output (their
  fractional offsets come out quantized to 1/4 px, and half-pixel mixes
  round to 127 where mix gives 128); they bound the geometry by pixie's
  diff score (images.nim:138-166) instead.
* rock_minified.png rounds the 2x2 mean down where minifyBy2 adds 2 before
  the division (images.nim:200-205, and the exact minifiedMandrill master
  confirms +2); the test checks that every difference is such a rounding.
* Closed forms: minifyBy2 of an even image is the rounded 2x2 mean;
  magnifyBy2 repeats pixels; a draw by an integer translation equals the
  integer composite; getRgbaSmooth at integer coordinates reads the pixel,
  at a half step the rounded mean of neighbours (mix at t = 0.5 rounds to
  x = 128).
"""
import math

import numpy as np
import pytest

from pixie_amd.paths import f32, sinf, cosf

ROOT = __file__.rsplit("/tests/", 1)[0]


def _input(orc, name):
    from PIL import Image as PILImage
    a = np.asarray(PILImage.open("%s/tests/golden/inputs/%s" % (ROOT, name)).convert("RGBA")).copy()
    return orc.to_premultiplied_alpha(a)


def m3(*v):
    return np.array(v, np.float32)


def T(x, y):
    return m3(1, 0, 0, 0, 1, 0, x, y, 1)


def S(x, y):
    return m3(x, 0, 0, 0, y, 0, 0, 0, 1)


def R(angle):
    """vmath rotate(angle) with the angle already float32."""
    a = f32(angle)
    s, c = sinf(a), cosf(a)
    return m3(c, s, 0, -s, c, 0, 0, 0, 1)


def mul(a, b):
    """vmath Mat3 product in float32, each entry summed left to right."""
    A = a.reshape(3, 3)  # A[i, j] = a[3 i + j]
    B = b.reshape(3, 3)
    r = np.zeros((3, 3), np.float32)
    for i in range(3):
        for j in range(3):
            acc = np.float32(B[i, 0] * A[0, j])
            acc = np.float32(acc + np.float32(B[i, 1] * A[1, j]))
            acc = np.float32(acc + np.float32(B[i, 2] * A[2, j]))
            r[i, j] = acc
    return r.reshape(9)


def rad(deg):
    """toRadians on an integer literal: PI * deg / 180 in float64, used as
    a float32 angle."""
    return f32(math.pi * deg / 180.0)


def _cmp(masters, name, img):
    # pixie writes masters in straight alpha; compare there, since the
    # premultiplied -> straight -> premultiplied round trip is lossy
    import oracle
    m = masters(name, premultiply=False)
    assert m.shape == img.shape, (m.shape, img.shape)
    s = oracle.to_straight_alpha(img)
    d = np.abs(m.astype(int) - s.astype(int))
    assert d.max() == 0, "%s: %d pixels differ (max %d), first %s" % (
        name, (d.max(-1) > 0).sum(), d.max(), np.argwhere(d.max(-1) > 0)[0])


def _canvas(orc, w, h, color):
    a = orc.new_image(w, h)
    orc.fill(a, color)
    return a


RED, GREEN, WHITE, BLACK = (255, 0, 0, 255), (0, 255, 0, 255), (255, 255, 255, 255), (0, 0, 0, 255)


def _rotated(orc, deg):
    a = _canvas(orc, 1000, 1000, RED)
    b = _canvas(orc, 500, 500, GREEN)
    m = T(250, 250) if deg == 0 else mul(T(250, 250), R(f32(f32(-deg) * f32(math.pi)) / f32(180)))
    orc.draw_transform(a, b, m)
    return a


@pytest.mark.parametrize("deg", [0, 90, 270])
def test_rotate_multiples_of_90(orc, masters, deg):
    # test_images_draw.nim:3-51
    _cmp(masters, "tests/images/rotate%d.png" % deg, _rotated(orc, deg))


@pytest.mark.parametrize("deg,row,cols", [(180, 250, (0, 250)), (360, 249, (250, 750))])
def test_rotate_half_and_full_turn(orc, masters, deg, row, cols):
    a = orc.to_straight_alpha(_rotated(orc, deg))
    m = masters("tests/images/rotate%d.png" % deg, premultiply=False)
    d = np.any(a != m, axis=-1)
    ys, xs = np.nonzero(d)
    assert set(ys.tolist()) == {row} and xs.min() == cols[0] and xs.max() == cols[1] - 1
    assert (a[row, cols[0]:cols[1]] == RED).all() and (m[row, cols[0]:cols[1]] == GREEN).all()


def test_scale_half(orc, masters):
    # :121-129
    a = _canvas(orc, 1000, 1000, RED)
    b = _canvas(orc, 500, 500, GREEN)
    orc.draw_transform(a, b, mul(T(250, 250), S(0.5, 0.5)))
    _cmp(masters, "tests/images/scaleHalf.png", a)


SMOOTH = [
    # (master, canvas colour, (b size or input file, b colour), transform)   :131-243
    ("smooth1", (99, 99), BLACK, lambda: T(0.5, 0.5)),
    ("smooth2", (50, 50), BLACK, lambda: mul(T(0, 50), R(-rad(45)))),
    ("smooth3", (50, 50), BLACK, lambda: T(25.2, 25)),
    ("smooth4", (50, 50), BLACK, lambda: T(25.2, 25.6)),
    ("smooth5", (10, 10), RED, lambda: mul(T(50, 50), R(-rad(30)))),
    ("smooth6", "turtle.png", None, lambda: mul(T(50, 50), R(-rad(30)))),
    ("smooth7", "turtle@10x.png", None, lambda: mul(mul(T(50, 50), R(-rad(30))), S(0.1, 0.1))),
    ("smooth8", "turtle.png", None, lambda: S(2, 2)),
    ("smooth9", "turtle.png", None, lambda: mul(T(1, 1), S(2, 2))),
    ("smooth10", "turtle.png", None, lambda: mul(T(0.5, 0.5), S(2, 2))),
    ("smooth11", "turtle.png", None,
     lambda: mul(mul(T(-43.29, -103.87), R(rad(15))), S(263.86 / 40, 263.86 / 40))),
]


EXACT_SMOOTH = ("smooth8", "smooth9")


@pytest.mark.parametrize("name,src,color,mat", SMOOTH, ids=[s[0] for s in SMOOTH])
def test_smooth(orc, masters, name, src, color, mat):
    a = _canvas(orc, 100, 100, WHITE)
    if isinstance(src, str):
        b = _input(orc, src)
    else:
        b = _canvas(orc, src[0], src[1], color)
    orc.draw_transform(a, b, mat())
    if name in EXACT_SMOOTH:
        _cmp(masters, "tests/images/masters/%s.png" % name, a)
    else:
        score = orc.diff_score(masters("tests/images/masters/%s.png" % name), a)
        assert score < 0.5, score


def test_smooth12_four_draws(orc, masters):
    # :233-243
    a = _canvas(orc, 100, 100, WHITE)
    b = _input(orc, "turtle.png")
    m = mul(T(50, 50), R(rad(5)))
    for tx, ty in [(0, 0), (-40, 0), (-40, -40), (0, -40)]:
        orc.draw_transform(a, b, mul(m, T(tx, ty)))
    assert orc.diff_score(masters("tests/images/masters/smooth12.png"), a) < 0.5


def test_minify_odd(orc, masters):
    # :245-252
    a = _canvas(orc, 100, 100, WHITE)
    b = _canvas(orc, 99, 99, BLACK)
    orc.draw_transform(a, b, S(0.5, 0.5))
    _cmp(masters, "tests/images/masters/minify_odd.png", a)


@pytest.mark.parametrize("src,power,name", [
    ("flipped1.png", 1, "minifiedBy2"), ("flipped1.png", 2, "minifiedBy4"),
    ("mandrill.png", 1, "minifiedMandrill")])
def test_minify_masters(orc, masters, src, power, name):
    # test_images.nim:90-94, 102-106, 114-118 (flipped1 is 101x101: odd edges)
    _cmp(masters, "tests/images/%s.png" % name, orc.minify_by2(_input(orc, src), power))


@pytest.mark.parametrize("src,power,name", [("minifiedBy2", 1, "magnifiedBy2"),
                                            ("minifiedBy4", 2, "magnifiedBy4")])
def test_magnify_masters(orc, masters, src, power, name):
    # test_images.nim:96-100, 108-112 (their inputs are the minify masters)
    _cmp(masters, "tests/images/%s.png" % name,
         orc.magnify_by2(masters("tests/images/%s.png" % src), power))


def test_rock_minified_rounds_down(orc, masters):
    # test_images_draw.nim:254-258
    src = _input(orc, "rock.png")
    got = orc.minify_by2(src)
    m = masters("tests/images/masters/rock_minified.png")
    # rock.png is 400x177: the even rows give 88 full 2x2 rows
    h, w = src.shape[0] // 2, src.shape[1] // 2
    sums = src[:2 * h, :2 * w].astype(int).reshape(h, 2, w, 2, 4).sum(axis=(1, 3))
    assert np.array_equal(got[:h], ((sums + 2) // 4).astype(np.uint8))
    # the master is the floor of the mean: premultiplied it differs by at
    # most 1, and only where the sum's remainder mod 4 is 2 or 3 (or where
    # straight -> premultiplied rounding moves a low-alpha channel)
    d = got[:h].astype(int) - m[:h].astype(int)
    assert np.abs(d).max() <= 1
    rem = sums % 4
    assert (rem[..., 3][d[..., 3] != 0] >= 2).all()


@pytest.mark.parametrize("name,tiled,mat,opacity", [
    ("paintImage", False, (0.2, 0.2), 1.0),
    ("paintImageOpacity", False, (0.2, 0.2), 0.5),
    ("paintImageTiled", True, (0.02, 0.02), 1.0),
    ("paintImageTiledOpacity", True, (0.02, 0.02), 0.5),
])
def test_image_paint_masters(orc, masters, name, tiled, mat, opacity):
    # test_paints.nim:20-55: the heart filled with mandrill.png
    from test_oracle_paints import HEART
    paint = orc.make_image_paint(_input(orc, "mandrill.png"), S(*mat), tiled=tiled,
                                 opacity=opacity)
    img = orc.new_image(100, 100)
    orc.fill_path_paint(img, HEART, paint)
    _cmp(masters, "tests/paths/%s.png" % name, img)


# ---- closed forms ----

def _rand_img(seed, w, h):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    a[..., :3] = np.minimum(a[..., :3], a[..., 3:4])
    return a


def test_minify_even_is_rounded_mean(orc):
    a = _rand_img(1, 64, 48)
    got = orc.minify_by2(a)
    s = a.astype(int).reshape(24, 2, 32, 2, 4).sum(axis=(1, 3))
    assert np.array_equal(got, ((s + 2) // 4).astype(np.uint8))


def test_magnify_repeats(orc):
    a = _rand_img(2, 13, 7)
    got = orc.magnify_by2(a, 2)
    assert np.array_equal(got, np.repeat(np.repeat(a, 4, axis=0), 4, axis=1))


def test_integer_translation_is_composite(orc):
    for mode in (0, 2, 10, 16, 17):
        base = _rand_img(3, 80, 60)
        b = _rand_img(4, 30, 20)
        got = base.copy()
        orc.draw_transform(got, b, T(17, -5), mode)
        want = base.copy()
        orc.draw(want, b, 17, -5, mode)
        assert np.array_equal(got, want), mode


def test_smooth_sampling_points(orc):
    b = _rand_img(5, 9, 7)
    for x, y in [(0, 0), (3, 4), (8, 6)]:
        assert orc.get_rgba_smooth(b, x, y) == tuple(b[y, x])
    # half a pixel right: mix at t = 0.5 uses weight round(127.5) = 128
    x, y = 2, 3
    p, q = b[y, x].astype(int), b[y, x + 1].astype(int)
    want = tuple(int(v) for v in (p * 127 + q * 128 + 127) // 255) if not np.array_equal(p, q) \
        else tuple(int(v) for v in p)
    assert orc.get_rgba_smooth(b, x + 0.5, y) == want
    # outside the image: transparent; wrapped: the pixel modulo the size
    assert orc.get_rgba_smooth(b, -5, -5) == (0, 0, 0, 0)
    assert orc.get_rgba_smooth(b, 9 + 3, 7 + 2, wrapped=True) == tuple(b[2, 3])


def test_sub_pixel_tall_draw_on_a_scanline_draws_nothing(orc):
    # drawSmooth keeps an intersection only if it is not the segment's end
    # point (images.nim:583).  An image 10x1 under T(3, 5) * S(1, 0.75) has
    # corners (3, 5), (13, 5), (13, 5.75), (3, 5.75): row 5's scanline y = 5
    # meets the right edge at its start (kept) and the left edge at its end
    # (dropped), y = 6 meets nothing, so the row's range is [13, 13) and it
    # is skipped; without the rule the row would span [3, 13)
    a = _canvas(orc, 20, 10, WHITE)
    before = a.copy()
    b = _canvas(orc, 10, 1, BLACK)
    orc.draw_transform(a, b, mul(T(3, 5), S(1, 0.75)))
    assert np.array_equal(a, before)
    # half a pixel lower the left edge's end is off the scanline and the row
    # is drawn
    orc.draw_transform(a, b, mul(T(3, 5.5), S(1, 0.75)))
    assert (a[5, 3:13, 3] == 255).all() and (a[5, 3:13, 0] < 255).all()


def test_shadow_fractional_offset_is_smooth_draw(orc):
    # shadow (images.nim:760-776) with a fractional offset: the mask is the
    # image drawn with OverwriteBlend through drawSmooth, then spread, blur
    # and the colour masked by it
    src = _rand_img(7, 90, 70)
    color = (20, 40, 60, 200)
    got = orc.shadow(src, (3.25, -1.5), 2, 4, color)
    mask = orc.new_image(90, 70)
    orc.draw_transform(mask, src, T(3.25, -1.5), 17)
    orc.spread(mask, 2)
    orc.blur(mask, 4)
    want = _canvas(orc, 90, 70, color)
    orc.draw(want, mask, 0, 0, 16)
    assert np.array_equal(got, want)
    # the sub-pixel shift changes the shadow's edge
    assert not np.array_equal(got, orc.shadow(src, (3, -1), 2, 4, color))
