"""Pins the oracle's gradient paints (paints.nim:72-248) and fillPath's route
for non-solid paints (paths.nim:2115-2142).

* The reference's gradient masters (tests/test_paints.nim:57-134): the heart
  filled with a linear (two directions), radial, angular and angular-at-half-
  opacity gradient, compared with the oracle bit for bit.
* Closed forms over the whole canvas, in float64 numpy: the linear parameter
  is the projection onto the handle axis, the radial one the elliptical
  distance in the handle frame, the angular one the clockwise angle from the
  handle direction; the colour is the straight mix of the two stops around t,
  rounded to bytes and premultiplied.  Each is checked to 1 LSB (float32
  against float64) on pixels whose t is not within 1e-4 of a stop.
* Stop selection: a single stop gives a constant canvas; t outside the stops
  clamps to the end stops.
"""
import math

import numpy as np
import pytest

HEART = """
    M 10,30
    A 20,20 0,0,1 50,30
    A 20,20 0,0,1 90,30
    Q 90,60 50,90
    Q 10,60 10,30 z
  """
STOPS = [((1, 0, 0, 1), 0.0), ((1, 0, 0, 0.15625), 1.0)]


@pytest.mark.parametrize("name,kind,handles,opacity", [
    ("tests/paths/gradientLinear.png", 3, [(0, 50), (100, 50)], 1.0),
    ("tests/paths/gradientLinear2.png", 3, [(50, 0), (50, 100)], 1.0),
    ("tests/paths/gradientRadial.png", 4, [(50, 50), (100, 50), (50, 100)], 1.0),
    ("tests/paths/gradientAngular.png", 5, [(50, 50), (100, 50), (50, 100)], 1.0),
    ("tests/paths/gradientAngularOpacity.png", 5, [(50, 50), (100, 50), (50, 100)], 0.5),
])
@pytest.mark.parametrize("variant", ["scalar", "avx2"])
def test_gradient_masters(orc, masters, name, kind, handles, opacity, variant):
    m = masters(name)
    img = orc.new_image(100, 100)
    paint = orc.make_paint(kind, handles, STOPS, opacity)
    if variant == "avx2":
        with orc.simd_mode():
            orc.fill_path_paint(img, HEART, paint)
    else:
        orc.fill_path_paint(img, HEART, paint)
    d = np.abs(m.astype(int) - img.astype(int))
    assert d.max() == 0, "%s: %d pixels differ (max %d)" % (name, (d.max(-1) > 0).sum(), d.max())


def _expected(t, stops):
    """Colour of each t by the stop rule, straight float mix, bytes,
    premultiplied; t near a stop position is flagged for exclusion."""
    pos = np.array([p for _, p in stops])
    col = np.array([c for c, _ in stops], np.float64)
    out = np.empty(t.shape + (4,), np.float64)
    idx = np.searchsorted(pos, t, side="left") - 1   # last stop strictly below t
    lo = np.clip(idx, 0, len(stops) - 1)
    hi = np.clip(idx + 1, 0, len(stops) - 1)
    span = np.where(hi > lo, pos[hi] - pos[lo], 1.0)
    v = np.where((idx >= 0) & (hi > lo), (t - pos[lo]) / span, 0.0)
    out[:] = col[lo] * (1 - v[..., None]) + col[hi] * v[..., None]
    b = np.round(np.clip(out, 0, 1) * 255)
    a = b[..., 3:4]
    prem = np.where(a == 255, b, (b * a + 127) // 255)
    prem[..., 3] = b[..., 3]
    near = np.min(np.abs(t[..., None] - pos), axis=-1) < 1e-4
    return prem.astype(int), near


def _check_canvas(orc, kind, handles, stops, tfun, w=120, h=90):
    img = orc.new_image(w, h)
    orc.fill_gradient(img, orc.make_paint(kind, handles, stops))
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    t = tfun(xx, yy)
    want, near = _expected(t, stops)
    d = np.abs(img.astype(int) - want).max(-1)
    d[near] = 0
    assert d.max() <= 1, "max diff %d at %s" % (d.max(), np.argwhere(d > 1)[0])
    return img


STOPS3 = [((0.1, 0.9, 0.3, 1.0), 0.2), ((1.0, 0.5, 0.0, 0.5), 0.6), ((0.0, 0.2, 1.0, 0.8), 0.9)]


def test_linear_closed_form(orc):
    at, to = (13.0, 7.0), (97.0, 61.0)
    dx, dy = to[0] - at[0], to[1] - at[1]
    _check_canvas(orc, 3, [at, to], STOPS3,
                  lambda x, y: ((x - at[0]) * dx + (y - at[1]) * dy) / (dx * dx + dy * dy))


def test_radial_closed_form(orc):
    # an ellipse in a frame rotated by the handle direction: t is the
    # distance to the centre with the along-axis and across-axis components
    # divided by the two radii
    c, e, s = (60.0, 45.0), (100.0, 75.0), (45.0, 65.0)
    rx = math.hypot(e[0] - c[0], e[1] - c[1])
    ry = math.hypot(s[0] - c[0], s[1] - c[1])
    # the frame's angle is that of centre - edge (paints.nim:187-190)
    ang = math.atan2(c[1] - e[1], c[0] - e[0])
    ca, sa = math.cos(ang), math.sin(ang)

    def t(x, y):
        u = (x - c[0]) * ca + (y - c[1]) * sa
        v = -(x - c[0]) * sa + (y - c[1]) * ca
        return np.hypot(u / rx, v / ry)
    _check_canvas(orc, 4, [c, e, s], STOPS3, t)


def test_angular_closed_form(orc):
    c, e = (55.0, 40.0), (90.0, 20.0)
    ga = math.atan2(e[1] - c[1], e[0] - c[0])

    def t(x, y):
        a = np.arctan2(y - c[1], x - c[0]) + ga + math.pi / 2
        a = (a + math.pi) % (2 * math.pi) - math.pi
        return a / (2 * math.pi) + 0.5
    img = orc.new_image(120, 90)
    orc.fill_gradient(img, orc.make_paint(5, [c, e, (0, 0)], STOPS3))
    yy, xx = np.mgrid[0:90, 0:120].astype(np.float64)
    tt = t(xx, yy)
    want, near = _expected(tt, STOPS3)
    # the seam where the angle wraps (t = 0 / 1) and the centre pixel are
    # decided by float32 rounding; leave them out
    near |= (np.abs(tt) < 1e-4) | (np.abs(tt - 1) < 1e-4) | ((xx == c[0]) & (yy == c[1]))
    d = np.abs(img.astype(int) - want).max(-1)
    d[near] = 0
    assert d.max() <= 1


def test_single_stop_constant(orc):
    img = orc.new_image(40, 30)
    orc.fill_gradient(img, orc.make_paint(3, [(0, 0), (40, 30)], [((0.2, 0.4, 0.6, 0.5), 0.3)]))
    # rgba (51, 102, 153, 128) premultiplied
    want = np.array([(51 * 128 + 127) // 255, (102 * 128 + 127) // 255,
                     (153 * 128 + 127) // 255, 128])
    assert (img.reshape(-1, 4) == want).all()


def test_no_stops_raises(orc):
    img = orc.new_image(8, 8)
    with pytest.raises(Exception, match="at least 1 color stop"):
        orc.fill_gradient(img, orc.make_paint(3, [(0, 0), (8, 0)], []))


def test_fill_path_paint_is_mask_and_draw(orc):
    # the route itself: coverage from a white fill, gradient masked by it,
    # then drawn; built here from the oracle's own primitives at another
    # opacity and blend mode than the masters use
    w = h = 64
    paint = orc.make_paint(4, [(30, 30), (60, 30), (30, 50)], STOPS3, opacity=0.7, mode=2)
    base = np.random.default_rng(5).integers(0, 256, (h, w, 4), dtype=np.uint8)
    base[..., :3] = np.minimum(base[..., :3], base[..., 3:4])
    got = base.copy()
    orc.fill_path_paint(got, "M 5 5 L 60 12 L 40 58 Z", paint)
    mask = orc.new_image(w, h)
    orc.fill_path(mask, "M 5 5 L 60 12 L 40 58 Z", (255, 255, 255, 255))
    orc.apply_opacity(mask, 0.7)
    fill = orc.new_image(w, h)
    orc.fill_gradient(fill, orc.make_paint(4, [(30, 30), (60, 30), (30, 50)], STOPS3))
    orc.draw(fill, mask, 0, 0, 16)
    want = base.copy()
    orc.draw(want, fill, 0, 0, 2)
    assert np.array_equal(got, want)


def test_hard_stop_takes_first_duplicate(orc):
    # two stops at one position make a hard edge; at exactly that position the
    # stop rule (last stop strictly below t, then the next one, paints.nim:
    # 77-82) lands on the first of the pair: t = 0.5 mixes stops 0 and 1 with
    # weight 1, which is stop 1's colour, not stop 2's
    stops = [((1, 0, 0, 1), 0.0), ((0, 1, 0, 1), 0.5), ((0, 0, 1, 1), 0.5), ((1, 1, 1, 1), 1.0)]
    img = orc.new_image(101, 1)
    orc.fill_gradient(img, orc.make_paint(3, [(0, 0), (100, 0)], stops))
    assert tuple(img[0, 50]) == (0, 255, 0, 255)
    assert tuple(img[0, 51]) == (5, 5, 255, 255)   # 0.02 of the way to white: 5.1
    assert tuple(img[0, 0]) == (255, 0, 0, 255) and tuple(img[0, 100]) == (255, 255, 255, 255)
