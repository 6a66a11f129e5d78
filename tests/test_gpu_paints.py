"""Gradient paints on the GPU (paints.nim:72-248 through fillPath's mask
route, paths.nim:2115-2142) against the oracle and the reference's masters.

The device evaluates the gradient per pixel in the same float32 operation
order as the oracle (no contraction, correctly rounded division and square
root); the radial matrix and the angular handle angle are computed on the
host with the C library, as the oracle does.  Only the angular parameter
calls a device atan2f, whose last-bit rounding may differ from the host's, so
the angular bar is the coverage bar (<= 1/255 on <= 0.05 % of pixels); linear
and radial results must be bit-exact wherever the coverage is.
"""
import numpy as np
import pytest

import pixie_amd as px

pytestmark = pytest.mark.gpu

HEART = """
    M 10,30
    A 20,20 0,0,1 50,30
    A 20,20 0,0,1 90,30
    Q 90,60 50,90
    Q 10,60 10,30 z
  """
STOPS = [((1, 0, 0, 1), 0.0), ((1, 0, 0, 0.15625), 1.0)]
STOPS3 = [((0.1, 0.9, 0.3, 1.0), 0.2), ((1.0, 0.5, 0.0, 0.5), 0.6), ((0.0, 0.2, 1.0, 0.8), 0.9)]


@pytest.fixture(scope="module")
def gpu():
    px.init()
    return px


def _paint(kind, handles, stops, opacity=1.0, mode=0):
    return px.Paint(kind, blendMode=mode, opacity=opacity, gradientHandlePositions=handles,
                    gradientStops=[px.ColorStop(c, p) for c, p in stops])


def _diff(got, want):
    d = np.abs(got.astype(int) - want.astype(int)).max(-1)
    return int(d.max()), float(np.mean(d > 0))


@pytest.mark.parametrize("name,kind,handles,opacity", [
    ("tests/paths/gradientLinear.png", 3, [(0, 50), (100, 50)], 1.0),
    ("tests/paths/gradientLinear2.png", 3, [(50, 0), (50, 100)], 1.0),
    ("tests/paths/gradientRadial.png", 4, [(50, 50), (100, 50), (50, 100)], 1.0),
    ("tests/paths/gradientAngular.png", 5, [(50, 50), (100, 50), (50, 100)], 1.0),
    ("tests/paths/gradientAngularOpacity.png", 5, [(50, 50), (100, 50), (50, 100)], 0.5),
])
def test_gradient_masters(gpu, masters, name, kind, handles, opacity):
    img = px.Image(100, 100)
    img.fillPath(HEART, _paint(kind, handles, STOPS, opacity))
    mx, frac = _diff(img.download(), masters(name))
    print("%s: max diff %d on %.4f%% of pixels" % (name, mx, 100 * frac))
    assert mx == 0


def _random_path(rng, W, H, n=9):
    pts = rng.uniform([-20, -20], [W + 20, H + 20], (n, 2))
    s = "M %f %f " % tuple(pts[0])
    for p in pts[1:]:
        s += "L %f %f " % tuple(p)
    return s + "Z"


def _backdrop(rng, W, H):
    b = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    b[..., :3] = np.minimum(b[..., :3], b[..., 3:4])
    b[rng.random((H, W)) < 0.2] = 0
    return b


CASES = [
    (3, [(30.5, 20.0), (480.0, 350.25)]),
    (4, [(250.0, 190.0), (400.0, 260.0), (200.0, 330.0)]),
    (5, [(260.5, 170.25), (100.0, 40.0), (0.0, 0.0)]),
]


@pytest.mark.parametrize("mode", range(20))
@pytest.mark.parametrize("case", range(3))
def test_random_paths_every_mode(gpu, orc, mode, case):
    W, H = 512, 384
    kind, handles = CASES[case]
    rng = np.random.default_rng(100 + 20 * case + mode)
    base = _backdrop(rng, W, H)
    path = _random_path(rng, W, H)
    winding = int(rng.integers(0, 2))
    opacity = [1.0, 0.6, 0.25][mode % 3]
    img = px.Image.from_array(base)
    img.fillPath(path, _paint(kind, handles, STOPS3, opacity, mode), windingRule=winding)
    got = img.download()
    want = base.copy()
    st = orc.fill_path_paint(want, path, orc.make_paint(kind, handles, STOPS3, opacity, mode),
                             winding=winding)
    assert st["dropped"] == 0
    mx, frac = _diff(got, want)
    print("kind %d mode %d opacity %g: max diff %d on %.5f%%" % (kind, mode, opacity, mx,
                                                                  100 * frac))
    assert mx <= 1 and frac <= (5e-4 if kind == 5 else 1e-4)


def test_transform_and_large_canvas(gpu, orc):
    # a 2048x1536 canvas, a transformed heart, a radial gradient with an
    # elliptical frame and 8 stops
    W, H = 2048, 1536
    rng = np.random.default_rng(7)
    stops = [((float(a), float(b), float(c), float(d)), p) for (a, b, c, d), p in
             zip(rng.random((8, 4)), np.sort(rng.random(8)))]
    m = np.array([18.0, 3.0, 0.0, -4.0, 15.0, 0.0, 100.0, 80.0, 1.0], np.float32)
    paint = (4, [(900.0, 700.0), (1700.0, 900.0), (850.0, 1400.0)])
    img = px.Image(W, H)
    img.fill((10, 20, 30, 255))
    img.fillPath(HEART, _paint(paint[0], paint[1], stops, 0.8, 0), transform=m)
    want = orc.new_image(W, H)
    orc.fill(want, (10, 20, 30, 255))
    st = orc.fill_path_paint(want, HEART, orc.make_paint(paint[0], paint[1], stops, 0.8, 0),
                             mat=m)
    assert st["dropped"] == 0
    mx, frac = _diff(img.download(), want)
    print("2048x1536 radial, 8 stops: max diff %d on %.5f%%" % (mx, 100 * frac))
    assert mx <= 1 and frac <= 1e-4


def test_errors_and_no_ops(gpu):
    img = px.Image(64, 64)
    img.fill((1, 2, 3, 4))
    before = img.download()
    img.fillPath("M 0 0 L 64 0 L 64 64 Z", _paint(3, [(0, 0), (64, 0)], STOPS, opacity=0.0))
    assert np.array_equal(img.download(), before)          # opacity 0 draws nothing
    with pytest.raises(px.PixieError, match="at least 1 color stop"):
        img.fillPath("M 0 0 L 64 0 L 64 64 Z", _paint(3, [(0, 0), (64, 0)], []))
    with pytest.raises(px.PixieError, match="requires 2 handles"):
        img.fillPath("M 0 0 L 64 0 L 64 64 Z", _paint(3, [(0, 0), (64, 0), (1, 1)], STOPS))
    with pytest.raises(px.PixieError):
        img.fillPath("M 0 0 L 64 0 L 64 64 Z", _paint(3, [(0, 0), (64, 0)], STOPS * 17))
    assert np.array_equal(img.download(), before)
