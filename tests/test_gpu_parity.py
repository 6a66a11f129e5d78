"""GPU parity of libpixie_hip against the oracle (pixie's scalar semantics).

Bars (from the reference's own numerics): blends, masks, composites, blur,
spread and shadow are bit-exact on RGBA8; path coverage may differ by at most
COV_MAX_DIFF per channel on at most COV_MAX_FRAC of the pixels (float32
accumulation order).  The implementation replicates pixie's float32 order, so
the path tests also report how many pixels differ at all.
"""
import numpy as np
import pytest

import pixie_amd as px
from pixie_amd import workloads as WL
from pixie_amd.paths import Path, translate, scale

pytestmark = pytest.mark.gpu

COV_MAX_DIFF = 1        # 1/255 per channel
COV_MAX_FRAC = 1e-4     # 0.01 % of pixels


def _cov_check(got, want, what=""):
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    px_diff = d.max(axis=-1) if d.ndim == 3 else d
    frac = float(np.mean(px_diff > 0))
    print("%s: max diff %d, %.6f%% of pixels differ" % (what, int(d.max()), 100 * frac))
    assert d.max() <= COV_MAX_DIFF, what
    assert frac <= COV_MAX_FRAC, what


def _exact(got, want, what=""):
    if not np.array_equal(got, want):
        d = np.abs(got.astype(np.int32) - want.astype(np.int32))
        idx = np.argwhere(d.max(axis=-1) > 0 if d.ndim == 3 else d > 0)
        raise AssertionError("%s: %d pixels differ, first %s got %s want %s" % (
            what, len(idx), idx[0], got[tuple(idx[0])], want[tuple(idx[0])]))


def _rand_premul(rng, n):
    a = rng.integers(0, 256, size=n)
    out = np.empty((n, 4), np.uint8)
    for k in range(3):
        out[:, k] = np.floor(rng.random(n) * (a + 1)).astype(np.uint8)
    out[:, 3] = a
    return out


@pytest.fixture(scope="module")
def gpu():
    px.init()
    return px


# ---------------------------------------------------------------- blends --

@pytest.mark.parametrize("mode", range(20))
def test_blend_modes_bit_exact(gpu, orc, mode):
    rng = np.random.default_rng(100 + mode)
    n = 1 << 18
    b = _rand_premul(rng, n)
    s = _rand_premul(rng, n)
    # edge values: fully transparent / opaque / extremes, plus invalid
    # (channel > alpha) pixels to exercise Nim's wrapping arithmetic
    edge = np.array([[0, 0, 0, 0], [255, 255, 255, 255], [0, 0, 0, 255], [1, 2, 3, 4],
                     [255, 0, 128, 128], [200, 100, 50, 30]], np.uint8)
    b = np.concatenate([b, np.repeat(edge, len(edge), 0), rng.integers(0, 256, (4096, 4), dtype=np.uint8)])
    s = np.concatenate([s, np.tile(edge, (len(edge), 1)), rng.integers(0, 256, (4096, 4), dtype=np.uint8)])
    got = px.blend_pixels(mode, b, s)
    want = orc.blend_pixels(mode, b, s)
    _exact(got, want, "blend mode %d" % mode)


def test_normal_and_mask_exhaustive_alpha_planes(gpu, orc):
    # every (backdrop, source) pair over grey premultiplied pixels (v,v,v,v)
    v = np.arange(256, dtype=np.uint8)
    bb, ss = np.meshgrid(v, v, indexing="ij")
    b = np.repeat(bb.reshape(-1, 1), 4, 1)
    s = np.repeat(ss.reshape(-1, 1), 4, 1)
    for mode in (0, 16, 17, 18, 19):
        _exact(px.blend_pixels(mode, b, s), orc.blend_pixels(mode, b, s), "mode %d" % mode)


# ----------------------------------------------------------------- fills --

FILL_CASES = [
    ("M 10 10 H 90 V 90 H 10 L 10 10", 100, 100, {}),
    ("M 10.1 10.1 H 60.1 V 60.1 H 10.1 z", 100, 100, {}),
    ("M30 60 A 20 20 0 0 0 90 60 L 30 60", 100, 100, {}),
    ("M 10,30 A 20,20 0,0,1 50,30 A 20,20 0,0,1 90,30 Q 90,60 50,90 Q 10,60 10,30 z", 100, 100, {}),
    ("M 20 50 A 20 10 45 1 1 80 50 L 20 50", 100, 100, {}),
    ("M 40 40 L 40 80 L 80 80 L 80 40 C 80 -20 40 100 40 40", 100, 100, {}),
    ("M99 99 L999 99 L999 100 L99 100 Z", 100, 100, {}),
    ("M100 100 L999 100 L999 101 L100 101 Z", 100, 100, {}),
    ("M0 0 L0 0 L60 0 L60 60 L0 60", 60, 60, {}),
    ("M 5 5 L 95 20 L 30 95 Z M 50 10 L 90 90 L 10 60 Z", 100, 100, {"winding": 1}),
    ("M 5 5 L 95 20 L 30 95 Z M 50 10 L 90 90 L 10 60 Z", 100, 100, {"winding": 0}),
    ("M -20 -20 L 130 30 L 40 140 Z", 100, 100, {}),
]


@pytest.mark.parametrize("case", range(len(FILL_CASES)))
@pytest.mark.parametrize("mode", [0, 16, 17, 18, 19, 2, 11])
def test_fill_path_cases(gpu, orc, case, mode):
    path, w, h, kw = FILL_CASES[case]
    rng = np.random.default_rng(case * 31 + mode)
    back = WL.random_premultiplied_image(rng, w, h)
    color = WL.premul(200, 40, 90, 180)
    want = back.copy()
    orc.fill_path(want, path, color, winding=kw.get("winding", 0), mode=mode)
    img = px.Image.from_array(back)
    img.fillPath(path, color, windingRule=kw.get("winding", 0), blendMode=mode)
    _cov_check(img.download(), want, "case %d mode %d" % (case, mode))


@pytest.mark.parametrize("cap,join", [(0, 0), (1, 1), (2, 2)])
def test_stroke_path(gpu, orc, cap, join):
    want = orc.new_image(100, 100)
    orc.stroke_path(want, "M 15 10 L 30 90 60 30 90 90", (255, 255, 0, 255), width=10, cap=cap, join=join)
    img = px.Image(100, 100)
    img.strokePath("M 15 10 L 30 90 60 30 90 90", (255, 255, 0, 255), strokeWidth=10, lineCap=cap,
                   lineJoin=join)
    _cov_check(img.download(), want, "stroke cap %d join %d" % (cap, join))


def test_masters_through_gpu(gpu, orc, masters):
    # the same calls as the reference's xray tests, checked against the master
    img = px.Image(100, 100)
    img.fillPath("M 10 10 H 60 V 60 H 10 z", (255, 0, 0, 255))
    img.fillPath("M 30 30 H 80 V 80 H 30 z", (0, 255, 0, 255), blendMode=16)
    _exact(img.download(), masters("tests/paths/rectMask.png"), "rectMask")
    img = px.Image(60, 120)
    img.fill((255, 255, 255, 255))
    for ty, d in [(5, ()), (25, (2, 2)), (45, (4, 4)), (65, (2, 4, 6, 2)), (85, (1,)),
                  (105, (1, 2, 3, 4, 5, 6, 7, 8, 9))]:
        img.strokePath("M 0 0 L 50 0", (0, 0, 0, 255), transform=translate(5, ty), strokeWidth=10,
                       lineCap=0, lineJoin=2, dashes=d)
    _exact(img.download(), masters("tests/paths/dashes.png"), "dashes")
    for sides in range(3, 9):
        p = Path()
        p.polygon(50, 50, 30, sides)
        img = px.Image(100, 100)
        img.fillPath(p, (255, 255, 255, 255))
        _exact(img.download(), masters("tests/paths/polygon%d.png" % sides), "polygon%d" % sides)


def test_star512(gpu, orc):
    cmd, w, h = WL.star512()
    want = orc.new_image(w, h)
    st = orc.fill_path(want, cmd.cmds, cmd.color)
    img = px.Image(w, h)
    img.fillPath(cmd.cmds, cmd.color)
    got = img.download()
    _cov_check(got, want, "star512")
    assert st["touched"] > 100000
    assert st["dropped"] == 0


@pytest.mark.parametrize("seed", [2, 12])
def test_mixed_scene_small(gpu, orc, seed):
    cmds, canvas = WL.mixed_scene(seed=seed, n_paths=60, W=512, H=512)
    want = canvas.copy()
    img = px.Image.from_array(canvas)
    px.touched_pixels(reset=True)
    touched = dropped = 0
    for c in cmds:
        if c.stroke:
            st = orc.stroke_path(want, c.cmds, c.color, width=c.width, cap=c.cap,
                                 join=c.join, miter_limit=c.miter, mode=c.mode)
            img.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                           miterLimit=c.miter, blendMode=c.mode)
        else:
            st = orc.fill_path(want, c.cmds, c.color, winding=c.winding, mode=c.mode)
            img.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
        touched += st["touched"]
        dropped += st["dropped"]
    assert dropped == 0
    _cov_check(img.download(), want, "mixed seed %d" % seed)
    # the GPU's N_touch counter (the bench's algorithmic-byte unit) matches
    assert px.touched_pixels(reset=True) == touched


def test_glyphs_into_a8(gpu, orc):
    rng = np.random.default_rng(4)
    img = px.Image(256, 256, a8=True)
    want = orc.new_image(256, 256)
    for _ in range(8):
        p = WL.glyph_path(rng, 200, 200)
        img.fillPath(p, (255, 255, 255, 255))
        assert orc.fill_path(want, p.array(), (255, 255, 255, 255))["dropped"] == 0
    got = img.download()
    assert np.array_equal(want[..., 0], want[..., 3])
    _cov_check(got, want[..., 3], "glyphs a8")


# ----------------------------------------------------------- image ops --

@pytest.mark.parametrize("mode", range(20))
@pytest.mark.parametrize("dx,dy", [(0, 0), (13, -7), (-40, 25), (90, 90), (-200, 0)])
def test_draw_blend_rect(gpu, orc, mode, dx, dy):
    rng = np.random.default_rng(mode * 7 + abs(dx) * 3 + abs(dy))
    a = WL.random_premultiplied_image(rng, 100, 80)
    b = WL.random_premultiplied_image(rng, 60, 70)
    want = a.copy()
    orc.draw(want, b, dx, dy, mode)
    ia, ib = px.Image.from_array(a), px.Image.from_array(b)
    ia.draw(ib, dx, dy, mode)
    _exact(ia.download(), want, "draw mode %d at (%d,%d)" % (mode, dx, dy))


@pytest.mark.parametrize("mode", [16, 18, 19])
def test_draw_a8_mask(gpu, orc, mode):
    rng = np.random.default_rng(mode)
    a = WL.random_premultiplied_image(rng, 128, 96)
    m = rng.integers(0, 256, size=(96, 128), dtype=np.uint8)
    mrgba = np.repeat(m[..., None], 4, axis=2)
    want = a.copy()
    orc.draw(want, mrgba, 0, 0, mode)
    ia = px.Image.from_array(a)
    im = px.Image.from_array(m)
    ia.draw(im, 0, 0, mode)
    _exact(ia.download(), want, "a8 mask mode %d" % mode)


@pytest.mark.parametrize("radius", [1, 2, 5, 8, 20, 32])
@pytest.mark.parametrize("oob", [(0, 0, 0, 0), (0, 0, 0, 255), (10, 20, 30, 40)])
def test_blur(gpu, orc, radius, oob):
    rng = np.random.default_rng(radius)
    a = WL.random_premultiplied_image(rng, 257, 130)
    want = a.copy()
    orc.blur(want, radius, oob)
    img = px.Image.from_array(a)
    img.blur(radius, oob)
    _exact(img.download(), want, "blur r=%d" % radius)


@pytest.mark.parametrize("s", [1, 4, -3])
def test_spread(gpu, orc, s):
    rng = np.random.default_rng(s + 10)
    a = WL.random_premultiplied_image(rng, 130, 77)
    want = a.copy()
    orc.spread(want, s)
    img = px.Image.from_array(a)
    img.spread(s)
    _exact(img.download(), want, "spread %d" % s)


@pytest.mark.parametrize("offset,spread,blur", [((12, 12), 4, 16), ((0, 0), 0, 5), ((-3, 7), 0, 0),
                                                ((2, 2), 2, 10), ((5, -5), -2, 3)])
def test_shadow(gpu, orc, offset, spread, blur):
    rng = np.random.default_rng(99)
    a = WL.random_premultiplied_image(rng, 150, 120)
    color = WL.premul(10, 20, 30, 200)
    want = orc.shadow(a, offset, spread, blur, color)
    img = px.Image.from_array(a)
    got = img.shadow(offset, spread, blur, color).download()
    _exact(got, want, "shadow %s %s %s" % (offset, spread, blur))


@pytest.mark.parametrize("o", [0.0, 0.25, 0.5, 0.999, 1.0])
def test_apply_opacity(gpu, orc, o):
    rng = np.random.default_rng(5)
    a = WL.random_premultiplied_image(rng, 64, 33)
    want = a.copy()
    orc.apply_opacity(want, o)
    img = px.Image.from_array(a)
    img.applyOpacity(o)
    _exact(img.download(), want, "opacity %s" % o)


def test_errors_are_raised(gpu):
    img = px.Image(10, 10)
    with pytest.raises(px.PixieError):
        img.blur(-3)
    before = img.download()
    img.fillPath("M 0 0 L 10 0 Z", (255, 0, 0, 255))  # no segment: nothing drawn, no error
    _exact(img.download(), before, "all-horizontal fill")
    with pytest.raises(px.PixieError):
        img.strokePath("M 0 0 L 10 10", (255, 0, 0, 255), dashes=(1, 0))
