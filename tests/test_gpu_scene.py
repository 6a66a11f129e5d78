"""GPU parity of the batched form (pxh_scene_*) against the per-call form and
the oracle.

A scene replays pixie's fillPath / strokePath calls (paths.nim:2093-2176) in
order with device geometry and two raster phases.  Its device geometry
(flattening, stroking, segments, bounds, bands) must equal the host geometry
of the per-call form bit for bit, its image must equal the image the same
calls leave through the per-call form bit for bit, with the same N_touch, and
both stay within the coverage bar of the oracle.
"""
import numpy as np
import pytest

import pixie_amd as px
from pixie_amd import workloads as WL
from pixie_amd.paths import Path, translate, scale, mat_mul

pytestmark = pytest.mark.gpu

COV_MAX_DIFF = 1        # 1/255 per channel against the oracle
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


def rotate(a):
    c, s = np.float32(np.cos(a)), np.float32(np.sin(a))
    return np.array([c, s, 0, -s, c, 0, 0, 0, 1], np.float32)


@pytest.fixture(scope="module")
def gpu():
    px.init()
    return px


# Commands without elliptical arcs, covering every canonicalisation rule:
# relative / H / V forms, smooth cubics and quads (reflection and its absence),
# implicit moves after Close, the relative-move no-close quirk, degenerate
# pieces, paths leaving the canvas.
PATHS = [
    "M 10 10 H 90 V 90 H 10 L 10 10",
    "M 10.1 10.1 H 60.1 V 60.1 H 10.1 z",
    "m 12 15 h 50 v 40 h -50 z m 60 0 l 20 30 l -25 5 z",
    "M 10 50 C 20 10 40 10 50 50 S 80 90 90 50 Z",
    "M 10 50 S 30 10 50 50 S 80 90 90 50",
    "M 10 80 Q 30 10 50 80 T 90 80 Z",
    "M 10 80 T 50 20 T 90 80",
    "M 10 40 c 10 -30 30 -30 40 0 s 30 40 40 0 q 10 30 -20 40 t -40 0 z",
    "M 10 10 l 30 0 l 0 30 m 20 20 l 30 0 l 0 30 z",
    "M 10 10 L 50 10 L 50 50 Z L 80 80 L 10 80",
    "M 40 40 L 40 80 L 80 80 L 80 40 C 80 -20 40 100 40 40",
    "M99 99 L999 99 L999 100 L99 100 Z",
    "M0 0 L0 0 L60 0 L60 60 L0 60",
    "M 5 5 L 95 20 L 30 95 Z M 50 10 L 90 90 L 10 60 Z",
    "M -20 -20 L 130 30 L 40 140 Z",
    "M 20 20 L 20 20 M 30 30 L 70 30 L 70 70 Z M 5 90",
    "M 10 10 v 0 h 0 L 60 12 L 30 70 Z z",
]

XFORMS = [None, translate(3.25, -7.5), mat_mul(scale(1.7, 0.6), rotate(0.4)),
          mat_mul(translate(50, 50), mat_mul(rotate(-1.1), translate(-50, -50)))]


def _calls(rng, n, W, H):
    """Random (stroke, path, xform, kwargs) calls drawn from PATHS."""
    out = []
    for i in range(n):
        path = PATHS[int(rng.integers(len(PATHS)))]
        xf = XFORMS[int(rng.integers(len(XFORMS)))]
        off = translate(float(rng.uniform(-20, W - 60)), float(rng.uniform(-20, H - 60)))
        xf = off if xf is None else mat_mul(off, xf)
        rgb = rng.integers(0, 256, 3)
        color = WL.premul(int(rgb[0]), int(rgb[1]), int(rgb[2]), int(rng.integers(40, 256)))
        mode = int(rng.integers(0, 20)) if rng.random() < 0.5 else 0
        if mode == 16 and rng.random() < 0.7:
            mode = 0  # keep masks rare: they clear everything outside them
        if rng.random() < 0.35:
            out.append((True, path, xf, dict(color=color, strokeWidth=float(rng.uniform(0.5, 9)),
                                             lineCap=int(rng.integers(0, 3)),
                                             lineJoin=int(rng.integers(0, 3)),
                                             miterLimit=float(rng.choice([1.5, 4.0, 10.0])),
                                             blendMode=mode)))
        else:
            out.append((False, path, xf, dict(color=color, windingRule=int(rng.integers(0, 2)),
                                              blendMode=mode)))
    return out


def _per_call(img, calls):
    for stroke, path, xf, kw in calls:
        kw = dict(kw)
        color = kw.pop("color")
        if stroke:
            img.strokePath(path, color, transform=xf, **kw)
        else:
            img.fillPath(path, color, transform=xf, **kw)


def _scene(calls):
    sc = px.Scene()
    for stroke, path, xf, kw in calls:
        kw = dict(kw)
        color = kw.pop("color")
        if stroke:
            sc.strokePath(path, color, transform=xf, **kw)
        else:
            sc.fillPath(path, color, transform=xf, **kw)
    return sc


def _oracle(orc, canvas, calls):
    want = canvas.copy()
    dropped = 0
    for stroke, path, xf, kw in calls:
        c = px._cmds(path)
        if stroke:
            st = orc.stroke_path(want, c, kw["color"], width=kw["strokeWidth"], cap=kw["lineCap"],
                                 join=kw["lineJoin"], miter_limit=kw["miterLimit"],
                                 dashes=kw.get("dashes", ()), mode=kw["blendMode"], mat=xf)
        else:
            st = orc.fill_path(want, c, kw["color"], winding=kw["windingRule"],
                               mode=kw["blendMode"], mat=xf)
        dropped += st["dropped"]
    # the oracle confines a fill's writes to its span widened by 16 columns;
    # pixie has no such clip, so the comparison holds only if nothing fell
    # outside it
    assert dropped == 0, "oracle write clip dropped %d writes" % dropped
    return want


def _compare(gpu, calls, canvas, what, orc=None):
    ref = px.Image.from_array(canvas)
    px.synchronize()
    px.touched_pixels(reset=True)
    _per_call(ref, calls)
    want_touch = px.touched_pixels(reset=True)
    want = ref.download()
    sc = _scene(calls)
    img = px.Image.from_array(canvas)
    px.touched_pixels(reset=True)
    sc.render(img)
    got = img.download()
    got_touch = px.touched_pixels(reset=True)
    print("%s: %d calls, N_touch %d" % (what, len(sc), got_touch), flush=True)
    _exact(got, want, what + " scene vs per-call")
    assert got_touch == want_touch, (got_touch, want_touch)
    # a second render reuses the validated pools without a host round trip
    img2 = px.Image.from_array(canvas)
    sc.render(img2)
    _exact(img2.download(), want, what + " second render")
    if orc is not None:
        _cov_check(got, _oracle(orc, canvas, calls), what + " vs oracle")
    return sc


def test_scene_symbols_and_empty(gpu):
    sc = px.Scene()
    assert len(sc) == 0
    img = px.Image(16, 16)
    sc.render(img)  # nothing recorded: nothing to do
    assert not img.download().any()


@pytest.mark.parametrize("k", range(len(PATHS)))
def test_scene_geometry_matches_host(gpu, k):
    path = PATHS[k]
    calls = []
    for j, xf in enumerate(XFORMS):
        calls.append((False, path, xf, dict(color=(255, 0, 0, 255), windingRule=j & 1,
                                            blendMode=0)))
        calls.append((True, path, xf, dict(color=(0, 0, 255, 255), strokeWidth=1.0 + 2.5 * j,
                                           lineCap=j % 3, lineJoin=(j + 1) % 3, miterLimit=4.0,
                                           blendMode=0)))
    sc = _scene(calls)
    img = px.Image(130, 110)
    sc.render(img)
    st = sc.status()
    for i, (stroke, p, xf, kw) in enumerate(calls):
        info, parts, ents, wind = px.debug_fill_geometry(
            p, 130, 110, transform=xf, stroke=stroke, width=kw.get("strokeWidth", 1.0),
            cap=kw.get("lineCap", 0), join=kw.get("lineJoin", 0), miter=kw.get("miterLimit", 4.0))
        dinfo, dparts, dents, dwind = sc.debug_geometry(i)
        what = "path %d call %d" % (k, i)
        assert st[i] == 0, what
        keep = [0, 1, 2, 3, 4, 5, 10, 11, 12, 13, 14]  # bounds, bands, clip, status, size
        assert np.array_equal(dinfo[keep], info[keep]), (what, dinfo, info)
        if info[2] <= 0:
            continue
        assert np.array_equal(dparts, parts), what
        assert np.array_equal(dents.view(np.int32), ents.view(np.int32)), what
        assert np.array_equal(dwind, wind), what


@pytest.mark.parametrize("seed,W,H", [(1, 256, 192), (2, 301, 157), (3, 96, 400)])
def test_scene_random_calls(gpu, orc, seed, W, H):
    rng = np.random.default_rng(seed)
    calls = _calls(rng, 80, W, H)
    canvas = WL.random_premultiplied_image(rng, W, H)
    _compare(gpu, calls, canvas, "random calls seed %d %dx%d" % (seed, W, H), orc)


@pytest.mark.parametrize("mode", range(20))
def test_scene_blend_modes(gpu, orc, mode):
    rng = np.random.default_rng(40 + mode)
    calls = []
    for i in range(6):
        path = PATHS[(3 * i + mode) % len(PATHS)]
        xf = translate(float(rng.uniform(0, 60)), float(rng.uniform(0, 60)))
        color = WL.premul(200, 40 + 30 * i, 90, 255 if i % 2 else 150)
        if i == 5:
            color = (0, 0, 0, 0)  # only drawn by OverwriteBlend
        calls.append((i % 3 == 2, path, xf, dict(color=color, strokeWidth=3.0, lineCap=1,
                                                  lineJoin=i % 3, miterLimit=4.0, blendMode=mode)
                      if i % 3 == 2 else dict(color=color, windingRule=i & 1, blendMode=mode)))
    canvas = WL.random_premultiplied_image(rng, 160, 150)
    _compare(gpu, calls, canvas, "mode %d" % mode, orc)


@pytest.mark.parametrize("seed", [2, 12])
def test_scene_mixed_small(gpu, orc, seed):
    cmds, canvas = WL.mixed_scene(seed=seed, n_paths=60, W=512, H=512)
    calls = []
    for c in cmds:
        if c.stroke:
            calls.append((True, c.cmds, None, dict(color=c.color, strokeWidth=c.width,
                                                   lineCap=c.cap, lineJoin=c.join,
                                                   miterLimit=c.miter, blendMode=c.mode)))
        else:
            calls.append((False, c.cmds, None, dict(color=c.color, windingRule=c.winding,
                                                    blendMode=c.mode)))
    _compare(gpu, calls, canvas, "mixed seed %d" % seed, orc)


def test_scene_a8(gpu, orc):
    rng = np.random.default_rng(4)
    sc = px.Scene()
    img = px.Image(300, 260, a8=True)
    ref = px.Image(300, 260, a8=True)
    want = orc.new_image(300, 260)
    for i in range(10):
        p = WL.glyph_path(rng, 200, 200)
        mode = 0 if i != 7 else 18  # one SubtractMask call
        sc.fillPath(p, (255, 255, 255, 255), blendMode=mode)
        ref.fillPath(p, (255, 255, 255, 255), blendMode=mode)
        orc.fill_path(want, p.array(), (255, 255, 255, 255), mode=mode)
    sc.render(img)
    _exact(img.download(), ref.download(), "a8 scene vs per-call")
    _cov_check(img.download(), want[..., 3], "a8 scene vs oracle")


def test_scene_wide_bands_and_pool_growth(gpu, orc):
    # a band with more entries than LDS holds (global row scratch), more tasks
    # than the first pool estimate, and a flattening that needs more points
    # than the first estimate (huge transformed circle)
    W, H = 3000, 2300
    zig = Path()
    zig.move_to(0, 100)
    for i in range(700):
        zig.line_to(i * 4 + 2, 110 if i % 2 == 0 else 100)
    zig.line_to(2800, 140)
    zig.line_to(0, 140)
    zig.close_path()
    circle = Path()
    circle.ellipse(0, 0, 1, 1)
    calls = [
        (False, zig.array(), None, dict(color=(255, 0, 0, 255), windingRule=1, blendMode=0)),
        (False, circle.array(), mat_mul(translate(1500, 1150), scale(2.5e5)),
         dict(color=(0, 120, 0, 200), windingRule=0, blendMode=0)),
        (False, "M 10 10 L 2990 20 L 2900 2290 L 30 2200 Z", None,
         dict(color=(0, 0, 200, 128), windingRule=0, blendMode=3)),
    ]
    canvas = np.zeros((H, W, 4), np.uint8)
    _compare(gpu, calls, canvas, "wide bands + pool growth", orc)


def test_scene_status_and_errors(gpu):
    sc = px.Scene()
    with pytest.raises(px.PixieError):
        sc.fillPath("M 0 0 L 10 0 L 10 10 Z", (255, 0, 0, 255), blendMode=25)
    assert len(sc) == 0
    sc.fillPath("M 0 0 L 10 0 L 10 10 Z", (0, 0, 0, 0))            # transparent: skipped
    sc.strokePath("M 0 0 L 10 0", (255, 0, 0, 255), strokeWidth=0)  # zero width: skipped
    assert len(sc) == 0
    sc.fillPath("M 1 1 L 30 1 L 30 30 Z", (255, 0, 0, 255))
    sc.fillPath("M 1e30 0 L -1e30 0 L 0 10 Z", (0, 255, 0, 255))    # int overflow
    sc.fillPath("M 5 5 L 20 5 L 20 20 Z", (0, 0, 255, 255))
    img = px.Image(40, 40)
    sc.render(img)
    st = sc.status()
    ref = px.Image(40, 40)
    ref.fillPath("M 1 1 L 30 1 L 30 30 Z", (255, 0, 0, 255))
    with pytest.raises(px.PixieError) as e:
        ref.fillPath("M 1e30 0 L -1e30 0 L 0 10 Z", (0, 255, 0, 255))
    ref.fillPath("M 5 5 L 20 5 L 20 20 Z", (0, 0, 255, 255))
    assert st[0] == 0 and st[2] == 0
    assert st[1] == e.value.code
    _exact(img.download(), ref.download(), "status scene")


def test_scene_band_over_capacity_renders(gpu):
    # more than 8,192 segments crossing one row: the batched form grows its
    # global row scratch and draws what the per-call form draws
    zig = Path()
    zig.move_to(0, 10)
    for i in range(9000):
        zig.line_to(i * 0.25, 20 if i % 2 == 0 else 10)
    zig.line_to(0, 30)
    zig.close_path()
    calls = [(False, zig.array(), None, dict(color=(255, 0, 0, 255), windingRule=0, blendMode=0)),
             (False, "M 5 2 L 60 2 L 60 38 Z", None,
              dict(color=(0, 0, 255, 160), windingRule=1, blendMode=11))]
    canvas = np.zeros((40, 2300, 4), np.uint8)
    _compare(gpu, calls, canvas, "band of 9,000 entries")


def test_scene_big_calls(gpu, orc):
    # calls of more than 256 commands take the 1024-thread geometry
    # workgroups: 300 circles in one fill (1,800 commands), a 400-segment
    # polyline stroked with round joins and dashes, a 600-command EvenOdd
    # fill, each between small calls
    rng = np.random.default_rng(31)
    W, H = 700, 500
    circles = Path()
    for _ in range(300):
        cx, cy = rng.uniform(0, W), rng.uniform(0, H)
        circles.circle(cx, cy, rng.uniform(3, 60))
    poly = Path()
    poly.move_to(20, 20)
    for i in range(400):
        poly.line_to(20 + i * 1.6, 250 + 200 * np.sin(i * 0.11) + rng.uniform(-5, 5))
    star = Path()
    star.move_to(350, 250)
    for i in range(600):
        a = i * 2.39996
        r = 40 + 200 * ((i * 7) % 13) / 13.0
        star.line_to(350 + r * np.cos(a), 250 + r * np.sin(a))
    star.close_path()
    small = (False, "M 10 10 L 120 30 L 60 110 Z", None,
             dict(color=(200, 40, 40, 200), windingRule=0, blendMode=0))
    calls = [small,
             (False, circles.array(), None, dict(color=(255, 255, 255, 255), windingRule=0,
                                                 blendMode=0)),
             (True, poly.array(), None, dict(color=(20, 100, 200, 220), strokeWidth=6.5,
                                             lineCap=1, lineJoin=1, miterLimit=4.0,
                                             dashes=(9, 4, 2, 4), blendMode=2)),
             small,
             (False, star.array(), None, dict(color=(40, 160, 60, 180), windingRule=1,
                                              blendMode=11))]
    for c in calls:
        assert c[0] or len(px._cmds(c[1])) > 0
    assert sum(1 for c in calls if len(px._cmds(c[1])) > 3 * 257) >= 3
    base = np.random.default_rng(32).integers(0, 256, (H, W, 4), dtype=np.uint8)
    base[..., :3] = np.minimum(base[..., :3], base[..., 3:4])
    _compare(gpu, calls, base, "big calls", orc)


# Elliptical arcs (flattened when recorded) in every form pixie's tests use:
# small / large arc, sweep, rotation, relative arcs, arcs inside longer paths.
ARC_PATHS = [
    "M30 60 A 20 20 0 0 0 90 60 L 30 60",
    "M 10,30 A 20,20 0,0,1 50,30 A 20,20 0,0,1 90,30 Q 90,60 50,90 Q 10,60 10,30 z",
    "M 20 50 A 20 10 45 1 1 80 50 L 20 50",
    "M 0 50 A 50 50 0 0 0 50 0 L 50 50 L 0 50",
    "M 10 50 a 30 15 -30 1 0 60 10 l -20 30 z m 70 -30 a 8 16 60 0 1 10 10",
]
DASHES = [(), (2, 2), (4, 1, 1), (5,), (1, 2, 3, 4, 5, 6, 7, 8, 9), (0.75, 0.5)]


def test_scene_arcs_and_dashes(gpu, orc):
    rng = np.random.default_rng(8)
    W, H = 360, 280
    calls = []
    for i in range(48):
        path = ARC_PATHS[i % len(ARC_PATHS)] if i % 3 else PATHS[i % len(PATHS)]
        xf = XFORMS[i % len(XFORMS)]
        off = translate(float(rng.uniform(-10, W - 90)), float(rng.uniform(-10, H - 90)))
        xf = off if xf is None else mat_mul(off, xf)
        rgb = rng.integers(0, 256, 3)
        color = WL.premul(int(rgb[0]), int(rgb[1]), int(rgb[2]), int(rng.integers(60, 256)))
        mode = int(rng.integers(0, 16))
        if i % 2:
            calls.append((True, path, xf, dict(color=color, strokeWidth=float(rng.uniform(0.8, 7)),
                                               lineCap=i % 3, lineJoin=(i // 3) % 3,
                                               miterLimit=4.0, dashes=DASHES[i % len(DASHES)],
                                               blendMode=mode)))
        else:
            calls.append((False, path, xf, dict(color=color, windingRule=i % 4 // 2, blendMode=mode)))
    canvas = WL.random_premultiplied_image(np.random.default_rng(9), W, H)
    _compare(gpu, calls, canvas, "arcs and dashes", orc)


def test_scene_invalid_dash_is_a_call_status(gpu):
    # a dash value <= 0 makes strokePath raise (paths.nim:2029-2031); in a
    # scene that call draws nothing and reports the error, the others draw
    sc = px.Scene()
    sc.fillPath("M 1 1 L 30 1 L 30 30 Z", (255, 0, 0, 255))
    sc.strokePath("M 5 20 L 35 20", (0, 255, 0, 255), strokeWidth=4, dashes=(2, 0))
    sc.strokePath("M 5 30 L 35 30", (0, 0, 255, 255), strokeWidth=4, dashes=(3,))
    img = px.Image(40, 40)
    sc.render(img)
    st = sc.status()
    ref = px.Image(40, 40)
    ref.fillPath("M 1 1 L 30 1 L 30 30 Z", (255, 0, 0, 255))
    with pytest.raises(px.PixieError) as e:
        ref.strokePath("M 5 20 L 35 20", (0, 255, 0, 255), strokeWidth=4, dashes=(2, 0))
    ref.strokePath("M 5 30 L 35 30", (0, 0, 255, 255), strokeWidth=4, dashes=(3,))
    assert e.value.code == 3
    assert list(st) == [0, 3, 0]
    _exact(img.download(), ref.download(), "invalid dash scene")


def test_scene_repeated_renders_are_identical(gpu):
    # phase-A waves take tasks from work queues in a different order on every
    # render; every render must still fit the pools and give the same image
    cmds, canvas = WL.mixed_scene(seed=7, n_paths=150, W=640, H=480)
    ref = px.Image.from_array(canvas)
    sc = px.Scene()
    for c in cmds:
        if c.stroke:
            ref.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                           miterLimit=c.miter, blendMode=c.mode)
            sc.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                          miterLimit=c.miter, blendMode=c.mode)
        else:
            ref.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
            sc.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
    px.synchronize()
    want = ref.download()
    touches = []
    for k in range(8):
        img = px.Image.from_array(canvas)
        px.touched_pixels(reset=True)
        sc.render(img)
        _exact(img.download(), want, "render %d" % k)
        touches.append(px.touched_pixels(reset=True))
    assert len(set(touches)) == 1, touches
    # renders after the first replay a captured graph (kernel nodes only)
    assert sc.graph_active()
    # renders accumulate on the same image exactly like repeated calls
    img = px.Image.from_array(canvas)
    sc.render(img)
    sc.render(img)
    ref2 = px.Image.from_array(want)
    for c in cmds:
        if c.stroke:
            ref2.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                            miterLimit=c.miter, blendMode=c.mode)
        else:
            ref2.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
    _exact(img.download(), ref2.download(), "two renders")


def _comb(x0, y0, teeth, pitch, height, slant):
    """A comb of slanted teeth: every row of it crosses 2 * teeth sloped
    edges, so its row windows write hundreds of area and delta runs."""
    p = Path()
    p.move_to(x0, y0 + height)
    for i in range(teeth):
        x = x0 + i * pitch
        p.line_to(x + slant, y0)
        p.line_to(x + slant + pitch * 0.55, y0)
        p.line_to(x + pitch * 0.55, y0 + height)
    p.close_path()
    return p.array()


def test_scene_heavy_runs_replayed(gpu):
    # row windows that each write ~1000 runs (the most one task can write),
    # replayed from the captured graph several times: phase-A waves take
    # tasks in a new order on every render, and every render must fit the
    # run pool sized by the first one and give the per-call image
    W, H = 1536, 384
    calls = []
    rng = np.random.default_rng(5)
    for k in range(24):
        y0 = float(rng.uniform(0, H - 140))
        calls.append(_comb(float(rng.uniform(0, 40)), y0, 240, 6.03, 120.0, 2.37))
    ref = px.Image(W, H)
    sc = px.Scene()
    for i, c in enumerate(calls):
        col = (30 + 9 * i, 200 - 5 * i, 120, 255)
        ref.fillPath(c, col, blendMode=i % 16)
        sc.fillPath(c, col, blendMode=i % 16)
    want = ref.download()
    for k in range(6):
        img = px.Image(W, H)
        sc.render(img)
        px.synchronize()
        _exact(img.download(), want, "heavy-run render %d" % k)
    assert sc.graph_active()


@pytest.mark.parametrize("shape", range(6))
def test_scene_every_phase_a_shape(gpu, shape):
    # every phase-A launch shape the autotuner may pick (one launch, or the
    # large-band tasks in a second launch beside or before the rest) draws
    # the per-call image, on a mixed scene with large bands and on combs
    cmds, canvas = WL.mixed_scene(seed=13, n_paths=120, W=768, H=512)
    ref = px.Image.from_array(canvas)
    sc = px.Scene()
    for c in cmds:
        if c.stroke:
            for t in (ref, sc):
                t.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                             miterLimit=c.miter, blendMode=c.mode)
        else:
            for t in (ref, sc):
                t.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
    for k in range(4):
        comb = _comb(5.0 + 3 * k, 40.0 + 100 * k, 110, 6.03, 90.0, 2.37)
        ref.fillPath(comb, (40 * k, 90, 200, 220), blendMode=k)
        sc.fillPath(comb, (40 * k, 90, 200, 220), blendMode=k)
    want = ref.download()
    sc.set_shape(shape)
    for rep in range(3):
        img = px.Image.from_array(canvas)
        sc.render(img)
        px.synchronize()
        _exact(img.download(), want, "shape %d render %d" % (shape, rep))


def test_scene_replay_overflow_is_reported(gpu):
    # a replayed render that runs out of its run pool (forced from the host by
    # lowering the capacity its kernels see) stops before compose: the canvas
    # is untouched, the next synchronize raises a capacity error, and the
    # next render re-validates the pools and draws the same image as before
    cmds, canvas = WL.mixed_scene(seed=9, n_paths=40, W=256, H=256)
    sc = px.Scene()
    for c in cmds:
        if c.stroke:
            sc.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                          miterLimit=c.miter, blendMode=c.mode)
        else:
            sc.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
    img = px.Image.from_array(canvas)
    sc.render(img)
    want = img.download()
    sc.debug_limit_runs(16)
    img2 = px.Image.from_array(canvas)
    sc.render(img2)
    with pytest.raises(px.PixieError) as e:
        px.synchronize()
    assert e.value.code == 9
    _exact(img2.download(), canvas, "overflowed render leaves the canvas")
    px.synchronize()  # the error word was cleared
    img3 = px.Image.from_array(canvas)
    sc.render(img3)
    px.synchronize()
    _exact(img3.download(), want, "render after re-validation")


def test_scene_glyph_atlas(gpu, orc):
    # the glyph-atlas workload at small scale: every glyph filled NonZero into
    # its own cell of one 8-bit atlas, in one batched render
    cells, aw, ah = WL.glyph_atlas(seed=11, n=300, atlas_w=1024)
    sc = px.Scene()
    ref = px.Image(aw, ah, a8=True)
    want = orc.new_image(aw, ah)
    for cmds, x, y, w, h in cells:
        sc.fillPath(cmds, (255, 255, 255, 255))
        ref.fillPath(cmds, (255, 255, 255, 255))
        orc.fill_path(want, cmds, (255, 255, 255, 255))
    img = px.Image(aw, ah, a8=True)
    sc.render(img)
    got = img.download()
    _exact(got, ref.download(), "atlas scene vs per-call")
    _cov_check(got, want[..., 3], "atlas vs oracle")
    # every cell holds only its own glyph: nothing is written in the gutters
    mask = np.zeros((ah, aw), bool)
    for cmds, x, y, w, h in cells:
        mask[y:y + h, x:x + w] = True
    assert not got[~mask].any()


def test_scene_autotune_keeps_results(gpu):
    cmds, canvas = WL.mixed_scene(seed=9, n_paths=80, W=700, H=500)
    sc = px.Scene()
    ref = px.Image.from_array(canvas)
    for c in cmds:
        if c.stroke:
            sc.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                          miterLimit=c.miter, blendMode=c.mode)
            ref.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                           miterLimit=c.miter, blendMode=c.mode)
        else:
            sc.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
            ref.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
    img = px.Image.from_array(canvas)
    shape, ms = sc.autotune(img, reps=2)
    assert 0 <= shape < 6 and ms > 0
    _exact(img.download(), canvas, "autotune leaves the image alone")
    sc.render(img)
    _exact(img.download(), ref.download(), "render after autotune")


def test_scene_full_mixed_4k(gpu, orc):
    # the benchmark's own workload at full size (4096x4096, 2,000 paths,
    # 2,606 calls): batched form against the per-call form (bit for bit,
    # same N_touch) and against the oracle (coverage bar)
    cmds, canvas = WL.mixed_scene(n_paths=2000)
    calls = []
    for c in cmds:
        if c.stroke:
            calls.append((True, c.cmds, None, dict(color=c.color, strokeWidth=c.width,
                                                   lineCap=c.cap, lineJoin=c.join,
                                                   miterLimit=c.miter, blendMode=c.mode)))
        else:
            calls.append((False, c.cmds, None, dict(color=c.color, windingRule=c.winding,
                                                    blendMode=c.mode)))
    _compare(gpu, calls, canvas, "mixed 4096x4096 scene", orc)


def test_scene_full_glyph_atlas(gpu, orc):
    # the full 10,000-glyph atlas: batched render against one launch per glyph
    cells, aw, ah = WL.glyph_atlas()
    sc = px.Scene()
    ref = px.Image(aw, ah, a8=True)
    for cmds, x, y, w, h in cells:
        sc.fillPath(cmds, (255, 255, 255, 255))
        ref.fillPath(cmds, (255, 255, 255, 255))
    img = px.Image(aw, ah, a8=True)
    sc.render(img)
    got = img.download()
    _exact(got, ref.download(), "10k-glyph atlas")
    # and against the oracle: each cell is a white NonZero fill onto
    # transparent, whose alpha is the coverage the A8 atlas holds
    want = orc.new_image(aw, ah)
    dropped = 0
    for cmds, x, y, w, h in cells:
        dropped += orc.fill_path(want, cmds, (255, 255, 255, 255))["dropped"]
    assert dropped == 0
    _cov_check(got.reshape(ah, aw), want[..., 3], "10k-glyph atlas vs oracle")


def test_per_call_big_paths_through_device_geometry(gpu, orc):
    # per-call fills and strokes of 4096 floats or more are flattened on the
    # device (a one-call scene): same image as the oracle, and the host
    # path's error when the device reports one
    rng = np.random.default_rng(41)
    W, H = 600, 400
    circles = Path()
    for _ in range(200):
        circles.circle(rng.uniform(0, W), rng.uniform(0, H), rng.uniform(2, 80))
    cmds = circles.array()
    assert len(cmds) >= 4096
    base = np.random.default_rng(42).integers(0, 256, (H, W, 4), dtype=np.uint8)
    base[..., :3] = np.minimum(base[..., :3], base[..., 3:4])
    img = px.Image.from_array(base)
    img.fillPath(cmds, (30, 60, 90, 200), windingRule=px.EvenOdd, blendMode=3)
    img.strokePath(cmds, (200, 100, 0, 255), strokeWidth=3.5, lineJoin=px.RoundJoin,
                   dashes=(6, 3))
    want = base.copy()
    st = orc.fill_path(want, cmds, (30, 60, 90, 200), winding=1, mode=3)
    st2 = orc.stroke_path(want, cmds, (200, 100, 0, 255), width=3.5, join=1, dashes=(6, 3))
    assert st["dropped"] == 0 and st2["dropped"] == 0
    _cov_check(img.download(), want, "big per-call fill + dashed stroke")
    # a huge coordinate: pixie raises "Path int overflow detected"
    bad = cmds.copy()
    bad[1] = 1e30
    before = img.download()
    with pytest.raises(px.PixieError, match="overflow") as e:
        img.fillPath(bad, (255, 0, 0, 255))
    assert e.value.code == 2
    _exact(img.download(), before, "image untouched by a failing big call")
