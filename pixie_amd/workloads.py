"""Seeded synthetic workloads of the raster hot path.

  star512      one closed 64-vertex star, 512x512, NonZero, Normal (opaque)
  mixed4k      2,000 random closed cubic paths on 4096x4096, 16 blend modes,
               30% also stroked
  blurmask4k   4096x4096 random premultiplied image for blur / shadow / mask
               modes, plus a 500-circle mask
  glyphs10k    10,000 quadratic glyph-like paths, each into its own atlas cell
  serve        scenes for the serving loop (mixed4k-like, random sizes)

Every generator is deterministic for its seed.  Paint colours are RGBA8 and
premultiplied with (c*a + 127) div 255 (chroma rgbx, pinned by pixie's
test_images.nim:204-228).
"""
import math

import numpy as np

from .paths import Path, f32

MITER, ROUND_JOIN, BEVEL = 0, 1, 2
BUTT, ROUND_CAP, SQUARE = 0, 1, 2


def premul(r, g, b, a):
    if a == 255:
        return (r, g, b, a)
    return ((r * a + 127) // 255, (g * a + 127) // 255, (b * a + 127) // 255, a)


def random_premultiplied_image(rng, w, h):
    """Uniform random premultiplied pixels: a ~ U{0..255}, c ~ U{0..a}."""
    a = rng.integers(0, 256, size=(h, w), dtype=np.int32)
    img = np.empty((h, w, 4), np.uint8)
    for k in range(3):
        img[..., k] = np.floor(rng.random((h, w)) * (a + 1)).astype(np.uint8)
    img[..., 3] = a
    return img


class FillCmd:
    """One fillPath / strokePath call of a scene."""
    __slots__ = ("cmds", "color", "mode", "winding", "stroke", "width", "cap", "join",
                 "miter", "dashes", "mat")

    def __init__(self, cmds, color, mode=0, winding=0, stroke=False, width=1.0, cap=0, join=0,
                 miter=4.0, dashes=(), mat=None):
        self.cmds = np.asarray(cmds, np.float32)
        self.color = tuple(int(c) for c in color)
        self.mode = int(mode)
        self.winding = int(winding)
        self.stroke = bool(stroke)
        self.width = float(width)
        self.cap = int(cap)
        self.join = int(join)
        self.miter = float(miter)
        self.dashes = tuple(dashes)
        self.mat = mat


def star512(seed=1, n=64, size=512):
    """One closed 64-vertex star: vertex radii U[0.3, 0.5]*512 at evenly spaced
    angles around the canvas middle; opaque seeded colour; Normal, NonZero."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(0.3, 0.5, size=n) * size
    c = size / 2.0
    p = Path()
    for i in range(n):
        th = 2 * math.pi * i / n
        x, y = f32(c + radii[i] * math.cos(th)), f32(c + radii[i] * math.sin(th))
        if i == 0:
            p.move_to(x, y)
        else:
            p.line_to(x, y)
    p.close_path()
    rgb = rng.integers(0, 256, size=3)
    return FillCmd(p.array(), (int(rgb[0]), int(rgb[1]), int(rgb[2]), 255), 0, 0), size, size


def _rand_paint(rng):
    rgb = rng.integers(0, 256, size=3)
    a = int(rng.integers(51, 256))
    return premul(int(rgb[0]), int(rgb[1]), int(rgb[2]), a)


def random_cubic_path(rng, W, H, side):
    """A closed path of 4..32 cubics with control points uniform in a bbox of
    the given side placed uniformly inside the canvas."""
    side = min(side, W, H)
    x0 = rng.uniform(0, W - side)
    y0 = rng.uniform(0, H - side)
    n = int(rng.integers(4, 33))
    pts = rng.uniform(0, side, size=(n, 3, 2)) + (x0, y0)
    p = Path()
    sx, sy = f32(rng.uniform(0, side) + x0), f32(rng.uniform(0, side) + y0)
    p.move_to(sx, sy)
    for i in range(n):
        c1, c2, to = pts[i]
        if i == n - 1:
            to = (sx, sy)
        p.bezier_curve_to(c1[0], c1[1], c2[0], c2[1], to[0], to[1])
    p.close_path()
    return p


def mixed_scene(seed=2, n_paths=2000, W=4096, H=4096, stroke_p=0.3):
    """Mixed fill+stroke scene: returns (commands, canvas) where canvas is the
    seeded random premultiplied backdrop."""
    rng = np.random.default_rng(seed)
    cmds = []
    for i in range(n_paths):
        side = float(np.clip(rng.lognormal(math.log(64), 1.0), 4, 2048))
        p = random_cubic_path(rng, W, H, side)
        color = _rand_paint(rng)
        winding = int(rng.integers(0, 2))
        mode = i % 16
        cmds.append(FillCmd(p.array(), color, mode, winding))
        if rng.random() < stroke_p:
            width = float(rng.uniform(1, 8))
            join = int(rng.integers(0, 3))
            cap = int(rng.integers(0, 3))
            cmds.append(FillCmd(p.array(), color, mode, 0, stroke=True, width=width, cap=cap,
                                join=join, miter=4.0))
    canvas = random_premultiplied_image(np.random.default_rng(seed + 1000), W, H)
    return cmds, canvas


def circles_mask_cmds(seed=3, n=500, W=4096, H=4096):
    """500 random circles (centres uniform, radius U[8, 256]) as one path,
    filled white NonZero into a mask."""
    rng = np.random.default_rng(seed + 7)
    p = Path()
    for _ in range(n):
        cx, cy = rng.uniform(0, W), rng.uniform(0, H)
        r = rng.uniform(8, 256)
        p.circle(cx, cy, r)
    return FillCmd(p.array(), (255, 255, 255, 255), 0, 0)


def blurmask_image(seed=3, W=4096, H=4096):
    return random_premultiplied_image(np.random.default_rng(seed), W, H)


def glyph_path(rng, w, h):
    """A glyph-like path: 1..4 closed quadratic contours (outer contour plus
    holes of reversed orientation), 10..120 quads split across them."""
    n_quads = int(rng.integers(10, 121))
    n_cont = int(rng.integers(1, 5))
    n_cont = min(n_cont, n_quads // 3) or 1
    splits = np.sort(rng.choice(np.arange(1, n_quads), size=n_cont - 1, replace=False)) \
        if n_cont > 1 else np.array([], int)
    counts = np.diff(np.concatenate([[0], splits, [n_quads]])).astype(int)
    p = Path()
    cx, cy = w / 2.0, h / 2.0
    for ci, k in enumerate(counts):
        # outer contour spans the cell, holes shrink inward and reverse
        # direction; every point and control point stays inside the cell
        scale = 0.48 if ci == 0 else 0.12 + 0.18 * rng.random()
        ox = cx + (0 if ci == 0 else rng.uniform(-0.15, 0.15) * w)
        oy = cy + (0 if ci == 0 else rng.uniform(-0.15, 0.15) * h)
        direction = 1 if ci == 0 else -1
        angles = np.sort(rng.uniform(0, 2 * math.pi, size=k)) * direction
        rad = rng.uniform(0.6, 1.0, size=k)
        pts = [(ox + math.cos(a) * r * scale * w, oy + math.sin(a) * r * scale * h)
               for a, r in zip(angles, rad)]
        p.move_to(*pts[0])
        for j in range(k):
            a0 = angles[j]
            a1 = angles[(j + 1) % k] + (2 * math.pi * direction if j == k - 1 else 0)
            am = (a0 + a1) / 2
            rc = rng.uniform(0.5, 1.0) * scale
            ctrl = (ox + math.cos(am) * rc * w, oy + math.sin(am) * rc * h)
            to = pts[(j + 1) % k]
            p.quadratic_curve_to(ctrl[0], ctrl[1], to[0], to[1])
        p.close_path()
    return p


def glyph_atlas(seed=4, n=10000, gutter=1, atlas_w=8192):
    """10,000 glyph paths with shelf-packed cells.  Returns (cells, atlas_w,
    atlas_h) where cells = [(cmds, x, y, w, h)], cmds already translated into
    the atlas."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(rng.lognormal(math.log(32), 0.8, size=(n, 2)), 8, 256).astype(int)
    cells = []
    x = y = shelf_h = 0
    for i in range(n):
        w, h = int(sizes[i, 0]), int(sizes[i, 1])
        if x + w + gutter > atlas_w:
            x = 0
            y += shelf_h + gutter
            shelf_h = 0
        p = glyph_path(rng, w, h)
        cmds = p.array().copy()
        _translate_cmds(cmds, x, y)
        cells.append((cmds, x, y, w, h))
        x += w + gutter
        shelf_h = max(shelf_h, h)
    atlas_h = y + shelf_h
    return cells, atlas_w, atlas_h


def _translate_cmds(c, dx, dy):
    """Translate absolute Move/Line/Cubic/Quad/Close commands in place."""
    from .paths import PARAM_COUNT, MOVE, LINE, CUBIC, QUAD
    i = 0
    dx, dy = f32(dx), f32(dy)
    while i < len(c):
        k = int(c[i])
        n = PARAM_COUNT[k]
        if k in (MOVE, LINE, CUBIC, QUAD):
            for j in range(0, n, 2):
                c[i + 1 + j] = f32(c[i + 1 + j] + dx)
                c[i + 2 + j] = f32(c[i + 2 + j] + dy)
        elif k != 0:
            raise ValueError("unsupported command kind in glyph path")
        i += 1 + n


def serve_scene(rng, max_paths=500):
    """One serving request: 1..500 paths (log-uniform) as in the mixed scene,
    canvas side log-uniform in [256, 4096], blur radius 0 or 8."""
    n = int(round(math.exp(rng.uniform(0, math.log(max_paths)))))
    side = int(round(math.exp(rng.uniform(math.log(256), math.log(4096)))))
    cmds = []
    for i in range(n):
        s = float(np.clip(rng.lognormal(math.log(64), 1.0), 4, min(2048, side)))
        p = random_cubic_path(rng, side, side, s)
        color = _rand_paint(rng)
        cmds.append(FillCmd(p.array(), color, int(rng.integers(0, 16)), int(rng.integers(0, 2))))
        if rng.random() < 0.3:
            cmds.append(FillCmd(p.array(), color, int(rng.integers(0, 16)), 0, stroke=True,
                                width=float(rng.uniform(1, 8)), cap=int(rng.integers(0, 3)),
                                join=int(rng.integers(0, 3))))
    blur = 8 if rng.random() < 0.5 else 0
    return cmds, side, blur


def poisson_arrivals(pool, n, rate, seed=6):
    """n requests cycled from `pool` (serve.Request objects) with fresh
    Poisson arrival times at `rate` per second."""
    from .serve import Request
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate)) if rate > 0 else 0.0
        r = pool[i % len(pool)]
        out.append(Request(r.calls, r.side, r.blur, t))
    return out


def serve_requests(seed=5, n=100, rate=100.0, max_paths=500, max_side=4096):
    """Poisson arrivals at `rate` scenes/s of serving requests (serve_scene),
    as serve.Request objects."""
    from .serve import Request
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += float(rng.exponential(1.0 / rate)) if rate > 0 else 0.0
        cmds, side, blur = serve_scene(rng, max_paths=max_paths)
        side = min(side, max_side)
        # paths were placed for the drawn side; a smaller cap clips them
        out.append(Request(cmds, side, blur, t))
    return out
