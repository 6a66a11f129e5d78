#!/usr/bin/env python3
"""bench.py -- throughput of pixie's fillPath + blend raster path on one MI355X.

Workload (one step): the mixed fill+stroke scene -- 2,000 random closed paths
of 4-32 cubic Beziers (bbox side lognormal(ln 64, 1) clipped to [4, 2048]),
50/50 NonZero / EvenOdd, paint alpha in [0.2, 1], the 16 blend modes cycled,
30% of paths also stroked (width 1-8, random joins and caps) -- rendered onto
a 4096x4096 premultiplied RGBA8 canvas that is resident in HBM.

Forms: "batched" (default) records the calls once as a device-resident scene
and renders them per step in one device pass (geometry, bands, row programs,
in-register compose); "plain" issues one fillPath / strokePath per call.

value = N_touch / step time, in Mpix/s, where N_touch counts the pixels the
fills write (the unit pixie's fillShapes does its per-pixel work in).
Algorithmic bytes = 8 B per touched pixel (RGBA8 read + write).

The CPU baseline is the oracle (a C restatement of pixie's scalar path) on a
bounded prefix of the same scene, one core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--form batched|plain]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpix/s fillPath+blend 4K RGBA8 (%HBM peak); glyphs/s batched; serve p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_plain(px, img, cmds):
    for c in cmds:
        if c.stroke:
            img.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                           miterLimit=c.miter, blendMode=c.mode)
        else:
            img.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)


def record_scene(px, cmds):
    sc = px.Scene()
    for c in cmds:
        if c.stroke:
            sc.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                          miterLimit=c.miter, blendMode=c.mode)
        else:
            sc.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
    return sc


def timed(px, fn, steps, profile=True):
    """Runs fn() `steps` times between device synchronisations; returns
    (wall seconds, N_touch, {stage: ms}).  With profile, every launch is
    bracketed by HIP events on its stream (renders then run eagerly rather
    than as graph replays)."""
    px.synchronize()
    px.touched_pixels(reset=True)
    if profile:
        px.profile_begin()
    t0 = time.perf_counter()
    for k in range(steps):
        fn()
        log("step %d queued" % k)
    px.synchronize()
    wall = time.perf_counter() - t0
    stages = {}
    if profile:
        px.profile_end()
        for name in STAGES:
            ms, n = px.profile_kernel(name)
            if n:
                stages[name] = (ms, n)
    touched = px.touched_pixels(reset=True)
    return wall, touched, stages


STAGES = ("fill_rows_kernel", "scene_geometry", "scene_bands", "scene_bins", "scene_rows_kernel",
          "scene_compose_kernel")


def glyph_extra(px, WL, reps=5, cpu_seconds=5.0):
    """10,000-glyph atlas: batched (one scene render) vs one fillPath per glyph."""
    import oracle
    from pixie_amd.paths import parse_command_count
    t = time.perf_counter()
    cells, aw, ah = WL.glyph_atlas()
    gen = time.perf_counter() - t
    white = (255, 255, 255, 255)
    sc = px.Scene()
    for cmds, x, y, w, h in cells:
        sc.fillPath(cmds, white)
    atlas = px.Image(aw, ah, a8=True)
    shape, _ = sc.autotune(atlas)
    sc.render(atlas)
    wall, touched, _ = timed(px, lambda: sc.render(atlas), reps, profile=False)
    _, _, stages = timed(px, lambda: sc.render(atlas), reps, profile=True)
    n = len(cells)
    kname, (kms, klaunch) = max(stages.items(), key=lambda kv: kv[1][0])
    # input points: the command array minus one kind per command, two floats each
    points = sum((len(c[0]) - parse_command_count(c[0])) // 2 for c in cells)
    plain = px.Image(aw, ah, a8=True)
    pw, pt, _ = timed(px, lambda: [plain.fillPath(c[0], white) for c in cells], 1)
    # CPU: the oracle on a bounded prefix of the glyphs
    cimg = oracle.new_image(aw, ah)
    t1 = time.perf_counter()
    k = 0
    for cmds, x, y, w, h in cells:
        oracle.fill_path(cimg, cmds, white)
        k += 1
        if time.perf_counter() - t1 > cpu_seconds:
            break
    cdt = time.perf_counter() - t1
    algo_bytes = aw * ah + 8 * points
    return {"workload": "glyph atlas: %d glyphs of 10-120 quads, 1-4 contours, NonZero, into one "
                        "%dx%d 8-bit atlas (seed 4; generated in %.1f s)" % (n, aw, ah, gen),
            "glyphs_per_s": round(n * reps / wall, 1), "ms_per_atlas": round(wall / reps * 1e3, 3),
            "n_touch_per_atlas": int(touched / reps),
            "stages_ms": {k2: round(v[0] / reps, 4) for k2, v in stages.items()},
            "algorithmic_gbs": round(algo_bytes / (wall / reps) / 1e9, 2),
            "dominant_kernel": kname, "phase_a_shape": shape,
            "plain_glyphs_per_s": round(n / pw, 1), "batched_speedup": round(n * reps / wall / (n / pw), 2),
            "cpu_glyphs_per_s": round(k / cdt, 1), "cpu_sample": "first %d glyphs, %.1f s, 1 core" % (k, cdt)}


def image_ops_extra(px, WL):
    """Blur / shadow / mask pipeline on a 4096x4096 random premultiplied image:
    each op on a fresh copy, timed with HIP events on the library stream."""
    import oracle
    import numpy as np
    src_arr = WL.blurmask_image()
    H, W = src_arr.shape[:2]
    src = px.Image.from_array(src_arr)
    work = px.Image(W, H)
    npx = W * H
    out = {"workload": "blur/shadow/mask pipeline, %dx%d RGBA8, alpha U{0..255}, channels <= alpha "
                       "(seed 3); each op on a fresh copy" % (W, H)}

    def timed_op(fn, reps=5):
        ms = []
        for _ in range(reps):
            work_reset()
            px.event_record(0)
            fn()
            px.event_record(1)
            ms.append(px.event_elapsed_ms(0, 1))
        return min(ms), sorted(ms)[len(ms) // 2]

    def work_reset():
        lib = px.lib()
        lib.pxh_image_copy(work.handle, src.handle)

    for r in (2, 8, 32):
        best, med = timed_op(lambda: work.blur(r))
        out["blur_r%d" % r] = {"ms": round(med, 4), "algorithmic_gbs": round(16.0 * npx / (med / 1e3) / 1e9, 1),
                               "bytes_per_px": 16}
    shadow_color = WL.premul(20, 30, 40, 200)
    res = {}

    def do_shadow():
        res["img"] = work.shadow((12, 12), 4, 16, shadow_color)
    best, med = timed_op(do_shadow)
    out["shadow_o12_s4_b16"] = {"ms": round(med, 4), "algorithmic_gbs": round(33.0 * npx / (med / 1e3) / 1e9, 1),
                                "bytes_per_px": 33}
    # the mask: 500 random circles filled NonZero into an 8-bit plane
    circ = WL.circles_mask_cmds()
    mask = px.Image(W, H, a8=True)
    fills = []
    for k in range(4):  # the first call also sizes the device pools
        px.lib().pxh_image_fill(mask.handle, 0)
        px.event_record(0)
        mask.fillPath(circ.cmds, circ.color)
        px.event_record(1)
        fills.append(px.event_elapsed_ms(0, 1))
    out["mask_fill_500_circles_ms"] = round(sorted(fills[1:])[1], 4)
    out["mask_fill_500_circles_first_call_ms"] = round(fills[0], 4)
    for name, mode in (("subtract", px.BlendMode.SubtractMask), ("intersect", px.BlendMode.Mask),
                       ("exclude", px.BlendMode.ExcludeMask)):
        best, med = timed_op(lambda: work.draw(mask, 0, 0, mode))
        out["mask_%s" % name] = {"ms": round(med, 4), "algorithmic_gbs": round(9.0 * npx / (med / 1e3) / 1e9, 1),
                                 "bytes_per_px": 9}
    # CPU: the oracle's blur r=8 on a 1024x1024 crop, scaled per pixel
    crop = np.ascontiguousarray(src_arr[:1024, :1024])
    t = time.perf_counter()
    oracle.blur(crop, 8)
    cdt = time.perf_counter() - t
    out["cpu_blur_r8_mpix_s"] = round(1024 * 1024 / cdt / 1e6, 2)
    out["gpu_blur_r8_mpix_s"] = round(npx / (out["blur_r8"]["ms"] / 1e3) / 1e6, 1)
    return out


def serve_extra(px, WL, rates=(50.0, 200.0, 800.0), n=500, pool_n=120):
    """Serving loop: Poisson arrivals of scenes (1-500 paths, side 256-4096,
    blur 0 or 8), n requests per rate cycled from a pool of pool_n distinct
    scenes.  Three forms serve the same arrivals: continuous batching over
    the execution lanes driven from Python threads, the library's own C++
    server (pxh_server_*), and the plain per-call form one request at a
    time."""
    from pixie_amd import serve
    out = {"workload": "serving: Poisson arrivals, scenes of 1-500 paths (log-uniform) as in the mixed "
                       "scene, canvas side log-uniform 256-4096 px RGBA8, blur 0 or 8 (seed 5); "
                       "%d requests per rate from %d distinct scenes" % (n, pool_n),
           "lanes": px.lanes() - 1, "hbm_budget_gib": 4}
    pool = WL.serve_requests(seed=5, n=pool_n, rate=0)
    server = serve.Server(hbm_budget=4 << 30)
    native = serve.NativeServer(hbm_budget=4 << 30)
    warm = WL.serve_requests(seed=99, n=8, rate=0)
    server.run(warm)
    native.run(warm)
    for rate in rates:
        reqs = WL.poisson_arrivals(pool, n, rate)
        out["batched_rate_%d" % rate] = serve.latency_stats(server.run(reqs), reqs)
        out["native_rate_%d" % rate] = serve.latency_stats(native.run(reqs), reqs)
        out["plain_rate_%d" % rate] = serve.latency_stats(serve.serve_plain(reqs), reqs)
        log("serving at %g/s: %s" % (rate, out["native_rate_%d" % rate]))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(cmds, canvas, seconds):
    """The C restatement of pixie's raster path on one host core, in two
    builds: the scalar one (pixie -d:pixieNoSimd semantics, -O2) and the
    model of pixie's AVX2 build (-O3 -mavx2, SIMD coverage path).  The
    touched pixels are counted by the scalar build, once; each build then
    renders the same prefix of the scene."""
    import oracle
    half = seconds / 2

    def run(limit_calls=None):
        img = canvas.copy()
        touched = n = 0
        t0 = time.perf_counter()
        for c in cmds:
            if limit_calls is not None and n >= limit_calls:
                break
            if c.stroke:
                st = oracle.stroke_path(img, c.cmds, c.color, width=c.width, cap=c.cap,
                                        join=c.join, miter_limit=c.miter, mode=c.mode)
            else:
                st = oracle.fill_path(img, c.cmds, c.color, winding=c.winding, mode=c.mode)
            touched += st["touched"]
            n += 1
            if limit_calls is None and time.perf_counter() - t0 > half:
                break
        return touched, n, time.perf_counter() - t0

    touched, n, dt = run()
    with oracle.simd_mode():
        _, _, dt_simd = run(limit_calls=n)
    scalar = touched / dt / 1e6
    simd = touched / dt_simd / 1e6
    return {"value": max(scalar, simd), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": "first %d of %d fill/stroke calls of the same scene (N_touch %d); scalar "
                      "build %.2f s, AVX2-model build %.2f s; one thread" % (n, len(cmds), touched, dt,
                                                                          dt_simd),
            "scalar_mpix_s": round(scalar, 2), "avx2_model_mpix_s": round(simd, 2),
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


def draw_extra(px, WL, reps=5):
    """Transformed draws and paints on the 4096x4096 canvas (images.nim:
    531-683, paths.nim:2115-2142): a 3000x3000 image rotated and scaled,
    translated by a fractional offset, drawn tiled, and a heart-shaped path
    filled with a radial gradient and with an image paint, each timed with
    HIP events on the library stream (median of `reps`)."""
    import numpy as np
    rng = np.random.default_rng(8)
    W = H = 4096
    base = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    base[..., :3] = np.minimum(base[..., :3], base[..., 3:4])
    pic = rng.integers(0, 256, (3000, 3000, 4), dtype=np.uint8)
    pic[..., :3] = np.minimum(pic[..., :3], pic[..., 3:4])
    src = px.Image.from_array(base)
    work = px.Image(W, H)
    b = px.Image.from_array(pic)
    c30, s30 = np.float32(np.cos(np.pi / 6)), np.float32(np.sin(np.pi / 6))
    rot = np.array([1.2 * c30, 1.2 * s30, 0, -1.2 * s30, 1.2 * c30, 0, 2048, -500, 1], np.float32)
    frac = np.array([1, 0, 0, 0, 1, 0, 500.5, 300.25, 1], np.float32)
    tile = np.array([0.8 * 0.985, 0.8 * 0.174, 0, -0.8 * 0.174, 0.8 * 0.985, 0, 0, 0, 1], np.float32)
    heart = ("M 10,30 A 20,20 0,0,1 50,30 A 20,20 0,0,1 90,30 Q 90,60 50,90 Q 10,60 10,30 z")
    hm = np.array([40, 0, 0, 0, 40, 0, 48, 248, 1], np.float32)
    grad = px.Paint(px.PaintKind.RadialGradientPaint, blendMode=px.BlendMode.Multiply, opacity=0.8,
                    gradientHandlePositions=[(2048, 2048), (3500, 2048), (2048, 3000)],
                    gradientStops=[px.ColorStop((1, 0, 0, 1), 0), px.ColorStop((0, 0, 1, 0.5), 0.7),
                                   px.ColorStop((0, 1, 0, 1), 1)])
    ipaint = px.Paint(px.PaintKind.ImagePaint, image=b, imageMat=rot)

    def timed_op(fn):
        ms = []
        for _ in range(reps):
            px.lib().pxh_image_copy(work.handle, src.handle)
            px.event_record(0)
            fn()
            px.event_record(1)
            ms.append(px.event_elapsed_ms(0, 1))
        return sorted(ms)[len(ms) // 2]

    out = {"workload": "draws and paints on a 4096x4096 RGBA8 canvas: a 3000x3000 source; "
                       "algorithmic bytes 8 per canvas pixel written plus the source once"}
    for name, fn, npx_written in (
            ("draw_rot30_scale1.2", lambda: work.draw(b, transform=rot), None),
            ("draw_translate_frac", lambda: work.draw(b, transform=frac), 3000 * 3000),
            ("draw_tiled_rot10_scale0.8", lambda: work.drawTiled(b, tile), W * H),
            ("fill_radial_gradient", lambda: work.fillPath(heart, grad, transform=hm), W * H),
            ("fill_image_paint", lambda: work.fillPath(heart, ipaint, transform=hm), W * H)):
        ms = timed_op(fn)
        rec = {"ms": round(ms, 4)}
        if npx_written:
            rec["algorithmic_gbs"] = round((8.0 * npx_written + 4.0 * 3000 * 3000) / (ms / 1e3) / 1e9, 1)
        out[name] = rec
    return out


def ingest_extra(px, reps=5):
    """Text and SVG ingestion (fonts.nim typeset + fillText, svg.nim parseSvg
    + newImage): a 36,540-glyph page (NotoSans 12 px, 52 lines) and a
    608-glyph paragraph (Roboto 16 px on 1000x150) filled on the device; the
    reference's Ghostscript tiger (240 elements, 900x900) and dragon (1,377
    elements, 1568x940) drawn on the device.  Host times are wall clock;
    device times are HIP events around the call on the library stream
    (median of `reps`).  cpu_baseline: the oracle (a C restatement of pixie's
    scalar path, one core) drawing the same span paths / documents."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    golden = os.path.join(ROOT, "tests", "golden")
    paras = ("She had come to the conclusion that you could tell a lot about a person by their "
             "ears. The way they stuck out and the size of the earlobes could give you wonderful "
             "insights into the person. Of course, she couldn't scientifically prove any of this, "
             "but that didn't matter to her. ")

    def wall(fn):
        ts = []
        for _ in range(reps):
            px.synchronize()
            t = time.perf_counter()
            fn()
            px.synchronize()
            ts.append(time.perf_counter() - t)
        return sorted(ts)[len(ts) // 2] * 1e3

    def dev(fn):
        ms = []
        for _ in range(reps):
            px.event_record(0)
            fn()
            px.event_record(1)
            ms.append(px.event_elapsed_ms(0, 1))
        return sorted(ms)[len(ms) // 2]

    out = {}
    for key, font, size, text, (w, h) in (
            ("text_page", "NotoSans-Regular_4.ttf", 12, paras * 120, (4096, 4096)),
            ("text_paragraph", "Roboto-Regular_1.ttf", 16, paras * 3, (1000, 150))):
        f = px.readFont(os.path.join(golden, "fonts", font))
        f.size = size
        t = time.perf_counter()
        a = px.typeset(f, text, bounds=(w - 96, h))
        ts_ms = (time.perf_counter() - t) * 1e3
        img = px.Image(w, h)
        img.fill((255, 255, 255, 255))
        img.fillText(a)  # glyph paths cached, pools sized
        fw = wall(lambda: img.fillText(a))
        rec = {"glyphs": len(a.runes), "canvas": [w, h], "typeset_ms": round(ts_ms, 3),
               "fill_text_ms": round(fw, 3), "device_ms": round(dev(lambda: img.fillText(a)), 3),
               "glyphs_per_s": round(len(a.runes) / (fw / 1e3))}
        try:
            import oracle as orc
            (path,) = a.computePaths()
            o = orc.new_image(w, h)
            t = time.perf_counter()
            orc.fill_path(o, path, (0, 0, 0, 255))
            cms = (time.perf_counter() - t) * 1e3
            rec["cpu_baseline"] = {"ms": round(cms, 2), "cores": 1, "kind": "port",
                                   "sample": "the oracle's fillPath of the same span path"}
        except Exception as e:  # the checker is optional here
            rec["cpu_baseline"] = {"error": str(e)}
        out[key] = rec
    for name in ("Ghostscript_Tiger", "dragon2"):
        with open(os.path.join(golden, "svg", name + ".svg"), "rb") as fh:
            doc = fh.read()
        t = time.perf_counter()
        s = px.parseSvg(doc)
        parse_ms = (time.perf_counter() - t) * 1e3
        img = px.svg.newImage(s)
        render = lambda: px._check(px.lib().pxh_svg_render(s._h, img._h))  # noqa: E731
        rw = wall(render)
        rec = {"elements": len(s), "canvas": [s.width, s.height], "parse_ms": round(parse_ms, 3),
               "render_ms": round(rw, 3), "device_ms": round(dev(render), 3)}
        try:
            import oracle as orc
            from oracle import svg as osvg
            od = osvg.parse_svg(orc, doc)
            t = time.perf_counter()
            osvg.render(orc, od)
            rec["cpu_baseline"] = {"ms": round((time.perf_counter() - t) * 1e3, 2), "cores": 1,
                                   "kind": "port", "sample": "the oracle's newImage(svg)"}
        except Exception as e:
            rec["cpu_baseline"] = {"error": str(e)}
        out["svg_" + name] = rec
    return out


def calibrate(px, reps=3):
    """A plain device copy of a 64 MiB image, for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    a = px.Image(4096, 4096)
    b = px.Image(4096, 4096)
    for _ in range(reps):
        px.lib().pxh_image_copy(b.handle, a.handle)
    px.synchronize()
    print(json.dumps({"calibration": "pxh_image_copy of 67108864 bytes", "copies": reps}), flush=True)


def star_extra(px, WL, reps=50):
    import oracle
    cmd, w, h = WL.star512()
    img = px.Image(w, h)
    img.fillPath(cmd.cmds, cmd.color)
    px.synchronize()
    px.touched_pixels(reset=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        img.fillPath(cmd.cmds, cmd.color)
    px.synchronize()
    dt = (time.perf_counter() - t0) / reps
    touched = px.touched_pixels(reset=True) / reps
    want = oracle.new_image(w, h)
    t1 = time.perf_counter()
    st = oracle.fill_path(want, cmd.cmds, cmd.color)
    cdt = time.perf_counter() - t1
    return {"workload": "star 64-vertex 512x512 NonZero Normal (one fillPath call)",
            "gpu_ms_per_call": dt * 1e3, "gpu_mpix_s": touched / dt / 1e6,
            "cpu_ms_per_call": cdt * 1e3, "cpu_mpix_s": st["touched"] / cdt / 1e6}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="default 20 batched, 3 plain")
    ap.add_argument("--warmup", type=int, default=None, help="default 2 batched, 1 plain")
    ap.add_argument("--form", default="batched", choices=["batched", "plain"])
    ap.add_argument("--paths", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-glyphs", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="only copy a 64 MiB image (counter calibration under rocprofv3)")
    ap.add_argument("--only", default=None,
                    choices=["image_ops", "draw", "glyphs", "serving", "ingest"],
                    help="run one extra section and print it (profiling)")
    ap.add_argument("--shape", type=int, default=None,
                    help="fixed phase-A launch shape instead of the autotuner's choice (profiling)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 20 if args.form == "batched" else 3
    if args.warmup is None:
        args.warmup = 2 if args.form == "batched" else 1
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has one MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pixie_amd as px
    from pixie_amd import workloads as WL

    px.init()
    if args.calibrate:
        calibrate(px)
        return
    if args.only:
        fn = {"image_ops": image_ops_extra, "draw": draw_extra, "glyphs": glyph_extra,
              "serving": serve_extra, "ingest": lambda px, WL: ingest_extra(px)}[args.only]
        print(json.dumps({args.only: fn(px, WL)}), flush=True)
        return
    t = time.perf_counter()
    cmds, canvas = WL.mixed_scene(n_paths=args.paths)
    log("scene: %d calls generated in %.1fs" % (len(cmds), time.perf_counter() - t))
    img = px.Image.from_array(canvas)

    shape = None
    if args.form == "batched":
        t = time.perf_counter()
        sc = record_scene(px, cmds)
        log("recorded %d calls in %.2fs" % (len(sc), time.perf_counter() - t))
        if args.shape is None:
            shape, tune_ms = sc.autotune(img)
            log("autotuner: phase-A shape %d (%.3f ms per render)" % (shape, tune_ms))
        else:
            shape = sc.set_shape(args.shape)
        step = lambda: sc.render(img)
    else:
        step = lambda: run_plain(px, img, cmds)
    for _ in range(args.warmup):
        step()
    px.synchronize()
    # the value: K renders as they run in production (graph replays for the
    # batched form); the stage times: the same K renders, each launch timed
    wall, touched, _ = timed(px, step, args.steps, profile=False)
    _, _, stages = timed(px, step, args.steps, profile=True)

    per_step_touch = touched / args.steps
    ms_per_step = wall / args.steps * 1e3
    value = per_step_touch / (wall / args.steps) / 1e6
    bytes_per_step = 8.0 * per_step_touch
    # dominant kernel: the stage with the most device time
    kname, (kms, klaunch) = max(stages.items(), key=lambda kv: kv[1][0])
    avg_launch_s = kms / klaunch / 1e3
    bytes_per_launch = bytes_per_step * args.steps / klaunch
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "mixed fill+stroke scene 4096x4096 RGBA8: %d cubic paths, "
                               "%d fill/stroke calls, seed 2" % (args.paths, len(cmds)),
                   "form": args.form, "n_touch_per_step": int(per_step_touch),
                   "phase_a_shape": shape if args.form == "batched" else None,
                   "graph_replay": sc.graph_active() if args.form == "batched" else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": kname, "launches_per_step": klaunch // args.steps,
                     "avg_launch_ms": round(kms / klaunch, 4),
                     "stages_ms_per_step": {k: round(v[0] / args.steps, 4)
                                            for k, v in stages.items()}},
    }
    # HBM traffic of the dominant kernel, as measured with rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE (copy-calibrated) and recorded beside the profile
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("kernel") == kname and args.form == "batched" and tr.get("shape") == shape:
            result["roofline"]["traffic"] = tr["traffic_bytes_per_launch"]
            result["roofline"]["traffic_source"] = "profiles/traffic.json: " + tr["source"]
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cmds, canvas, args.cpu_seconds)
    if not args.no_extra:
        extra = {"star512": star_extra(px, WL)}
        if not args.no_glyphs:
            extra["glyph_atlas"] = glyph_extra(px, WL)
        extra["image_ops"] = image_ops_extra(px, WL)
        extra["draw"] = draw_extra(px, WL)
        extra["ingest"] = ingest_extra(px)
        extra["serving"] = serve_extra(px, WL)
        if args.form == "batched":
            pimg = px.Image.from_array(canvas)
            pw, pt, pst = timed(px, lambda: run_plain(px, pimg, cmds), 1)
            extra["plain_form"] = {"mpix_s": round(pt / pw / 1e6, 2), "ms_per_step": round(pw * 1e3, 2),
                                   "kernel_ms": round(sum(v[0] for v in pst.values()), 3),
                                   "batched_speedup": round(value / (pt / pw / 1e6), 2)}
        result["extra"] = extra
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
