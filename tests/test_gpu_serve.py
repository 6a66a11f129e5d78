"""The serving loop (pixie_amd.serve): requests served concurrently on
several execution lanes give exactly the images they give alone, and the
images the plain per-call form gives."""
import numpy as np
import pytest

import pixie_amd as px
from pixie_amd import serve
from pixie_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    px.init()
    return px


def _same(a, b, what):
    if not np.array_equal(a, b):
        d = np.argwhere(np.any(a != b, axis=-1))
        raise AssertionError("%s: %d pixels differ, first %s" % (what, len(d), d[0]))


def test_serve_alone_vs_mix(gpu):
    reqs = WL.serve_requests(seed=21, n=14, rate=0, max_paths=80, max_side=768)
    alone = []
    solo = serve.Server(n_lanes=1, hbm_budget=64 << 20, keep_pixels=True)
    for r in reqs:
        one = serve.Request(r.calls, r.side, r.blur, 0.0)
        alone.append(solo.run([one])[0].pixels)
    # 6 MiB of canvas memory: at most two 768x768 canvases are in flight at once
    mix = serve.Server(n_lanes=px.lanes() - 1, hbm_budget=6 << 20, keep_pixels=True)
    res = mix.run(reqs)  # all arrive at once: admitted as lanes and memory free up
    for i, r in enumerate(res):
        _same(r.pixels, alone[i], "request %d in the mix" % i)
    plain = serve.serve_plain(reqs, keep_pixels=True)
    for i, r in enumerate(plain):
        _same(r.pixels, alone[i], "request %d per-call" % i)


def test_serve_poisson_latency(gpu):
    reqs = WL.serve_requests(seed=22, n=30, rate=200.0, max_paths=120, max_side=1024)
    res = serve.Server(hbm_budget=64 << 20).run(reqs)
    st = serve.latency_stats(res, reqs)
    print(st)
    assert st["scenes"] == 30 and all(r.latency > 0 for r in res)


def _oracle_scene(orc, r):
    want = orc.new_image(r.side, r.side)
    dropped = 0
    for c in r.calls:
        if c.stroke:
            st = orc.stroke_path(want, c.cmds, c.color, width=c.width, cap=c.cap, join=c.join,
                                 miter_limit=c.miter, mode=c.mode)
        else:
            st = orc.fill_path(want, c.cmds, c.color, winding=c.winding, mode=c.mode)
        dropped += st["dropped"]
    assert dropped == 0
    if r.blur:
        orc.blur(want, int(r.blur))
    return want


def test_serve_full_distribution(gpu, orc):
    # requests drawn from the whole serving distribution: 1-500 paths
    # (log-uniform), canvas side 256-4096 (log-uniform), blur 0 or 8.  Each
    # result is the same alone, in the mix and per call, and two of them
    # (the largest canvas, blurred, and the largest unblurred one) match the
    # oracle
    reqs = WL.serve_requests(seed=23, n=16, rate=0, max_paths=500, max_side=4096)
    # and one at the top of the range: a 4096x4096 canvas, 500 paths, blur 8
    top, _ = WL.mixed_scene(seed=25, n_paths=500, W=4096, H=4096)
    reqs.insert(5, serve.Request(top, 4096, 8.0, 0.0))
    assert max(r.side for r in reqs) == 4096 and max(len(r.calls) for r in reqs) >= 500
    assert any(r.blur for r in reqs) and not all(r.blur for r in reqs)
    solo = serve.Server(n_lanes=1, hbm_budget=80 << 20, keep_pixels=True)
    alone = [solo.run([serve.Request(r.calls, r.side, r.blur, 0.0)])[0] for r in reqs]
    for a in alone:
        assert a.error is None
    # 72 MiB of canvas: the 64 MiB canvas waits for memory while small ones pass
    mix = serve.Server(n_lanes=px.lanes() - 1, hbm_budget=72 << 20, keep_pixels=True)
    res = mix.run(reqs)
    for i, r in enumerate(res):
        assert r.error is None
        _same(r.pixels, alone[i].pixels, "request %d in the mix" % i)
    plain = serve.serve_plain(reqs, keep_pixels=True)
    for i, r in enumerate(plain):
        _same(r.pixels, alone[i].pixels, "request %d per-call" % i)
    big = max(range(len(reqs)), key=lambda i: reqs[i].side)
    sharp = max((i for i in range(len(reqs)) if not reqs[i].blur), key=lambda i: reqs[i].side)
    assert reqs[big].blur and big != sharp
    for i in (big, sharp):
        want = _oracle_scene(orc, reqs[i])
        d = np.abs(alone[i].pixels.astype(int) - want.astype(int))
        frac = float(np.mean(d.max(-1) > 0))
        print("request %d (side %d, %d calls, blur %g) vs oracle: max diff %d, %.6f%% differ" % (
            i, reqs[i].side, len(reqs[i].calls), reqs[i].blur, d.max(), 100 * frac))
        if reqs[i].blur:
            # blur spreads a coverage difference over its window; the bar is
            # the same per pixel
            assert d.max() <= 1 and frac <= 1e-3
        else:
            assert d.max() <= 1 and frac <= 1e-4


def test_serve_failing_request_is_answered_alone(gpu):
    # a request that raises PixieError (here a negative blur radius,
    # images.nim:310-311) is answered with its error; the requests before and
    # after it are served as usual, on the same lanes
    good = WL.serve_requests(seed=24, n=6, rate=0, max_paths=60, max_side=512)
    bad = serve.Request(good[0].calls, 300, -3.0, 0.0)
    reqs = good[:3] + [bad] + good[3:]
    solo = serve.Server(n_lanes=1, hbm_budget=32 << 20, keep_pixels=True)
    alone = [solo.run([serve.Request(r.calls, r.side, r.blur, 0.0)])[0] for r in good]
    for mix in (serve.Server(hbm_budget=64 << 20, keep_pixels=True), None):
        res = mix.run(reqs) if mix else serve.serve_plain(reqs, keep_pixels=True)
        assert res[3].error is not None and res[3].error.code == 4
        for k, i in enumerate([0, 1, 2, 4, 5, 6]):
            assert res[i].error is None
            _same(res[i].pixels, alone[k].pixels, "request %d around a failing one" % i)


def test_native_server_alone_vs_mix(gpu):
    # the library's own server (pxh_server_*, the C-ABI served form): every
    # result equals the request served alone and the per-call form
    reqs = WL.serve_requests(seed=26, n=24, rate=0, max_paths=200, max_side=1536)
    solo = serve.Server(n_lanes=1, hbm_budget=16 << 20, keep_pixels=True)
    alone = [solo.run([serve.Request(r.calls, r.side, r.blur, 0.0)])[0] for r in reqs]
    # 12 MiB of arena: canvases wait for memory as well as for lanes
    mix = serve.NativeServer(hbm_budget=12 << 20, keep_pixels=True)
    res = mix.run(reqs)
    for i, r in enumerate(res):
        assert r.error is None, r.error
        _same(r.pixels, alone[i].pixels, "request %d in the native mix" % i)
    assert all(r.latency > 0 for r in res)


def test_native_server_failing_request_and_arrivals(gpu):
    # Poisson arrivals with one failing request (negative blur): it is done
    # with its status, the others are served and equal their alone images
    good = WL.serve_requests(seed=27, n=10, rate=300.0, max_paths=80, max_side=700)
    bad = serve.Request(good[0].calls, 256, -2.0, good[4].arrival)
    reqs = good[:5] + [bad] + good[5:]
    solo = serve.Server(n_lanes=1, hbm_budget=8 << 20, keep_pixels=True)
    alone = [solo.run([serve.Request(r.calls, r.side, r.blur, 0.0)])[0] for r in good]
    res = serve.NativeServer(n_lanes=2, hbm_budget=32 << 20, keep_pixels=True).run(reqs)
    assert res[5].error is not None and res[5].error.code == 4
    for k, i in enumerate([0, 1, 2, 3, 4, 6, 7, 8, 9, 10]):
        assert res[i].error is None
        _same(res[i].pixels, alone[k].pixels, "request %d (native, Poisson)" % i)
    st = serve.latency_stats(res, reqs)
    assert st["failed"] == 1 and st["scenes"] == 10
