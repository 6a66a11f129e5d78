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
