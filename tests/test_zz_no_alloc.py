"""The allocation-free blur entry point (pxh_blur_noalloc) against pxh_blur.

pxh_blur_noalloc keeps its intermediate image and Gaussian table in a static
device array shared by every call, so two calls in a row, the second on a
smaller image, must each equal pxh_blur bit for bit (the second must not see
the first's leftovers), and inputs beyond the array must be refused.
"""
import ctypes

import numpy as np
import pytest

import pixie_amd as px

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    px.init()
    L = px.lib()
    L.pxh_blur_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_uint32]
    L.pxh_blur_noalloc.restype = ctypes.c_int
    return L


def _premul(rng, h, w):
    a = rng.integers(0, 256, (h, w), dtype=np.int64)
    c = (rng.random((h, w, 3)) * (a[..., None] + 1)).astype(np.int64)
    c = np.minimum(c, a[..., None])
    return np.concatenate([c, a[..., None]], -1).astype(np.uint8)


def _noalloc(lib, img, radius, oob):
    rc = lib.pxh_blur_noalloc(img.handle, radius, px.pack(oob))
    if rc != 0:
        raise px.PixieError(rc, lib.pxh_last_error().decode())


@pytest.mark.parametrize("oob", [(0, 0, 0, 0), (10, 20, 30, 255)])
def test_noalloc_blur_matches_blur_twice(lib, oob):
    rng = np.random.default_rng(77)
    for (h, w, r) in [(1024, 2048, 8), (300, 517, 3)]:   # second call smaller
        src = _premul(rng, h, w)
        want = px.Image.from_array(src)
        want.blur(r, oob)
        got = px.Image.from_array(src)
        _noalloc(lib, got, float(r), oob)
        px.synchronize()
        assert np.array_equal(got.download(), want.download()), (h, w, r)


def test_noalloc_blur_full_scratch_and_radius_32(lib):
    rng = np.random.default_rng(78)
    src = _premul(rng, 2048, 2048)                  # exactly the scratch size
    want = px.Image.from_array(src)
    want.blur(32)
    got = px.Image.from_array(src)
    _noalloc(lib, got, 32.0, (0, 0, 0, 0))
    px.synchronize()
    assert np.array_equal(got.download(), want.download())


def test_noalloc_blur_limits(lib):
    big = px.Image(2049, 2048)
    with pytest.raises(px.PixieError) as e:
        _noalloc(lib, big, 4.0, (0, 0, 0, 0))
    assert e.value.code == 9
    img = px.Image(16, 16)
    with pytest.raises(px.PixieError) as e:
        _noalloc(lib, img, -2.0, (0, 0, 0, 0))
    assert e.value.code == 4
    _noalloc(lib, img, 0.2, (0, 0, 0, 0))           # rounds to 0: no-op
    px.synchronize()
