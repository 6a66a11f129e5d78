"""The blur / shadow / mask pipeline at its full size, on the GPU against the
oracle: a 4096x4096 premultiplied RGBA8 image with alpha U{0..255} and every
channel <= alpha (seed 3), blurred at radii 2, 8 and 32 (images.nim:304-365),
shadowed at offset (12, 12) with spread 4 and blur 16 (images.nim:760-776),
and masked with an 8-bit plane of 500 random circles (radius U[8, 256],
filled white NonZero, paths.nim:2093-2113) under SubtractMask, Mask and
ExcludeMask (images.nim:468-529).  Each op runs on a fresh copy, as the
benchmark does.  Blur, shadow and the mask modes are integer arithmetic and
must be bit-exact; the circle mask is coverage and must meet the coverage
bar (<= 1/255 on <= 0.01 % of pixels).
"""
import numpy as np
import pytest

import pixie_amd as px
from pixie_amd import workloads as WL

pytestmark = pytest.mark.gpu

W = H = 4096


@pytest.fixture(scope="module")
def gpu():
    px.init()
    return px


@pytest.fixture(scope="module")
def src():
    a = WL.blurmask_image()
    assert a.shape == (H, W, 4)
    assert (a[..., :3] <= a[..., 3:4]).all()
    return a


@pytest.fixture(scope="module")
def circle_mask(gpu, orc):
    circ = WL.circles_mask_cmds()
    got = px.Image(W, H, a8=True)
    got.fillPath(circ.cmds, circ.color)
    want = orc.new_image(W, H)
    st = orc.fill_path(want, circ.cmds, circ.color)
    assert st["dropped"] == 0
    return got, got.download(), want


def _exact(got, want, what):
    if not np.array_equal(got, want):
        d = np.argwhere(np.any(got != want, axis=-1) if got.ndim == 3 else got != want)
        raise AssertionError("%s: %d pixels differ, first %s got %s want %s" % (
            what, len(d), d[0], got[tuple(d[0])], want[tuple(d[0])]))


@pytest.mark.parametrize("r", [2, 8, 32])
def test_blur_4k(gpu, orc, src, r):
    img = px.Image.from_array(src)
    img.blur(r)
    got = img.download()
    want = src.copy()
    orc.blur(want, r)
    _exact(got, want, "blur r=%d 4096x4096" % r)


def test_shadow_4k(gpu, orc, src):
    color = WL.premul(20, 30, 40, 200)
    img = px.Image.from_array(src)
    got = img.shadow((12, 12), 4, 16, color).download()
    want = orc.shadow(src, (12, 12), 4, 16, color)
    _exact(got, want, "shadow (12,12) spread 4 blur 16, 4096x4096")


def test_circle_mask_4k(gpu, circle_mask):
    _, got, want = circle_mask
    assert np.array_equal(want[..., 0], want[..., 3])
    d = np.abs(got.astype(int) - want[..., 3].astype(int))
    frac = float(np.mean(d > 0))
    print("500-circle mask: max diff %d, %.6f%% of pixels differ" % (d.max(), 100 * frac))
    assert d.max() <= 1 and frac <= 1e-4
    # the circles cover a real share of the canvas
    assert 0.2 < float(np.mean(got > 0)) < 0.99


@pytest.mark.parametrize("mode", [px.BlendMode.SubtractMask, px.BlendMode.Mask,
                                  px.BlendMode.ExcludeMask])
def test_mask_modes_4k(gpu, orc, src, circle_mask, mode):
    mask_img, mask_a8, _ = circle_mask
    img = px.Image.from_array(src)
    img.draw(mask_img, 0, 0, mode)
    got = img.download()
    # the A8 plane composites as rgbx(v, v, v, v); the oracle gets the GPU's
    # own plane, so this checks the composite alone, bit for bit
    m = np.repeat(mask_a8[..., None], 4, axis=-1)
    want = src.copy()
    orc.draw(want, m, 0, 0, mode)
    _exact(got, want, "mask mode %d 4096x4096" % mode)
