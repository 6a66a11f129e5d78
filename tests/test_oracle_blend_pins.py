"""Pins the oracle's blend table (blends.nim:18-299) with properties that
follow from pixie's formulas, chosen so that a plausible slip (a swapped
operand, a dropped term, the wrong alpha, a wrong rounding) fails at least one.

No reference fixture covers most of these modes (only Exclusion, through
tests/contexts/blendmode_1.png, pinned in test_oracle_contexts.py), so each
mode gets:
  * identities with a transparent operand (every separable pixie formula
    reduces to the other operand);
  * the closed form on opaque pairs (the formula with both alphas at 255
    reduces to a plain function of the two colours);
  * symmetry where pixie's formula is symmetric, and the swap relations the
    reference defines (Overlay(b, s) is HardLight with its operands swapped,
    blends.nim:170-174 against :179-183);
  * special values (a white or black operand) where the formula collapses.
The float modes (SoftLight, Hue, Saturation, Color, Luminosity) go through
chroma, which is absent; they are pinned by the W3C compositing spec's
special cases that blends.nim:5-6 cites, within 1 LSB.
"""
import numpy as np
import pytest

NORMAL, DARKEN, MULTIPLY, COLORBURN, LIGHTEN, SCREEN, COLORDODGE, OVERLAY = range(8)
SOFTLIGHT, HARDLIGHT, DIFFERENCE, EXCLUSION, HUE, SATURATION, COLOR, LUMINOSITY = range(8, 16)
MASK, OVERWRITE, SUBTRACTMASK, EXCLUDEMASK = range(16, 20)

N = 200_000


def _premul(rng, n, alpha=None):
    """n valid premultiplied pixels (every channel <= alpha); alpha drawn from
    a mix of edge values and uniform ones unless given."""
    if alpha is None:
        edge = np.array([0, 1, 2, 127, 128, 254, 255])
        a = np.where(rng.random(n) < 0.3, edge[rng.integers(0, len(edge), n)],
                     rng.integers(0, 256, n))
    else:
        a = np.full(n, alpha)
    c = (rng.random((n, 3)) * (a[:, None] + 1)).astype(np.int64)
    c = np.minimum(c, a[:, None])
    return np.concatenate([c, a[:, None]], axis=1).astype(np.uint8)


def _opaque(rng, n):
    return _premul(rng, n, alpha=255)


def _alpha(b, s):
    """blendAlpha, blends.nim:41-43."""
    b = b.astype(np.int64)
    s = s.astype(np.int64)
    return (s + (b * (255 - s)) // 255).astype(np.uint8)


@pytest.fixture(scope="module")
def rng():
    return np.random.default_rng(20240601)


def _bl(orc, mode, b, s):
    return orc.blend_pixels(mode, b, s)


SEPARABLE_INT = [NORMAL, DARKEN, MULTIPLY, LIGHTEN, SCREEN, OVERLAY, HARDLIGHT, DIFFERENCE,
                 EXCLUSION]


@pytest.mark.parametrize("mode", SEPARABLE_INT)
def test_transparent_operand_identities(orc, rng, mode):
    b = _premul(rng, N)
    s = _premul(rng, N)
    t = np.zeros_like(b)
    assert np.array_equal(_bl(orc, mode, t, s), s), "transparent backdrop must give the source"
    assert np.array_equal(_bl(orc, mode, b, t), b), "transparent source must leave the backdrop"


@pytest.mark.parametrize("mode", [DARKEN, MULTIPLY, LIGHTEN, SCREEN, DIFFERENCE, EXCLUSION])
def test_symmetric_modes(orc, rng, mode):
    # each of these formulas is symmetric in (backdrop, source), alpha included:
    # s + floor(b(255-s)/255) == s + b - ceil(bs/255)
    b = _premul(rng, N)
    s = _premul(rng, N)
    assert np.array_equal(_bl(orc, mode, b, s), _bl(orc, mode, s, b))


@pytest.mark.parametrize("mode", SEPARABLE_INT + [COLORBURN, COLORDODGE])
def test_alpha_is_blend_alpha(orc, rng, mode):
    # every separable mode composites alpha as source-over (blendAlpha), and
    # ColorBurn/Dodge's alphaFix gives sa + ba(255-sa) div 255 too
    b = _premul(rng, N)
    s = _premul(rng, N)
    out = _bl(orc, mode, b, s)
    assert np.array_equal(out[:, 3], _alpha(b[:, 3], s[:, 3]))


def test_opaque_closed_forms(orc, rng):
    b = _opaque(rng, N)
    s = _opaque(rng, N)
    bc = b[:, :3].astype(np.int64)
    sc = s[:, :3].astype(np.int64)
    full = np.full((N, 1), 255)

    def chk(mode, want):
        out = _bl(orc, mode, b, s)
        assert np.array_equal(out[:, :3].astype(np.int64), want), "mode %d" % mode
        assert (out[:, 3] == 255).all()

    screen = bc + sc - (bc * sc) // 255
    chk(NORMAL, sc)
    chk(DARKEN, np.minimum(bc, sc))
    chk(LIGHTEN, np.maximum(bc, sc))
    chk(MULTIPLY, (bc * sc) // 255)
    chk(SCREEN, screen)
    chk(DIFFERENCE, np.abs(bc - sc))
    chk(EXCLUSION, bc + sc - (2 * bc * sc) // 255)
    # HardLight keys on the source, Overlay on the backdrop
    chk(HARDLIGHT, np.where(2 * sc <= 255, (2 * sc * bc) // 255, screen))
    chk(OVERLAY, np.where(2 * bc <= 255, (2 * sc * bc) // 255, screen))
    del full


def test_overlay_is_swapped_hardlight(orc, rng):
    b = _premul(rng, N)
    s = _premul(rng, N)
    ov = _bl(orc, OVERLAY, b, s)
    hl = _bl(orc, HARDLIGHT, s, b)
    assert np.array_equal(ov[:, :3], hl[:, :3])


def test_darken_le_lighten_and_bracket(orc, rng):
    b = _premul(rng, N)
    s = _premul(rng, N)
    dk = _bl(orc, DARKEN, b, s).astype(int)
    lt = _bl(orc, LIGHTEN, b, s).astype(int)
    nm = _bl(orc, NORMAL, b, s).astype(int)
    assert (dk[:, :3] <= lt[:, :3]).all()
    # Normal is one of the two terms Darken/Lighten choose between
    # (source + backdrop(255-sa)/255), so it lies between them
    assert (dk[:, :3] <= nm[:, :3]).all() and (nm[:, :3] <= lt[:, :3]).all()


def test_multiply_and_screen_special_backdrops(orc, rng):
    s = _premul(rng, N)
    white = np.full_like(s, 255)
    black = np.zeros_like(s)
    black[:, 3] = 255
    # Multiply onto opaque white is the source composited over white
    assert np.array_equal(_bl(orc, MULTIPLY, white, s), _bl(orc, NORMAL, white, s))
    # Screen / Exclusion onto opaque black keep the source's colour, alpha 255
    for mode in (SCREEN, EXCLUSION):
        out = _bl(orc, mode, black, s)
        assert np.array_equal(out[:, :3], s[:, :3]) and (out[:, 3] == 255).all()
    # Screen onto opaque white is white
    assert (_bl(orc, SCREEN, white, s) == 255).all()
    # Exclusion onto opaque white inverts an opaque source
    so = _opaque(rng, N)
    out = _bl(orc, EXCLUSION, white, so)
    assert np.array_equal(out[:, :3], 255 - so[:, :3])


def test_difference_of_equal_opaque_is_black(orc, rng):
    x = _opaque(rng, N)
    out = _bl(orc, DIFFERENCE, x, x)
    assert (out[:, :3] == 0).all() and (out[:, 3] == 255).all()


def test_color_burn_dodge_special_values(orc, rng):
    b = _opaque(rng, N)
    s = _opaque(rng, N)
    white = np.full_like(b, 255)
    black = np.zeros_like(b)
    black[:, 3] = 255
    # a white source leaves ColorBurn's backdrop, a black one ColorDodge's
    assert np.array_equal(_bl(orc, COLORBURN, b, white), b)
    assert np.array_equal(_bl(orc, COLORDODGE, b, black), b)
    # a white backdrop stays white under ColorBurn, a black one black under Dodge
    assert np.array_equal(_bl(orc, COLORBURN, white, s), white)
    assert np.array_equal(_bl(orc, COLORDODGE, black, s), black)
    # opposite extremes: burn by black keeps only saturated channels, dodge by
    # white lifts every non-zero channel to 255
    burn = _bl(orc, COLORBURN, b, black)
    assert np.array_equal(burn[:, :3], np.where(b[:, :3] == 255, 255, 0))
    dodge = _bl(orc, COLORDODGE, b, white)
    assert np.array_equal(dodge[:, :3], np.where(b[:, :3] == 0, 0, 255))
    # opaque pairs against the straight-colour formulas of blends.nim:115-121, 162-168
    bc = b[:, :3].astype(np.int64)
    sc = s[:, :3].astype(np.int64)
    with np.errstate(divide="ignore"):
        burn_want = np.where(bc == 255, 255, np.where(sc == 0, 0,
                             255 - np.minimum(255, (255 * (255 - bc)) // np.maximum(sc, 1))))
        dodge_want = np.where(bc == 0, 0, np.where(sc == 255, 255,
                              np.minimum(255, (255 * bc) // np.maximum(255 - sc, 1))))
    assert np.array_equal(_bl(orc, COLORBURN, b, s)[:, :3].astype(np.int64), burn_want)
    assert np.array_equal(_bl(orc, COLORDODGE, b, s)[:, :3].astype(np.int64), dodge_want)


def test_color_burn_dodge_transparent_backdrop(orc, rng):
    # with a transparent backdrop alphaFix returns the straight source, which
    # premultiplies back to the source within the un-premultiply round trip
    s = _premul(rng, N)
    t = np.zeros_like(s)
    for mode in (COLORBURN, COLORDODGE):
        out = _bl(orc, mode, t, s).astype(int)
        assert np.array_equal(out[:, 3], s[:, 3])
        assert np.abs(out - s.astype(int)).max() <= 1


def test_mask_modes_closed_forms(orc, rng):
    b = _premul(rng, N)
    s = _premul(rng, N)
    so = _opaque(rng, N)
    t = np.zeros_like(b)
    # Mask keeps the backdrop under an opaque source, clears it under none
    assert np.array_equal(_bl(orc, MASK, b, so), b)
    assert (_bl(orc, MASK, b, t) == 0).all()
    # SubtractMask is the inverse
    assert (_bl(orc, SUBTRACTMASK, b, so) == 0).all()
    # blends.nim:233-239 scales the colour by the new alpha, not by 255 - sa,
    # so a transparent source keeps alpha but multiplies colour by it once more
    sub = _bl(orc, SUBTRACTMASK, b, t)
    assert np.array_equal(sub[:, 3], b[:, 3])
    assert np.array_equal(sub[:, :3].astype(int),
                          (b[:, :3].astype(int) * b[:, 3:4].astype(int)) // 255)
    ob = _opaque(rng, N)
    assert np.array_equal(_bl(orc, SUBTRACTMASK, ob, t), ob)
    # ExcludeMask of a pixel with itself is empty; it reads the source's colour
    assert (_bl(orc, EXCLUDEMASK, b, b) == 0).all()
    ob = _opaque(rng, N)
    assert np.array_equal(_bl(orc, EXCLUDEMASK, t, so), so)
    assert (_bl(orc, EXCLUDEMASK, ob, so) == 0).all()
    # Mask and SubtractMask alphas add up to the backdrop's within 1
    ma = _bl(orc, MASK, b, s)[:, 3].astype(int)
    sa = _bl(orc, SUBTRACTMASK, b, s)[:, 3].astype(int)
    assert (np.abs(ma + sa - b[:, 3].astype(int)) <= 1).all()
    # Overwrite is the source
    assert np.array_equal(_bl(orc, OVERWRITE, b, s), s)


# ---- float modes (W3C compositing-1 special cases) ----

def _lum(c):
    c = c.astype(np.float64) / 255
    return 0.3 * c[:, 0] + 0.59 * c[:, 1] + 0.11 * c[:, 2]


@pytest.mark.parametrize("mode", [SOFTLIGHT, HUE, SATURATION, COLOR, LUMINOSITY])
def test_float_modes_transparent_and_self(orc, rng, mode):
    b = _opaque(rng, 20_000)
    out = _bl(orc, mode, b, b).astype(int)
    if mode == SOFTLIGHT:
        # B(cb, cb) is not the identity for soft light; only alpha is fixed
        assert (out[:, 3] == 255).all()
    else:
        # SetLum(SetSat(C, Sat(C)), Lum(C)) == C for the non-separable modes
        assert np.abs(out - b.astype(int)).max() <= 1
    s = _premul(rng, 20_000)
    t = np.zeros_like(s)
    out = _bl(orc, mode, t, s).astype(int)
    assert np.array_equal(out[:, 3], s[:, 3])
    assert np.abs(out - s.astype(int)).max() <= 1
    out = _bl(orc, mode, s, t).astype(int)
    assert np.abs(out - s.astype(int)).max() <= 1


def test_soft_light_special_sources(orc, rng):
    b = _opaque(rng, 20_000)
    cb = b[:, :3].astype(np.float64) / 255
    black = np.zeros_like(b); black[:, 3] = 255
    white = np.full_like(b, 255)
    # cs = 0: B = cb - cb(1 - cb) = cb^2
    out = _bl(orc, SOFTLIGHT, b, black)[:, :3].astype(np.float64)
    assert np.abs(out - cb * cb * 255).max() <= 1.0
    # cs = 1: B = D(cb), sqrt(cb) above 1/4, ((16cb - 12)cb + 4)cb below
    d = np.where(cb <= 0.25, ((16 * cb - 12) * cb + 4) * cb, np.sqrt(cb))
    out = _bl(orc, SOFTLIGHT, b, white)[:, :3].astype(np.float64)
    assert np.abs(out - d * 255).max() <= 1.0
    # cs = 128/255 is within 0.4 % of mid grey, where B = cb
    grey = np.full_like(b, 128); grey[:, 3] = 255
    out = _bl(orc, SOFTLIGHT, b, grey)[:, :3].astype(int)
    assert np.abs(out - b[:, :3].astype(int)).max() <= 1


def test_non_separable_with_grey_and_swaps(orc, rng):
    b = _opaque(rng, 20_000)
    s = _opaque(rng, 20_000)
    g = rng.integers(0, 256, 20_000)
    grey = np.stack([g, g, g, np.full_like(g, 255)], 1).astype(np.uint8)
    # Color(b, s) = SetLum(Cs, Lum(Cb)) = Luminosity(s, b) on opaque pairs
    assert np.array_equal(_bl(orc, COLOR, b, s), _bl(orc, LUMINOSITY, s, b))
    # a grey source has no saturation: Saturation, Hue and Color all give the
    # grey of the backdrop's luminosity
    for mode in (SATURATION, HUE, COLOR):
        out = _bl(orc, mode, b, grey).astype(int)
        spread = out[:, :3].max(1) - out[:, :3].min(1)
        assert spread.max() <= 1, "mode %d" % mode
        assert np.abs(out[:, :3].mean(1) - _lum(b) * 255).max() <= 1.5, "mode %d" % mode
    # Luminosity onto a grey backdrop takes the source's luminosity
    out = _bl(orc, LUMINOSITY, grey, s)
    assert np.abs(_lum(out[:, :3]) - _lum(s[:, :3])).max() * 255 <= 1.5
    # Luminosity keeps the backdrop's hue ordering (SetLum shifts all channels
    # equally before clipping)
    out = _bl(orc, LUMINOSITY, b, grey).astype(int)
    assert np.abs(_lum(out[:, :3]) - g / 255).max() * 255 <= 1.5


def test_div255_identity_exhaustive():
    # (x * 0x8081) >> 23 == x div 255 for every product of two bytes; pixie's
    # SIMD blendNormal/blendMask/applyOpacity (sse2.nim, avx2.nim) rely on it,
    # so their floor results equal the scalar `div 255` ones
    x = np.arange(0, 255 * 255 + 1, dtype=np.int64)
    assert np.array_equal((x * 0x8081) >> 23, x // 255)


def test_spec_separable_modes_within_rounding(orc, rng):
    # For these six modes pixie's integer formula is the W3C premultiplied
    # form  B(cb, cs) * sa * ba + cs (1 - ba) + cb (1 - sa)  with integer
    # floors; evaluated exactly on translucent pairs it lands within 2 LSB.
    b = _premul(rng, N)
    s = _premul(rng, N)
    bc, ba = b[:, :3] / 255.0, b[:, 3:4] / 255.0
    sc, sa = s[:, :3] / 255.0, s[:, 3:4] / 255.0
    over = sc * (1 - ba) + bc * (1 - sa)
    spec = {
        DARKEN: np.minimum(sc * ba, bc * sa) + over,
        MULTIPLY: sc * bc + over,
        LIGHTEN: np.maximum(sc * ba, bc * sa) + over,
        SCREEN: sc + bc - sc * bc,
        DIFFERENCE: sc + bc - 2 * np.minimum(sc * ba, bc * sa),
        EXCLUSION: sc + bc - 2 * sc * bc,
    }
    for mode, want in spec.items():
        got = _bl(orc, mode, b, s)[:, :3].astype(np.float64)
        assert np.abs(got - want * 255).max() <= 2.0, "mode %d" % mode


def test_non_separable_channel_order(orc, rng):
    # SetSat and SetLum keep the order of their colour argument's channels,
    # so Saturation and Luminosity keep the backdrop's channel order and Hue
    # and Color the source's (W3C compositing-1, non-separable modes)
    b = _opaque(rng, 50_000)
    s = _opaque(rng, 50_000)

    def order_kept(out, ref):
        o = out[:, :3].astype(int)
        r = ref[:, :3].astype(int)
        ok = np.ones(len(o), bool)
        for i, j in ((0, 1), (1, 2), (0, 2)):
            # compare only where both sides are clearly apart
            sep = (np.abs(o[:, i] - o[:, j]) >= 3) & (np.abs(r[:, i] - r[:, j]) >= 3)
            ok &= ~sep | (np.sign(o[:, i] - o[:, j]) == np.sign(r[:, i] - r[:, j]))
        return ok.mean()

    assert order_kept(_bl(orc, SATURATION, b, s), b) == 1.0
    assert order_kept(_bl(orc, LUMINOSITY, b, s), b) == 1.0
    assert order_kept(_bl(orc, HUE, b, s), s) == 1.0
    assert order_kept(_bl(orc, COLOR, b, s), s) == 1.0
