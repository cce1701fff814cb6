"""Pins of the display-post oracle (oracle/post_oracle.c, the restatement of src/demo/post.js), each chosen so
that a plausible slip fails it:

* the sampler equals scipy's order-1 map_coordinates (clamp-to-edge at uv * size - 0.5) on random images;
* the downsample of a linear ramp is the ramp at the destination pixel centre, away from the borders
  (a wrong tap offset, sign or weight breaks the symmetry), and of a constant image is the constant;
* the Karis weighting equals the plain 13-tap filter on a constant image and damps a single firefly more
  than the plain filter does;
* the upsample adds the 3x3 tent: a constant level c lands c on top of the destination;
* AgX: grey in, grey out (every matrix row sums to 1: a transposed matrix breaks it), the closed forms at
  both ends of the log range (black and very bright, where the log clamps to 0 and 1), monotone along grey;
* sRGB encode: the standard OETF values (continuity at 0.0031308, 0.5 -> 0.7353569830, 1 -> 1);
* the whole pass on a constant image: levels 1..6 after the upsamples are 6c, 5c, ..., c and the final
  bloom term is the closed form c (1 - m) + 6 c m;
* the composite picks left | mid | right at split / split2 and `same` shows each image's middle third.
"""
import numpy as np
import pytest

from oracle import ho

scipy_ndimage = pytest.importorskip("scipy.ndimage")


def test_sampler_matches_scipy_bilinear():
    rng = np.random.default_rng(1)
    img = rng.random((13, 17, 4)).astype(np.float32)
    for _ in range(300):
        u, v = rng.uniform(-0.2, 1.2, 2).astype(np.float32)
        got = ho.post_sample(img, u, v)
        x, y = np.float64(np.float32(u) * 17 - 0.5), np.float64(np.float32(v) * 13 - 0.5)
        want = [scipy_ndimage.map_coordinates(img[..., c].astype(np.float64), [[y], [x]], order=1, mode="nearest")[0]
                for c in range(3)]
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6)


@pytest.mark.parametrize("which", [0, 1])
def test_downsample_constant_and_ramp(which):
    c = np.full((32, 48, 4), 0.375, np.float32)
    out = ho.post_pass(which, c, 24, 16)
    assert np.allclose(out[..., :3], 0.375, rtol=0, atol=1e-3)
    if which == 1:                                    # the plain filter is linear: a ramp stays a ramp
        ys, xs = np.mgrid[0:32, 0:48].astype(np.float32)
        ramp = np.stack([xs * 0.01, ys * 0.02, xs * 0.005 + ys * 0.003, np.ones_like(xs)], -1).astype(np.float32)
        out = ho.post_pass(1, ramp, 24, 16)
        oy, ox = np.mgrid[0:16, 0:24].astype(np.float64)
        sx, sy = 2 * ox + 0.5, 2 * oy + 0.5           # the destination centre in source texel coordinates
        want = np.stack([sx * 0.01, sy * 0.02, sx * 0.005 + sy * 0.003], -1)
        inner = (slice(2, -2), slice(2, -2))
        assert np.abs(out[inner][..., :3] - want[inner]).max() < 2e-3


def test_karis_damps_a_firefly():
    # output texel (7, 7) samples source texels 12..17; a firefly at (12, 12) falls in the corner tap A only,
    # so only the first of the five Karis groups sees it and that group's weight drops by ~1 / (1 + 62)
    img = np.full((32, 32, 4), 0.1, np.float32)
    img[12, 12, :3] = 1000.0
    plain = ho.post_pass(1, img, 16, 16)[7, 7, 0]
    karis = ho.post_pass(0, img, 16, 16)[7, 7, 0]
    assert abs(plain - (0.1 + 250 * 0.03125)) < 0.05
    g = 0.1 * 0.75 + 250 * 0.25                      # group A, B, D, E: A = 0.75 * 0.1 + 0.25 * 1000
    w0, w = 0.125 / (1 + g), 0.125 / 1.1
    assert abs(karis - (g * w0 + 0.1 * (3 * w + 0.5 / 1.1)) / (w0 + 3 * w + 0.5 / 1.1)) < 0.01
    assert karis < plain / 10


def test_upsample_adds_the_tent():
    src = np.full((8, 12, 4), 0.5, np.float32)
    dst = np.full((16, 24, 4), 0.25, np.float32)
    out = ho.post_pass(2, src, 24, 16, dst)
    assert np.allclose(out[..., :3], 0.75, atol=1e-3)
    assert np.allclose(out[..., 3], 1.25)


def test_agx_grey_axis_and_closed_forms():
    # black: the log range clamps to 0, contrast(0) = -0.00232, the look gives -0.15 l, outset keeps grey
    l = -0.00232 * (0.2126 + 0.7152 + 0.0722)
    rows = np.array([1.6605 - 0.5876 - 0.0728, -0.1246 + 1.1329 - 0.0083, -0.0182 - 0.1006 + 1.1187])  # toSrgb
    np.testing.assert_allclose(ho.post_agx([0.0, 0.0, 0.0]), (-0.15 * l) ** 2.2 * rows, rtol=2e-3)
    # very bright: the log range clamps to 1, contrast(1) = the coefficient sum
    k = 15.5 - 40.14 + 31.96 - 6.868 + 0.4298 + 0.1191 - 0.00232
    bright = (k + (k ** 1.22 - k) * 1.15) ** 2.2 * rows
    np.testing.assert_allclose(ho.post_agx([1e6, 1e6, 1e6]), np.minimum(bright, 1.0), rtol=2e-6)
    prev = -1.0
    for g in np.geomspace(1e-4, 1e3, 60):
        o = ho.post_agx([g, g, g])
        assert np.ptp(o) < 2e-3 * max(o.max(), 1e-3), (g, o)     # grey stays grey
        assert o.mean() >= prev - 1e-6                             # monotone
        prev = o.mean()
    assert ho.post_agx([0.18, 0.18, 0.18]).mean() < ho.post_agx([1.0, 1.0, 1.0]).mean() < 1.0
    # a saturated primary keeps its hue ordering
    r = ho.post_agx([4.0, 0.1, 0.1])
    assert r[0] > r[1] and r[0] > r[2]


def test_srgb_oetf():
    assert ho.post_srgb(0.0) == 0.0
    assert abs(ho.post_srgb(1.0) - 1.0) < 1e-6
    assert abs(ho.post_srgb(0.5) - 0.7353569830524495) < 1e-6
    lo, hi = ho.post_srgb(0.0031308), ho.post_srgb(np.nextafter(np.float32(0.0031308), np.float32(1)))
    assert abs(lo - 0.0031308 * 12.92) < 1e-7 and abs(hi - lo) < 1e-5


def test_constant_image_closed_form():
    c = 0.25
    img = np.full((64, 96, 4), c, np.float16)
    target, hdr, levels = ho.post_encode(img, img, split=0.5, line=0.0, bloom=0.03, vignette=0.0)
    assert np.all(hdr[..., :3] == np.float32(c)) and np.all(hdr[..., 3] == 1.0)
    for i, lv in enumerate(levels):                   # after the upsamples: levels[i] holds (6 - i) c
        assert np.allclose(lv[..., :3], (6 - i) * c, rtol=4e-3), (i, lv[..., :3].min(), lv[..., :3].max())
    mixed = c * (1 - np.float32(0.03)) + 6 * c * np.float32(0.03)
    t = ho.post_agx([mixed] * 3)
    want = np.rint(np.clip([ho.post_srgb(x) for x in t], 0, 1) * 255)
    assert np.abs(target[..., :3].astype(int) - want).max() <= 1


def test_composite_split_and_same():
    h, w = 8, 30
    left = np.zeros((h, w, 4), np.float16); left[..., 0] = 1
    mid = np.zeros((h, w, 4), np.float16); mid[..., 1] = 1
    right = np.zeros((h, w, 4), np.float16); right[..., 2] = 1
    _, hdr, _ = ho.post_encode(left, right, mid, split=0.2, split2=0.7, exposure=2.0, line=0.0)
    u = (np.arange(w) + 0.5) / w
    want = np.where(u < 0.2, 0, np.where(u < 0.7, 1, 2))
    assert np.array_equal(np.argmax(hdr[0, :, :3], -1), want)
    assert np.all(hdr[..., :3].max(-1) == 2.0)
    # same: each third shows the middle third of its own image
    ramp = np.zeros((h, w, 4), np.float16); ramp[..., 0] = np.arange(w)
    _, hdr, _ = ho.post_encode(ramp, ramp, ramp, split=1 / 3, split2=2 / 3, same=True, line=0.0)
    assert np.array_equal(hdr[0, :, 0], np.tile(np.arange(10, 20), 3).astype(np.float32))
    # negative radiance is clamped at 0 before the exposure
    neg = np.full((h, w, 4), -3.0, np.float16)
    _, hdr, _ = ho.post_encode(neg, neg, line=0.0)
    assert np.all(hdr[..., :3] == 0.0)


def test_divider_line_and_formats():
    img = np.full((20, 40, 4), 0.05, np.float16)
    t_rgba, _, _ = ho.post_encode(img, img, split=0.5, line=1.5, bgra=False)
    t_bgra, _, _ = ho.post_encode(img, img, split=0.5, line=1.5, bgra=True)
    assert np.array_equal(t_rgba[..., [2, 1, 0, 3]], t_bgra)
    # columns whose centre is within 1.5 px of the divider at x = 20 are lightened towards white
    lit = t_rgba[10, :, 1].astype(int)
    assert lit[19] > lit[16] + 100 and lit[20] > lit[23] + 100 and np.all(t_rgba[..., 3] == 255)
