"""The Hush state machine on the GPU through the Python API (hush_amd.Hush over the C ABI), end to end
against the FP32-mirror oracle run through the same calls: resize in the middle of a sequence (hush.js:
764-794: reallocates and resets; same size is a no-op), reset (hush.js:801), an exposure change between
frames, sizes down to 1x1 and sizes one texel past a tile edge, MLP hidden widths other than 16
(hiddenOf, hush.js:22-25: widths up to 64 zero-padded to a register form, wider ones in chunks of 32
units), the shader-f16 MLP variant (hush.js:348-356 with f16: true), and the reference's input errors
(hush.js:813-822).

Tolerance (as the end-to-end parity test): HDR output within 1e-3 relative (max(|ref|, 1e-3)) for
>= 99.9% of values, none beyond 1e-2; reprojection masks bit-exact.  A reset frame must be bit-identical
to the first frame of a fresh context."""
import numpy as np
import pytest

import hush_amd
from oracle.ho import Oracle
from tests.parity import compare, log_parity, upload_frame

pytestmark = pytest.mark.gpu


def _scene(w, h, seed, kind="boxes"):
    from hush_amd.scene import SceneGenerator
    return SceneGenerator(kind, w, h, seed)


def _check(stats, what):
    log_parity(f"api-{what}", stats)
    for st in stats:
        assert st["mask_exact"] == 1.0, (what, st)
        assert st["frac_rel_gt_1e-3"] <= 1e-3, (what, st)
        assert st["max_rel"] <= 1e-2, (what, st)


def _step(hs, o, g, surfs):
    surfs = upload_frame(g, surfs)
    out = hs.denoise(surfs).download().astype(np.float32)
    ref = o.denoise(g)
    st = compare(out[..., :3], ref[..., :3])
    st["mask_exact"] = float(np.mean(hs.surface("mask").download()[..., 0] == o.mask()))
    return st, surfs


def _pair(w, h, **opts):
    weights = opts.get("weights")
    hs = hush_amd.Hush(0, dict({"width": w, "height": h}, **opts))
    hs.set_debug_mask(True)
    o = Oracle(w, h, hush_amd.default_weights() if weights is None else weights,
               exposure=opts.get("exposure", 1.0), f16=opts.get("precision") == "f16")
    return hs, o


def test_resize_midsequence_matches_oracle():
    hs, o = _pair(64, 48)
    stats, surfs = [], None
    g1 = _scene(64, 48, 31)
    for f in range(3):
        st, surfs = _step(hs, o, g1.render(f), surfs)
        stats.append(st)
    hs.resize(97, 61)
    o.resize(97, 61)
    g2 = _scene(97, 61, 32)
    surfs = None
    for f in range(3):
        st, surfs = _step(hs, o, g2.render(f), surfs)
        stats.append(st)
    n = hs.frame
    hs.resize(97, 61)                      # same size: no reallocation, history kept
    o.resize(97, 61)
    assert hs.frame == n
    for f in range(3, 5):
        st, surfs = _step(hs, o, g2.render(f), surfs)
        stats.append(st)
    hs.destroy()
    _check(stats, "resize")


def test_reset_equals_fresh_first_frame():
    w, h = 80, 56
    gen = _scene(w, h, 33)
    frames = [gen.render(f) for f in range(5)]
    a = hush_amd.Hush(0, {"width": w, "height": h})
    surfs = None
    for g in frames[:4]:
        surfs = upload_frame(g, surfs)
        a.denoise(surfs)
    a.reset()
    surfs = upload_frame(frames[4], surfs)
    out_a = a.denoise(surfs).download()
    b = hush_amd.Hush(0, {"width": w, "height": h})
    out_b = b.denoise(surfs).download()
    a.destroy()
    b.destroy()
    assert np.array_equal(out_a.view(np.uint16), out_b.view(np.uint16))


@pytest.mark.parametrize("w,h", [(96, 64), (333, 197)])
def test_wide_history_gathers_give_the_same_bits(w, h):
    """the temporal pass gathers its history through 32-bit buffer offsets below 2 GiB per surface, and the
    temporal and fine+final passes address their surfaces with 32-bit byte offsets there, and through 64-bit
    addresses above; forced onto the 64-bit forms, a sequence gives the same bits in every output"""
    gen = _scene(w, h, 41)
    frames = [gen.render(f) for f in range(5)]
    a = hush_amd.Hush(0, {"width": w, "height": h})
    b = hush_amd.Hush(0, {"width": w, "height": h})
    b.set_wide_gathers(True)
    surfs = None
    for g in frames:
        surfs = upload_frame(g, surfs)
        out_a = a.denoise(surfs).download()
        out_b = b.denoise(surfs).download()
        assert np.array_equal(out_a.view(np.uint16), out_b.view(np.uint16))
        for name in ("A0", "X", "F0", "F3", "key0", "key1", "hist0", "hist1"):
            assert np.array_equal(a.surface(name).download(), b.surface(name).download()), name
    a.destroy()
    b.destroy()


@pytest.mark.parametrize("w,h", [(64, 24), (63, 17), (65, 9), (1, 5)])
def test_paired_history_taps_at_the_image_edges(w, h):
    """the temporal pass reads a row's two horizontal colour / moment history taps as one 16-byte load (the
    histories keep a texel of padding per row; widths 64 and 63 put that padding in a new 64-texel block or
    fill the row's last texel): a motion field of -4 .. 4 px sends taps past both image edges (bix = -1 and
    bix = W - 1 among them), and every output must equal the 64-bit form's, which loads each clamped tap
    separately"""
    gen = _scene(w, h, 43)
    rng = np.random.default_rng(7)
    frames = []
    for f in range(5):
        g = gen.render(f)
        mv = g["motion"].view(np.float16).astype(np.float32)       # binary16 bits as generated
        mv[..., 0] = rng.uniform(-4.0, 4.0, (h, w)).round(2)
        mv[..., 1] = rng.uniform(-2.0, 2.0, (h, w)).round(2)
        g["motion"] = mv.astype(np.float16).view(np.uint16)
        frames.append(g)
    a = hush_amd.Hush(0, {"width": w, "height": h})
    b = hush_amd.Hush(0, {"width": w, "height": h})
    b.set_wide_gathers(True)
    a.set_debug_mask(True)
    b.set_debug_mask(True)
    assert a.surface("mom0").pitch > w * 8 and a.surface("hist0").pitch > w * 8
    surfs, used = None, 0
    for g in frames:
        surfs = upload_frame(g, surfs)
        out_a = a.denoise(surfs).download()
        out_b = b.denoise(surfs).download()
        assert np.array_equal(out_a.view(np.uint16), out_b.view(np.uint16))
        for name in ("A0", "X", "F0", "F1", "F2", "F3", "mom0", "mom1", "mask"):
            assert np.array_equal(a.surface(name).download(), b.surface(name).download()), name
        used += int(np.count_nonzero(a.surface("mask").download()[..., 0] & 0x0f))
    if w > 1:
        assert used > 0                  # some bilinear taps passed their checks (at 1 x 5 none can)
    a.destroy()
    b.destroy()


def test_exposure_change_matches_oracle():
    hs, o = _pair(96, 64, exposure=1.0)
    gen = _scene(96, 64, 34, "spheres")
    stats, surfs = [], None
    for f in range(4):
        if f == 2:
            hs.exposure = 2.5
            o.set_exposure(2.5)
        st, surfs = _step(hs, o, gen.render(f), surfs)
        stats.append(st)
    assert abs(hs.exposure - 2.5) < 1e-12
    hs.destroy()
    _check(stats, "exposure")


@pytest.mark.parametrize("w,h", [(1, 1), (2, 1), (1, 5), (3, 3), (33, 9), (65, 17)])
def test_small_and_ragged_sizes_match_oracle(w, h):
    hs, o = _pair(w, h)
    gen = _scene(w, h, 35)
    stats, surfs = [], None
    for f in range(3):
        st, surfs = _step(hs, o, gen.render(f), surfs)
        stats.append(st)
    hs.destroy()
    _check(stats, f"{w}x{h}")


@pytest.mark.parametrize("hidden", [1, 8, 12, 32, 40, 72, 256])
def test_other_hidden_widths_match_oracle(hidden):
    """widths 8 and 32 run the register form at their own width, 1, 12 and 40 the width-8, -16 and -64 forms
    over zero-padded weights, 72 and 256 (the largest hiddenOf admits) the chunked form over weights padded to
    96 and 256"""
    from hush_amd.train import random_weights, to_f16_blob
    wts = to_f16_blob(random_weights(40 + hidden, hidden))
    assert hush_amd.hiddenOf(wts.size) == hidden
    hs, o = _pair(96, 64, weights=wts)
    assert hs.hidden == hidden
    gen = _scene(96, 64, 36)
    stats, surfs = [], None
    for f in range(3):
        st, surfs = _step(hs, o, gen.render(f), surfs)
        stats.append(st)
    hs.destroy()
    _check(stats, f"hidden {hidden}")


def test_f16_mlp_variant_matches_oracle():
    hs, o = _pair(96, 64, precision="f16")
    gen = _scene(96, 64, 37)
    stats, surfs = [], None
    for f in range(3):
        st, surfs = _step(hs, o, gen.render(f), surfs)
        stats.append(st)
    hs.destroy()
    _check(stats, "f16")


@pytest.mark.parametrize("hidden", [1, 8, 12, 24])
def test_f16_mlp_other_widths_match_oracle(hidden):
    """the f16 option at widths 1, 8 and 12 (the packed width-16 form over zero-padded weights) and 24 (the
    generic f16 form)"""
    from hush_amd.train import random_weights, to_f16_blob
    wts = to_f16_blob(random_weights(60 + hidden, hidden))
    hs, o = _pair(96, 64, precision="f16", weights=wts)
    gen = _scene(96, 64, 38)
    stats, surfs = [], None
    for f in range(3):
        st, surfs = _step(hs, o, gen.render(f), surfs)
        stats.append(st)
    hs.destroy()
    _check(stats, f"f16 hidden {hidden}")


def test_input_errors_match_reference():
    from hush_amd.hip import Surface
    from tests.parity import INPUT_KEYS
    hs = hush_amd.Hush(0, {"width": 64, "height": 48})
    good = {k: Surface(64, 48) for k in INPUT_KEYS}
    bad = dict(good, motion=Surface(32, 48))
    with pytest.raises(hush_amd.HushError, match=r"^Hush\.denoise: 'motion' is 32x48, the denoiser 64x48 \(call resize\)$"):
        hs.denoise(bad)
    missing = {k: v for k, v in good.items() if k != "material"}
    with pytest.raises(hush_amd.HushError, match=r"^Hush\.denoise: missing input 'material'$"):
        hs.denoise(missing)
    # a pitch whose rows would run past 2^31 texels (the kernels index texels in 32 bits): refused before
    # anything is launched, so the view's memory is never read
    r = good["radiance"]
    wide = dict(good, radiance=Surface(64, 48, ptr=r.ptr, pitch=(1 << 23) * 8, owner=r))
    with pytest.raises(hush_amd.HushError, match=r"^Hush\.denoise: 'radiance' has pitch 67108864 bytes; pitch x height"):
        hs.denoise(wide)
    hs.denoise(good)                                   # the context is still usable
    hs.destroy()
    with pytest.raises(hush_amd.HushError, match=r"bad size 16385x8"):
        hush_amd.Hush(0, {"width": 16385, "height": 8})
