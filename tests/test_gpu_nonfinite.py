"""GPU regression for motion vectors that overflowed binary16 (+-inf) or are NaN (tests/nonfinite.py):
every kernel that reprojects (Hush temporal and TAA, SVGF temporal and resolve, the differentiable frame)
keeps its index arithmetic defined and in bounds, and matches the oracles as on ordinary frames.
Tolerances as in test_gpu_parity.py / test_gpu_svgf.py.  Also non-finite radiance (+inf from an overflowed
firefly; NaN, with +inf and zero albedo) through every Hush and SVGF pass, teacher forced (stores clamp NaN to
-65504 on both sides), and the TAA tone map's zero divisor: a colour whose luminance is exactly -1 tone-maps
to infinities, which the stores clamp like the oracle's."""
import numpy as np
import pytest

from oracle.ho import Oracle, SvgfOracle
from tests.nonfinite import inject, inject_radiance
from tests.parity import compare, force_oracle, gpu_F, gpu_surface, pass_outputs, upload_frame

pytestmark = pytest.mark.gpu


def test_hush_passes_with_nonfinite_motion():
    import hush_amd
    from hush_amd.scene import SceneGenerator
    w, h = 96, 64
    gen = SceneGenerator("boxes", w, h, 9)
    hs = hush_amd.Hush(0, {"width": w, "height": h})
    hs.set_debug_mask(True)
    o = Oracle(w, h, hush_amd.default_weights())
    surfs, bad = None, []
    for f in range(3):
        g = inject(gen.render(f), f)
        surfs = upload_frame(g, surfs)
        o.set_inputs(g)
        for p in range(8):
            hs.set_state(f, f == 0)
            o.set_state(f, f == 0)
            force_oracle(hs, o)
            hs.run_pass(p, surfs)
            o.run_pass(p)
            for name in pass_outputs(p, f & 1):
                st = compare(gpu_F(hs) if name == "F" else gpu_surface(hs, name), o.surface(name))
                if st["max_ulp"] > 1.0 or st["exact"] < 0.999:
                    bad.append((f, p, name, st))
            if p == 0:
                assert np.array_equal(hs.surface("mask").download()[..., 0], o.mask()), f
    hs.destroy()
    assert not bad, bad[:4]


@pytest.mark.parametrize("kind", ["inf", "nan"])
def test_hush_passes_with_infinite_radiance(kind):
    """binary16 +inf radiance (overflowed fireflies), or NaN radiance with +inf and zero albedo, through every
    pass, teacher forced, against the oracle"""
    import hush_amd
    from hush_amd.scene import SceneGenerator
    w, h = 96, 64
    gen = SceneGenerator("boxes", w, h, 19)
    hs = hush_amd.Hush(0, {"width": w, "height": h})
    o = Oracle(w, h, hush_amd.default_weights())
    surfs, bad = None, []
    for f in range(3):
        g = inject_radiance(gen.render(f), f, kind)
        surfs = upload_frame(g, surfs)
        o.set_inputs(g)
        for p in range(8):
            hs.set_state(f, f == 0)
            o.set_state(f, f == 0)
            force_oracle(hs, o)
            hs.run_pass(p, surfs)
            o.run_pass(p)
            for name in pass_outputs(p, f & 1):
                st = compare(gpu_F(hs) if name == "F" else gpu_surface(hs, name), o.surface(name))
                if st["max_ulp"] > 1.0 or st["exact"] < 0.999:
                    bad.append((f, p, name, st))
    hs.destroy()
    assert not bad, bad[:4]


def test_svgf_with_nonfinite_motion():
    import hush_amd
    from hush_amd.hip import Surface
    from hush_amd.scene import SceneGenerator
    w, h = 96, 64
    gen = SceneGenerator("boxes", w, h, 10)
    sv = hush_amd.SVGF(0, w, h)
    o = SvgfOracle(w, h)
    keys = ("radiance", "normalDepth", "motion", "material", "emission")
    prev = None
    for f in range(4):
        g = inject(gen.render(f), f)
        src = prev if prev is not None else g
        g2 = dict(g, prevNormalDepth=src["normalDepth"], prevMotion=src["motion"])
        prev = g
        surfs = {k: Surface(w, h).upload(np.ascontiguousarray(g2[k]).view(np.float16))
                 for k in keys + ("prevNormalDepth", "prevMotion")}
        out = sv.encode(surfs).download().astype(np.float32)
        st = compare(out, o.encode(g2))
        assert st["frac_rel_gt_1e-3"] <= 1e-3 and st["max_rel"] <= 1e-2, (f, st)
    sv.destroy()


@pytest.mark.parametrize("kind", ["inf", "nan"])
def test_svgf_passes_with_nonfinite_radiance(kind):
    """+inf radiance, or NaN radiance with +inf and zero albedo, through every SVGF pass, teacher forced (the
    single-ulp flips of ordinary frames compound through the history next to such texels, so each kernel is
    checked on identical inputs)"""
    import hush_amd
    from hush_amd.hip import Surface
    from hush_amd.scene import SceneGenerator
    from tests.test_gpu_svgf import NAMES, outputs_of
    w, h = 96, 64
    gen = SceneGenerator("boxes", w, h, 10)
    sv = hush_amd.SVGF(0, w, h)
    o = SvgfOracle(w, h)
    iters = sv.iterations
    keys = ("radiance", "normalDepth", "motion", "material", "emission")
    prev, bad = None, []
    for f in range(3):
        g = inject_radiance(gen.render(f), f, kind)
        src = prev if prev is not None else g
        g2 = dict(g, prevNormalDepth=src["normalDepth"], prevMotion=src["motion"])
        prev = g
        surfs = {k: Surface(w, h).upload(np.ascontiguousarray(g2[k]).view(np.float16))
                 for k in keys + ("prevNormalDepth", "prevMotion")}
        o.set_inputs(g2)
        for p in range(3 + iters):
            sv.set_state(f, f == 0)
            o.set_state(f, f == 0)
            for name in NAMES:
                o.surface(name)[...] = sv.surface(name).download().astype(np.float32)
            sv.run_pass(p, surfs)
            o.run_pass(p)
            for name in outputs_of(p, f & 1, iters):
                st = compare(sv.surface(name).download().astype(np.float32), o.surface(name))
                if st["max_ulp"] > 1.0 or st["exact"] < 0.999:
                    bad.append((f, p, name, st))
    sv.destroy()
    assert not bad, bad[:4]


def test_train_frame_with_nonfinite_motion():
    import torch
    import hush_amd
    from hush_amd import train as T
    from hush_amd.scene import SceneGenerator
    w, h = 96, 64
    gen = SceneGenerator("boxes", w, h, 11)
    hs = hush_amd.Hush(0, {"width": w, "height": h})
    fr = T.DiffFrame(hs)
    wt = torch.tensor(T.weights_to_float(hush_amd.default_weights()), dtype=torch.float32, device="cuda")
    prev = None
    for f in range(4):
        g = inject(gen.render(f, clean=True), f)
        fr.prepare(T.upload(g))
        out = fr.forward(wt)
        tgt = T.target_of(g, torch.float32)
        l, gr = fr.loss(out, tgt, prev[0] if prev else None, prev[1] if prev else None, 0.5 if prev else 0.0)
        gw = fr.backward(wt, gr)
        assert torch.isfinite(out).all() and torch.isfinite(gw).all() and np.isfinite(float(l)), f
        fr.commit(out)
        prev = (out, tgt)
    fr.destroy()
    hs.destroy()


def test_taa_tone_map_zero_divisor():
    """(-7.75390625, 1.5419921875, -6.29296875) has fp32 luminance exactly -1, so tm() divides by 0: the
    reset frame's TAA output is itm(tm(c)) = (-inf, +inf, -inf) / 1e-4, stored as -65504, 65504, -65504
    (a NaN in place of an infinity would store -65504 in every channel)."""
    import hush_amd
    from hush_amd.scene import SceneGenerator
    w, h = 64, 32
    gen = SceneGenerator("boxes", w, h, 5)
    hs = hush_amd.Hush(0, {"width": w, "height": h})
    o = Oracle(w, h, hush_amd.default_weights())
    g = gen.render(0)
    surfs = upload_frame(g)
    o.set_inputs(g)
    for p in range(7):
        hs.set_state(0, True)
        o.set_state(0, True)
        force_oracle(hs, o)
        hs.run_pass(p, surfs)
        o.run_pass(p)
    x = hs.surface("X").download()
    spots = [(3, 5), (10, 40), (31, 63)]
    for (yy, xx) in spots:
        x[yy, xx, :3] = np.array([-7.75390625, 1.5419921875, -6.29296875], np.float16)
    hs.surface("X").upload(x)
    hs.set_state(0, True)
    o.set_state(0, True)
    force_oracle(hs, o)
    hs.run_pass(7, surfs)
    o.run_pass(7)
    gpu, ref = gpu_surface(hs, "taa0"), o.surface("taa0")
    for (yy, xx) in spots:
        assert list(ref[yy, xx, :3]) == [-65504.0, 65504.0, -65504.0], ref[yy, xx]
    st = compare(gpu, ref)
    assert np.array_equal(gpu[..., :3], ref[..., :3]) or (st["max_ulp"] <= 1.0 and st["exact"] >= 0.999), st
    for (yy, xx) in spots:
        assert np.array_equal(gpu[yy, xx, :3], ref[yy, xx, :3]), (gpu[yy, xx], ref[yy, xx])
    hs.destroy()
