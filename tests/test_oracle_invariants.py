"""Invariants the mathematics fixes, checked on the C oracle (closed forms that a dropped term, a wrong
index or a transposed operand would break)."""
import numpy as np
import pytest

import hush_amd
from hush_amd.scene import SceneGenerator
from oracle.ho import Oracle

W16 = hush_amd.default_weights()


def f16(a):
    return np.ascontiguousarray(np.asarray(a, np.float16)).view(np.uint16)


def static_scene(w, h, albedo=0.5, c=2.0, depth=10.0, sky_cols=0):
    """a fronto-parallel plane with constant illumination c, static camera (mv = 0)"""
    rad = np.zeros((h, w, 4)); rad[..., :3] = albedo * c; rad[..., 3] = 1
    nd = np.zeros((h, w, 4)); nd[..., 2] = 1.0; nd[..., 3] = depth
    mv = np.zeros((h, w, 4)); mv[..., 2] = 7.0; mv[..., 3] = depth
    mat = np.zeros((h, w, 4)); mat[..., :3] = albedo; mat[..., 3] = 0.5
    emi = np.zeros((h, w, 4))
    if sky_cols:
        nd[:, :sky_cols, 3] = -1.0
        emi[:, :sky_cols, :3] = 0.75
        mat[:, :sky_cols, :3] = 1.0
    return {"radiance": f16(rad), "normalDepth": f16(nd), "motion": f16(mv), "material": f16(mat), "emission": f16(emi)}


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_constant_static_scene_is_a_fixed_point(mode):
    """constant illumination on a plane: every filter level and TAA preserve it; output = albedo * c"""
    w, h = 40, 24
    g = static_scene(w, h)
    o = Oracle(w, h, W16, mode=mode)
    tol = 1e-9 if mode == "fp64" else 2e-3
    for f in range(5):
        out = o.denoise(g)
        assert np.allclose(out[..., :3], 1.0, rtol=tol, atol=0), (f, out[..., :3].min(), out[..., :3].max())
        assert np.all(out[..., 3] == 1.0)


def test_history_length_counts_and_saturates():
    """len1 = min(previous len1 + 1, 32) with a static camera (hush.js:196): 1, 2, ..., 32, 32"""
    w, h = 16, 16
    g = static_scene(w, h)
    o = Oracle(w, h, W16, mode="fp64")
    for f in range(36):
        o.denoise(g)
        mom = o.surface(f"mom{f & 1}")
        assert np.all(mom[..., 3] == min(f + 1, 32)), (f, np.unique(mom[..., 3]))
        assert np.all((o.mask() & 0x20) != 0) == (f > 0)


def test_sky_passes_emission_and_zero_history():
    w, h = 32, 16
    g = static_scene(w, h, sky_cols=5)
    o = Oracle(w, h, W16, mode="fp64")
    for f in range(3):
        out = o.denoise(g)
    assert np.allclose(out[:, :5, :3], 0.75)           # sky: X = emission (hush.js:576), TAA fixed point
    assert np.all(o.surface("A0")[:, :5] == 0)
    assert np.all(o.mask()[:, :5] == 0x80)


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_reset_equals_fresh_first_frame(mode):
    """after reset() the next frame equals a first frame of a new denoiser on the same inputs"""
    w, h = 48, 32
    gen = SceneGenerator("spheres", w, h, 9)
    a = Oracle(w, h, W16, mode=mode)
    for f in range(3):
        a.denoise(gen.render(f))
    a.reset()
    g = gen.render(3)
    out_a = a.denoise(g)
    b = Oracle(w, h, W16, mode=mode)
    out_b = b.denoise(g)
    assert np.array_equal(out_a, out_b)


def test_resize_resets_and_same_size_is_noop():
    w, h = 24, 16
    gen = SceneGenerator("boxes", w, h, 2)
    o = Oracle(w, h, W16)
    o.denoise(gen.render(0))
    o.resize(w, h)
    assert o.frame == 1
    o.resize(w + 2, h)
    assert o.frame == 0
