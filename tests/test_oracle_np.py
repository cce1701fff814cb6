"""Pins the C oracle with a second, independent restatement (oracle/hush_np.py, vectorised numpy written
separately from the WGSL): every surface of every pass must agree after each frame.

* FP64 ideal (no binary16 storage in either): agreement to ~1e-12 relative — pins the algorithm
  (indices, formulas, operand order of each pass) exactly.
* FP64 with binary16 storage (RTZ stores, qf16): agreement except for rare one-ulp storage flips where the
  two restatements' last-bit differences straddle a binary16 boundary.
"""
import numpy as np
import pytest

import hush_amd
from hush_amd.scene import SceneGenerator
from oracle import hush_np
from oracle.ho import Oracle

SURF = ["A0", "X", "Bh0", "Bh1", "mom0", "mom1", "key0", "key1", "hist0", "hist1", "taa0", "taa1", "par0_0",
        "par0_1", "par1", "par2", "F", "C0", "C1", "CG"]


def np_surface(S, name):
    if name[-1] in "01" and name[:-1] in ("Bh", "mom", "key", "hist", "taa"):
        return getattr(S, name[:-1])[int(name[-1])]
    if name.startswith("par0_"):
        return S.par0[int(name[-1])]
    if name in ("C0", "C1"):
        return S.C[int(name[-1])]
    return getattr(S, name)


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-6)


@pytest.mark.parametrize("kind,w,h,frames,seed", [("spheres", 48, 32, 4, 1), ("boxes", 63, 37, 4, 2),
                                                  ("train", 40, 40, 3, 3)])
def test_ideal_fp64_restatements_agree(kind, w, h, frames, seed):
    gen = SceneGenerator(kind, w, h, seed)
    wts = hush_amd.default_weights()
    c = Oracle(w, h, wts, mode="fp64")
    n = hush_np.Hush(w, h, wts, quant=False)
    for f in range(frames):
        g = gen.render(f)
        oc = c.denoise(g)
        on = n.denoise(g)
        assert np.array_equal(c.mask(), n.S.mask), f"mask differs at frame {f}"
        for name in SURF:
            a, b = c.surface(name), np_surface(n.S, name)
            r = rel(a, b)
            assert np.quantile(r, 0.999) < 1e-9 and r.max() < 1e-6, (f, name, float(r.max()))
        assert np.quantile(rel(oc, on), 0.999) < 1e-9


@pytest.mark.parametrize("kind,w,h,frames,seed", [("spheres", 48, 32, 3, 4), ("boxes", 63, 37, 3, 5)])
def test_storage_fp64_restatements_agree(kind, w, h, frames, seed):
    gen = SceneGenerator(kind, w, h, seed)
    wts = hush_amd.default_weights()
    c = Oracle(w, h, wts, mode="fp64q")
    n = hush_np.Hush(w, h, wts, quant=True)
    for f in range(frames):
        g = gen.render(f)
        c.denoise(g)
        n.denoise(g)
        assert np.array_equal(c.mask(), n.S.mask), f"mask differs at frame {f}"
        for name in SURF:
            a, b = c.surface(name), np_surface(n.S, name)
            same = np.mean(a == b)
            assert same >= 0.995, (f, name, same)
            assert np.quantile(rel(a, b), 0.999) <= 2e-3, (f, name)
