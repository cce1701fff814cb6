"""The oracles on frames with non-finite motion vectors (tests/nonfinite.py): defined behaviour, no
history for those pixels, finite outputs elsewhere."""
import numpy as np

import hush_amd
from hush_amd.scene import SceneGenerator
from oracle.ho import Oracle, SvgfOracle
from tests.nonfinite import inject, nonfinite_mask


def test_hush_oracle_treats_nonfinite_motion_as_no_history():
    w, h = 64, 48
    gen = SceneGenerator("boxes", w, h, 3)
    o = Oracle(w, h, hush_amd.default_weights())
    for f in range(3):
        g = inject(gen.render(f), f)
        out = o.denoise(g)
        assert np.isfinite(out).all()
        if f:
            bad = nonfinite_mask(g)
            m = o.mask()
            sky = (m & 0x80) != 0
            assert not (m[bad & ~sky] & 0x3F).any()      # no tap, no fallback, no history
            assert ((m[~bad & ~sky] & 0x20) != 0).mean() > 0.5


def test_svgf_oracle_treats_nonfinite_motion_as_no_history():
    w, h = 64, 48
    gen = SceneGenerator("boxes", w, h, 4)
    o = SvgfOracle(w, h)
    prev = None
    for f in range(3):
        g = inject(gen.render(f), f)
        src = prev if prev is not None else g
        g2 = dict(g, prevNormalDepth=src["normalDepth"], prevMotion=src["motion"])
        out = o.encode(g2)
        prev = g
        assert np.isfinite(out).all()
        if f:
            bad = nonfinite_mask(g) & (g["normalDepth"].view(np.float16)[..., 3] >= 0)
            assert np.all(o.surface("histMoments%d" % (f & 1))[..., 2][bad] == 1.0)   # history length restarts
