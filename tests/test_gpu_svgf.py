"""GPU parity of the SVGF baseline (src/denoise/svgf.js) against the FP32-mirror SVGF oracle.

Teacher forcing: before each pass the oracle receives the GPU's nine internal surfaces, so every kernel is
checked in isolation on identical inputs; the end-to-end test runs whole sequences independently.

Tolerances:
  * teacher forced, per stored value: at most 1 binary16 ULP and >= 99.9% bit-exact (round-toward-zero
    boundary flips from last-bit differences of expf / powf between the GPU and glibc);
  * end to end, HDR output: |gpu - oracle| <= 1e-3 * max(|oracle|, 1e-3) for >= 99.9% of values, the rest
    within 1e-2 relative (the history feedback carries single-ULP flips from frame to frame).
"""
import numpy as np
import pytest

from oracle.ho import SvgfOracle
from tests.parity import compare

pytestmark = pytest.mark.gpu

KEYS = ("radiance", "normalDepth", "motion", "material", "emission", "prevNormalDepth", "prevMotion")
NAMES = ["histColor0", "histColor1", "histMoments0", "histMoments1", "tmpColor", "atrous0", "atrous1", "taa0", "taa1"]


def frames(kind, w, h, seed, n):
    from hush_amd.scene import SceneGenerator
    gen = SceneGenerator(kind, w, h, seed)
    prev = None
    for f in range(n):
        g = dict(gen.render(f))
        src = prev if prev is not None else g
        g["prevNormalDepth"], g["prevMotion"] = src["normalDepth"], src["motion"]
        prev = g
        yield g


def upload(g, surfs=None):
    from hush_amd.hip import Surface
    h, w = g["radiance"].shape[:2]
    surfs = surfs or {k: Surface(w, h) for k in KEYS}
    for k in KEYS:
        surfs[k].upload(np.ascontiguousarray(g[k]).view(np.float16))
    return surfs


def outputs_of(p, cur, iterations):
    if p == 0:
        return ["tmpColor", f"histMoments{cur}"]
    if p == 1:
        return ["atrous0"]
    if p < 2 + iterations:
        i = p - 2
        return [f"atrous{(i & 1) ^ 1}"] + ([f"histColor{cur}"] if i == 0 else [])
    return [f"taa{cur}"]


def teacher_forced(kind, w, h, n, seed, **opts):
    import hush_amd
    sv = hush_amd.SVGF(0, w, h, opts)
    o = SvgfOracle(w, h, **opts)
    iters = sv.iterations
    stats, surfs = [], None
    for f, g in enumerate(frames(kind, w, h, seed, n)):
        surfs = upload(g, surfs)
        o.set_inputs(g)
        for p in range(3 + iters):
            sv.set_state(f, f == 0)
            o.set_state(f, f == 0)
            for name in NAMES:
                o.surface(name)[...] = sv.surface(name).download().astype(np.float32)
            sv.run_pass(p, surfs)
            o.run_pass(p)
            for name in outputs_of(p, f & 1, iters):
                st = compare(sv.surface(name).download().astype(np.float32), o.surface(name))
                st.update(frame=f, pass_=p, surface=name)
                stats.append(st)
    sv.destroy()
    return stats


@pytest.mark.parametrize("kind,w,h,n,seed,opts", [
    ("boxes", 96, 64, 4, 11, {}),
    ("spheres", 77, 53, 3, 12, {"iterations": 3, "sigmaN": 64.5, "taa": False}),
    ("train", 128, 72, 3, 13, {"sigmaL": 2.0, "alphaColor": 0.1}),
    # sigmaL, sigmaZ above 2^60: the a-trous passes take IEEE quotients in place of shared-reciprocal ones
    ("boxes", 64, 48, 3, 14, {"sigmaL": 3e18, "sigmaZ": 3e18}),
])
def test_teacher_forced_passes(kind, w, h, n, seed, opts):
    stats = teacher_forced(kind, w, h, n, seed, **opts)
    bad = [s for s in stats if s["max_ulp"] > 1.0 or s["exact"] < 0.999]
    assert not bad, bad[:4]


@pytest.mark.parametrize("kind,w,h,n,seed", [("boxes", 160, 90, 8, 21), ("spheres", 131, 67, 6, 22)])
def test_end_to_end(kind, w, h, n, seed):
    import hush_amd
    sv = hush_amd.SVGF(0, w, h)
    o = SvgfOracle(w, h)
    surfs = None
    for f, g in enumerate(frames(kind, w, h, seed, n)):
        surfs = upload(g, surfs)
        out = sv.encode(surfs).download().astype(np.float32)
        ref = o.encode(g)
        st = compare(out, ref)
        assert st["frac_rel_gt_1e-3"] <= 1e-3, (f, st)
        assert st["max_rel"] <= 1e-2, (f, st)
    sv.destroy()


def test_output_getter_and_frame_counter():
    """the `output` getter reads taa[frame & 1] after encode advanced the frame (svgf.js:327, 376-377)"""
    import hush_amd
    w, h = 64, 32
    sv = hush_amd.SVGF(0, w, h)
    surfs = None
    for f, g in enumerate(frames("boxes", w, h, 3, 3)):
        surfs = upload(g, surfs)
        out = sv.encode(surfs)
        assert sv.frame == f + 1
        assert out.ptr == sv.surface(f"taa{f & 1}").ptr
        assert sv.output.ptr == sv.surface(f"taa{(f + 1) & 1}").ptr
    sv.resize(w, h)                 # unchanged size: a no-op
    assert sv.frame == 3
    sv.resize(32, 16)
    assert (sv.width, sv.height, sv.frame) == (32, 16, 0)
    sv.destroy()


def test_input_errors():
    import hush_amd
    from hush_amd.hip import Surface
    sv = hush_amd.SVGF(0, 32, 16)
    surfs = {k: Surface(32, 16) for k in KEYS}
    surfs.pop("prevMotion")
    with pytest.raises(hush_amd.SVGFError, match="missing input 'prevMotion'"):
        sv.encode(surfs)
    surfs["prevMotion"] = Surface(16, 16)
    with pytest.raises(hush_amd.SVGFError, match=r"'prevMotion' is 16x16, the denoiser 32x16"):
        sv.encode(surfs)
    assert sv.frame == 0
    sv.destroy()


def test_timer_labels():
    import hush_amd
    w, h = 64, 48
    sv = hush_amd.SVGF(0, w, h, {"iterations": 4})
    t = hush_amd.SVGFTimer()
    g = next(frames("boxes", w, h, 5, 1))
    sv.encode(upload(g), t)
    rec = t.read()
    assert [r[0] for r in rec] == ["temporal", "variance"] + ["atrous"] * 4 + ["resolve"]
    assert all(ms >= 0 for _, ms in rec)
    sv.destroy()
