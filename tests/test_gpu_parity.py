"""GPU parity: every HIP pass against the FP32-mirror oracle on identical inputs.

Teacher forcing: before each pass the oracle receives the GPU's full state, so each kernel is checked in
isolation; the end-to-end test then runs whole sequences independently on both sides.

Both math policies are checked.  Tolerances (FP32 path):
  * reprojection validity masks (tap bits, fallback, has, sky): bit-exact under either policy;
  * IEEE policy, teacher forced: at most 1 binary16 ULP per stored value (RTZ boundary flips from the
    1-ULP differences of expf / logf between the GPU's ocml and glibc) and >= 99.9% bit-exact;
  * end-to-end HDR output, IEEE, 8 frames at 640x360: |gpu - oracle| <= 1e-3 * max(|oracle|, 1e-3) for
    >= 99.99% of values, all within 8e-3 relative, 99.99% within one binary16 ulp (measured: 4.1e-5 beyond
    1e-3, max 3.8e-3; the residual is one binary16 round-toward-zero ulp, 4.9e-4 .. 9.8e-4 relative, flipped
    by single-ulp expf / logf differences between the GPU's math library and the host's, and carried by the
    history feedback);
  * FAST policy (hardware rcp / sqrt / exp2 / log2, 1-2 fp32 ulp, in the spatial filters only) is an
    opt-in outside the drop-in tolerance: teacher forced <= 8 binary16 ULP with <= 1% of values beyond
    1e-3 relative; end to end <= 2% of values beyond 1e-3 and none beyond 0.25.
"""

import numpy as np
import pytest

from oracle.ho import Oracle
from tests.parity import (INPUT_KEYS, compare, force_oracle, gpu_F, gpu_surface, log_parity, pass_outputs,
                          upload_frame)

pytestmark = pytest.mark.gpu

PASS_NAMES = ["temporal", "infer+mid1", "mid2", "mid3", "coarse1", "coarse2", "fine+final", "taa"]


def _gen(kind, w, h, seed):
    from hush_amd.scene import SceneGenerator
    return SceneGenerator(kind, w, h, seed)


def _oracle_out(o, name):
    return o.surface(name)


def teacher_forced(kind, w, h, frames, seed, exposure=1.0, ieee=True):
    import hush_amd
    gen = _gen(kind, w, h, seed)
    wts = hush_amd.default_weights()
    hs = hush_amd.Hush(0, {"width": w, "height": h, "exposure": exposure, "ieee_math": ieee})
    hs.set_debug_mask(True)
    o = Oracle(w, h, wts, exposure=exposure)
    stats = []
    surfs = None
    for f in range(frames):
        g = gen.render(f)
        surfs = upload_frame(g, surfs)
        o.set_inputs(g)
        reset = f == 0
        cur = f & 1
        for p in range(8):
            hs.set_state(f, reset)
            o.set_state(f, reset)
            force_oracle(hs, o)
            hs.run_pass(p, surfs)
            o.run_pass(p)
            for name in pass_outputs(p, cur):
                gv = gpu_F(hs) if name == "F" else gpu_surface(hs, name)
                st = compare(gv, o.surface(name))
                st.update(frame=f, pass_=PASS_NAMES[p], surface=name)
                stats.append(st)
            if p == 0:
                gm = hs.surface("mask").download()[..., 0]
                om = o.mask()
                stats.append({"frame": f, "pass_": "temporal", "surface": "mask", "n": int(gm.size),
                              "exact": float(np.mean(gm == om)), "max_ulp": float(np.sum(gm != om)), "max_rel": 0.0,
                              "mask_has_frac": float(np.mean((om & 0x20) != 0)),
                              "mask_fallback_frac": float(np.mean((om & 0x10) != 0))})
    hs.destroy()
    return stats


def _report(tag, stats):
    log_parity(tag, stats)
    worst = sorted(stats, key=lambda s: -s["max_ulp"])[:6]
    for s in worst:
        print(tag, s)


@pytest.mark.parametrize("ieee", [True, False], ids=["ieee", "fast"])
@pytest.mark.parametrize("kind,w,h,frames,seed", [("spheres", 640, 360, 4, 1), ("boxes", 160, 96, 4, 2),
                                                  ("train", 63, 37, 3, 3)])
def test_teacher_forced_passes(kind, w, h, frames, seed, ieee):
    stats = teacher_forced(kind, w, h, frames, seed, ieee=ieee)
    _report(f"tf-{kind}-{w}x{h}-{'ieee' if ieee else 'fast'}", stats)
    masks = [s for s in stats if s["surface"] == "mask"]
    assert all(s["exact"] == 1.0 for s in masks), masks
    vals = [s for s in stats if s["surface"] != "mask"]
    if ieee:
        bad = [s for s in vals if s["max_ulp"] > 1.0 or s["exact"] < 0.999]
    else:
        bad = [s for s in vals if s["max_ulp"] > 8.0 or s["frac_rel_gt_1e-3"] > 1e-2]
    assert not bad, bad


@pytest.mark.parametrize("ieee", [True, False], ids=["ieee", "fast"])
def test_end_to_end_sequence(ieee):
    import hush_amd
    w, h, frames = 640, 360, 8
    gen = _gen("spheres", w, h, 7)
    hs = hush_amd.Hush(0, {"width": w, "height": h, "ieee_math": ieee})
    hs.set_debug_mask(True)
    o = Oracle(w, h, hush_amd.default_weights())
    surfs = None
    stats = []
    for f in range(frames):
        g = gen.render(f)
        surfs = upload_frame(g, surfs)
        out = hs.denoise(surfs).download().astype(np.float32)
        ref = o.denoise(g)
        st = compare(out[..., :3], ref[..., :3])
        st["frame"] = f
        st["mask_exact"] = float(np.mean(hs.surface("mask").download()[..., 0] == o.mask()))
        stats.append(st)
    _report(f"e2e-spheres-640x360-{'ieee' if ieee else 'fast'}", stats)
    hs.destroy()
    for st in stats:
        assert st["mask_exact"] == 1.0, st
        if ieee:
            # measured: <= 4.1e-5 of the values beyond 1e-3 relative, max 3.8e-3, 99.99% within one ulp
            assert st["frac_rel_gt_1e-3"] <= 1e-4, st
            assert st["max_rel"] <= 8e-3, st
            assert st["p9999_ulp"] <= 1.0, st
        else:
            assert st["frac_rel_gt_1e-3"] <= 2e-2, st
            assert st["max_rel"] <= 0.25, st


def _oct_dec_np(ex, ey):
    """octDec (hush.js:40-46) in numpy float32: IEEE-exact, same operation order as the kernels' oct_dec"""
    ex = ex.astype(np.float32)
    ey = ey.astype(np.float32)
    vx, vy = ex.copy(), ey.copy()
    vz = (np.float32(1) - np.abs(ex)) - np.abs(ey)
    t = np.maximum(-vz, np.float32(0))
    vx = vx + np.where(vx >= 0, -t, t)
    vy = vy + np.where(vy >= 0, -t, t)
    l = np.maximum(np.sqrt((vx * vx + vy * vy) + vz * vz), np.float32(1e-20))
    return np.stack([vx / l, vy / l, vz / l], axis=-1)


@pytest.mark.parametrize("ieee", [True, False], ids=["ieee", "fast"])
def test_decoded_guides_match_reference_surfaces(ieee):
    """The decoded-guide surfaces the kernels read (history key, half-res and coarse guides) hold exactly the
    IEEE octDec of the reference's stored surfaces (key, F3, CG) and the same binary16 depth / id / count."""
    import hush_amd
    w, h = 200, 120
    gen = _gen("boxes", w, h, 5)
    hs = hush_amd.Hush(0, {"width": w, "height": h, "ieee_math": ieee})
    surfs = None
    for f in range(3):
        surfs = upload_frame(gen.render(f), surfs)
        hs.denoise(surfs)
        cur = f & 1
        key = hs.surface(f"key{cur}").download()
        kd = hs.surface(f"KD{cur}").download()
        assert np.array_equal(kd[..., :3], _oct_dec_np(key[..., 0], key[..., 1]))
        assert np.array_equal(kd[..., 3:].view(np.uint32)[..., 0], key[..., 2:].copy().view(np.uint32)[..., 0])
        f3 = hs.surface("F3").download()
        gh = hs.surface("GH").download()
        o = f3[..., :1].copy().view(np.float16)
        assert np.array_equal(gh[..., :3], _oct_dec_np(o[..., 0], o[..., 1]))
        assert np.array_equal(gh[..., 3].view(np.uint32), f3[..., 1])
        cg = hs.surface("CG").download()
        cgd = hs.surface("CGD").download()
        assert np.array_equal(cgd[..., :3], _oct_dec_np(cg[..., 0], cg[..., 1]))
        assert np.array_equal(cgd[..., 3], cg[..., 2].astype(np.float32))
    hs.destroy()


def test_compensated_sums_error_vs_fp64():
    """Compensated moment sums (a precision form): the output stays within the drop-in tolerance of the
    FP32 oracle, and its error against the FP64 oracle with the same binary16 storage is no larger than the
    plain FP32 path's (storage rounding dominates both)"""
    import hush_amd
    w, h, frames = 160, 96, 6
    gen = _gen("boxes", w, h, 31)
    gs = [gen.render(f) for f in range(frames)]
    oq = Oracle(w, h, hush_amd.default_weights(), mode="fp64q")
    refs = [oq.denoise(g) for g in gs]
    o32 = Oracle(w, h, hush_amd.default_weights())
    r32 = [o32.denoise(g) for g in gs]
    errs = {}
    for comp in (False, True):
        hs = hush_amd.Hush(0, {"width": w, "height": h, "compensated": comp})
        assert hs.compensated == comp
        surfs, e = None, []
        for g, r, r2 in zip(gs, refs, r32):
            surfs = upload_frame(g, surfs)
            out = hs.denoise(surfs).download().astype(np.float64)
            e.append(np.abs(out[..., :3] - r[..., :3]) / np.maximum(np.abs(r[..., :3]), 1e-3))
            st = compare(out[..., :3].astype(np.float32), r2[..., :3])
            assert st["frac_rel_gt_1e-3"] <= 1e-2, (comp, st)
        hs.destroy()
        errs[comp] = float(np.quantile(np.concatenate([x.ravel() for x in e]), 0.999))
    print("p99.9 relative error vs FP64 (binary16 storage): plain", errs[False], "compensated", errs[True])
    assert errs[True] <= errs[False] * 1.1 + 1e-6, errs
