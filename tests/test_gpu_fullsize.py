"""Size-independent properties at the benchmark sizes (1920x1080, 3840x2160), where the oracle is too slow
for a frame-by-frame comparison:

* a constant static scene (a plane under constant illumination, static camera) is a fixed point of every
  filter level and TAA: output = albedo x c within 2e-3 (the FP32 oracle's own tolerance for this
  invariant), alpha exactly 1, sky columns pass their emission through (same tolerance);
* the history length counts 1, 2, ..., 32 and saturates, and the reprojection mask reports history on
  every pixel from the second frame on (bit-exact);
* a frame after reset() is bit-identical to the first frame of a fresh context;
* the same sequence run twice gives bit-identical outputs (no order-dependent reductions, no races);
* the default schedule (coarse1 and coarse2 in the launches of infer + mid1 and mid3) gives bit-identical
  outputs to the eight separate passes of a timed frame."""
import numpy as np
import pytest

import hush_amd
from hush_amd.hip import Surface
from tests.parity import INPUT_KEYS, log_parity, upload_frame

pytestmark = pytest.mark.gpu

SIZES = [(1920, 1080), (3840, 2160)]


def _f16(a):
    return np.ascontiguousarray(np.asarray(a, np.float16)).view(np.uint16)


def _static(w, h, albedo=0.5, c=2.0, depth=10.0, sky_cols=0):
    rad = np.zeros((h, w, 4), np.float32); rad[..., :3] = albedo * c; rad[..., 3] = 1
    nd = np.zeros((h, w, 4), np.float32); nd[..., 2] = 1.0; nd[..., 3] = depth
    mv = np.zeros((h, w, 4), np.float32); mv[..., 2] = 7.0; mv[..., 3] = depth
    mat = np.zeros((h, w, 4), np.float32); mat[..., :3] = albedo; mat[..., 3] = 0.5
    emi = np.zeros((h, w, 4), np.float32)
    if sky_cols:
        nd[:, :sky_cols, 3] = -1.0
        emi[:, :sky_cols, :3] = 0.75
        mat[:, :sky_cols, :3] = 1.0
    return {"radiance": _f16(rad), "normalDepth": _f16(nd), "motion": _f16(mv), "material": _f16(mat),
            "emission": _f16(emi)}


def _gpu_frames(w, h, n, seed, kind="boxes"):
    from hush_amd.scene import GpuSceneGenerator
    gen = GpuSceneGenerator(kind, w, h, seed)
    frames = []
    for f in range(n):
        d = {k: Surface(w, h) for k in INPUT_KEYS}
        gen.render(f, d)
        frames.append(d)
    hush_amd.hip.synchronize()
    gen.close()
    return frames


@pytest.mark.parametrize("w,h", SIZES)
def test_constant_static_scene_is_a_fixed_point(w, h):
    g = _static(w, h, sky_cols=7)
    hs = hush_amd.Hush(0, {"width": w, "height": h})
    surfs = upload_frame(g)
    for f in range(5):
        out = hs.denoise(surfs).download().astype(np.float32)
        lit = out[:, 7:, :3]
        assert np.allclose(lit, 1.0, rtol=2e-3, atol=0), (f, float(lit.min()), float(lit.max()))
        assert np.allclose(out[:, :7, :3], 0.75, rtol=2e-3, atol=0), (f, float(out[:, :7, :3].min()), float(out[:, :7, :3].max()))
        assert np.all(out[..., 3] == 1.0)
    hs.destroy()


@pytest.mark.parametrize("w,h", SIZES[:1])
def test_history_length_counts_and_saturates(w, h):
    g = _static(w, h)
    hs = hush_amd.Hush(0, {"width": w, "height": h})
    hs.set_debug_mask(True)
    surfs = upload_frame(g)
    for f in range(36):
        hs.denoise(surfs)
        mom = hs.surface(f"mom{f & 1}").download().astype(np.float32)
        assert np.all(mom[..., 3] == min(f + 1, 32)), (f, np.unique(mom[..., 3])[:8])
        assert np.all((hs.surface("mask").download()[..., 0] & 0x20) != 0) == (f > 0), f
    hs.destroy()


@pytest.mark.parametrize("w,h", SIZES)
def test_reset_equals_fresh_first_frame(w, h):
    frames = _gpu_frames(w, h, 5, seed=11)
    a = hush_amd.Hush(0, {"width": w, "height": h})
    for d in frames[:4]:
        a.denoise(d)
    a.reset()
    out_a = a.denoise(frames[4]).download()
    b = hush_amd.Hush(0, {"width": w, "height": h})
    out_b = b.denoise(frames[4]).download()
    a.destroy()
    b.destroy()
    assert np.array_equal(out_a.view(np.uint16), out_b.view(np.uint16))


@pytest.mark.parametrize("w,h", SIZES[:1])
def test_runs_are_deterministic_and_schedules_agree(w, h):
    frames = _gpu_frames(w, h, 6, seed=12)
    outs = []
    for timed in (False, False, True):
        hs = hush_amd.Hush(0, {"width": w, "height": h})
        timer = hush_amd.HushTimer() if timed else None
        seq = [hs.denoise(d, timer=timer).download().view(np.uint16).copy() for d in frames]
        outs.append(seq)
        hs.destroy()
    for f in range(len(frames)):
        assert np.array_equal(outs[0][f], outs[1][f]), ("rerun", f)
        assert np.array_equal(outs[0][f], outs[2][f]), ("serial schedule", f)


def test_benchmark_sequence_matches_oracle_and_schedules():
    """The benchmark workload itself: the 64-frame 1920x1080 boxes sequence (seed 0: 200 boxes, 40 moving,
    ~1% emissive pixels, ~10% sky, 3 px/frame pan) from a reset, end to end against the FP32-mirror oracle on
    every frame, where drift over the recurrence would show: validity masks bit-exact; per frame at most 1e-4
    of the HDR values beyond 1e-3 relative, 1e-5 beyond 1e-2, and 99.99% within one binary16 ulp.  A rerun
    and the timed eight-launch schedule run in lock-step, bit-identical to the six-launch frame on all 64
    frames."""
    from hush_amd.scene import SceneGenerator
    from oracle.ho import Oracle
    from tests.parity import compare
    w, h, F = 1920, 1080, 64
    gen = SceneGenerator("boxes", w, h, seed=0)
    o = Oracle(w, h, hush_amd.default_weights())
    ctx = [hush_amd.Hush(0, {"width": w, "height": h}) for _ in range(3)]
    ctx[0].set_debug_mask(True)
    timer = hush_amd.HushTimer()
    surfs = None
    stats = []
    for f in range(F):
        g = gen.render(f)
        surfs = upload_frame(g, surfs)
        outs = [c.denoise(surfs, timer=timer if i == 2 else None).download() for i, c in enumerate(ctx)]
        ref = o.denoise(g)
        st = compare(outs[0].astype(np.float32)[..., :3], ref[..., :3])
        st["frame"] = f
        st["mask_exact"] = float(np.mean(ctx[0].surface("mask").download()[..., 0] == o.mask()))
        print(f"frame {f}: exact {st['exact']:.5f} max_rel {st['max_rel']:.3e} frac>1e-3 {st['frac_rel_gt_1e-3']:.2e} "
              f"p99.99 ulp {st['p9999_ulp']:.1f} mask {st['mask_exact']}", flush=True)
        if st["max_rel"] > 1e-2:
            a, r = outs[0].astype(np.float64)[..., :3], ref[..., :3].astype(np.float64)
            rel = np.abs(a - r) / np.maximum(np.abs(r), 1e-3)
            for k in np.argsort(-rel.ravel())[:3]:
                y, x, c = np.unravel_index(k, rel.shape)
                lum = lambda v: float(v[0] * 0.2126 + v[1] * 0.7152 + v[2] * 0.0722)
                print(f"    ({x},{y}) gpu {a[y, x]} oracle {r[y, x]} lum {lum(a[y, x]):.4g} / {lum(r[y, x]):.4g}", flush=True)
        stats.append(st)
        assert np.array_equal(outs[0].view(np.uint16), outs[1].view(np.uint16)), ("rerun", f)
        assert np.array_equal(outs[0].view(np.uint16), outs[2].view(np.uint16)), ("timed schedule", f)
    for c in ctx:
        c.destroy()
    log_parity("e2e-boxes-1920x1080-64", stats)
    # measured over the 64 frames: at most 4.4e-5 of the values beyond 1e-3 relative and 3.9e-6 beyond 1e-2,
    # 99.99% within one binary16 ulp.  The largest single deviations (up to ~5x relative at one pixel) sit at
    # the reference's own singular points: TAA's inverse tone map c / max(1 - lum, 1e-4) of a blended value
    # near luminance 1 (an emissive pixel of a few hundred to 1e4), and dark values that are small differences
    # of the detail recombination uB / sum + kappa (f - uA / sum), where one binary16 ulp of an input is a
    # large part of the result.
    for st in stats:
        assert st["mask_exact"] == 1.0, st
        assert st["frac_rel_gt_1e-3"] <= 1e-4 and st["frac_rel_gt_1e-2"] <= 1e-5 and st["p9999_ulp"] <= 1.0, st


def test_uhd_mlp_precisions_against_fp64_oracle():
    """The precision configuration at 3840x2160: six frames of the GPU-generated boxes sequence through the FP32
    path and the FP16, BF16 and FP8 MFMA forms of the MLP.  PSNR of the tone-mapped output against the
    FP64-ideal oracle on the same inputs: FP16 within 0.05 dB of the FP32 path, BF16 within 0.1 dB, FP8
    within 1 dB; each form reruns bit-identically."""
    from oracle.ho import Oracle
    from tests.parity import psnr
    w, h, F = 3840, 2160, 6
    frames = _gpu_frames(w, h, F, seed=2)
    o = Oracle(w, h, hush_amd.default_weights(), mode="fp64")
    refs = [o.denoise({k: d[k].download().view(np.uint16) for k in INPUT_KEYS}) for d in frames]
    o.close()
    res = {}
    for p in ("fp32", "fp16-mfma", "bf16-mfma", "fp8-mfma"):
        a = hush_amd.Hush(0, {"width": w, "height": h, "precision": p})
        b = hush_amd.Hush(0, {"width": w, "height": h, "precision": p})
        vals = []
        for d, r in zip(frames, refs):
            oa = a.denoise(d).download()
            ob = b.denoise(d).download()
            assert np.array_equal(oa.view(np.uint16), ob.view(np.uint16)), p
            vals.append(psnr(oa.astype(np.float32), r))
        a.destroy()
        b.destroy()
        res[p] = float(np.mean(vals))
    print("PSNR vs the FP64 oracle at 3840x2160 (dB):", res)
    assert res["fp32"] > 30, res
    assert res["fp16-mfma"] >= res["fp32"] - 0.05, res
    assert res["bf16-mfma"] >= res["fp32"] - 0.1, res
    assert res["fp8-mfma"] >= res["fp32"] - 1.0, res
