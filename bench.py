#!/usr/bin/env python3
"""Benchmark of the Hush denoiser hot path on one MI355X.

Workload: 1920x1080, 64-frame sequence of the seeded "boxes" scene (200 boxes, 40 moving, 1% emissive
pixels, ~10% sky, camera pan 3 px/frame), FP32 plain path, built-in weights.  All 64 frames of inputs are
generated on the host and resident in HBM before timing.  A step is one denoised frame (the 8 passes).

Prints one JSON line: ms/frame (lower is better), the per-pass times, the roofline of the dominant kernel
(temporal) from its algorithmic bytes and its hipEvent-measured duration, and the CPU baseline (the
oracle's FP32-mirror restatement, OpenMP on the host cores, on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/frame @1080p (plain FP32 vs FP16-MFMA MLP) at stated PSNR; %HBM peak"
HBM_PEAK_GBS = 8000.0
INPUT_KEYS = ("radiance", "normalDepth", "motion", "material", "emission")


def algorithmic_bytes(w: int, h: int) -> dict:
    """Bytes per pass at the reference's pass boundaries: each bound texture once per texel per pass,
    rgba16float 8 B, rgba32uint 16 B (reads and writes)."""
    w2, h2 = (w + 1) // 2, (h + 1) // 2
    w4, h4 = (w2 + 1) // 2, (h2 + 1) // 2
    N, N2, N4 = w * h, w2 * h2, w4 * h4
    return {"temporal": 80 * N + 72 * N2, "infer+mid1": 96 * N2, "mid2": 48 * N2, "mid3": 48 * N2,
            "coarse1": 32 * N2 + 16 * N4, "coarse2": 24 * N4, "fine+final": 48 * N + 64 * N2 + 16 * N4,
            "taa": 32 * N}


def measured_traffic(kernel: str, grid: int, precision: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    this workload (counters scaled by the copy-kernel calibration recorded beside them); None if absent."""
    path = os.path.join(ROOT, "profiles", f"pmc_{grid}_{precision}.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        d = json.load(fh)
    for k, v in d["kernels"].items():
        if k.startswith(kernel + "/"):
            return int(v["traffic_bytes"]), os.path.relpath(path, ROOT)
    return None, None


def mlp_stage(hs, hip, W, H, reps=20, replays=3):
    """The MLP alone (hush_mlp_forward) over this frame's real half-res feature rows, per precision: `reps`
    launches captured in one HIP graph and replayed `replays` times between hipEvents (a ~10 us kernel
    launched one call at a time from Python would be timed with the host's issue gaps), the mean duration
    per launch, and achieved TFLOP/s at 1536 FLOP per row.  Every replay's output is checked against an
    eager launch's, bit for bit."""
    F = np.concatenate([hs.surface(f"F{k}").download().view(np.float16) for k in range(3)], -1)
    rows = np.ascontiguousarray(F.reshape(-1, 24)[:, :20])
    n = rows.shape[0]
    fin, fout = hip.DeviceBuffer(rows.nbytes), hip.DeviceBuffer(n * 48)
    hip.check(hip.lib().hipMemcpy(hip.ctypes.c_void_p(fin.ptr), rows.ctypes.data_as(hip.ctypes.c_void_p), rows.nbytes,
                                  hip.hipMemcpyHostToDevice), "features")

    def download():
        out = np.empty(n * 12, np.float32)
        hip.check(hip.lib().hipMemcpy(out.ctypes.data_as(hip.ctypes.c_void_p), hip.ctypes.c_void_p(fout.ptr), out.nbytes,
                                      hip.hipMemcpyDeviceToHost), "outputs")
        return out

    res = {}
    stream = hip.Stream()
    e0, e1 = hip.Event(), hip.Event()
    for p in ("fp32", "f16", "fp16-mfma", "bf16-mfma", "fp8-mfma"):
        hs.mlp_forward(fin, fout, n, p, stream=stream.handle)
        stream.synchronize()
        ref = download()
        g = hip.Graph(stream, lambda s: [hs.mlp_forward(fin, fout, n, p, stream=s) for _ in range(reps)])
        g.launch()                                          # untimed replay
        stream.synchronize()
        tot = 0.0
        for _ in range(replays):
            hip.check(hip.lib().hipMemset(hip.ctypes.c_void_p(fout.ptr), 0, n * 48), "clear outputs")
            hip.synchronize()
            e0.record(stream.handle)
            g.launch()
            e1.record(stream.handle)
            e1.synchronize()
            tot += e0.elapsed_ms(e1)
            if not np.array_equal(download().view(np.uint32), ref.view(np.uint32)):
                raise RuntimeError(f"MLP stage {p}: a graph replay differs from the eager launch")
        g.destroy()
        us = tot / (reps * replays) * 1e3
        res[p] = {"us": round(us, 2), "TFLOP/s": round(1536.0 * n / (us * 1e-6) / 1e12, 2)}
    stream.destroy()
    res["rows"] = n
    res["fp16_mfma_speedup_vs_fp32"] = round(res["fp32"]["us"] / res["fp16-mfma"]["us"], 2)
    return res


def psnr_check(precisions, w=640, h=360, frames=8, seed=1):
    """Error of each MLP precision's output against the FP64 oracle over a sequence of the boxes scene:
    PSNR (tone-mapped, peak 1, mean over frames) against the FP64 ideal oracle (no binary16 storage) and its
    change from the FP32 path; and, against the FP64 oracle with the same binary16 storage (which isolates
    the arithmetic), the 99.9th percentile and the maximum of the HDR output's relative error
    |gpu - ref| / max(|ref|, 1e-3) and the largest absolute error of the tone-mapped output (the maxima are
    set by the few pixels where a threshold decision flips between FP32 and FP64, or where TAA's inverse
    tone map saturates)"""
    from oracle.ho import Oracle
    from tests.parity import psnr, tonemap, upload_frame
    import hush_amd
    from hush_amd.scene import SceneGenerator
    gen = SceneGenerator("boxes", w, h, seed=seed)
    gs = [gen.render(f) for f in range(frames)]
    o = Oracle(w, h, hush_amd.default_weights(), mode="fp64")
    refs = [o.denoise(g) for g in gs]
    o.close()
    oq = Oracle(w, h, hush_amd.default_weights(), mode="fp64q")
    refq = [oq.denoise(g)[..., :3] for g in gs]
    oq.close()
    out = {}
    for p in precisions:
        prec, comp = (p[:-len("+compensated")], True) if p.endswith("+compensated") else (p, False)
        hs = hush_amd.Hush(0, {"width": w, "height": h, "precision": prec, "compensated": comp})
        surfs, vals, rels, tmax = None, [], [], 0.0
        for g, r, rq in zip(gs, refs, refq):
            surfs = upload_frame(g, surfs)
            o_ = hs.denoise(surfs).download().astype(np.float64)[..., :3]
            vals.append(psnr(o_.astype(np.float32), r))
            rels.append((np.abs(o_ - rq) / np.maximum(np.abs(rq), 1e-3)).ravel().astype(np.float32))
            tmax = max(tmax, float(np.abs(tonemap(o_) - tonemap(rq)).max()))
        hs.destroy()
        rel = np.concatenate(rels)
        out[p] = {"psnr_db": round(float(np.mean(vals)), 3), "p999_rel_hdr_vs_fp64_storage": round(float(np.quantile(rel, 0.999)), 5),
                  "max_rel_hdr_vs_fp64_storage": round(float(rel.max()), 5),
                  "max_abs_tonemapped_vs_fp64_storage": round(tmax, 5)}
    for p in precisions:
        out[p]["psnr_delta_vs_fp32_db"] = round(out[p]["psnr_db"] - out["fp32"]["psnr_db"], 3)
    out["frames"] = frames
    return out


def uhd_leg(hip, stream, steps=32, warmup=4, frames=32):
    """3840x2160, a 32-frame sequence of the same scene type rendered on the GPU (include/hush_scene.h):
    ms/frame with the FP32 and the FP16-MFMA MLP, and the MLP stage alone on that frame's feature rows"""
    import hush_amd
    from hush_amd.scene import GpuSceneGenerator
    W, H = 3840, 2160
    gen = GpuSceneGenerator("boxes", W, H, seed=2)
    dev = []
    for f in range(frames):
        d = {k: hip.Surface(W, H) for k in INPUT_KEYS}
        gen.render(f, d, stream=stream.handle)
        dev.append(d)
    stream.synchronize()
    gen.close()
    res = {}
    e0, e1 = hip.Event(), hip.Event()
    for p in ("fp32", "fp16-mfma"):
        hs = hush_amd.Hush(0, {"width": W, "height": H, "precision": p})
        t_r, i = time.perf_counter(), 0
        while i < warmup or time.perf_counter() - t_r < 0.5:   # warm-up and clock ramp, untimed
            if i % frames == 0:
                hs.reset()
            hs.encode(dev[i % frames], stream=stream.handle)
            i += 1
            if i % 16 == 0:
                stream.synchronize()
        stream.synchronize()
        e0.record(stream.handle)
        for j in range(i, i + steps):                          # the sequence plays on from the warm-up
            if j % frames == 0:
                hs.reset()
            hs.encode(dev[j % frames], stream=stream.handle)
        e1.record(stream.handle)
        e1.synchronize()
        res[f"{p}_ms_per_frame"] = round(e0.elapsed_ms(e1) / steps, 5)
        if p == "fp32":
            res["mlp_stage"] = mlp_stage(hs, hip, W, H)
        hs.destroy()
    for d in dev:
        for v in d.values():
            v._buf.free()
    return res


def post_bytes(w, h, tw, th):
    """algorithmic bytes of one Post.encode (src/demo/post.js:179-209): every texture a pass binds once per
    texel, reads and writes (rgba16float 8 B, the 8-bit target 4 B); the additive upsample reads and
    writes its destination"""
    n = [max(1, w >> i) * max(1, h >> i) for i in range(7)]
    composite = 16 * n[0]
    down = sum(8 * n[i - 1] + 8 * n[i] for i in range(1, 7))
    up = sum(8 * n[i + 1] + 16 * n[i] for i in range(1, 6))
    final = 8 * n[0] + 8 * n[1] + 4 * tw * th
    return composite + down + up + final


def post_leg(dev, stream, hip, W, H, steps):
    """the display post (src/demo/post.js) of two resident HDR images into a W x H bgra8 target"""
    import hush_amd
    from hush_amd.hip import Surface
    post = hush_amd.Post(0, "bgra8unorm")
    tgt = Surface(W, H, np.uint8, 4)
    left, right = dev[0]["radiance"], dev[1]["radiance"]
    for _ in range(4):
        post.encode(left, right, split=0.5, target=tgt, stream=stream.handle)
    stream.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream.handle)
    for _ in range(steps):
        post.encode(left, right, split=0.5, target=tgt, stream=stream.handle)
    e1.record(stream.handle)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    post.destroy()
    b = post_bytes(W, H, W, H)
    return {"ms_per_frame": round(ms, 5), "kernels": 13, "algorithmic_MB": round(b / 1e6, 2),
            "GB/s": round(b / (ms * 1e-3) / 1e9, 1), "hbm_frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}


def svgf_leg(dev, stream, hip, W, H, steps, warmup):
    """the SVGF baseline (src/denoise/svgf.js) over the same resident frames: ms/frame and per-pass ms"""
    import hush_amd
    F = len(dev)
    sv = hush_amd.SVGF(0, W, H)
    ins = []
    for f in range(F):
        g = dict(dev[f])
        p = dev[f - 1] if f > 0 else dev[f]
        g["prevNormalDepth"], g["prevMotion"] = p["normalDepth"], p["motion"]
        ins.append(g)

    def run(n):
        for i in range(n):
            if i % F == 0:
                sv.reset()
            sv.encode(ins[i % F], stream=stream.handle)

    run(warmup)
    stream.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream.handle)
    run(steps)
    e1.record(stream.handle)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    timer = hush_amd.SVGFTimer()
    acc, n_t = None, min(steps, 32)
    for i in range(n_t):
        if i % F == 0:
            sv.reset()
        sv.encode(ins[i % F], timer=timer, stream=stream.handle)
        r = timer.read()
        v = np.array([m for _, m in r])
        acc = v if acc is None else acc + v
    labels = [l for l, _ in r]
    sv.destroy()
    passes = [[l, round(float(m) / n_t, 5)] for l, m in zip(labels, acc)]
    return {"ms_per_frame": round(ms, 5), "passes": passes, "iterations": 5}


def train_leg(dev, W, H, reps=8):
    """the differentiable frame (include/hush_train.h), FP32: (1) one 1920x1080 frame forward + loss +
    backward + commit on resident inputs; (2) one training step of 8 sequences x 16 frames of 512x512
    crops (the gradient of every frame accumulated)"""
    import torch
    import hush_amd
    from hush_amd import train as T
    from hush_amd.scene import SceneGenerator
    w = torch.tensor(T.weights_to_float(hush_amd.default_weights()), dtype=torch.float32, device="cuda")
    res = {}
    hs = hush_amd.Hush(0, {"width": W, "height": H})
    fr = T.DiffFrame(hs)
    tgt = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")

    def frame(f, prev):
        fr.prepare(dev[f % len(dev)])
        out = fr.forward(w)
        l, g = fr.loss(out, tgt, prev, tgt if prev is not None else None, 0.5 if prev is not None else 0.0)
        fr.backward(w, g)
        fr.commit(out)
        return out

    prev = None
    for f in range(3):
        prev = frame(f, prev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for f in range(reps):
        prev = frame(3 + f, prev)
    e1.record()
    e1.synchronize()
    res["fwd_bwd_ms_per_frame_1080p"] = round(e0.elapsed_time(e1) / reps, 3)
    fr.destroy()
    hs.destroy()
    # 8 sequences x 16 frames of 512x512
    S, F, C = 8, 16, 512
    gen = SceneGenerator("train", C, C, 123)
    seq = [T.upload(gen.render(f, clean=True)) for f in range(F)]
    tg = [torch.from_numpy(gen.render(f, clean=True)["clean"]).cuda() for f in range(F)]
    ctx = [hush_amd.Hush(0, {"width": C, "height": C}) for _ in range(S)]
    frs = [T.DiffFrame(c) for c in ctx]
    gw = torch.zeros(812, dtype=torch.float32, device="cuda")

    def step():
        gw.zero_()
        for c, d in zip(ctx, frs):
            c.reset()
            p_out = p_tg = None
            for f in range(F):
                d.prepare(seq[f])
                out = d.forward(w)
                l, g = d.loss(out, tg[f], p_out, p_tg, 0.5 if p_out is not None else 0.0)
                gw.add_(d.backward(w, g))
                d.commit(out)
                p_out, p_tg = out, tg[f]

    step()
    torch.cuda.synchronize()
    e0.record()
    step()
    e1.record()
    e1.synchronize()
    res["step_ms_8seq_x_16frames_512"] = round(e0.elapsed_time(e1), 2)
    res["step_frames"] = S * F
    for d in frs:
        d.destroy()
    for c in ctx:
        c.destroy()
    return res


def scene_leg(W, H, frames=8):
    """the seeded scene generator: one frame rendered on the GPU (include/hush_scene.h, host set-up
    included) against the host renderer"""
    import torch
    from hush_amd.scene import GpuSceneGenerator
    g = GpuSceneGenerator("boxes", W, H, 0)
    out = {k: torch.empty((H, W, 4), dtype=torch.float16, device="cuda") for k in INPUT_KEYS}
    clean = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    g.render(0, out, clean=clean)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for f in range(frames):
        g.render(f, out, clean=clean)
    torch.cuda.synchronize()
    gpu_ms = (time.perf_counter() - t) / frames * 1e3
    t = time.perf_counter()
    for f in range(2):
        g.cpu.render(f, clean=True)
    cpu_ms = (time.perf_counter() - t) / 2 * 1e3
    g.close()
    return {"gpu_ms_per_frame": round(gpu_ms, 3), "host_ms_per_frame": round(cpu_ms, 2)}


def cpu_baseline(frames_g, w, h, n_frames):
    """The oracle (FP32 mirror, C + OpenMP) timed on the host on the first `n_frames` frames."""
    from oracle.ho import Oracle
    import hush_amd
    o = Oracle(w, h, hush_amd.default_weights())
    o.denoise(frames_g[0])        # warm-up (page-in)
    o.reset()
    o.set_state(0, True)
    t = time.perf_counter()
    for f in range(n_frames):
        o.denoise(frames_g[f])
    dt = time.perf_counter() - t
    threads = o.threads()
    return {"value": round(dt / n_frames * 1e3, 3), "unit": "ms/frame", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{n_frames} frames of the same 1920x1080 boxes sequence, FP32 mirror oracle (C, OpenMP, "
                      f"{threads} threads)"}


def cpu_model():
    """the host CPU's model name (/proc/cpuinfo), or the platform's processor string"""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def c0_leg(hip, stream, reps=16):
    """The reference's CPU-runnable case: 640x360 spheres (16 spheres on a ground plane, orbiting camera),
    an 8-frame sequence, FP32, built-in weights.  GPU ms/frame over `reps` passes of the sequence (reset at
    each start), and the FP32-mirror oracle's ms/frame on the host cores over the same 8 frames."""
    import hush_amd
    from hush_amd.scene import SceneGenerator
    from oracle.ho import Oracle
    W, H, F = 640, 360, 8
    gen = SceneGenerator("spheres", W, H, seed=0)
    gs = [gen.render(f) for f in range(F)]
    dev = [{k: hip.Surface(W, H).upload(g[k].view(np.float16)) for k in INPUT_KEYS} for g in gs]
    hs = hush_amd.Hush(0, {"width": W, "height": H})
    for d in dev:
        hs.encode(d, stream=stream.handle)
    e0, e1 = hip.Event(), hip.Event()
    stream.synchronize()
    e0.record(stream.handle)
    for r in range(reps):
        hs.reset()
        for d in dev:
            hs.encode(d, stream=stream.handle)
    e1.record(stream.handle)
    e1.synchronize()
    gpu_ms = e0.elapsed_ms(e1) / (reps * F)
    hs.destroy()
    o = Oracle(W, H, hush_amd.default_weights())
    o.denoise(gs[0])
    o.reset()
    o.set_state(0, True)
    t = time.perf_counter()
    for g in gs:
        o.denoise(g)
    cpu_ms = (time.perf_counter() - t) / F * 1e3
    for d in dev:
        for v in d.values():
            v._buf.free()
    return {"gpu_ms_per_frame": round(gpu_ms, 5), "cpu_oracle_ms_per_frame": round(cpu_ms, 3),
            "cpu_cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "frames": F, "gpu_passes": reps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--math", default=os.environ.get("HUSH_BENCH_MATH", "ieee"), choices=["ieee", "fast"],
                    help="math policy: ieee (the oracle's operation-exact mirror) or fast")
    ap.add_argument("--cpu-frames", type=int, default=6)
    ap.add_argument("--ramp-s", type=float, default=0.5,
                    help="untimed frames before the warm-up steps: at least one pass over the sequence and this many "
                         "seconds, so the GPU reaches its steady clock (reported as clock_ramp)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the MLP-stage, FP16-MFMA frame, SVGF, training and PSNR legs")
    ap.add_argument("--mlp-stage-only", action="store_true",
                    help="after the warm-up frames, time only the MLP stage (hush_mlp_forward over the frame's half-res "
                         "feature rows, every precision, --steps launches each) and print it as one JSON line; for "
                         "profiling the MLP kernels on their own")
    ap.add_argument("--train-only", action="store_true",
                    help="after generating the inputs, time only the differentiable frame (the training leg: 1080p "
                         "forward + loss + backward, and the 8 x 16 x 512^2 step) and print it as one JSON line; for "
                         "profiling the training kernels on their own")
    ap.add_argument("--svgf-only", action="store_true",
                    help="only the SVGF baseline over the resident sequence (ms/frame and per-pass ms), one JSON line")
    ap.add_argument("--post-only", action="store_true",
                    help="time only the display post (Post.encode of two resident seeded HDR images into a "
                         "width x height bgra8 target, --steps encodes) and print it as one JSON line")
    ap.add_argument("--uhd-error-frames", type=int, default=8,
                    help="frames of the 3840x2160 error check against the FP64 oracle (CPU-bound: ~4 s per frame)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run device-to-device copies of this size (the runtime's copy kernel calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters against a known byte count)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the denoiser path runs whole frames on one device); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import hush_amd
    from hush_amd import hip
    from hush_amd.scene import SceneGenerator

    W, H, F = args.width, args.height, args.frames
    if args.post_only:
        rng = np.random.default_rng(0)
        imgs = [{"radiance": hip.Surface(W, H).upload(rng.lognormal(-1.0, 1.0, (H, W, 4)).astype(np.float16))}
                for _ in range(2)]
        print(json.dumps({"post": post_leg(imgs, hip.Stream(), hip, W, H, args.steps), "width": W, "height": H}),
              flush=True)
        return
    t0 = time.perf_counter()
    gen = SceneGenerator("boxes", W, H, seed=0)
    frames_g = [gen.render(f) for f in range(F)]
    print(f"[bench] generated {F} frames {W}x{H} in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    dev = []
    for g in frames_g:
        dev.append({k: hip.Surface(W, H).upload(g[k].view(np.float16)) for k in INPUT_KEYS})
    hip.synchronize()
    print(f"[bench] inputs resident in HBM ({F * 5 * W * H * 8 / 1e9:.2f} GB)", file=sys.stderr, flush=True)

    if args.train_only:
        print(json.dumps({"train_fp32": train_leg(dev, W, H), "width": W, "height": H}), flush=True)
        return
    if args.svgf_only:
        leg = svgf_leg(dev, hip.Stream(), hip, W, H, min(args.steps, 64), args.warmup)
        print(json.dumps({"svgf": leg, "width": W, "height": H}), flush=True)
        return
    hs = hush_amd.Hush(0, {"width": W, "height": H, "precision": args.precision, "ieee_math": args.math == "ieee"})
    stream = hip.Stream()

    # one playback position across every loop below: the sequence plays on from where the previous loop
    # stopped and restarts (reset) only at its frame 0, so a run of K frames holds the sequence's own share of
    # reset frames (one in F) and the short timed run that follows the warm-up holds none
    pos = [0]

    def next_frame(h):
        f = pos[0] % F
        pos[0] += 1
        if f == 0:
            h.reset()
        return f

    def run(n):
        for _ in range(n):
            hs.encode(dev[next_frame(hs)], stream=stream.handle)

    # clock ramp: the GPU leaves its idle clock only under sustained load, and a few warm-up frames (~1 ms of
    # work) end before it gets there; play untimed frames for one pass over the sequence and at least
    # --ramp-s seconds, then the --warmup steps, then the timed steps
    def clock_ramp(h):
        n, t_r = 0, time.perf_counter()
        while n < F or time.perf_counter() - t_r < args.ramp_s:
            for _ in range(16):
                h.encode(dev[next_frame(h)], stream=stream.handle)
            stream.synchronize()
            n += 16
        return {"frames": n, "s": round(time.perf_counter() - t_r, 3)}

    ramp = clock_ramp(hs)
    run(args.warmup)
    stream.synchronize()
    if args.mlp_stage_only:
        print(json.dumps({"mlp_stage": mlp_stage(hs, hip, W, H, reps=args.steps), "width": W, "height": H}), flush=True)
        return
    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    e0.record(stream.handle)
    t = time.perf_counter()
    run(args.steps)
    e1.record(stream.handle)
    e1.synchronize()
    wall = time.perf_counter() - t
    ms_frame = e0.elapsed_ms(e1) / args.steps
    print(f"[bench] timed {args.steps} frames: {ms_frame:.4f} ms/frame (wall {wall * 1e3 / args.steps:.4f})",
          file=sys.stderr, flush=True)

    # per-pass durations with hipEvents on the launch stream: the context's per-pass timer (each pass's kernel
    # records a start and a stop event of its own execution on the launch stream), averaged over 64 frames
    timer = hush_amd.HushTimer()
    acc = np.zeros(8)
    n_t = min(args.steps, 64)
    for i in range(n_t):
        hs.encode(dev[next_frame(hs)], timer=timer, stream=stream.handle)
        acc += np.array([ms for _, ms in timer.read()])
    pass_ms = acc / n_t
    names = ["temporal", "infer+mid1", "mid2", "mid3", "coarse1", "coarse2", "fine+final", "taa"]
    ab = algorithmic_bytes(W, H)
    total_bytes = sum(ab.values())
    passes = {n: {"ms": round(float(m), 5), "GB/s": round(ab[n] / (m * 1e-3) / 1e9, 1),
                  "hbm_frac": round(ab[n] / (m * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)} for n, m in zip(names, pass_ms)}
    tk = float(pass_ms[0])
    achieved = ab["temporal"] / (tk * 1e-3) / 1e9
    traffic, traffic_src = (measured_traffic("k_temporal", "1080p_boxes", args.precision)
                            if (W, H) == (1920, 1080) else (None, None))
    roofline = {"kernel": "temporal", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": traffic_src, "algorithmic_bytes": ab["temporal"],
                "algorithmic_bytes_rule": "80*N + 72*N2 (N = W*H, N2 = ceil(W/2)*ceil(H/2))"}

    # per-frame durations: an event after every frame of a separate 64-frame loop, median and spread
    n_m = min(args.steps, 64)
    evs = [hip.Event() for _ in range(n_m + 1)]
    stream.synchronize()
    evs[0].record(stream.handle)
    for i in range(n_m):
        hs.encode(dev[next_frame(hs)], stream=stream.handle)
        evs[i + 1].record(stream.handle)
    evs[-1].synchronize()
    per_frame = np.array([evs[i].elapsed_ms(evs[i + 1]) for i in range(n_m)])
    frame_stats = {"median_ms": round(float(np.median(per_frame)), 5), "p10_ms": round(float(np.quantile(per_frame, 0.1)), 5),
                   "p90_ms": round(float(np.quantile(per_frame, 0.9)), 5), "frames": n_m}

    extras = {f"frame_time_{W}x{H}": frame_stats}
    if not args.no_extras:
        extras["mlp_stage_1080p"] = mlp_stage(hs, hip, W, H)
        print(f"[bench] MLP stage: {extras['mlp_stage_1080p']}", file=sys.stderr, flush=True)
        # the whole frame with the FP16-MFMA MLP, with the f16 MLP, and with the FAST math policy (the other
        # policy than --math)
        def frame_ms(opts):
            hm = hush_amd.Hush(0, dict({"width": W, "height": H}, **opts))
            pos[0] = 0
            clock_ramp(hm)
            for i in range(args.warmup):
                hm.encode(dev[next_frame(hm)], stream=stream.handle)
            stream.synchronize()
            e0.record(stream.handle)
            for i in range(args.steps):
                hm.encode(dev[next_frame(hm)], stream=stream.handle)
            e1.record(stream.handle)
            e1.synchronize()
            hm.destroy()
            return round(e0.elapsed_ms(e1) / args.steps, 5)

        extras["fp16_mfma_ms_per_frame"] = frame_ms({"precision": "fp16-mfma", "ieee_math": args.math == "ieee"})
        # the reference's f16 option (binary16 arithmetic in the MLP, hush.js:690)
        extras["f16_ms_per_frame"] = frame_ms({"precision": "f16", "ieee_math": args.math == "ieee"})
        other = "fast" if args.math == "ieee" else "ieee"
        extras[f"{other}_math_ms_per_frame"] = frame_ms({"precision": args.precision, "ieee_math": other == "ieee"})
        extras[f"{other}_math_fp16_mfma_ms_per_frame"] = frame_ms({"precision": "fp16-mfma", "ieee_math": other == "ieee"})
        extras["svgf_1080p"] = svgf_leg(dev, stream, hip, W, H, min(args.steps, 64), args.warmup)
        extras["post_1080p"] = post_leg(dev, stream, hip, W, H, 100)
        print(f"[bench] post {extras['post_1080p']}", file=sys.stderr, flush=True)
        extras["train_fp32"] = train_leg(dev, W, H)
        extras["scene_generator_1080p"] = scene_leg(W, H)
        print(f"[bench] train {extras['train_fp32']}", file=sys.stderr, flush=True)
        extras["c0_640x360_spheres"] = c0_leg(hip, stream)
        print(f"[bench] 640x360 {extras['c0_640x360_spheres']}", file=sys.stderr, flush=True)
        extras["uhd_2160p"] = uhd_leg(hip, stream)
        print(f"[bench] 2160p {extras['uhd_2160p']}", file=sys.stderr, flush=True)
        extras["error_vs_fp64_oracle_640x360"] = psnr_check(["fp32", "fp16-mfma", "bf16-mfma", "fp8-mfma", "fp32+compensated",
                                                              "fp16-mfma+compensated"])
        # the same at 3840x2160 on a bounded sample (the first 8 frames of the 2160p sequence's scene type):
        # the FP64 oracle takes ~2 s per 2160p frame and mode on the host cores
        t0 = time.perf_counter()
        extras["error_vs_fp64_oracle_3840x2160"] = psnr_check(["fp32", "fp16-mfma", "bf16-mfma", "fp8-mfma"], 3840, 2160,
                                                               frames=args.uhd_error_frames, seed=2)
        print(f"[bench] 2160p error check {time.perf_counter() - t0:.0f}s", file=sys.stderr, flush=True)
        print(f"[bench] extras {extras}", file=sys.stderr, flush=True)

    if args.copy_probe_mb:
        nb = args.copy_probe_mb << 20
        src, dst = hip.DeviceBuffer(nb), hip.DeviceBuffer(nb)
        for _ in range(4):
            hip.check(hip.lib().hipMemcpy(hip.ctypes.c_void_p(dst.ptr), hip.ctypes.c_void_p(src.ptr), nb,
                                          hip.hipMemcpyDeviceToDevice), "copy probe")
        hip.synchronize()
        print(f"[bench] copy probe: 4 x {nb} bytes device-to-device", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(frames_g, W, H, args.cpu_frames)
        print(f"[bench] cpu baseline {cpu['value']} ms/frame", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC, "value": round(ms_frame, 5), "unit": "ms/frame", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_frame, 5), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{W}x{H} boxes scene, {F}-frame sequence, {args.precision} MLP, built-in weights",
                   "width": W, "height": H, "frames": F, "mlp": args.precision, "math": args.math},
        "frame_GB/s_algorithmic": round(total_bytes / (ms_frame * 1e-3) / 1e9, 1),
        "frame_hbm_frac": round(total_bytes / (ms_frame * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "clock_ramp": ramp, "passes": passes, "roofline": roofline, "cpu_baseline": cpu, **extras,
    }
    print(json.dumps(line), flush=True)
    hs.destroy()


if __name__ == "__main__":
    main()
