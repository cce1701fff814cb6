"""Training the Hush MLP on MI355X: the differentiable frame of include/hush_train.h as torch.autograd
functions, and a small Adam trainer over seeded procedural sequences with clean targets.

    frame = DiffFrame(hush)                       # attaches to a Hush context (its size and history)
    frame.prepare(inputs)                         # temporal pass + coarse level of this frame
    out = frame_apply(weights, frame)             # torch (H, W, 4); differentiable in `weights` (812)
    loss = loss_apply(out, target, prev_out, prev_target, lam, frame)   # FP64-reduced, differentiable in out
    loss.backward(); frame.commit(out)            # history for the next frame

The trained weights keep the layout of hush-weights.bin (W1 b1 W2 b2 W3 b3, hush.js:20-21);
`save_weights` writes them as 812 binary16 values that the reference's Hush (and this one) load unchanged.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from . import Hush, HushError, default_weights

N_WEIGHTS = 812


def _load():
    L = _lib.load()
    if getattr(L, "_train_declared", False):
        return L
    P, i32 = ctypes.c_void_p, ctypes.c_int32
    L.hush_train_create.argtypes = [P, i32, ctypes.POINTER(P)]; L.hush_train_create.restype = i32
    L.hush_train_destroy.argtypes = [P]; L.hush_train_destroy.restype = None
    L.hush_train_last_error.argtypes = [P]; L.hush_train_last_error.restype = ctypes.c_char_p
    L.hush_train_prepare.argtypes = [P, ctypes.POINTER(_lib.hush_inputs), P]; L.hush_train_prepare.restype = i32
    L.hush_train_forward.argtypes = [P, P, P, P]; L.hush_train_forward.restype = i32
    L.hush_train_loss.argtypes = [P, P, P, P, P, ctypes.c_double, P, P, P]; L.hush_train_loss.restype = i32
    L.hush_train_backward.argtypes = [P, P, P, P, P, P]; L.hush_train_backward.restype = i32
    L.hush_train_commit.argtypes = [P, P, P]; L.hush_train_commit.restype = i32
    L.hush_train_a0.argtypes = [P, ctypes.POINTER(P)]; L.hush_train_a0.restype = i32
    L.hush_train_set_a0_fixed.argtypes = [P, i32]; L.hush_train_set_a0_fixed.restype = i32
    L.hush_train_set_forward_precision.argtypes = [P, i32]; L.hush_train_set_forward_precision.restype = i32
    L.hush_train_mlp_out.argtypes = [P, ctypes.POINTER(P)]; L.hush_train_mlp_out.restype = i32
    L._train_declared = True
    return L


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class DiffFrame:
    """The differentiable frame attached to a Hush context; fp64=True for gradient checks.  `forward`
    selects the MLP arithmetic of the forward ("fp32", or "fp16-mfma" / "bf16-mfma" in an FP32 context:
    mixed-precision training with an FP32 backward)."""

    def __init__(self, hush: Hush, fp64: bool = False, forward: str = "fp32"):
        import torch
        self._L = _load()
        self.hush = hush
        self.fp64 = bool(fp64)
        self.dtype = torch.float64 if fp64 else torch.float32
        p = ctypes.c_void_p()
        r = self._L.hush_train_create(ctypes.c_void_p(hush._h), int(self.fp64), ctypes.byref(p))
        if r != 0:
            raise HushError(self._L.hush_train_last_error(None).decode(), r)
        self._t = p.value
        self.width, self.height = hush.width, hush.height
        self._keep = None
        if forward != "fp32":
            from . import PRECISIONS
            self._check(self._L.hush_train_set_forward_precision(self._t, PRECISIONS[forward]))

    def _check(self, r):
        if r != 0:
            raise HushError(self._L.hush_train_last_error(ctypes.c_void_p(self._t)).decode(), r)

    def prepare(self, inputs) -> None:
        ins = self.hush._inputs(inputs)
        self._keep = inputs                 # the surfaces stay referenced until the next prepare
        self._check(self._L.hush_train_prepare(self._t, ctypes.byref(ins), _stream()))

    def forward(self, weights):
        import torch
        assert weights.dtype == self.dtype and weights.is_cuda and weights.numel() == N_WEIGHTS
        out = torch.empty((self.height, self.width, 4), dtype=self.dtype, device=weights.device)
        self._check(self._L.hush_train_forward(self._t, _ptr(weights.contiguous()), _ptr(out), _stream()))
        return out

    def backward(self, weights, grad_out, want_a0: bool = False):
        import torch
        gw = torch.empty(N_WEIGHTS, dtype=self.dtype, device=weights.device)
        ga0 = torch.empty((self.height, self.width, 4), dtype=self.dtype, device=weights.device) if want_a0 else None
        self._check(self._L.hush_train_backward(self._t, _ptr(weights.contiguous()), _ptr(grad_out.contiguous()), _ptr(gw),
                                                _ptr(ga0), _stream()))
        return (gw, ga0) if want_a0 else gw

    def loss(self, out, target, prev_out=None, prev_target=None, lam: float = 0.0):
        """(loss as a 0-d float64 tensor, dloss/dout)"""
        import torch
        g = torch.empty_like(out)
        l = torch.empty((), dtype=torch.float64, device=out.device)
        self._check(self._L.hush_train_loss(self._t, _ptr(out.contiguous()), _ptr(target.contiguous()),
                                            _ptr(prev_out), _ptr(prev_target), float(lam), _ptr(g), _ptr(l), _stream()))
        return l, g

    def commit(self, out) -> None:
        self._check(self._L.hush_train_commit(self._t, _ptr(out.contiguous()), _stream()))

    def set_a0_fixed(self, fixed: bool) -> None:
        """True: forwards take A0 as set (set_a0) instead of re-blending it from the temporal history"""
        self._check(self._L.hush_train_set_a0_fixed(self._t, int(bool(fixed))))

    def a0_copy(self):
        """the prepared A0 as a tensor copy"""
        import torch
        p = ctypes.c_void_p()
        self._check(self._L.hush_train_a0(self._t, ctypes.byref(p)))
        t = torch.empty((self.height, self.width, 4), dtype=self.dtype, device="cuda")
        _memcpy_d2d(t.data_ptr(), p.value, t.numel() * t.element_size())
        return t

    def mlp_outputs(self):
        """the last forward's MLP outputs before the activations, (H/2, W/2, 12) numpy copy"""
        import torch
        p = ctypes.c_void_p()
        self._check(self._L.hush_train_mlp_out(self._t, ctypes.byref(p)))
        h2, w2 = (self.height + 1) // 2, (self.width + 1) // 2
        t = torch.empty((h2, w2, 12), dtype=self.dtype, device="cuda")
        _memcpy_d2d(t.data_ptr(), p.value, t.numel() * t.element_size())
        return t.cpu().numpy()

    def set_a0(self, a0) -> None:
        p = ctypes.c_void_p()
        self._check(self._L.hush_train_a0(self._t, ctypes.byref(p)))
        a0 = a0.contiguous()
        _memcpy_d2d(p.value, a0.data_ptr(), a0.numel() * a0.element_size())

    def destroy(self):
        if getattr(self, "_t", None):
            import torch
            torch.cuda.synchronize()
            self._L.hush_train_destroy(self._t)
            self._t = None

    def __del__(self):  # pragma: no cover
        try:
            self.destroy()
        except Exception:
            pass


def _memcpy_d2d(dst: int, src: int, n: int) -> None:
    import torch
    from .hip import check, lib
    torch.cuda.synchronize()
    check(lib().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(n), 3), "hipMemcpy D2D")


def _autograd():
    import torch

    class FrameFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, weights, frame):
            ctx.frame = frame
            ctx.save_for_backward(weights)
            return frame.forward(weights)

        @staticmethod
        def backward(ctx, grad_out):
            (weights,) = ctx.saved_tensors
            return ctx.frame.backward(weights, grad_out), None

    class LossFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, out, target, prev_out, prev_target, lam, frame):
            l, g = frame.loss(out, target, prev_out, prev_target, lam)
            ctx.save_for_backward(g)
            return l

        @staticmethod
        def backward(ctx, gl):
            (g,) = ctx.saved_tensors
            return g * gl.to(g.dtype), None, None, None, None, None

    return FrameFn, LossFn


_FNS = None


def frame_apply(weights, frame: DiffFrame):
    global _FNS
    _FNS = _FNS or _autograd()
    return _FNS[0].apply(weights, frame)


def loss_apply(out, target, prev_out, prev_target, lam: float, frame: DiffFrame):
    global _FNS
    _FNS = _FNS or _autograd()
    return _FNS[1].apply(out, target, prev_out, prev_target, lam, frame)


# ------------------------------------------------------------------------------------------------ weights
def weights_to_float(w16: np.ndarray) -> np.ndarray:
    return np.asarray(w16, np.uint16).view(np.float16).astype(np.float64)


def random_weights(seed: int, hidden: int = 16) -> np.ndarray:
    """seeded He-normal init at 20-16-16-12, biases zero, in the hush-weights.bin layout (float64)"""
    rng = np.random.default_rng(seed)
    h = hidden
    w1 = rng.normal(0, np.sqrt(2.0 / 20), (h, 20))
    w2 = rng.normal(0, np.sqrt(2.0 / h), (h, h))
    w3 = rng.normal(0, np.sqrt(1.0 / h), (12, h))
    return np.concatenate([w1.ravel(), np.zeros(h), w2.ravel(), np.zeros(h), w3.ravel(), np.zeros(12)])


def to_f16_blob(w: np.ndarray) -> np.ndarray:
    """812 binary16 values (round to nearest) in the hush-weights.bin layout"""
    return np.asarray(w, np.float64).astype(np.float16).view(np.uint16)


def save_weights(w, path: str) -> None:
    arr = w.detach().cpu().numpy() if hasattr(w, "detach") else np.asarray(w)
    to_f16_blob(arr).tofile(path)


# ------------------------------------------------------------------------------------------------ trainer
def upload(g: dict, keys=("radiance", "normalDepth", "motion", "material", "emission")):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(g[k]).view(np.float16)).cuda() for k in keys}


def target_of(g: dict, dtype):
    import torch
    return torch.from_numpy(g["clean"]).to(device="cuda", dtype=dtype)


def train(steps: int = 100, sequences: int = 2, frames: int = 8, crop=(256, 256), scene_size=(640, 360), lr: float = 1e-3,
          lam: float = 0.5, seed: int = 0, init: str = "random", log=print, gpu_scenes: bool = True, forward: str = "fp32",
          ckpt: str | None = None, ckpt_every: int = 0, resume: str | None = None, time_limit: float | None = None):
    """Adam over seeded procedural sequences ('train' scenes: 300 primitives, 30% moving, camera cuts every 48
    frames, which reset the history).  Each step: `sequences` crops of `scene_size` frames x `frames` frames;
    the gradient of every frame accumulates (the history enters each frame as an input).  `forward` selects
    the MLP arithmetic of the forward ("fp32", "fp16-mfma", "bf16-mfma"; the backward is FP32).  One denoiser
    and training context per sequence slot is reused across steps (reset at every sequence start); with
    gpu_scenes the crops and their clean targets are rendered on the GPU straight into device tensors.
    `ckpt` (+ `ckpt_every`) saves weights, Adam state and step; `resume` continues from such a file;
    `time_limit` (seconds) stops early after a whole step.  Returns (weights tensor, [(step, loss)])."""
    import torch
    w0 = random_weights(seed) if init == "random" else weights_to_float(default_weights())
    weights = torch.tensor(w0, dtype=torch.float32, device="cuda", requires_grad=True)
    opt = torch.optim.Adam([weights], lr=lr)
    hist, start = [], 0
    if resume:
        st = torch.load(resume, map_location="cuda", weights_only=True)
        with torch.no_grad():
            weights.copy_(st["weights"])
        opt.load_state_dict(st["opt"])
        hist = [tuple(x) for x in st["hist"].tolist()]
        start = int(st["step"]) + 1

    def save(step):
        torch.save({"weights": weights.detach(), "opt": opt.state_dict(), "step": step,
                    "hist": torch.tensor(hist, dtype=torch.float64)}, ckpt)

    slots = make_slots(sequences, crop, scene_size, forward, gpu_scenes)
    t0 = time.perf_counter()
    for step in range(start, steps):
        opt.zero_grad(set_to_none=False)
        total, _ = run_batch(weights, slots, seed, step, frames, crop, scene_size, lam, gpu_scenes)
        opt.step()
        hist.append((step, total))
        last = step == steps - 1
        if log and (step % max(1, steps // 20) == 0 or last):
            log(f"[train] step {step} loss {hist[-1][1]:.5f} ({time.perf_counter() - t0:.1f}s)")
        if ckpt and ((ckpt_every and step % ckpt_every == ckpt_every - 1) or last):
            save(step)
        if time_limit is not None and time.perf_counter() - t0 > time_limit and not last:
            if log:
                log(f"[train] time limit: stopped after step {step}")
            if ckpt:
                save(step)
            break
    close_slots(slots)
    return weights.detach(), hist


def make_slots(sequences: int, crop, scene_size, forward: str = "fp32", gpu_scenes: bool = True):
    """one denoiser + training context (and GPU scene generator and input tensors) per sequence slot"""
    import torch
    from .scene import GpuSceneGenerator
    cw, ch = crop
    slots = []
    for _ in range(sequences):
        hs = Hush(0, {"width": cw, "height": ch})      # temporal / coarse passes only: no MLP weights used
        fr = DiffFrame(hs, forward=forward)
        gen = GpuSceneGenerator("train", scene_size[0], scene_size[1], 0) if gpu_scenes else None
        ins = {k: torch.empty((ch, cw, 4), dtype=torch.float16, device="cuda") for k in GpuSceneGenerator.KEYS}
        tgt = [torch.empty((ch, cw, 4), dtype=torch.float32, device="cuda") for _ in range(2)]
        slots.append((hs, fr, gen, ins, tgt))
    return slots


def close_slots(slots) -> None:
    for hs, fr, gen, _, _ in slots:
        fr.destroy()
        hs.destroy()
        if gen is not None:
            gen.close()


def run_batch(weights, slots, seed: int, step: int, frames: int, crop, scene_size, lam: float = 0.5,
              gpu_scenes: bool = True, backward: bool = True):
    """One step's batch: for each slot a seeded crop of a 'train' sequence (scene seed, crop origin and first
    frame drawn from (seed, step, slot)), `frames` frames from a reset history, each frame's loss (L1 + lam x
    temporal term) accumulated into weights.grad / (slots x frames) when `backward`.  Returns (mean loss,
    number of camera cuts inside the batch, each of which reset the history)."""
    from .scene import SceneGenerator
    cw, ch = crop
    total, cuts = 0.0, 0
    for s, (hs, fr, gen, ins, tgts) in enumerate(slots):
        rs = np.random.default_rng([seed, step, s])
        gseed = int(rs.integers(1 << 30))
        if gpu_scenes:
            gen.set_seed(gseed)
        else:
            gen = SceneGenerator("train", scene_size[0], scene_size[1], gseed)
        x0 = int(rs.integers(0, scene_size[0] - cw + 1))
        y0 = int(rs.integers(0, scene_size[1] - ch + 1))
        f0 = int(rs.integers(0, 96))
        hs.reset()
        prev_out = prev_tgt = None
        for f in range(frames):
            if gpu_scenes:
                tgt = tgts[f & 1]
                gen.render(f0 + f, ins, crop=(x0, y0, cw, ch), clean=tgt)
                fins = ins
            else:
                import torch
                g = gen.render(f0 + f, crop=(x0, y0, cw, ch), clean=True)
                fins, tgt = upload(g), target_of(g, torch.float32)
            if gen.is_cut(f0 + f):
                hs.reset()
                prev_out = prev_tgt = None
                cuts += f > 0
            fr.prepare(fins)
            out = frame_apply(weights, fr)
            loss = loss_apply(out, tgt, prev_out, prev_tgt, lam if prev_out is not None else 0.0, fr)
            if backward:
                (loss / (len(slots) * frames)).backward()
            total += float(loss.detach())
            fr.commit(out.detach())
            prev_out, prev_tgt = out.detach(), tgt
        if not gpu_scenes:
            gen.close()
    return total / (len(slots) * frames), cuts




# ------------------------------------------------------------------------------------------------ validation
def _tm(c):
    c = np.maximum(np.asarray(c, np.float64)[..., :3], 0.0)
    return c / (1.0 + (c[..., 0] * 0.2126 + c[..., 1] * 0.7152 + c[..., 2] * 0.0722))[..., None]


def evaluate(weight_sets: dict, seeds=(1001, 1002), size=(640, 360), frames=24, kind="train", warm=4):
    """PSNR (tone-mapped, peak 1) against the clean target and temporal flicker
    mean |(tm(out_t) - tm(out_t-1)) - (tm(clean_t) - tm(clean_t-1))| over held-out seeded sequences, for
    the Hush drop-in path with each weight set (name -> 812 binary16 values), the SVGF baseline and the
    noisy input; the first `warm` frames of each sequence are not scored (history warm-up)."""
    from . import SVGF
    from .hip import Surface
    from .scene import SceneGenerator
    W, H = size
    keys = ("radiance", "normalDepth", "motion", "material", "emission")
    acc = {n: {"mse": [], "flick": []} for n in list(weight_sets) + ["svgf", "noisy"]}
    for seed in seeds:
        gen = SceneGenerator(kind, W, H, seed)
        gs = [gen.render(f, clean=True) for f in range(frames)]
        surfs = [{k: Surface(W, H).upload(g[k].view(np.float16)) for k in keys} for g in gs]
        outs = {}
        for name, w16 in weight_sets.items():
            hs = Hush(0, {"width": W, "height": H, "weights": w16})
            outs[name] = []
            for f, s in enumerate(surfs):
                if gen.is_cut(f):
                    hs.reset()
                outs[name].append(hs.denoise(s).download().astype(np.float64))
            hs.destroy()
        sv = SVGF(0, W, H)
        outs["svgf"] = []
        for f, s in enumerate(surfs):
            if gen.is_cut(f):
                sv.reset()
            g = dict(s)
            p = surfs[f - 1] if f > 0 else s
            g["prevNormalDepth"], g["prevMotion"] = p["normalDepth"], p["motion"]
            outs["svgf"].append(sv.encode(g).download().astype(np.float64))
        sv.destroy()
        outs["noisy"] = [g["radiance"].view(np.float16).astype(np.float64) + g["emission"].view(np.float16).astype(np.float64)
                         for g in gs]
        for name, seq in outs.items():
            for f in range(warm, frames):
                t, r = _tm(seq[f]), _tm(gs[f]["clean"])
                acc[name]["mse"].append(float(np.mean((t - r) ** 2)))
                if not gen.is_cut(f):
                    dt = (t - _tm(seq[f - 1])) - (r - _tm(gs[f - 1]["clean"]))
                    acc[name]["flick"].append(float(np.mean(np.abs(dt))))
        for d in surfs:
            for v in d.values():
                v._buf.free()
    return {n: {"psnr_db": round(float(10 * np.log10(1.0 / np.mean(v["mse"]))), 3),
                "flicker": round(float(np.mean(v["flick"])), 6)} for n, v in acc.items()}


def main(argv=None):
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train the Hush MLP on seeded procedural sequences, then validate")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--sequences", type=int, default=4)
    ap.add_argument("--frames", type=int, default=8)
    ap.add_argument("--crop", type=int, nargs=2, default=(256, 256))
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", choices=["random", "builtin"], default="random")
    ap.add_argument("--out", default=None, help="write the trained weights (812 binary16, hush-weights.bin layout)")
    ap.add_argument("--eval-frames", type=int, default=24)
    ap.add_argument("--scene-size", type=int, nargs=2, default=(640, 360), help="size of the scenes the crops are taken from")
    ap.add_argument("--forward", choices=["fp32", "fp16-mfma", "bf16-mfma"], default="fp32",
                    help="MLP arithmetic of the forward (the backward is FP32)")
    ap.add_argument("--ckpt", default=None, help="checkpoint file (weights, Adam state, step, loss history)")
    ap.add_argument("--ckpt-every", type=int, default=0)
    ap.add_argument("--resume", default=None, help="continue from a checkpoint file")
    ap.add_argument("--time-limit", type=float, default=None, help="stop after the step that passes this many seconds")
    ap.add_argument("--no-eval", action="store_true", help="skip the validation (e.g. when the run will be resumed)")
    ap.add_argument("--eval-size", type=int, nargs=2, default=(640, 360), help="size of the held-out validation sequences")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    w, hist = train(args.steps, args.sequences, args.frames, tuple(args.crop), scene_size=tuple(args.scene_size), lr=args.lr,
                    seed=args.seed, init=args.init, log=log, forward=args.forward, ckpt=args.ckpt, ckpt_every=args.ckpt_every,
                    resume=args.resume, time_limit=args.time_limit)
    blob = to_f16_blob(w.cpu().numpy())
    if args.out:
        blob.tofile(args.out)
    res = None
    if not args.no_eval:
        sets = {"trained": blob, "builtin": default_weights(), "random_init": to_f16_blob(random_weights(args.seed))}
        res = evaluate(sets, size=tuple(args.eval_size), frames=args.eval_frames)
    print(json.dumps({"steps_done": hist[-1][0] + 1 if hist else 0, "steps": args.steps, "sequences": args.sequences,
                      "frames": args.frames, "crop": list(args.crop), "scene_size": list(args.scene_size),
                      "forward": args.forward, "loss_first": hist[0][1], "loss_last": hist[-1][1],
                      "validation": res}), flush=True)


if __name__ == "__main__":
    main()
