"""Static input scales of the fp8 (E4M3) MLP form.

The fp8 form stores each layer's input x as the E4M3 value of x / s, with s a power of two per layer, so that
the kernels need no max-reductions.  E4M3 keeps 3 mantissa bits over [2^-6, 448] (subnormals down to 2^-9), so
s should bring the largest |x| a layer sees just under 448: larger s pushes small activations into the
subnormal range, smaller s saturates.

`calibrate` derives s from feature rows of real frames (a float64 forward of the network, the largest |input|
of each layer times a headroom factor).  `python -m hush_amd.quant OUT.json` renders the procedural scenes
through the denoiser on the GPU, collects the half-res feature rows of every non-sky texel and writes the
scales for the built-in weights; the result ships as `data/fp8_scales.json` and `Hush` loads it for those
weights unless `fp8_scales` is given.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SCALES_PATH = os.path.join(HERE, "data", "fp8_scales.json")
E4M3_MAX = 448.0


def pow2_at_least(m: float) -> float:
    """the smallest power of two s with m / s <= 448 (1 for m == 0)"""
    if not m > 0.0:
        return 1.0
    f, e = np.frexp(m / E4M3_MAX)            # m / 448 = f 2^e, f in [0.5, 1)
    return float(2.0 ** (int(e) - 1 if f == 0.5 else int(e)))


def layer_input_maxima(weights16: np.ndarray, rows: np.ndarray) -> list:
    """largest |input| of each of the three layers of the 20-16-16-12 network over the feature rows
    (n x 20), in float64"""
    w = np.asarray(weights16, np.uint16).view(np.float16).astype(np.float64)
    W1, B1 = w[:320].reshape(16, 20), w[320:336]
    W2, B2 = w[336:592].reshape(16, 16), w[592:608]
    x = np.asarray(rows).astype(np.float64)
    h1 = np.maximum(x @ W1.T + B1, 0.0)
    h2 = np.maximum(h1 @ W2.T + B2, 0.0)
    return [float(np.abs(x).max(initial=0.0)), float(h1.max(initial=0.0)), float(h2.max(initial=0.0))]


def calibrate(weights16: np.ndarray, rows: np.ndarray, headroom: float = 2.0) -> list:
    """per-layer input scales (powers of two) from feature rows: each layer's largest |input| x headroom"""
    return [pow2_at_least(m * headroom) for m in layer_input_maxima(weights16, rows)]


def weights_digest(weights16: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(weights16, np.uint16).tobytes()).hexdigest()


def load_scales(weights16: np.ndarray, path: str = SCALES_PATH):
    """the calibrated scales stored for these weights, or None"""
    try:
        with open(path) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    if d.get("weights_sha256") != weights_digest(weights16):
        return None
    s = [float(v) for v in d.get("scales", [])]
    return s if len(s) == 3 else None


def collect_rows(scenes=("boxes", "spheres", "train"), width=640, height=360, frames=6, seed=7):
    """half-res feature rows (F0, F1, F2.xy) of every non-sky texel of the procedural scenes, rendered and
    denoised on the GPU"""
    from . import Hush, hip
    from .scene import SceneGenerator
    keys = ("radiance", "normalDepth", "motion", "material", "emission")
    out = []
    for name in scenes:
        gen = SceneGenerator(name, width, height, seed)
        hs = Hush(0, {"width": width, "height": height})
        surfs = {k: hip.Surface(width, height) for k in keys}
        for f in range(frames):
            g = gen.render(f)
            for k in keys:
                surfs[k].upload(g[k].view(np.float16))
            hs.denoise(surfs)
            F = np.concatenate([hs.surface(f"F{k}").download().view(np.float16) for k in range(3)], -1)
            Z = hs.surface("F3").download().view(np.float16)[..., 2]
            out.append(F.reshape(-1, 24)[:, :20][Z.reshape(-1) >= 0])
        hs.destroy()
    return np.concatenate(out)


def main(argv=None) -> int:
    import argparse
    from . import default_weights
    ap = argparse.ArgumentParser(description="calibrate the fp8 input scales of the built-in weights")
    ap.add_argument("out", help="JSON file to write")
    ap.add_argument("--headroom", type=float, default=2.0)
    args = ap.parse_args(argv)
    w = default_weights()
    rows = collect_rows()
    maxima = layer_input_maxima(w, rows)
    scales = calibrate(w, rows, args.headroom)
    rec = {"weights_sha256": weights_digest(w), "scales": scales, "layer_input_max": maxima,
           "headroom": args.headroom, "rows": int(rows.shape[0]),
           "source": "procedural scenes boxes, spheres, train at 640x360, 6 frames each, seed 7"}
    with open(args.out, "w") as fh:
        json.dump(rec, fh, indent=1)
    print(json.dumps(rec))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
