"""Helpers shared by the GPU parity tests: teacher forcing (GPU state -> oracle) and comparisons."""
from __future__ import annotations

import numpy as np

from hush_amd.hip import Surface

COMMON = ["A0", "X", "Bh0", "Bh1", "mom0", "mom1", "key0", "key1", "hist0", "hist1", "taa0", "taa1", "par0_0",
          "par0_1", "par1", "par2", "C0", "C1", "CG"]
INPUT_KEYS = ("radiance", "normalDepth", "motion", "material", "emission")


def upload_frame(g: dict, surfs: dict | None = None) -> dict:
    """binary16 input arrays -> device surfaces (reusing `surfs` when given)."""
    h, w = g["radiance"].shape[:2]
    surfs = surfs or {k: Surface(w, h) for k in INPUT_KEYS}
    for k in INPUT_KEYS:
        surfs[k].upload(g[k].view(np.float16))
    return surfs


def gpu_surface(h, name) -> np.ndarray:
    """a GPU surface as float32 (F planes as (H2, W2, 8) binary16 values)"""
    a = h.surface(name).download()
    if a.dtype == np.uint32:
        return a.view(np.float16).astype(np.float32)
    return a.astype(np.float32)


def gpu_F(h) -> np.ndarray:
    return np.concatenate([gpu_surface(h, f"F{k}") for k in range(4)], axis=-1)


def force_oracle(h, o) -> None:
    """copy every GPU surface into the oracle (teacher forcing)"""
    for n in COMMON:
        o.surface(n)[...] = gpu_surface(h, n)
    o.surface("F")[...] = gpu_F(h)


def pass_outputs(p: int, cur: int) -> list:
    return {0: [f"A0", f"mom{cur}", f"key{cur}", "F"], 1: [f"par0_{cur}", "par1", "par2", "Bh0"], 2: ["Bh1"],
            3: ["Bh0"], 4: ["C1", "CG"], 5: ["C0"], 6: [f"hist{cur}", "X"], 7: [f"taa{cur}"]}[p]


def log_parity(case: str, stats: list) -> None:
    """appends per-frame comparison statistics to the JSON-lines file HUSH_PARITY_LOG names (if set)"""
    import json
    import os
    path = os.environ.get("HUSH_PARITY_LOG")
    if path:
        with open(path, "a") as fh:
            fh.write(json.dumps({"case": case, "stats": stats}) + "\n")


def ulp16(x: np.ndarray) -> np.ndarray:
    """spacing of binary16 at |x| (normal range), 2^-24 below it"""
    ax = np.abs(x.astype(np.float64))
    e = np.floor(np.log2(np.maximum(ax, 2.0 ** -14)))
    return np.where(ax < 2.0 ** -14, 2.0 ** -24, 2.0 ** (e - 10))


def compare(gpu: np.ndarray, ref: np.ndarray) -> dict:
    """exact-match fraction, max difference in binary16 ULPs of the reference, max relative error
    |g - r| / max(|r|, 1e-3)"""
    g = gpu.astype(np.float64)
    r = ref.astype(np.float64)
    d = np.abs(g - r)
    ulps = d / ulp16(r)
    rel = d / np.maximum(np.abs(r), 1e-3)
    return {"n": int(d.size), "exact": float(np.mean(d == 0)), "max_ulp": float(ulps.max(initial=0)),
            "p9999_ulp": float(np.quantile(ulps, 0.9999)) if d.size else 0.0, "max_rel": float(rel.max(initial=0)),
            "frac_rel_gt_1e-3": float(np.mean(rel > 1e-3)), "frac_rel_gt_1e-2": float(np.mean(rel > 1e-2)),
            "frac_ulp_gt_1": float(np.mean(ulps > 1.0))}


def tonemap(c: np.ndarray) -> np.ndarray:
    """c / (1 + lum(c)) (the TAA tone map, hush.js:635): maps HDR into [0, 1) for PSNR"""
    c = np.maximum(c[..., :3].astype(np.float64), 0.0)
    l = c[..., 0] * 0.2126 + c[..., 1] * 0.7152 + c[..., 2] * 0.0722
    return c / (1.0 + l)[..., None]


def psnr(a: np.ndarray, ref: np.ndarray) -> float:
    """PSNR (dB) of the tone-mapped images, peak 1"""
    mse = float(np.mean((tonemap(a) - tonemap(ref)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(1.0 / mse)
