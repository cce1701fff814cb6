"""Frames whose motion vectors overflowed binary16 (+-inf) or are NaN: the reprojection must treat them as
having no history, with defined index arithmetic (the reference's WGSL saturates f32 -> i32 and wraps
i32 addition)."""
import numpy as np

PINF, NINF, QNAN = 0x7C00, 0xFC00, 0x7E00


def inject(g: dict, frame: int) -> dict:
    """copy of the frame with non-finite motion.xy in three blocks (+inf, -inf, NaN) and one column of
    +inf in x only"""
    g = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in g.items()}
    mv = g["motion"]
    h, w = mv.shape[:2]
    if frame == 0:
        return g
    mv[h // 4:h // 4 + 5, w // 4:w // 4 + 7, 0:2] = PINF
    mv[h // 2:h // 2 + 4, w // 2:w // 2 + 6, 0:2] = NINF
    mv[h - 6:h - 3, 2:5, 0:2] = QNAN
    mv[:, w - 3, 0] = PINF
    mv[1:3, 1:4, 1] = NINF
    return g


def nonfinite_mask(g: dict) -> np.ndarray:
    mv = g["motion"].view(np.float16)[..., :2].astype(np.float32)
    return ~np.isfinite(mv).all(-1)


def inject_radiance(g: dict, frame: int, kind: str = "inf") -> dict:
    """copy of the frame with binary16 +inf radiance (a firefly that overflowed the renderer's binary16
    store): a few isolated texels, one 3x3 block, and one texel with +inf in a single channel.
    kind "nan": NaN radiance at those texels instead, plus +inf and zero albedo at a few others."""
    g = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in g.items()}
    rad = g["radiance"]
    h, w = rad.shape[:2]
    bad = PINF if kind == "inf" else QNAN
    for k in range(6):
        rad[(7 * k + 3 * frame) % h, (11 * k + 5 * frame) % w, 0:3] = bad
    rad[h // 2:h // 2 + 3, w // 3:w // 3 + 3, 0:3] = bad
    rad[h // 3, 2 * w // 3, 1] = bad
    if kind == "nan":
        mat = g["material"]
        mat[h // 4, w // 5:w // 5 + 4, 0:3] = PINF
        mat[3 * h // 4, w // 2:w // 2 + 4, 0:3] = 0
    return g
