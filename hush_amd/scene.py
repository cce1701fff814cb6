"""Seeded procedural G-buffer sequences (ctypes over libhush_scene.so, hush_amd/csrc/scene.c).

Kinds: "spheres" (16 spheres, orbiting camera), "boxes" (200 boxes, 40 moving, panning camera),
"train" (300 primitives, 30% moving, camera cut every 48 frames).  Frames are rgba16float arrays of shape
(H, W, 4) in the layout of src/index.d.ts:2-18.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import SCENE_LIB_PATH

KINDS = {"spheres": 0, "boxes": 1, "train": 2}
_L = None


def _load():
    global _L
    if _L is None:
        if not os.path.exists(SCENE_LIB_PATH):
            raise ImportError(f"{SCENE_LIB_PATH} is not built (python hush_amd/build.py)")
        L = ctypes.CDLL(SCENE_LIB_PATH)
        L.hs_create.restype = ctypes.c_void_p
        L.hs_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong]
        L.hs_destroy.argtypes = [ctypes.c_void_p]
        L.hs_render.restype = ctypes.c_int
        L.hs_render.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + [ctypes.c_void_p] * 6
        L.hs_is_cut.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.hs_num_prims.argtypes = [ctypes.c_void_p]
        _L = L
    return _L


class SceneGenerator:
    def __init__(self, kind: str, width: int, height: int, seed: int = 0):
        self._L = _load()
        self.kind, self.width, self.height, self.seed = kind, int(width), int(height), int(seed)
        self._s = self._L.hs_create(KINDS[kind], self.width, self.height, self.seed)
        if not self._s:
            raise ValueError(f"bad scene {kind} {width}x{height}")

    def is_cut(self, frame: int) -> bool:
        return bool(self._L.hs_is_cut(self._s, int(frame)))

    def render(self, frame: int, crop=None, clean: bool = False) -> dict:
        """One frame: {'radiance', 'normalDepth', 'motion', 'material', 'emission'} as (h, w, 4) uint16 (binary16
        bit patterns); with clean=True also 'clean' (h, w, 4) float32 noise-free expected radiance."""
        x0, y0, w, h = crop if crop is not None else (0, 0, self.width, self.height)
        out = {k: np.empty((h, w, 4), np.uint16) for k in ("radiance", "normalDepth", "motion", "material", "emission")}
        cl = np.empty((h, w, 4), np.float32) if clean else None
        r = self._L.hs_render(self._s, int(frame), x0, y0, w, h, *[out[k].ctypes.data for k in
                              ("radiance", "normalDepth", "motion", "material", "emission")],
                              cl.ctypes.data if clean else None)
        if r != 0:
            raise ValueError("hs_render failed (crop outside the frame?)")
        if clean:
            out["clean"] = cl
        return out

    def close(self):
        if self._s:
            self._L.hs_destroy(self._s)
            self._s = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
