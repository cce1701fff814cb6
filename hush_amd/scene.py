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
        L.hs_seed.argtypes = [ctypes.c_void_p]; L.hs_seed.restype = ctypes.c_ulonglong
        L.hs_frame_setup.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + [ctypes.c_void_p] * 4 + [ctypes.c_int]
        L.hs_frame_setup.restype = ctypes.c_int
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


class GpuSceneGenerator:
    """The same seeded scenes rendered on the GPU (include/hush_scene.h): the host computes each frame's
    set-up (cameras, primitive motion, per-tile primitive lists) and the GPU draws every pixel into device
    surfaces.  `render` fills `out` (dict of the five input names -> hush_amd.hip.Surface or (h, w, 4)
    float16 torch tensors) and optionally `clean` ((h, w, 4) float32 torch tensor or DeviceBuffer)."""

    KEYS = ("radiance", "normalDepth", "motion", "material", "emission")

    def __init__(self, kind: str, width: int, height: int, seed: int = 0, device: int = 0):
        from . import _lib
        self.cpu = SceneGenerator(kind, width, height, seed)
        self.width, self.height, self.kind = int(width), int(height), kind
        self._L = _lib.load()
        self._lib = _lib
        p = ctypes.c_void_p()
        r = self._L.hush_scene_create(int(device), ctypes.byref(p))
        if r != 0:
            raise RuntimeError(self._L.hush_scene_last_error(None).decode())
        self._g = p.value
        self._n = self.cpu._L.hs_num_prims(self.cpu._s)
        self._seed = self.cpu._L.hs_seed(self.cpu._s)

    def is_cut(self, frame: int) -> bool:
        return self.cpu.is_cut(frame)

    def set_seed(self, seed: int) -> None:
        """Switches to the scene of another seed (same kind and size), keeping the GPU renderer."""
        old = self.cpu
        self.cpu = SceneGenerator(self.kind, self.width, self.height, seed)
        old.close()
        self._n = self.cpu._L.hs_num_prims(self.cpu._s)
        self._seed = self.cpu._L.hs_seed(self.cpu._s)

    def setup(self, frame: int, crop=None):
        """the host half: (cams, prims, off, lst) numpy arrays of hs_frame_setup"""
        x0, y0, w, h = crop if crop is not None else (0, 0, self.width, self.height)
        nt = ((w + 15) // 16) * ((h + 15) // 16)
        cams, prims = np.zeros(28), np.zeros(self._n * 20)
        off = np.zeros(nt + 1, np.int32)
        cap = max(1024, nt * 8)
        while True:
            lst = np.zeros(cap, np.int32)
            n = self.cpu._L.hs_frame_setup(self.cpu._s, int(frame), x0, y0, w, h, cams.ctypes.data, prims.ctypes.data,
                                           off.ctypes.data, lst.ctypes.data, cap)
            if n < 0:
                raise ValueError("crop outside the frame")
            if n <= cap:
                return cams, prims, off, lst[:n].copy()
            cap = n

    def render(self, frame: int, out: dict, crop=None, clean=None, stream: int | None = None) -> None:
        from . import _surface_struct
        x0, y0, w, h = crop if crop is not None else (0, 0, self.width, self.height)
        cams, prims, off, lst = self.setup(frame, crop)
        fr = self._lib.hush_scene_frame()
        fr.cams, fr.prims, fr.off = cams.ctypes.data, prims.ctypes.data, off.ctypes.data
        fr.lst = lst.ctypes.data if lst.size else None
        fr.n_prims, fr.n_list, fr.seed, fr.frame = self._n, int(lst.size), self._seed, int(frame)
        fr.width, fr.height, fr.x0, fr.y0, fr.w, fr.h = self.width, self.height, x0, y0, w, h
        ins = self._lib.hush_inputs()
        for k, f in zip(self.KEYS, ("radiance", "normal_depth", "motion", "material", "emission")):
            setattr(ins, f, _surface_struct(out[k]))
        cptr, cpitch = None, 0
        if clean is not None:
            if hasattr(clean, "data_ptr"):
                cptr, cpitch = clean.data_ptr(), clean.stride(0) * clean.element_size()
            else:
                cptr, cpitch = clean.ptr, w * 16
        if stream is None:
            try:
                import torch
                stream = torch.cuda.current_stream().cuda_stream
            except Exception:
                stream = 0
        r = self._L.hush_scene_render(ctypes.c_void_p(self._g), ctypes.byref(fr), ctypes.byref(ins),
                                      ctypes.c_void_p(cptr) if cptr else None, cpitch, ctypes.c_void_p(stream))
        if r != 0:
            raise RuntimeError(self._L.hush_scene_last_error(ctypes.c_void_p(self._g)).decode())

    def close(self):
        if getattr(self, "_g", None):
            self._L.hush_scene_destroy(self._g)
            self._g = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
