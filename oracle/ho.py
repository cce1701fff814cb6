"""ctypes wrapper of the CPU oracle (oracle/libhush_oracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py import this module, and only as the
checker; the denoiser itself never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhush_oracle.so")

SURF = {"A0": 0, "X": 1, "Bh0": 2, "Bh1": 3, "mom0": 4, "mom1": 5, "key0": 6, "key1": 7, "hist0": 8, "hist1": 9,
        "taa0": 10, "taa1": 11, "par0_0": 12, "par0_1": 13, "par1": 14, "par2": 15, "F": 16, "C0": 17, "C1": 18,
        "CG": 19}
MODES = {"fp32": ("ho32_", ctypes.c_float, np.float32), "fp64q": ("ho64q_", ctypes.c_double, np.float64),
         "fp64": ("ho64_", ctypes.c_double, np.float64)}
_L = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        _L = ctypes.CDLL(LIB)
    return _L


class Oracle:
    """One FP32-mirror (or FP64) restatement of the Hush class state machine."""

    def __init__(self, width, height, weights, exposure=1.0, f16=False, mode="fp32", lib=None):
        L = load() if lib is None else lib
        self.mode = mode
        self._lib = L
        pre, self._ct, self._np = MODES[mode]
        self._f = {}
        for name in ("create", "destroy", "resize", "reset", "set_exposure", "set_state", "frame", "hidden",
                     "surface", "mask", "features_in", "set_inputs", "run_pass", "denoise", "output_id", "mlp",
                     "hidden_of", "f2h_rtz", "h2f", "qf16", "oct_enc", "oct_dec", "check"):
            self._f[name] = getattr(L, pre + name)
        F, ct = self._f, self._ct
        F["create"].restype = ctypes.c_void_p
        F["create"].argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_long, ct, ctypes.c_int]
        F["destroy"].argtypes = [ctypes.c_void_p]
        F["resize"].argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        F["reset"].argtypes = [ctypes.c_void_p]
        F["set_exposure"].argtypes = [ctypes.c_void_p, ctypes.c_double]
        F["set_state"].argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        F["frame"].argtypes = [ctypes.c_void_p]
        F["surface"].restype = ctypes.c_void_p
        F["surface"].argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
        F["mask"].restype = ctypes.c_void_p
        F["mask"].argtypes = [ctypes.c_void_p]
        F["features_in"].restype = ctypes.c_void_p
        F["features_in"].argtypes = [ctypes.c_void_p]
        F["set_inputs"].argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 5
        F["run_pass"].argtypes = [ctypes.c_void_p, ctypes.c_int]
        F["denoise"].argtypes = [ctypes.c_void_p]
        F["output_id"].argtypes = [ctypes.c_void_p]
        F["mlp"].argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        F["hidden_of"].argtypes = [ctypes.c_long]
        F["f2h_rtz"].restype = ctypes.c_uint16
        F["f2h_rtz"].argtypes = [ct]
        F["h2f"].restype = ct
        F["h2f"].argtypes = [ctypes.c_uint16]
        F["qf16"].restype = ct
        F["qf16"].argtypes = [ct]
        F["oct_enc"].argtypes = [ct, ct, ct, ctypes.c_void_p]
        F["oct_dec"].argtypes = [ct, ct, ctypes.c_void_p]
        F["check"].argtypes = [ctypes.c_void_p, ctypes.c_void_p, ct, ct]
        w = np.ascontiguousarray(weights, dtype=np.uint16)
        self._w = w
        self._c = F["create"](int(width), int(height), w.ctypes.data, int(w.size), exposure, int(bool(f16)))
        if not self._c:
            raise ValueError("oracle: bad weights or size")
        self.width, self.height = int(width), int(height)
        self._inputs = None

    def close(self):
        if self._c:
            self._f["destroy"](self._c)
            self._c = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    # ---- state machine
    def resize(self, w, h):
        r = self._f["resize"](self._c, int(w), int(h))
        if r != 0:
            raise ValueError(f"Hush: bad size {w}x{h}")
        self.width, self.height = int(w), int(h)

    def reset(self):
        self._f["reset"](self._c)

    def set_exposure(self, e):
        self._f["set_exposure"](self._c, float(e))

    def set_state(self, frame, reset):
        self._f["set_state"](self._c, int(frame), int(bool(reset)))

    @property
    def frame(self):
        return int(self._f["frame"](self._c))

    def set_inputs(self, g: dict):
        arrs = [np.ascontiguousarray(g[k], dtype=np.uint16) for k in ("radiance", "normalDepth", "motion", "material", "emission")]
        for a in arrs:
            assert a.shape == (self.height, self.width, 4), a.shape
        self._inputs = arrs
        self._f["set_inputs"](self._c, *[a.ctypes.data for a in arrs])

    def run_pass(self, p: int):
        self._f["run_pass"](self._c, int(p))

    def denoise(self, g: dict | None = None) -> np.ndarray:
        if g is not None:
            self.set_inputs(g)
        sid = self._f["denoise"](self._c)
        return self.surface(sid).copy()

    def output(self) -> np.ndarray:
        return self.surface(self._f["output_id"](self._c)).copy()

    # ---- surfaces (views into the oracle's memory; write through them to teacher-force)
    def surface(self, sid) -> np.ndarray:
        sid = SURF[sid] if isinstance(sid, str) else int(sid)
        w, h, ch = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        p = self._f["surface"](self._c, sid, ctypes.byref(w), ctypes.byref(h), ctypes.byref(ch))
        buf = (self._ct * (w.value * h.value * ch.value)).from_address(p)
        return np.ctypeslib.as_array(buf).reshape(h.value, w.value, ch.value)

    def mask(self) -> np.ndarray:
        p = self._f["mask"](self._c)
        buf = (ctypes.c_uint8 * (self.width * self.height)).from_address(p)
        return np.ctypeslib.as_array(buf).reshape(self.height, self.width)

    def features_in(self) -> np.ndarray:
        p = self._f["features_in"](self._c)
        buf = (self._ct * (self.width * self.height * 20)).from_address(p)
        return np.ctypeslib.as_array(buf).reshape(self.height, self.width, 20)

    # ---- scalar helpers
    def mlp(self, f: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(f, dtype=self._np)
        o = np.zeros(12, self._np)
        self._f["mlp"](self._c, f.ctypes.data, o.ctypes.data)
        return o

    def f2h_rtz(self, x) -> int:
        return int(self._f["f2h_rtz"](x))

    def h2f(self, h) -> float:
        return float(self._f["h2f"](int(h)))

    def qf16(self, x) -> float:
        return float(self._f["qf16"](x))

    def oct_enc(self, n):
        o = np.zeros(2, self._np)
        self._f["oct_enc"](float(n[0]), float(n[1]), float(n[2]), o.ctypes.data)
        return o

    def oct_dec(self, e):
        o = np.zeros(3, self._np)
        self._f["oct_dec"](float(e[0]), float(e[1]), o.ctypes.data)
        return o

    def check(self, k, n, zprev, obj) -> bool:
        k = np.ascontiguousarray(k, dtype=self._np)
        n = np.ascontiguousarray(n, dtype=self._np)
        return bool(self._f["check"](k.ctypes.data, n.ctypes.data, zprev, obj))

    def threads(self) -> int:
        """OpenMP threads the oracle's passes run on"""
        return int(getattr(self._lib, MODES[self.mode][0] + "threads")())

    def hidden_of(self, n) -> int:
        return int(self._f["hidden_of"](int(n)))


SVGF_DEFAULTS = dict(sigmaL=4.0, sigmaN=128.0, sigmaZ=1.0, alphaColor=0.2, alphaMoments=0.2, taaAlpha=0.1, taa=True,
                     iterations=5)
SVGF_SURF = {"histColor0": 0, "histColor1": 1, "histMoments0": 2, "histMoments1": 3, "tmpColor": 4, "atrous0": 5,
             "atrous1": 6, "taa0": 7, "taa1": 8}


class SvgfOracle:
    """CPU restatement of the SVGF class (src/denoise/svgf.js:273-379)."""

    def __init__(self, width, height, mode="fp32", **opts):
        L = load()
        pre, self._ct, self._np = MODES[mode]
        o = dict(SVGF_DEFAULTS, **opts)
        f = lambda n: getattr(L, pre + n)
        self._create, self._destroy = f("svgf_create"), f("svgf_destroy")
        self._create.restype = ctypes.c_void_p
        self._create.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 6 + [ctypes.c_int, ctypes.c_int]
        self._destroy.argtypes = [ctypes.c_void_p]
        self._reset = f("svgf_reset"); self._reset.argtypes = [ctypes.c_void_p]
        self._set = f("svgf_set_inputs"); self._set.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self._surf = f("svgf_surface"); self._surf.restype = ctypes.c_void_p; self._surf.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self._enc = f("svgf_encode"); self._enc.argtypes = [ctypes.c_void_p]
        self._frame = f("svgf_frame"); self._frame.argtypes = [ctypes.c_void_p]
        self._state = f("svgf_set_state"); self._state.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        self._pass = f("svgf_run_pass"); self._pass.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self.iterations = int(o["iterations"])
        self.width, self.height = int(width), int(height)
        self._c = self._create(self.width, self.height, o["sigmaL"], o["sigmaN"], o["sigmaZ"], o["alphaColor"],
                               o["alphaMoments"], o["taaAlpha"], int(bool(o["taa"])), int(o["iterations"]))

    def close(self):
        if self._c:
            self._destroy(self._c)
            self._c = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        self._reset(self._c)

    @property
    def frame(self):
        return int(self._frame(self._c))

    def surface(self, name):
        p = self._surf(self._c, SVGF_SURF[name] if isinstance(name, str) else int(name))
        buf = (self._ct * (self.width * self.height * 4)).from_address(p)
        return np.ctypeslib.as_array(buf).reshape(self.height, self.width, 4)

    def set_inputs(self, g: dict) -> None:
        """g: the five inputs plus prevNormalDepth / prevMotion (binary16 bit patterns)"""
        keys = ("radiance", "normalDepth", "motion", "material", "emission", "prevNormalDepth", "prevMotion")
        arrs = [np.ascontiguousarray(g[k], dtype=np.uint16) for k in keys]
        for a in arrs:
            assert a.shape == (self.height, self.width, 4), a.shape
        ptrs = (ctypes.c_void_p * 7)(*[a.ctypes.data for a in arrs])
        self._set(self._c, ptrs)

    def encode(self, g: dict | None = None) -> np.ndarray:
        if g is not None:
            self.set_inputs(g)
        sid = self._enc(self._c)
        return self.surface(sid).copy()

    def set_state(self, frame: int, reset: bool) -> None:
        self._state(self._c, int(frame), int(bool(reset)))

    def run_pass(self, p: int) -> None:
        self._pass(self._c, int(p))


def post_encode(left, right, mid=None, split=0.5, split2=None, same=False, exposure=1.0, bloom=0.03, vignette=0.35,
                line=1.5, width=None, height=None, bgra=False):
    """CPU restatement of Post.encode (src/demo/post.js:179-209).  left / right / mid: (h, w, 4) float16
    arrays.  Returns (target (height, width, 4) uint8, hdr (h, w, 4) float32, [level 1..6 arrays])."""
    L = load()
    f = L.po32_encode
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 2 + [ctypes.c_float] * 2 + [ctypes.c_int] + \
        [ctypes.c_float] * 4 + [ctypes.c_int] * 3 + [ctypes.c_void_p] * 3
    h, w = left.shape[:2]
    width = w if width is None else width
    height = h if height is None else height
    split2 = split if split2 is None else split2
    imgs = [np.ascontiguousarray(a, dtype=np.float16) for a in (left, right)]
    m = None if mid is None else np.ascontiguousarray(mid, dtype=np.float16)
    target = np.zeros((height, width, 4), np.uint8)
    hdr = np.zeros((h, w, 4), np.float32)
    sizes = [(max(1, h >> i), max(1, w >> i)) for i in range(1, 7)]
    lv = np.zeros(sum(a * b * 4 for a, b in sizes), np.float32)
    r = f(imgs[0].ctypes.data, imgs[1].ctypes.data, None if m is None else m.ctypes.data, w, h, split, split2,
          int(bool(same)), exposure, bloom, vignette, line, width, height, int(bool(bgra)), target.ctypes.data,
          hdr.ctypes.data, lv.ctypes.data)
    assert r == 0
    levels, o = [], 0
    for a, b in sizes:
        levels.append(lv[o:o + a * b * 4].reshape(a, b, 4))
        o += a * b * 4
    return target, hdr, levels


def post_agx(c):
    """The AgX transform with its look (post.js:99-119) of one linear colour -> display-linear [0, 1]^3."""
    L = load()
    f = L.po32_agx
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    i = np.ascontiguousarray(c, dtype=np.float32)
    o = np.zeros(3, np.float32)
    f(i.ctypes.data, o.ctypes.data)
    return o


def post_srgb(c):
    L = load()
    f = L.po32_srgb
    f.restype = ctypes.c_float
    f.argtypes = [ctypes.c_float]
    return float(f(c))


def post_sample(img, u, v):
    """The linear clamp-to-edge sampler of the post oracle at normalised (u, v) of img (h, w, 4)."""
    L = load()
    f = L.po32_sample
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
    a = np.ascontiguousarray(img, dtype=np.float32)
    o = np.zeros(3, np.float32)
    f(a.ctypes.data, a.shape[1], a.shape[0], u, v, o.ctypes.data)
    return o


def post_pass(which, src, w, h, dst=None):
    """One post pass alone: 0 Karis downsample, 1 downsample, 2 additive upsample into dst; returns dst."""
    L = load()
    f = L.po32_pass
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    a = np.ascontiguousarray(src, dtype=np.float32)
    d = np.zeros((h, w, 4), np.float32) if dst is None else np.ascontiguousarray(dst, dtype=np.float32).copy()
    f(which, a.ctypes.data, a.shape[1], a.shape[0], d.ctypes.data, w, h)
    return d
