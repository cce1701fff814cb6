"""Hush denoiser on MI355X: Python host mirror of the reference's `Hush` class over the C ABI.

Mirrors src/denoise/hush.js:725-875 (and src/index.d.ts:36-60): the constructor takes a device and the
options {width, height, exposure, weights, f16}; `denoise(inputs)` validates and runs one frame; `output`,
`resize`, `reset`, `exposure`, `destroy` and `hiddenOf` keep their meaning and error texts.  Device buffers
(`hush_amd.hip.Surface`, or torch float16 CUDA tensors of shape (H, W, 4)) stand in for GPUTextures.

All denoising runs in the HIP kernels of libhush.so; importing the package without the built library raises.
"""
from __future__ import annotations

import base64
import ctypes
import os
from types import MappingProxyType

import numpy as np

from . import _lib
from .hip import Surface

N_FEAT = 20
N_OUT = 12
INPUTS = ("radiance", "normalDepth", "motion", "material", "emission")
PASS_LABELS = ("temporal", "infer+mid", "infer+mid", "infer+mid", "coarse", "coarse", "fine+final", "taa")
PRECISIONS = {"fp32": 0, "f16": 1, "fp16-mfma": 2, "bf16-mfma": 3, "fp8-mfma": 4}
SURFACES = {"A0": 0, "X": 1, "Bh0": 2, "Bh1": 3, "mom0": 4, "mom1": 5, "key0": 6, "key1": 7, "hist0": 8, "hist1": 9,
            "taa0": 10, "taa1": 11, "par0_0": 12, "par0_1": 13, "par1": 14, "par2": 15, "F0": 16, "F1": 17,
            "F2": 18, "F3": 19, "C0": 20, "C1": 21, "CG": 22, "mask": 23,
            "KD0": 24, "KD1": 25, "GH": 26, "CGD": 27}

WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "hush-weights.bin")


class HushError(Exception):
    """Raised with the reference's messages (hush.js:735, 743, 765, 816, 818) or a HIP error."""

    def __init__(self, message: str, code: int = -1):
        super().__init__(message)
        self.code = code


def hiddenOf(n: int) -> int:
    """Hidden width of the MLP for `n` f16 weights, 0 if none fits (hush.js:22-25)."""
    return int(_lib.load().hush_hidden_of(int(n)))


hidden_of = hiddenOf


def default_weights() -> np.ndarray:
    """The built-in weights: 812 binary16 values (hush-weights.bin)."""
    return np.fromfile(WEIGHTS_PATH, dtype=np.uint16)


def decode_weights(w) -> np.ndarray:
    """decodeWeights (hush.js:713-717): base64 string, bytes-like, or an array of binary16 -> uint16 array."""
    if isinstance(w, str):
        return np.frombuffer(base64.b64decode(w), dtype=np.uint16).copy()
    if isinstance(w, np.ndarray):
        if w.dtype == np.float16:
            return w.view(np.uint16).ravel().copy()
        if w.dtype == np.uint16:
            return w.ravel().copy()
        return np.frombuffer(w.tobytes(), dtype=np.uint16).copy()
    return np.frombuffer(bytes(w), dtype=np.uint16).copy()


def _surface_struct(t) -> _lib.hush_surface:
    s = _lib.hush_surface()
    if t is None:
        return s
    if isinstance(t, Surface):
        s.ptr, s.pitch, s.width, s.height, s.format = t.ptr, t.pitch, t.width, t.height, 0
        return s
    if hasattr(t, "data_ptr"):   # torch tensor (H, W, 4) float16 on the GPU
        if t.dim() != 3 or t.shape[2] != 4 or t.stride(2) != 1 or t.stride(1) != 4:
            raise HushError("Hush.denoise: tensors must be (H, W, 4) with contiguous texels")
        s.ptr, s.pitch, s.width, s.height = t.data_ptr(), t.stride(0) * t.element_size(), t.shape[1], t.shape[0]
        s.format = 0 if str(t.dtype) == "torch.float16" else -1
        return s
    raise HushError(f"Hush.denoise: unsupported input type {type(t).__name__}")


class HushTimer:
    """Per-pass timer (the `timer` hook of hush.js:833-856): after `denoise(..., timer)`, `read()` returns
    [(label, ms)] for the 8 passes of that frame, in dispatch order."""

    def __init__(self):
        self._hush = None

    def read(self):
        if self._hush is None:
            return []
        ms = (ctypes.c_float * 8)()
        h = self._hush
        h._check(h._L.hush_pass_times(h._h, ms, None))
        return list(zip(PASS_LABELS, [float(x) for x in ms]))


class Hush:
    """new Hush(device, {width, height, exposure = 1, weights = built-in, f16 = false}) (hush.js:731-761).

    `device` is a HIP device ordinal, or an object with `ordinal` and an optional `limits` mapping that is
    checked against `Hush.requiredLimits` like the reference.  `precision` selects the MLP arithmetic
    ("fp32" drop-in, "f16", "fp16-mfma", "bf16-mfma", "fp8-mfma"); `f16=True` means "f16".  `fp8_scales`
    sets the fp8 form's static per-layer input scales (see `fp8_scales`).
    """

    requiredLimits = MappingProxyType({"maxStorageTexturesPerShaderStage": 8})

    def __init__(self, device=0, options=None, **kw):
        opts = dict(options or {})
        opts.update(kw)
        width, height = opts.get("width"), opts.get("height")
        exposure = float(opts.get("exposure", 1.0))
        weights = opts.get("weights")
        precision = opts.get("precision", "f16" if opts.get("f16") else "fp32")
        ieee = bool(opts.get("ieee_math", True))
        compensated = bool(opts.get("compensated", False))
        limits = getattr(device, "limits", None)
        if limits is not None:
            for k, v in self.requiredLimits.items():
                if not (limits.get(k, 0) >= v):
                    raise HushError(f"Hush needs the device limit {k} >= {v} (this device: {limits.get(k)}); "
                                    "create the device with requiredLimits: Hush.requiredLimits")
        ordinal = int(getattr(device, "ordinal", device))
        self._L = _lib.load()
        w = default_weights() if weights is None else decode_weights(weights)
        self._weights = w
        p = ctypes.c_void_p()
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        w_ptr = w.ctypes.data_as(ctypes.c_void_p) if w.size else ctypes.c_void_p(1)
        r = self._L.hush_create(ordinal, int(width or 0), int(height or 0), exposure, w_ptr, int(w.size), prec,
                                ctypes.byref(p))
        if r != 0:
            raise HushError(self._L.hush_last_error(None).decode(), r)
        self._h = p.value
        self.device = ordinal
        if not ieee:
            self.ieee_math = False
        if compensated:
            self.compensated = True
        fp8 = opts.get("fp8_scales")
        if fp8 is None and self.hidden == 16:
            from .quant import load_scales
            fp8 = load_scales(w)
        if fp8 is not None:
            self.fp8_scales = fp8
        self._timer_on = False

    # ---- helpers
    def _check(self, r: int) -> None:
        if r != 0:
            raise HushError(self._L.hush_last_error(ctypes.c_void_p(self._h)).decode(), r)

    def _inputs(self, inputs) -> _lib.hush_inputs:
        get = inputs.get if isinstance(inputs, dict) else (lambda k: getattr(inputs, k, None))
        ins = _lib.hush_inputs()
        ins.radiance = _surface_struct(get("radiance"))
        ins.normal_depth = _surface_struct(get("normalDepth"))
        ins.motion = _surface_struct(get("motion"))
        ins.material = _surface_struct(get("material"))
        ins.emission = _surface_struct(get("emission"))
        return ins

    def _wrap(self, s: _lib.hush_surface) -> Surface:
        return Surface(s.width, s.height, np.float16, 4, ptr=s.ptr, pitch=s.pitch, owner=self)

    # ---- reference surface
    @property
    def width(self) -> int:
        w, h = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._L.hush_get_size(self._h, ctypes.byref(w), ctypes.byref(h)))
        return w.value

    @property
    def height(self) -> int:
        w, h = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._L.hush_get_size(self._h, ctypes.byref(w), ctypes.byref(h)))
        return h.value

    @property
    def hidden(self) -> int:
        return int(self._L.hush_hidden(self._h))

    @property
    def frame(self) -> int:
        return int(self._L.hush_frame(self._h))

    @property
    def exposure(self) -> float:
        return float(self._L.hush_get_exposure(self._h))

    @exposure.setter
    def exposure(self, e: float) -> None:
        self._check(self._L.hush_set_exposure(self._h, float(e)))

    @property
    def precision(self) -> int:
        return int(self._L.hush_get_precision(self._h))

    @precision.setter
    def precision(self, p) -> None:
        self._check(self._L.hush_set_precision(self._h, PRECISIONS[p] if isinstance(p, str) else int(p)))

    @property
    def fp8_scales(self) -> tuple:
        """the fp8 form's static input scales of the three layers (powers of two): each layer's input x is
        stored as the E4M3 value of x / s.  Defaults: the calibration stored for the built-in weights, else
        bounds derived from the weights (hush_create)."""
        s = (ctypes.c_float * 3)()
        self._check(self._L.hush_get_fp8_scales(self._h, s))
        return tuple(float(v) for v in s)

    @fp8_scales.setter
    def fp8_scales(self, scales) -> None:
        v = [float(x) for x in scales]
        if len(v) != 3:
            raise HushError("fp8_scales: three scales, one per layer")
        self._check(self._L.hush_set_fp8_scales(self._h, (ctypes.c_float * 3)(*v)))

    @property
    def ieee_math(self) -> bool:
        """True (default): IEEE arithmetic everywhere, the oracle's operation-exact mirror.  False: hardware
        reciprocal / sqrt / exp2 / log2 in the spatial filters; IEEE for every threshold decision."""
        return bool(self._L.hush_get_ieee_math(self._h))

    @ieee_math.setter
    def ieee_math(self, on: bool) -> None:
        self._check(self._L.hush_set_ieee_math(self._h, int(bool(on))))

    @property
    def compensated(self) -> bool:
        """True: Neumaier-compensated moment sums in the temporal statistics (a precision form; False, the
        default, is the reference's plain FP32 summation)."""
        return bool(self._L.hush_get_compensated(self._h))

    @compensated.setter
    def compensated(self, on: bool) -> None:
        self._check(self._L.hush_set_compensated(self._h, int(bool(on))))

    def denoise(self, inputs, timer: HushTimer | None = None, stream: int | None = None) -> Surface:
        """Runs one frame (hush.js:813-822) and returns the output surface (rgba16float, linear HDR)."""
        return self._run(self._L.hush_denoise, inputs, timer, stream)

    def encode(self, inputs, timer: HushTimer | None = None, stream: int | None = None) -> Surface:
        """denoise() without the input checks (hush.js:835-874)."""
        return self._run(self._L.hush_encode, inputs, timer, stream)

    def _run(self, fn, inputs, timer, stream):
        want = timer is not None
        if want != self._timer_on:
            self._check(self._L.hush_set_timing(self._h, int(want)))
            self._timer_on = want
        ins = self._inputs(inputs)
        out = _lib.hush_surface()
        self._check(fn(self._h, ctypes.byref(ins), ctypes.c_void_p(stream or 0), ctypes.byref(out)))
        if timer is not None:
            timer._hush = self
        return self._wrap(out)

    @property
    def output(self) -> Surface:
        """Output of the last denoise() (hush.js:798); two surfaces alternate between frames."""
        out = _lib.hush_surface()
        self._check(self._L.hush_output(self._h, ctypes.byref(out)))
        return self._wrap(out)

    def resize(self, width: int, height: int) -> None:
        """Reallocates for a new size (no-op if unchanged) and resets the history (hush.js:764-794)."""
        self._check(self._L.hush_resize(self._h, int(width), int(height)))

    def reset(self) -> None:
        """Drops the temporal history (hush.js:801)."""
        self._check(self._L.hush_reset(self._h))

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            self._L.hush_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.destroy()
        except Exception:
            pass

    def mlp_forward(self, features, out, n: int, precision="fp32", stream: int | None = None) -> None:
        """The MLP alone: `features` a device buffer of n x 20 binary16, `out` of n x 12 float32."""
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        fp = features.ptr if hasattr(features, "ptr") else int(features)
        op = out.ptr if hasattr(out, "ptr") else int(out)
        self._check(self._L.hush_mlp_forward(self._h, prec, ctypes.c_void_p(fp), ctypes.c_void_p(op), int(n),
                                             ctypes.c_void_p(stream or 0)))

    # ---- parity hooks
    def surface(self, name) -> Surface:
        sid = SURFACES[name] if isinstance(name, str) else int(name)
        s = _lib.hush_surface()
        self._check(self._L.hush_debug_surface(self._h, sid, ctypes.byref(s)))
        if s.format == 1:
            return Surface(s.width, s.height, np.uint32, 4, ptr=s.ptr, pitch=s.pitch, owner=self)
        if s.format == 2:
            return Surface(s.width, s.height, np.uint8, 1, ptr=s.ptr, pitch=s.pitch, owner=self)
        if s.format == 3:
            return Surface(s.width, s.height, np.float32, 4, ptr=s.ptr, pitch=s.pitch, owner=self)
        return self._wrap(s)

    def set_state(self, frame: int, reset: bool) -> None:
        self._check(self._L.hush_debug_set_state(self._h, int(frame), int(bool(reset))))

    def run_pass(self, p: int, inputs, stream: int | None = None) -> None:
        ins = self._inputs(inputs)
        self._check(self._L.hush_debug_run_pass(self._h, int(p), ctypes.byref(ins), ctypes.c_void_p(stream or 0)))

    def set_debug_mask(self, on: bool) -> None:
        self._check(self._L.hush_set_debug_mask(self._h, int(bool(on))))

    def set_wide_gathers(self, on: bool) -> None:
        """debug: the temporal and fine+final passes on 64-bit addresses (their form for surfaces of 2 GiB and more)"""
        self._check(self._L.hush_debug_wide_gathers(self._h, int(bool(on))))


from .svgf import SVGF, SVGFError, SVGFTimer, SVGF_INPUTS  # noqa: E402
from .post import Post, PostError  # noqa: E402

__all__ = ["Hush", "HushError", "HushTimer", "hiddenOf", "hidden_of", "default_weights", "decode_weights", "INPUTS",
           "PASS_LABELS", "N_FEAT", "N_OUT", "Surface", "SVGF", "SVGFError", "SVGFTimer", "SVGF_INPUTS", "Post",
           "PostError"]
