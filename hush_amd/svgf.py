"""SVGF baseline denoiser (the `SVGF` class of src/denoise/svgf.js:273-379) over the C ABI of include/svgf.h.

    sv = SVGF(0, 1920, 1080, {"sigmaL": 4, "iterations": 5})
    out = sv.encode({"radiance": ..., "normalDepth": ..., "motion": ..., "material": ..., "emission": ...,
                     "prevNormalDepth": ..., "prevMotion": ...})
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .hip import Surface

SVGF_INPUTS = ("radiance", "normalDepth", "motion", "material", "emission", "prevNormalDepth", "prevMotion")
SVGF_SURFACES = {"histColor0": 0, "histColor1": 1, "histMoments0": 2, "histMoments1": 3, "tmpColor": 4,
                 "atrous0": 5, "atrous1": 6, "taa0": 7, "taa1": 8}
_OPT_FIELDS = {"sigmaL": "sigma_l", "sigmaN": "sigma_n", "sigmaZ": "sigma_z", "alphaColor": "alpha_color",
               "alphaMoments": "alpha_moments", "taaAlpha": "taa_alpha", "taa": "taa", "iterations": "iterations"}


class SVGFError(Exception):
    def __init__(self, message: str, code: int = -1):
        super().__init__(message)
        self.code = code


class SVGFTimer:
    """After `encode(..., timer)`, `read()` returns [(label, ms)] for that frame's passes."""

    def __init__(self):
        self._sv = None

    def read(self):
        if self._sv is None:
            return []
        sv = self._sv
        ms = (ctypes.c_float * 8)()
        labels = (ctypes.c_char_p * 8)()
        n = ctypes.c_int32()
        sv._check(sv._L.svgf_pass_times(sv._h, ms, labels, ctypes.byref(n)))
        return [(labels[i].decode(), float(ms[i])) for i in range(n.value)]


class SVGF:
    """new SVGF(device, width, height, opts) (svgf.js:274-295).  opts: sigmaL = 4, sigmaN = 128, sigmaZ = 1,
    alphaColor = 0.2, alphaMoments = 0.2, taaAlpha = 0.1, taa = true, iterations = 5 (0..5)."""

    def __init__(self, device=0, width=0, height=0, opts=None, **kw):
        self._L = _lib.load()
        o = _lib.svgf_options()
        self._L.svgf_default_options(ctypes.byref(o))
        given = dict(opts or {})
        given.update(kw)
        for k, v in given.items():
            if k not in _OPT_FIELDS:
                raise SVGFError(f"SVGF: unknown option '{k}'")
            f = _OPT_FIELDS[k]
            setattr(o, f, int(bool(v)) if f == "taa" else (int(v) if f == "iterations" else float(v)))
        p = ctypes.c_void_p()
        r = self._L.svgf_create(int(getattr(device, "ordinal", device)), int(width), int(height), ctypes.byref(o),
                                ctypes.byref(p))
        if r != 0:
            raise SVGFError(self._L.svgf_last_error(None).decode(), r)
        self._h = p.value
        self._timer_on = False

    def _check(self, r: int) -> None:
        if r != 0:
            raise SVGFError(self._L.svgf_last_error(ctypes.c_void_p(self._h)).decode(), r)

    def _inputs(self, g) -> _lib.svgf_inputs:
        from . import _surface_struct
        get = g.get if isinstance(g, dict) else (lambda k: getattr(g, k, None))
        ins = _lib.svgf_inputs()
        for k, f in zip(SVGF_INPUTS, ("radiance", "normal_depth", "motion", "material", "emission",
                                      "prev_normal_depth", "prev_motion")):
            setattr(ins, f, _surface_struct(get(k)))
        return ins

    def _wrap(self, s) -> Surface:
        return Surface(s.width, s.height, np.float16, 4, ptr=s.ptr, pitch=s.pitch, owner=self)

    @property
    def width(self) -> int:
        w, h = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._L.svgf_get_size(self._h, ctypes.byref(w), ctypes.byref(h)))
        return w.value

    @property
    def height(self) -> int:
        w, h = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._L.svgf_get_size(self._h, ctypes.byref(w), ctypes.byref(h)))
        return h.value

    @property
    def frame(self) -> int:
        return int(self._L.svgf_frame(self._h))

    @property
    def iterations(self) -> int:
        return self.params["iterations"]

    @property
    def params(self) -> dict:
        o = _lib.svgf_options()
        self._check(self._L.svgf_get_options(self._h, ctypes.byref(o)))
        return {k: (bool(getattr(o, f)) if f == "taa" else getattr(o, f)) for k, f in _OPT_FIELDS.items()}

    def encode(self, g, timer: SVGFTimer | None = None, stream: int | None = None) -> Surface:
        """One frame (svgf.js:343-378); returns this frame's output (rgba16float, linear HDR)."""
        want = timer is not None
        if want != self._timer_on:
            self._check(self._L.svgf_set_timing(self._h, int(want)))
            self._timer_on = want
        ins = self._inputs(g)
        out = _lib.hush_surface()
        self._check(self._L.svgf_encode(self._h, ctypes.byref(ins), ctypes.c_void_p(stream or 0), ctypes.byref(out)))
        if timer is not None:
            timer._sv = self
        return self._wrap(out)

    @property
    def output(self) -> Surface:
        """The `output` getter (svgf.js:327): taa[frame & 1] -- after encode() that is the previous frame's
        buffer, as in the reference."""
        out = _lib.hush_surface()
        self._check(self._L.svgf_output(self._h, ctypes.byref(out)))
        return self._wrap(out)

    def resize(self, width: int, height: int) -> None:
        self._check(self._L.svgf_resize(self._h, int(width), int(height)))

    def reset(self) -> None:
        self._check(self._L.svgf_reset(self._h))

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            self._L.svgf_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.destroy()
        except Exception:
            pass

    # ---- parity hooks
    def surface(self, name) -> Surface:
        s = _lib.hush_surface()
        self._check(self._L.svgf_debug_surface(self._h, SVGF_SURFACES[name] if isinstance(name, str) else int(name),
                                               ctypes.byref(s)))
        return self._wrap(s)

    def set_state(self, frame: int, reset: bool) -> None:
        self._check(self._L.svgf_debug_set_state(self._h, int(frame), int(bool(reset))))

    def run_pass(self, p: int, g, stream: int | None = None) -> None:
        ins = self._inputs(g)
        self._check(self._L.svgf_debug_run_pass(self._h, int(p), ctypes.byref(ins), ctypes.c_void_p(stream or 0)))
