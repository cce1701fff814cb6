"""The C ABI library loads, exports every symbol include/*.h declares, and reproduces the reference's
host-side error behaviour (hush.js:743, 765) without touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import hush_amd
from hush_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    inc = os.path.join(ROOT, "include")
    for name in sorted(os.listdir(inc)):
        if name.endswith(".h"):
            text = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, name)).read(), flags=re.S)
            syms |= set(re.findall(r"\b((?:hush|svgf|post)_[a-z_0-9]+)\s*\(", text))
    return sorted(syms)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [s for s in decl if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == decl


def test_hidden_of_closed_form():
    # hiddenOf (hush.js:22-25): h*h + 34h + 12 == n
    for h in range(1, 257):
        n = h * h + 34 * h + 12
        assert hush_amd.hiddenOf(n) == h
        assert hush_amd.hiddenOf(n + 1) == 0
    assert hush_amd.hiddenOf(812) == 16
    assert hush_amd.hiddenOf(0) == 0
    assert hush_amd.hiddenOf(257 * 257 + 34 * 257 + 12) == 0


def test_builtin_weights_blob():
    w = hush_amd.default_weights()
    assert w.size == 812 and hush_amd.hiddenOf(w.size) == 16
    f = w.view(np.float16).astype(np.float32)
    assert np.isfinite(f).all()
    # no subnormals in the blob (the reference's halfToFloat handles them anyway)
    assert not np.any(((w >> 10) & 31) == 0) or np.all(f[((w >> 10) & 31) == 0] == 0)


def test_decode_weights_forms():
    w = hush_amd.default_weights()
    import base64
    b64 = base64.b64encode(w.tobytes()).decode()
    assert np.array_equal(hush_amd.decode_weights(b64), w)
    assert np.array_equal(hush_amd.decode_weights(w.tobytes()), w)
    assert np.array_equal(hush_amd.decode_weights(w.view(np.float16)), w)
    assert np.array_equal(hush_amd.decode_weights(bytearray(w.tobytes())), w)


def test_constructor_errors_match_reference():
    with pytest.raises(hush_amd.HushError, match=r"^hush weights: 5 f16 values fit no MLP size$"):
        hush_amd.Hush(0, {"width": 8, "height": 8, "weights": np.zeros(5, np.uint16)})
    with pytest.raises(hush_amd.HushError, match=r"^Hush: bad size 0x64$"):
        hush_amd.Hush(0, {"width": 0, "height": 64})
    with pytest.raises(hush_amd.HushError, match=r"^Hush: bad size 16x-3$"):
        hush_amd.Hush(0, {"width": 16, "height": -3})


def test_required_limits_check():
    class Dev:
        ordinal = 0
        limits = {"maxStorageTexturesPerShaderStage": 4}
    with pytest.raises(hush_amd.HushError, match="Hush needs the device limit maxStorageTexturesPerShaderStage >= 8"):
        hush_amd.Hush(Dev(), {"width": 8, "height": 8})
    assert dict(hush_amd.Hush.requiredLimits) == {"maxStorageTexturesPerShaderStage": 8}


def test_null_context_errors():
    L = _lib.load()
    assert L.hush_reset(None) == -1
    assert b"null" in L.hush_last_error(None)
    out = _lib.hush_surface()
    assert L.hush_output(None, ctypes.byref(out)) == -1
    assert L.hush_get_precision(None) == -1


def test_svgf_host_errors():
    L = _lib.load()
    o = _lib.svgf_options()
    L.svgf_default_options(ctypes.byref(o))
    # the defaults of svgf.js:278-281
    assert (o.sigma_l, o.sigma_n, o.sigma_z, o.taa, o.iterations) == (4.0, 128.0, 1.0, 1, 5)
    assert abs(o.alpha_color - 0.2) < 1e-7 and abs(o.alpha_moments - 0.2) < 1e-7 and abs(o.taa_alpha - 0.1) < 1e-7
    p = ctypes.c_void_p()
    o.iterations = 6
    assert L.svgf_create(0, 8, 8, ctypes.byref(o), ctypes.byref(p)) == -6 and not p.value
    assert b"iterations" in L.svgf_last_error(None)
    o.iterations = 5
    assert L.svgf_create(0, 0, 8, ctypes.byref(o), ctypes.byref(p)) == -3
    assert L.svgf_last_error(None) == b"SVGF: bad size 0x8"
    assert L.svgf_reset(None) == -1
    assert L.svgf_frame(None) == -1
