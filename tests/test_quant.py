"""Host logic of the fp8 input-scale calibration (hush_amd/quant.py): no GPU."""
import json

import numpy as np

import hush_amd
from hush_amd import quant


def test_pow2_at_least():
    assert quant.pow2_at_least(0.0) == 1.0
    assert quant.pow2_at_least(448.0) == 1.0          # 448 / 1 <= 448
    assert quant.pow2_at_least(449.0) == 2.0
    assert quant.pow2_at_least(16.0) == 2.0 ** -4      # 16 / 2^-4 = 256; 16 / 2^-5 = 512 > 448
    for m in (1e-6, 0.3, 7.0, 1234.5, 1e9):
        s = quant.pow2_at_least(m)
        assert np.frexp(s)[0] == 0.5 and m / s <= 448.0 < m / (s / 2)


def _weights(seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 0.3, 812).astype(np.float16).view(np.uint16)


def test_layer_input_maxima_against_a_direct_forward():
    """the maxima are those of a float64 forward written out here unit by unit (ReLU hidden layers)"""
    w16 = _weights(1)
    w = w16.view(np.float16).astype(np.float64)
    rows = np.random.default_rng(2).normal(0.0, 2.0, (64, 20))
    m = quant.layer_input_maxima(w16, rows)
    h1 = np.array([[max(0.0, w[320 + i] + sum(w[i * 20 + j] * r[j] for j in range(20))) for i in range(16)] for r in rows])
    h2 = np.array([[max(0.0, w[592 + i] + sum(w[336 + i * 16 + j] * h[j] for j in range(16))) for i in range(16)] for h in h1])
    assert m[0] == np.abs(rows).max()
    assert abs(m[1] - h1.max()) <= 1e-12 * h1.max() and abs(m[2] - h2.max()) <= 1e-12 * h2.max()
    s = quant.calibrate(w16, rows, headroom=2.0)
    assert s == [quant.pow2_at_least(2.0 * v) for v in m]


def test_load_scales_is_keyed_by_the_weights(tmp_path):
    w16 = _weights(3)
    path = tmp_path / "scales.json"
    path.write_text(json.dumps({"weights_sha256": quant.weights_digest(w16), "scales": [0.25, 2.0, 8.0]}))
    assert quant.load_scales(w16, str(path)) == [0.25, 2.0, 8.0]
    assert quant.load_scales(_weights(4), str(path)) is None
    assert quant.load_scales(w16, str(tmp_path / "missing.json")) is None


def test_shipped_scales_belong_to_the_built_in_weights():
    s = quant.load_scales(hush_amd.default_weights())
    assert s is not None and len(s) == 3
    assert all(v > 0 and np.frexp(v)[0] == 0.5 and 2.0 ** -20 <= v <= 2.0 ** 20 for v in s)
