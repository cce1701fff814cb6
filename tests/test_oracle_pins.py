"""Pins of the CPU oracle against what the mathematics fixes (the reference ships no golden vectors; its
only fixture is the weight blob).  Each check is chosen so that a plausible restatement mistake fails it."""
import numpy as np
import pytest

import hush_amd
from oracle.ho import Oracle

W16 = hush_amd.default_weights()


@pytest.fixture(scope="module")
def o32():
    return Oracle(8, 8, W16)


@pytest.fixture(scope="module")
def o64():
    return Oracle(8, 8, W16, mode="fp64")


def rtz_reference(x: np.ndarray) -> np.ndarray:
    """binary16 round-toward-zero via numpy's RNE conversion stepped back toward zero when it rounded away;
    clamped to +-65504"""
    x = x.astype(np.float64)
    c = np.clip(x, -65504.0, 65504.0)
    r = c.astype(np.float16)
    away = np.abs(r.astype(np.float64)) > np.abs(c)
    r[away] = np.nextafter(r[away], np.float16(0))
    return r


def test_half_to_float_exhaustive(o32):
    # every finite binary16 pattern widens exactly (halfToFloat, hush.js:706-711)
    codes = np.arange(65536, dtype=np.uint32)
    finite = ((codes >> 10) & 31) != 31
    ref = codes.astype(np.uint16).view(np.float16).astype(np.float32)
    got = np.array([o32.h2f(int(c)) for c in codes[finite][::7]], np.float32)
    assert np.array_equal(got, ref[finite][::7])


def test_store_rtz_matches_numpy(o32):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-9, 6, 4000),
                        np.float32([0.0, -0.0, 1.0, 65504.0, 65519.0, 70000.0, -1e9, 2.0 ** -14, 2.0 ** -24,
                                    2.0 ** -25, 6.1e-5, 5.96e-8, 1.0 + 2.0 ** -11])]).astype(np.float32)
    got = np.array([o32.f2h_rtz(float(v)) for v in x], np.uint16)
    ref = rtz_reference(x).view(np.uint16)
    assert np.array_equal(got, ref)


def test_qf16_closed_form(o32):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(3000) * 10.0 ** rng.integers(-4, 5, 3000)).astype(np.float32)
    got = np.array([o32.qf16(float(v)) for v in x], np.float32)
    ref = (np.clip(x, -65504, 65504).view(np.uint32) & np.uint32(0xffffe000)).view(np.float32)
    assert np.array_equal(got, ref)
    # in the binary16 normal range qf16 equals the RTZ storage value (hush.js:98 states they agree)
    nrm = np.abs(x) >= 2.0 ** -14
    assert np.array_equal(got[nrm], rtz_reference(x[nrm]).astype(np.float32))


def test_oct_roundtrip_and_folds(o32, o64):
    rng = np.random.default_rng(2)
    n = rng.standard_normal((2000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    for o, tol in ((o32, 2e-6), (o64, 1e-12)):
        for v in n[:500]:
            e = o.oct_enc(v)
            assert np.all(np.abs(e) <= 1.0 + 1e-7)
            assert np.allclose(o.oct_dec(e), v, atol=tol)
    # closed forms: axes and the lower-hemisphere fold
    assert np.allclose(o64.oct_enc([0, 0, 1]), [0, 0])
    assert np.allclose(o64.oct_enc([0, 0, -1]), [1, 1])
    assert np.allclose(o64.oct_enc([1, 0, 0]), [1, 0])
    assert np.allclose(o64.oct_enc([0, -1, 0]), [0, -1])
    s = 1 / np.sqrt(3)
    assert np.allclose(o64.oct_enc([s, s, -s]), [2 / 3, 2 / 3])
    assert np.allclose(o64.oct_enc([-s, s, -s]), [-2 / 3, 2 / 3])


def test_check_cases(o64):
    n = np.array([0.0, 0.0, 1.0])
    k = np.array([0.0, 0.0, 10.0, 5.0])          # oct(0,0) = +z, depth 10, id 5
    assert o64.check(k, n, 10.0, 5.0)
    assert not o64.check(k, n, 10.0, 6.0)        # object id differs
    tol = lambda z: 0.05 * z + 0.01               # |k.z - zPrev| <= 0.05 zPrev + 0.01 (hush.js:104)
    assert o64.check(k, n, 10.0 / 1.05 + 1e-3, 5.0)
    assert not o64.check(k, n, 12.0, 5.0)
    assert abs(12.0 - 10.0) > tol(12.0)
    assert not o64.check(np.array([0.0, 0.0, -1.0, 5.0]), n, 10.0, 5.0)   # sky key
    # normal test: dot >= 0.9
    ang_ok, ang_bad = np.arccos(0.91), np.arccos(0.89)
    for a, want in ((ang_ok, True), (ang_bad, False)):
        m = np.array([np.sin(a), 0.0, np.cos(a)])
        assert o64.check(k, m, 10.0, 5.0) is want


def test_mlp_against_numpy_fp64(o32, o64):
    wf = W16.view(np.float16).astype(np.float64)
    h = 16
    W1 = wf[:320].reshape(16, 20); b1 = wf[320:336]
    W2 = wf[336:592].reshape(16, 16); b2 = wf[592:608]
    W3 = wf[608:800].reshape(12, 16); b3 = wf[800:812]
    assert W3.shape == (12, h)

    def ref(f):
        a = np.maximum(W1 @ f + b1, 0)
        a = np.maximum(W2 @ a + b2, 0)
        return W3 @ a + b3

    rng = np.random.default_rng(3)
    for _ in range(50):
        f = rng.standard_normal(20) * rng.uniform(0.1, 3)
        r = ref(f)
        assert np.allclose(o64.mlp(f), r, rtol=1e-12, atol=1e-12)
        assert np.allclose(o32.mlp(f.astype(np.float32)), r, rtol=1e-4, atol=1e-4)
    # f = 0: the biases alone
    assert np.allclose(o64.mlp(np.zeros(20)), W3 @ np.maximum(W2 @ np.maximum(b1, 0) + b2, 0) + b3)


def test_mlp_f16_variant_close_to_fp32():
    of = Oracle(8, 8, W16, f16=True)
    o = Oracle(8, 8, W16)
    rng = np.random.default_rng(4)
    for _ in range(20):
        f = (rng.standard_normal(20) * 0.5).astype(np.float32)
        a, b = of.mlp(f), o.mlp(f)
        assert np.all(np.abs(a - b) <= 2e-2 * (1 + np.abs(b)))
        # every value is a binary16 number
        assert np.array_equal(a.astype(np.float16).astype(np.float32), a)


def test_hidden_of_matches_reference_formula(o32):
    for n in list(range(0, 2000)) + [812, 1224, 256 * 256 + 34 * 256 + 12]:
        hs = [h for h in range(1, 257) if h * h + 34 * h + 12 == n]
        assert o32.hidden_of(n) == (hs[0] if hs else 0)
