"""The reduced-range square root, exponentials, reciprocal and logarithm the kernels use where their
arguments allow it (hush_device.h: Math<false>::sqrtn, expn, expm, rcpn, logn) against the toolchain's
sqrtf / expf / IEEE division / logf on the GPU, over every binary32 input of each domain
(hush_debug_math_selftest): bit-identical, so the IEEE policy stays the operation-exact mirror of the
oracle.  The domains are those the call sites guarantee: square roots of max(v, 1e-12), of |mv|^2 (zero
or at least 2^-48) and of the octahedral decode's squared norm (at least 1/3); exponentials of the
non-positive edge-stopping and joint-bilateral exponents and of the MLP outputs clamped to [-5, 5];
reciprocals of sums and counts bounded away from 0 and infinity (at least 1e-6, at most ~1.5e6);
logarithms of 1 + x (zero or at least 2^-24 in magnitude near x = -1, never subnormal).  The shared-reciprocal
division (Math<false>::divs) against IEEE division for every numerator of its domain over 24 divisors: the
special ones (+-0, +-inf, NaN), the normal extremes and hard significands.  The reduced sigmoid of the MLP's
first four outputs (hush_device.h: sigmoid<false>) through the activations as stored (1/32 + 31/32 s,
0.5 + 7.5 s and s itself, truncated to binary16) against 1 / (1 + expf(-x)) for every binary32 x.  SVGF's a-trous weight exponential
(Math<false>::expw) against expf over every finite x <= 0: the same bits from ln(2^-149) up, and below, a
value that any factor of at most 1/2 (the kernel weight times the normal term) takes to expf's 0."""
import ctypes

import pytest

from hush_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which,name", [(0, "sqrt, x >= 2^-96, +0, +inf, NaN"), (1, "exp, x <= 0, NaN"),
                                        (2, "exp, -103.28 <= x <= 88.72"), (3, "rcp, normal |x| < 2^126, NaN"),
                                        (4, "log, all but subnormals"),
                                        (5, "shared-reciprocal division, every numerator x 24 divisors"),
                                        (6, "sigmoid's stored activations, every x"),
                                        (7, "exp as a weight factor, finite x <= 0")])
def test_reduced_range_math_matches_toolchain(which, name):
    L = _lib.load()
    bad, first = ctypes.c_uint64(0), ctypes.c_uint32(0)
    assert L.hush_debug_math_selftest(which, ctypes.byref(bad), ctypes.byref(first)) == 0
    print(f"{name}: mismatches {bad.value}")
    assert bad.value == 0, (name, bad.value, hex(first.value))
