"""Closed-form pins of the oracle's passes (oracle/hush_oracle.c) on tiny hand-built scenes.

The reference ships no golden vectors (its only fixture is the weight blob), so each pass is pinned by
what the mathematics fixes on an input small enough to work out by hand: a constant plane, stripes and
checkerboards whose means and variances are known, a linear depth ramp, binomial-kernel moments, sky
neighbours, image corners, a depth step under a bilinear upsample, count-weighted means, a clip box of a
known 3x3 window.  Every expected value below is derived from the scene, not from the oracle's code: the
arithmetic behind each number is spelled out where it is used.

All pins run the FP64-ideal build (no binary16 storage), so the closed forms hold to rounding (1e-12).
`test_pins_catch_planted_mistakes` recompiles the oracle with one-token mistakes, the kind a restatement
can make (a dropped square, a wrong constant, a swapped operand, a bound made exclusive), and requires that
at least one pin fails for each.
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

import hush_amd
from oracle.ho import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
W16 = hush_amd.default_weights()
PASS = {"temporal": 0, "infer": 1, "mid2": 2, "mid3": 3, "coarse1": 4, "coarse2": 5, "fine": 6, "taa": 7}
TOL = 1e-12
LUM = (0.2126, 0.7152, 0.0722)


def _close(got, want, tol=TOL, what=""):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
    assert np.all(err <= tol), f"{what}: got {got}, want {want} (err {err.max():.3g})"


def _gbuf(W, H, rad=(0.25, 0.25, 0.25), nd=(0.0, 0.0, 1.0, 4.0), mv=(0.0, 0.0, 5.0, 4.0),
          mat=(0.5, 0.5, 0.5, 0.25), emi=(0.0, 0.0, 0.0)):
    """binary16 G-buffer, every channel broadcast from a scalar tuple or an (H, W, k) array"""
    def plane(v, k):
        a = np.zeros((H, W, 4), np.float64)
        a[..., :k] = np.broadcast_to(np.asarray(v, np.float64), (H, W, k)) if np.ndim(v) else v
        h = a.astype(np.float16)
        assert np.array_equal(h.astype(np.float64), a), "inputs must be binary16 values"
        return h.view(np.uint16)
    return {"radiance": plane(rad, 3), "normalDepth": plane(nd, 4), "motion": plane(mv, 4), "material": plane(mat, 4),
            "emission": plane(emi, 3)}


class Pins:
    """the pins, over an oracle library (the real one, or a build with a planted mistake)"""

    def __init__(self, lib=None):
        self.lib = lib

    def oracle(self, W, H, exposure=1.0):
        return Oracle(W, H, W16, exposure=exposure, mode="fp64", lib=self.lib)

    # ------------------------------------------------------------------ block() (hush.js:226-282)

    def block_constant_plane(self):
        """Reset frame of a constant grey plane: illumination 0.25 / 0.5 = 0.5, luminance l = 0.5, normal +z,
        depth 4, |mv| = |(0.75, -1)| = 1.25.  Without history A = illum, moments (l, l^2), len 1, so every
        block's 20 features are fixed: log(1 + l) for the luminance means (f0, f1, f18), 0 for every
        variance (f2, f3, f19), the variance ratio f7 and the lag f6, 1/32 for the history length, 0 for the
        validity v (f5), the octahedral code (0, 0) of +z, dz = dn = 0, the material, 1 - 4/4 = 0 and
        log(1 + 1.25).  F2.zw = the block's mean A, F3 = (oct, (Z, nv)) = ((0, 0), (4, 4))."""
        o = self.oracle(8, 8)
        o.set_inputs(_gbuf(8, 8, mv=(0.75, -1.0, 5.0, 4.0)))
        o.set_state(0, True)
        o.run_pass(PASS["temporal"])
        F = o.surface("F")
        l = 0.5 * sum(LUM)
        want = [math.log(1 + l), math.log(1 + l), 0, 0, 1 / 32, 0, 0, 0, 0, 0, 0, 0, 0.5, 0.5, 0.5, 0.25,
                0, math.log(2.25), math.log(1 + l), 0, 0.5, 0.5, 0.5, 0, 0, 0, 4, 4]
        for Y in range(4):
            for X in range(4):
                _close(F[Y, X, :28], want, what=f"block {X},{Y}")

    def block_partial_validity(self):
        """Blocks with 3, 2, 1 and 0 valid pixels (sky = depth < 0): f16 = 1 - nv/4 = 0.25, 0.5, 0.75;
        F3 = (oct, Z = mean depth of the valid pixels, nv); an all-sky block stores F0..F2 = 0 and
        F3 = (pk(0, 0), pk(-1, 0)) (hush.js:245-248)."""
        W, H = 8, 2
        depth = np.tile(np.array([4.0, 4.25, 4.5, 4.75, 5.0, 5.25, 5.5, 5.75]), (H, 1))
        sky = np.zeros((H, W), bool)
        sky[0, 0] = True                                     # block 0: 3 valid
        sky[0, 2:4] = True                                   # block 1: 2 valid
        sky[0, 4:6] = True
        sky[1, 4] = True                                     # block 2: 1 valid
        sky[:, 6:8] = True                                   # block 3: none
        d = np.where(sky, -1.0, depth)
        nd = np.zeros((H, W, 4))
        nd[..., 2] = 1.0
        nd[..., 3] = d
        o = self.oracle(W, H)
        o.set_inputs(_gbuf(W, H, nd=nd))
        o.set_state(0, True)
        o.run_pass(PASS["temporal"])
        F = o.surface("F")
        for X, nv in ((0, 3), (1, 2), (2, 1)):
            blk = d[:, 2 * X:2 * X + 2]
            _close(F[0, X, 16], 1 - nv / 4, what=f"1 - nv/4, block {X}")
            _close(F[0, X, 26:28], [blk[blk >= 0].mean(), nv], what=f"Z, nv, block {X}")
        _close(F[0, 3, :24], np.zeros(24), what="all-sky features")
        _close(F[0, 3, 24:28], [0, 0, -1, 0], what="all-sky guide")

    def block_depth_step_and_crease(self):
        """Depth 4 on even columns, 5 on odd ones: every block has dz = (5 - 4) / 4 = 0.25 (relative to the
        nearer depth).  Normals (0.5, 0, 0.75) on even rows and (-0.5, 0, 0.75) on odd rows: the mean normal
        is +z, oct (0, 0), and each pixel's unit normal makes cos = 0.75 / |n| with it, so dn = 1 - 0.75/|n|."""
        W = H = 8
        nd = np.zeros((H, W, 4))
        nd[..., 0] = np.where(np.arange(H)[:, None] % 2 == 0, 0.5, -0.5)
        nd[..., 2] = 0.75
        nd[..., 3] = np.where(np.arange(W)[None, :] % 2 == 0, 4.0, 5.0)
        o = self.oracle(W, H)
        o.set_inputs(_gbuf(W, H, nd=nd))
        o.set_state(0, True)
        o.run_pass(PASS["temporal"])
        F = o.surface("F")
        dn = 1 - 0.75 / math.sqrt(0.5 ** 2 + 0.75 ** 2)
        for Y in range(4):
            for X in range(4):
                _close(F[Y, X, 8:12], [0, 0, 0.25, dn], what=f"oct, dz, dn of block {X},{Y}")
                _close(F[Y, X, 24:27], [0, 0, 4.5], what=f"guide of block {X},{Y}")

    def block_luminance_statistics(self):
        """Exposure 2, grey illumination 1 on even columns and 3 on odd ones, rows 1 and 4 at 2.  Block (1, 1)
        (pixels 2..3 x 2..3) holds exposed luminances {2, 6, 2, 6}: Lacc = 4, VA = population variance 4;
        its 4x4 window (block +-1: pixels 1..4) holds 16 values whose mean and population variance give mu4
        and var4 (brute force).  On a reset frame P0.y = E l too, and P0.z = E^2 l^2 - (E l)^2 = 0."""
        W = H = 8
        E = 2.0
        g = np.where(np.arange(W)[None, :] % 2 == 0, 1.0, 3.0) * np.ones((H, 1))
        g[1, :] = 2.0
        g[4, :] = 2.0
        rad = np.repeat(g[..., None] * 0.5, 3, axis=2)           # albedo 0.5
        o = self.oracle(W, H, exposure=E)
        o.set_inputs(_gbuf(W, H, rad=rad))
        o.set_state(0, True)
        o.run_pass(PASS["temporal"])
        F = o.surface("F")[1, 1]
        s = sum(LUM)
        blk = E * s * g[2:4, 2:4].ravel()
        win = E * s * g[1:5, 1:5].ravel()
        _close([F[0], F[1], F[2], F[19]], [math.log1p(blk.mean()), math.log1p(blk.mean()), 0, math.log1p(blk.var())],
               what="Lacc, S0.y, S0.z, VA")
        _close([F[18], F[3]], [math.log1p(win.mean()), math.log1p(win.var())], what="mu4, var4")

    # ------------------------------------------------------------------ temporal with history (hush.js:143-211)

    def _history_scene(self, mv=(0.0, 0.0), ids=None):
        """frame 1 (history in the parity-0 surfaces): grey illumination 1 on even and 2 on odd columns,
        static camera, object id 5, depth 4, normal +z; history colour 3, moments (2.5, 7, 1, 3),
        alpha 0.1, gamma 0.5"""
        W = H = 8
        g = np.where(np.arange(W)[None, :] % 2 == 0, 1.0, 2.0) * np.ones((H, 1))
        rad = np.repeat(g[..., None] * 0.5, 3, axis=2)
        o = self.oracle(W, H)
        o.set_inputs(_gbuf(W, H, rad=rad, mv=(mv[0], mv[1], 5.0, 4.0)))
        key = o.surface("key0")
        key[...] = (0.0, 0.0, 4.0, 5.0)
        if ids is not None:
            key[..., 3] = ids
        hist = o.surface("hist0")
        hist[...] = (3.0, 3.0, 3.0, 0.0)
        o.surface("mom0")[...] = (2.5, 7.0, 1.0, 3.0)
        o.surface("par0_0")[...] = (0.1, 0.5, 0.0, 0.0)
        o.set_state(1, False)
        return o, g

    def lag_and_blend(self):
        """Pixel (3, 3) (illumination b = 2; its 3x3 columns hold 1, 2, 1): the static tap has weight 1, so
        hc = 3, hm = (2.5, 7, 1, 3), v = 1.  3x3 stats: mu = 4/3, luminance variance svar = (6 (1 - 4/3)^2
        + 3 (2 - 4/3)^2) / 9 = 2/9, per-channel deviation sqrt(2/9).  lag (hush.js:202) divides svar by the
        count a second time: lag = |4/3 - 3| / (sqrt(2/81) + 0.02 * 3 + 1e-3).  Blend: hc clamps to
        mu + 0.5 sqrt(2/9); len = 4, a = max(0.1, 1/4); colour = hcl + (2 - hcl) / 4; moments
        (2.5 + (2 - 2.5) / 4, 7 + (4 - 7) / 4) = (2.375, 6.25); nfast = 1 + (2 - 1) / 4; A.w = tvar =
        6.25 - 2.375^2 (len 4 takes all of it)."""
        o, g = self._history_scene()
        o.run_pass(PASS["temporal"])
        s = sum(LUM)
        mu = 4 / 3
        svar = 2 / 9
        lag = abs(mu * s - 3 * s) / (math.sqrt(svar * s * s / 9) + 0.02 * 3 * s + 1e-3)
        P = o.features_in()[3, 3]
        _close(P[4:6], [1.0, lag], tol=1e-9, what="v, lag")
        hcl = mu + 0.5 * math.sqrt(svar)
        col = hcl + (2 - hcl) * 0.25
        l = 2 * s
        m1, m2 = 2.5 + (l - 2.5) * 0.25, 7 + (l * l - 7) * 0.25
        _close(o.surface("A0")[3, 3], [col, col, col, max(m2 - m1 * m1, 0)], tol=1e-9, what="A0")
        _close(o.surface("mom1")[3, 3], [m1, m2, 1 + (l - 1) * 0.25, 4], tol=1e-9, what="moments")

    def fallback_reprojection(self):
        """mv = (0.5, 0): the bilinear taps are p and p + (1, 0), each of weight 1/2; both carry another
        object id, so the 3x3 search around floor(p + mv + 0.5) = p + (1, 0) takes its 7 other texels:
        v = 0.1 and hc = the mean of their history colours (history colour = x + 10 y here)."""
        ids = np.full((8, 8), 5.0)
        ids[3, 3] = ids[3, 4] = 9.0
        o, _ = self._history_scene(mv=(0.5, 0.0), ids=ids)
        yy, xx = np.mgrid[0:8, 0:8]
        o.surface("hist0")[..., 0] = xx + 10 * yy
        o.run_pass(PASS["temporal"])
        P = o.features_in()[3, 3]
        _close(P[4], 0.1, what="v of the 3x3 fallback")
        win = [(x, y) for y in (2, 3, 4) for x in (3, 4, 5) if (x, y) not in ((3, 3), (4, 3))]
        hc = np.mean([x + 10 * y for x, y in win])
        # the blend clamps hc to mu +- gamma sd only per channel; channel r carries hc, g and b carry 3
        mask = o.mask()[3, 3]
        assert mask & 0x10 and not mask & 0x0f, mask
        o2, _ = self._history_scene(mv=(0.5, 0.0), ids=ids)
        o2.surface("hist0")[..., 0] = xx + 10 * yy
        o2.surface("par0_0")[..., 1] = 1e6                      # gamma: no clamp, A0 shows hc itself
        o2.surface("par0_0")[..., 0] = 0.0
        o2.surface("mom0")[..., 3] = 1e6                         # len = 32: a = 1/32
        o2.run_pass(PASS["temporal"])
        a = 1 / 32
        _close(o2.surface("A0")[3, 3, 0], hc + (2 - hc) * a, tol=1e-9, what="fallback colour")

    # ------------------------------------------------------------------ infer activations (hush.js:357-366)

    def infer_activations(self):
        """Weights zero but the output bias b3 = o: the MLP returns o whatever the features, so the
        activations are closed forms: alpha = 1/32 + 31/32 sigma(o0), gamma = 0.5 + 7.5 sigma(o1),
        beta = sigma(o2), kappa = sigma(o3), s = 0.25 exp(clamp(o4..7, +-5)), c = exp(clamp(o8..11, +-4)).
        mid1 filters the block means with (s1, c1): over a uniform colour and variance v = 0.5 its variance is
        v ((c1/4)^2 + 5/64) / (3/4 + c1/4)^2.  A sky block (depth < 0) stores zeros everywhere."""
        o_b = np.array([0.5, -1.0, 2.0, -0.25, 4.5, -6.0, 1.0, 0.0, 3.0, 0.5, -5.0, 1.5])
        w = np.zeros(812)
        w[800:812] = o_b
        lib_w = w.astype(np.float16).view(np.uint16)
        o = Oracle(8, 8, lib_w, mode="fp64", lib=self.lib)
        o.set_inputs(_gbuf(8, 8))
        o.set_state(0, True)
        F = o.surface("F")
        F[...] = 0.0
        F[..., 20:24] = (0.5, 0.5, 0.5, 0.5)
        F[..., 26:28] = (4.0, 4.0)
        F[3, 3, 26] = -1.0
        o.run_pass(PASS["infer"])
        sg = lambda x: 1 / (1 + math.exp(-x))
        _close(o.surface("par0_0")[1, 2], [1 / 32 + 31 / 32 * sg(0.5), 0.5 + 7.5 * sg(-1.0), sg(2.0), sg(-0.25)],
               what="alpha, gamma, beta, kappa")
        _close(o.surface("par1")[1, 2], 0.25 * np.exp([4.5, -5.0, 1.0, 0.0]), what="s")
        _close(o.surface("par2")[1, 2], np.exp([3.0, 0.5, -4.0, 1.5]), what="c")
        c1 = math.exp(0.5)
        _close(o.surface("Bh0")[1, 1], [0.5, 0.5, 0.5, 0.5 * ((c1 / 4) ** 2 + 5 / 64) / (0.75 + c1 / 4) ** 2], tol=1e-9,
               what="mid1 with (s1, c1)")
        for sname in ("par0_0", "par1", "par2", "Bh0"):
            _close(o.surface(sname)[3, 3], np.zeros(4), what=f"sky block {sname}")

    # ------------------------------------------------------------------ filt (atrousFn, hush.js:431-476)

    def _fine(self, W, H, a0, key=(0.0, 0.0, 4.0, 5.0), s=1.0, c=1.0):
        """the fine level (R = 1, step 1, variance prefilter) over A0 with the key guide; returns hist[cur]"""
        o = self.oracle(W, H)
        o.set_inputs(_gbuf(W, H))
        o.set_state(1, False)
        o.surface("A0")[...] = a0
        o.surface("key1")[...] = key
        o.surface("par1")[..., 0] = s
        o.surface("par2")[..., 0] = c
        o.run_pass(PASS["fine"])
        return o.surface("hist1").copy()            # a copy: the view dies with the oracle

    def filt_equal_weights(self):
        """Flat guide, equal luminance everywhere, uniform variance 0.5: every tap weight is h(dx) h(dy), the
        centre h(0)^2 c (hush.js:468-471).  Colour r = x^2 (g compensates to keep the luminance): the
        binomial kernel [1/4, 1/2, 1/4] has second moment 1/2 along x, so r -> x0^2 + (1/2) / (3/4 + c/4)
        * (c = 1: 1/2; c = 3: 1/3).  Variance -> v (sum of squared weights) / (sum of weights)^2 =
        v ((c/4)^2 + 5/64) / (3/4 + c/4)^2."""
        W = H = 8
        yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
        t = xx * xx
        a0 = np.stack([0.5 + LUM[1] * t, 0.5 - LUM[0] * t, np.full_like(t, 0.5), np.full_like(t, 0.5)], -1)
        for c in (1.0, 3.0):
            out = self._fine(W, H, a0, c=c)[3, 3]
            sw = 0.75 + c / 4
            _close(out[0], 0.5 + LUM[1] * (9 + 0.5 / sw), tol=1e-9, what=f"binomial second moment, c={c}")
            _close(out[3], 0.5 * ((c / 4) ** 2 + 5 / 64) / sw ** 2, tol=1e-9, what=f"variance, c={c}")

    def filt_coarse_kernel(self):
        """coarse2: R = 2, step 2, prefilter, s = 0.25, c = 1 (hush.js:752, 866) on a 9x9 quarter-res level
        with equal luminances: taps at even offsets up to 4 with the binomial [1, 4, 6, 4, 1] / 16, whose
        second moment is 1 (in taps), i.e. 4 in texels at step 2: r = x^2 -> x0^2 + 4; the weights sum to
        1, so the variance -> v (sum of h^2)^2 = v (70/256)^2."""
        W = H = 36
        o = self.oracle(W, H)
        o.set_inputs(_gbuf(W, H))
        o.set_state(1, False)
        yy, xx = np.mgrid[0:9, 0:9].astype(np.float64)
        t = xx * xx
        o.surface("C1")[...] = np.stack([0.5 + LUM[1] * t, 0.5 - LUM[0] * t, np.full_like(t, 0.5), np.full_like(t, 0.5)], -1)
        o.surface("CG")[...] = (0.0, 0.0, 4.0, 0.0)
        o.run_pass(PASS["coarse2"])
        out = o.surface("C0")[4, 4]
        _close(out[0], 0.5 + LUM[1] * (16 + 4), tol=1e-9, what="coarse kernel second moment")
        _close(out[3], 0.5 * (70 / 256) ** 2, tol=1e-9, what="coarse variance")

    def _mid2(self, W2, src, guide, s, c):
        """mid2: R = 1, step 2, no prefilter, s = par1.z, c = par2.z, source Bh0, guide from F3"""
        o = self.oracle(2 * W2, 2 * W2)
        o.set_inputs(_gbuf(2 * W2, 2 * W2))
        o.set_state(1, False)
        o.surface("Bh0")[...] = src
        F = o.surface("F")
        F[..., 24:27] = guide
        F[..., 27] = 4.0
        o.surface("par1")[..., 2] = s
        o.surface("par2")[..., 2] = c
        o.run_pass(PASS["mid2"])
        return o.surface("Bh1").copy()

    def filt_luminance_edge(self):
        """Grey 1 (variance 1/4) on columns <= 5, grey 2 (variance 1) beyond; centre (4, 4), step 2, taps
        on columns 2, 4, 6.  sl = s / (sqrt(1/4) + 1e-6) with s = 0.5; column 6 is down-weighted by
        k = exp(-sl |2 - 1| lum(1,1,1)).  Weights: columns 2 and 4 (with the centre 1/4) carry 3/4, column 6
        k/4: colour (3/4 + 2 k/4) / (3/4 + k/4); variance (1/4 (3/128 + 3/32) + k^2 3/128) / (3/4 + k/4)^2."""
        W2 = 9
        yy, xx = np.mgrid[0:W2, 0:W2]
        right = xx >= 6
        src = np.where(right[..., None], [2.0, 2.0, 2.0, 1.0], [1.0, 1.0, 1.0, 0.25])
        out = self._mid2(W2, src, (0.0, 0.0, 4.0), s=0.5, c=1.0)[4, 4]
        sl = 0.5 / (0.5 + 1e-6)
        k = math.exp(-sl * sum(LUM))
        sw = 0.75 + k / 4
        _close(out[0], (0.75 + 2 * k / 4) / sw, tol=1e-9, what="edge-stopped colour")
        _close(out[3], (0.25 * (3 / 128 + 3 / 32) + k * k * 3 / 128) / sw ** 2, tol=1e-9, what="edge-stopped variance")

    def filt_depth_ramp(self):
        """Depth 4 + x/4 (a ramp along x), s = 0 (no luminance term), step 2.  zg comes from the +-1 texel
        neighbours (not step-scaled): zg = min(max(1/4, 0), z) = 1/4.  A tap at offset o = 2 (dx, dy) is
        |dz| = |dx| / 2 deep and |o| = 2 sqrt(dx^2 + dy^2) away: D = exp(-(|dx|/2) / (|o|/4 + 1e-3)),
        i.e. D1 = exp(-0.5/(0.5 + 1e-3)) across, D2 = exp(-0.5/(sqrt(2)/2 + 1e-3)) diagonally, 1 along y.
        Colour r = |x - 4| / 2 (1 on the side and corner taps), g = |y - 4| / 2 (1 on the vertical and
        corner taps): r = (D1/4 + D2/4) / S, g = (1/4 + D2/4) / S with S = 1/4 + 1/4 + D1/4 + D2/4; uniform
        variance v -> v (1/16 + 2/64 + 2 D1^2/64 + 4 D2^2/256) / S^2."""
        W2 = 9
        yy, xx = np.mgrid[0:W2, 0:W2].astype(np.float64)
        src = np.stack([np.abs(xx - 4) / 2, np.abs(yy - 4) / 2, np.zeros_like(xx), np.full_like(xx, 0.5)], -1)
        guide = np.stack([np.zeros_like(xx), np.zeros_like(xx), 4 + xx / 4], -1)
        out = self._mid2(W2, src, guide, s=0.0, c=1.0)[4, 4]
        D1 = math.exp(-0.5 / (0.5 + 1e-3))
        D2 = math.exp(-0.5 / (math.sqrt(2) / 2 + 1e-3))
        S = 0.5 + D1 / 4 + D2 / 4
        _close(out[:2], [(D1 + D2) / 4 / S, (1 + D2) / 4 / S], tol=1e-9, what="depth-ramp weights")
        _close(out[3], 0.5 * (1 / 16 + 2 / 64 + 2 * D1 ** 2 / 64 + 4 * D2 ** 2 / 256) / S ** 2, tol=1e-9,
               what="depth-ramp variance")

    def filt_sky_neighbours(self):
        """Sky +-1 neighbours (depth -1) enter zg as 1e9 (hush.js:464-465).  (a) All four are sky:
        zg = min(max(1e9, 1e9), z) = z = 4.  (b) Only the horizontal ones are sky and the vertical ones
        lie 0.5 deeper: zg = min(max(1e9, 0.5), 4) = 4 again.  Every step-2 tap is one unit deeper:
        D = exp(-1 / (4 |o| + 1e-3)), |o| = 2 (side) or 2 sqrt(2) (corner); r = 1 on the side taps:
        r = (4/8) Ds / (1/4 + 4/8 Ds + 4/16 Dc)."""
        W2 = 9
        yy, xx = np.mgrid[0:W2, 0:W2].astype(np.float64)
        side = ((np.abs(xx - 4) == 2) & (yy == 4)) | ((np.abs(yy - 4) == 2) & (xx == 4))
        src = np.stack([side.astype(np.float64), np.zeros_like(xx), np.zeros_like(xx), np.full_like(xx, 0.5)], -1)
        Ds = math.exp(-1 / (8 + 1e-3))
        Dc = math.exp(-1 / (8 * math.sqrt(2) + 1e-3))
        for vertical in (-1.0, 4.5):
            z = np.full((W2, W2), 5.0)
            z[4, 4] = 4.0
            z[4, 3] = z[4, 5] = -1.0
            z[3, 4] = z[5, 4] = vertical
            guide = np.stack([np.zeros_like(xx), np.zeros_like(xx), z], -1)
            out = self._mid2(W2, src, guide, s=0.0, c=1.0)[4, 4]
            _close(out[0], 0.5 * Ds / (0.25 + 0.5 * Ds + 0.25 * Dc), tol=1e-9,
                   what=f"sky-neighbour depth scale, vertical neighbours at {vertical}")

    def filt_corner(self):
        """Pixel (0, 0): the variance prefilter reads the clamped tile (hush.js:426-430), so its 3x3 Gaussian
        is gv = (9 v00 + 3 v10 + 3 v01 + v11) / 16; the taps outside the image are skipped, leaving (1, 0)
        and (0, 1) at 1/8 and (1, 1) at 1/16.  Centre grey 1, the rest grey 2: the three taps take
        k = exp(-sl lum(1,1,1)) with sl = s / (sqrt(gv) + 1e-6); colour (c/4 + 2 (5/16) k) / (c/4 + (5/16) k),
        variance ((c/4)^2 v00 + k^2 (v10 + v01) / 64 + k^2 v11 / 256) / (c/4 + (5/16) k)^2."""
        W = H = 8
        v = np.full((H, W), 8.0)
        v[0, 0], v[0, 1], v[1, 0], v[1, 1] = 0.5, 1.0, 2.0, 4.0
        grey = np.full((H, W), 2.0)
        grey[0, 0] = 1.0
        a0 = np.stack([grey, grey, grey, v], -1)
        s, c = 0.75, 2.0
        out = self._fine(W, H, a0, s=s, c=c)[0, 0]
        gv = (9 * 0.5 + 3 * 1.0 + 3 * 2.0 + 4.0) / 16
        k = math.exp(-s / (math.sqrt(gv) + 1e-6) * sum(LUM))
        sw = c / 4 + 5 / 16 * k
        _close(out[0], (c / 4 + 2 * 5 / 16 * k) / sw, tol=1e-9, what="corner colour")
        _close(out[3], ((c / 4) ** 2 * 0.5 + k * k * (1.0 + 2.0) / 64 + k * k * 4.0 / 256) / sw ** 2, tol=1e-9,
               what="corner variance")

    def filt_normal_term(self):
        """One neighbour's normal tilted by 0.1 rad: its weight takes cos(0.1)^128 (pow128 of the normals'
        dot product, hush.js:448); r = 1 there only: r = (1/8) C / (7/8 + (1/8) C), C = cos(0.1)^128."""
        W = H = 8
        th = 0.1
        n = np.array([math.sin(th), 0.0, math.cos(th)])
        oc = n[:2] / np.abs(n).sum()
        key = np.zeros((H, W, 4))
        key[..., 2:] = (4.0, 5.0)
        key[3, 4, :2] = oc
        r = np.zeros((H, W))
        r[3, 4] = 1.0
        a0 = np.stack([r, np.zeros_like(r), np.zeros_like(r), np.full_like(r, 0.5)], -1)
        out = self._fine(W, H, a0, key=key, s=0.0)[3, 3]
        C = math.cos(th) ** 128
        _close(out[0], C / 8 / (7 / 8 + C / 8), tol=1e-9, what="normal weight")

    # ------------------------------------------------------------------ jbw upsampling (hush.js:552-619)

    def _final(self, half_z, half_n=None, mid=(0.0, 0.0), uA=(0.0, 0.0), kappa=0.0, beta=0.0, quarter=None, a0=0.5):
        """fine+final at 8x4: pixel (3, 1) upsamples half-res texels x = 1, 2 of rows 0, 1 (bilinear 3/4, 1/4
        per axis) and quarter-res texels (0, 0), (1, 0) (row -1 is outside).  half_z: the two columns'
        depths; half_n: their oct codes; mid / uA: their mid level / block means (r channel).  Returns X."""
        W, H = 8, 4
        o = self.oracle(W, H)
        o.set_inputs(_gbuf(W, H, mat=(1.0, 1.0, 1.0, 0.0)))
        o.set_state(1, False)
        o.surface("A0")[...] = (a0, a0, a0, 0.5)
        o.surface("key1")[...] = (0.0, 0.0, 4.0, 5.0)
        o.surface("par0_1")[...] = (0.1, 1.0, beta, kappa)
        o.surface("par1")[...] = 0.25
        o.surface("par2")[...] = 1.0
        F = o.surface("F")
        F[...] = 0.0
        F[..., 27] = 4.0
        F[..., 26] = 4.0
        for i, X in enumerate((1, 2)):
            F[:, X, 26] = half_z[i]
            if half_n is not None:
                F[:, X, 24:26] = half_n[i]
            F[:, X, 20] = uA[i]
            o.surface("Bh0")[:, X, 0] = mid[i]
        CG = o.surface("CG")
        CG[...] = (0.0, 0.0, 4.0, 0.0)
        if quarter is not None:
            (z0, z1), (c0, c1) = quarter
            CG[0, 0, 2], CG[0, 1, 2] = z0, z1
            o.surface("C0")[0, 0, 0], o.surface("C0")[0, 1, 0] = c0, c1
        o.run_pass(PASS["fine"])
        return o.surface("X")[1, 3].copy()

    def jbw_depth_step(self):
        """Half-res column 1 at the pixel's depth 4, column 2 at 4.1: jbw = bilinear x exp(-|dz| / (0.05 z +
        1e-3)) with the pixel's z (hush.js:557), e = exp(-0.1 / 0.201).  Column weights 3/4 and e/4 (the row
        weights cancel): ill = (3/4 m1 + e/4 m2) / (3/4 + e/4) (kappa = beta = 0, albedo 1, X = ill)."""
        e = math.exp(-0.1 / (0.05 * 4 + 1e-3))
        x = self._final((4.0, 4.1), mid=(1.0, 3.0))
        _close(x[0], (0.75 * 1.0 + e / 4 * 3.0) / (0.75 + e / 4), tol=1e-9, what="jbw depth step")
        # kappa = 1/2 adds kappa (f - uA / sum) with the same weights; f = qf16(fine(A0)) = 0.5 here
        x = self._final((4.0, 4.1), mid=(1.0, 3.0), uA=(0.25, 2.0), kappa=0.5)
        uA = (0.75 * 0.25 + e / 4 * 2.0) / (0.75 + e / 4)
        _close(x[0], (0.75 + e / 4 * 3.0) / (0.75 + e / 4) + 0.5 * (0.5 - uA), tol=1e-9, what="jbw with kappa")

    def jbw_normal_and_fallback(self):
        """A normal tilted by 0.3 rad on column 2 weighs cos(0.3)^8 (dn^4 squared, hush.js:554-557); when
        every tap is far deeper the weights vanish (sum <= 1e-4) and ill falls back to the fine colour."""
        th = 0.3
        n = np.array([math.sin(th), 0.0, math.cos(th)])
        oc = n[:2] / np.abs(n).sum()
        x = self._final((4.0, 4.0), half_n=((0.0, 0.0), oc), mid=(1.0, 3.0))
        C = math.cos(th) ** 8
        _close(x[0], (0.75 + C / 4 * 3.0) / (0.75 + C / 4), tol=1e-9, what="jbw normal term")
        x = self._final((40.0, 40.0), mid=(1.0, 3.0), a0=0.75)
        _close(x[0], 0.75, what="jbw fallback to the fine colour")

    def jbw_coarse(self):
        """beta = 1 takes the quarter-res level: cpos = ((3, 1) + 0.5) / 4 - 0.5 = (0.375, -0.125), so row 0
        carries (5/8, 3/8) of the two columns (row -1 is outside); column 1 is 0.1 deeper:
        up = (5/8 c0 + 3/8 e c1) / (5/8 + 3/8 e)."""
        e = math.exp(-0.1 / (0.05 * 4 + 1e-3))
        x = self._final((4.0, 4.0), mid=(1.0, 1.0), beta=1.0, quarter=((4.0, 4.1), (2.0, 6.0)))
        _close(x[0], (5 / 8 * 2.0 + 3 / 8 * e * 6.0) / (5 / 8 + 3 / 8 * e), tol=1e-9, what="coarse upsample")

    # ------------------------------------------------------------------ down() (hush.js:385-402)

    def down_counts(self):
        """A 4x4 image has a 2x2 half level and a single quarter texel.  Half texels with counts 4, 2, 1, 0
        (N = 7; the empty one is skipped whatever it holds): colour = sum(n a) / 7, variance = sum(n v) / 7^2,
        depth = sum(n z) / 7, normal = normalise(sum(n u)).  The 5x5 filter of a one-texel level skips every
        tap and its prefilter sees the centre alone, so C1 is down()'s colour and variance."""
        o = self.oracle(4, 4)
        o.set_inputs(_gbuf(4, 4))
        o.set_state(1, False)
        F = o.surface("F")
        cnt = np.array([4.0, 2.0, 1.0, 0.0])
        col = np.array([[1.0, 2.0, 3.0], [5.0, 1.0, 0.5], [0.25, 4.0, 2.0], [9.0, 9.0, 9.0]])
        var = np.array([1.0, 2.0, 4.0, 100.0])
        dep = np.array([4.0, 6.0, 11.0, 50.0])
        th = np.array([0.0, 0.4, -0.2, 1.0])
        nrm = np.stack([np.sin(th), np.zeros(4), np.cos(th)], -1)
        for j in range(4):
            X, Y = j & 1, j >> 1
            F[Y, X, 20:23] = col[j]
            F[Y, X, 23] = var[j]
            F[Y, X, 24:26] = nrm[j, :2] / np.abs(nrm[j]).sum()
            F[Y, X, 26] = dep[j]
            F[Y, X, 27] = cnt[j]
        o.run_pass(PASS["coarse1"])
        N = cnt.sum()
        n = (cnt[:, None] * nrm).sum(0)
        n /= np.linalg.norm(n)
        _close(o.surface("C1")[0, 0], list((cnt[:, None] * col).sum(0) / N) + [(cnt * var).sum() / N ** 2], tol=1e-9,
               what="down colour and variance")
        _close(o.surface("CG")[0, 0, :3], [n[0] / np.abs(n).sum(), 0.0, (cnt * dep).sum() / N], tol=1e-9,
               what="down guide")

    # ------------------------------------------------------------------ TAA (hush.js:625-677)

    def _taa(self, X, prev, mv=(0.0, 0.0)):
        H, W = X.shape[:2]
        o = self.oracle(W, H)
        m = np.zeros((H, W, 4))
        m[..., 0], m[..., 1] = mv
        o.set_inputs(_gbuf(W, H, mv=m))
        o.set_state(1, False)
        o.surface("X")[...] = X
        o.surface("taa0")[...] = prev
        o.run_pass(PASS["taa"])
        return o.surface("taa1").copy()

    @staticmethod
    def _taa_expected(greys, cur, hist):
        """TAA of a grey pixel over a grey 3x3 window (values before tone mapping), grey history: YCoCg of a
        grey g is (g, 0, 0), so only Y clips, to mean +- 1.25 sd of tm(window); res = hr + (cur - hr) / 10"""
        tm = lambda g: g / (1 + g)
        y = tm(np.asarray(greys, np.float64))
        m, sd = y.mean(), math.sqrt(max(y.var(), 1e-12))
        hr = min(max(tm(hist), m - 1.25 * sd), m + 1.25 * sd)
        res = hr + (tm(cur) - hr) * 0.1
        return res / max(1 - res, 1e-4)

    def taa_clip(self):
        """Checkerboard of greys 1 and 3 (tone mapped 1/2 and 3/4), history grey 15 (15/16): the history's
        Y clips to the top of the box mean + 1.25 sd of the 3x3 window (five 1/2s, four 3/4s)."""
        W = H = 8
        yy, xx = np.mgrid[0:H, 0:W]
        g = np.where((xx + yy) % 2 == 0, 1.0, 3.0)
        X = np.stack([g, g, g, np.zeros_like(g)], -1)
        out = self._taa(X, (15.0, 15.0, 15.0, 1.0))[3, 3]
        want = self._taa_expected(g[2:5, 2:5].ravel(), 1.0, 15.0)
        _close(out, [want, want, want, 1.0], tol=1e-9, what="TAA clip")

    def taa_uv_bounds(self):
        """uv = (p + 0.5 + mv) / size is tested inclusively (hush.js:661): mv.x = +0.5 at x = 7 gives uv.x = 1
        and mv.x = -0.5 at x = 0 gives uv.x = 0, both take the history (bilinear of the clamped taps = 15);
        mv.x = -0.5 - 2^-11 at x = 0 leaves the image and keeps itm(tm(X)) = X.  The tile clamps at the
        edge, so the 3x3 window repeats the edge column."""
        W = H = 8
        yy, xx = np.mgrid[0:H, 0:W]
        g = np.where((xx + yy) % 2 == 0, 1.0, 3.0)
        X = np.stack([g, g, g, np.zeros_like(g)], -1)
        prev = (15.0, 15.0, 15.0, 1.0)
        cl = lambda a: np.clip(a, 0, 7)
        win = lambda x, y: [g[cl(y + dy), cl(x + dx)] for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
        out = self._taa(X, prev, mv=(0.5, 0.0))[3, 7]
        _close(out[0], self._taa_expected(win(7, 3), g[3, 7], 15.0), tol=1e-9, what="uv.x = 1 (inclusive)")
        out = self._taa(X, prev, mv=(-0.5, 0.0))[3, 0]
        _close(out[0], self._taa_expected(win(0, 3), g[3, 0], 15.0), tol=1e-9, what="uv.x = 0 (inclusive)")
        out = self._taa(X, prev, mv=(-0.5 - 2.0 ** -11, 0.0))[3, 0]
        _close(out[0], g[3, 0], tol=1e-9, what="uv.x < 0: no history")


PIN_NAMES = [n for n in dir(Pins) if not n.startswith("_") and callable(getattr(Pins, n)) and n != "oracle"]


@pytest.mark.parametrize("name", PIN_NAMES)
def test_pin(name):
    getattr(Pins(), name)()


# One-token mistakes a restatement of hush.js can make, each with the pins that must catch it
# (source text in oracle/hush_oracle.c -> replacement).
MUTANTS = [
    ("S0.w / HMAX", "S0.w / 16", ["block_constant_plane"]),
    ("1 - nv / 4, log1p_(MV)", "nv / 4, log1p_(MV)", ["block_constant_plane", "block_partial_validity"]),
    ("log1p_(var4), S0.w", "var4 + 1, S0.w", ["block_luminance_statistics"]),
    ("maxr(zmin, RX(1e-6))", "maxr(zmax, RX(1e-6))", ["block_depth_step_and_crease"]),
    ("real dn = clampr(1 - mind, 0, 2)", "real dn = clampr(1 - mind * mind, 0, 2)", ["block_depth_step_and_crease"]),
    ("VA *= k;", "VA *= k * k;", ["block_luminance_statistics"]),
    ("s1 += t.w * E * qf16", "s1 += t.w * qf16", ["block_luminance_statistics"]),
    ("F[26] = st(qf16(Z))", "F[26] = st(qf16(zs[0]))", ["block_partial_validity"]),
    ("SQRT(maxr(svar / cnt, RX(1e-12)))", "SQRT(maxr(svar, RX(1e-12)))", ["lag_and_blend"]),
    ("RX(0.02) * lh", "RX(0.02) * slm", ["lag_and_blend"]),
    ("maxr(RX(1.0) / RX(6.0), 1 / len1)", "minr(RX(1.0) / RX(6.0), 1 / len1)", ["lag_and_blend"]),
    ("v = RX(0.1); has = 1;", "v = RX(0.2); has = 1;", ["fallback_reprojection"]),
    ("int cx = f2i(FLOOR(ppx + RX(0.5)))", "int cx = f2i(FLOOR(ppx))", ["fallback_reprojection"]),
    ("sv += w * w * cq.w", "sv += w * cq.w", ["filt_equal_weights", "filt_luminance_edge"]),
    ("real cw = (real)(hk[R] * hk[R]) * cc", "real cw = (real)(hk[R]) * cc", ["filt_equal_weights"]),
    ("{1.0 / 16, 1.0 / 4, 3.0 / 8, 1.0 / 4, 1.0 / 16}", "{1.0 / 8, 1.0 / 4, 1.0 / 4, 1.0 / 4, 1.0 / 8}",
     ["filt_coarse_kernel"]),
    ("SQRT(maxr(gv, RX(1e-12))) + RX(1e-6)", "maxr(gv, RX(1e-12)) + RX(1e-6)", ["filt_luminance_edge", "filt_corner"]),
    ("zr > 0 ? FABS(zr - gp.w) : RX(1e9)", "zr > 0 ? FABS(zr - gp.w) : RX(0)", ["filt_sky_neighbours"]),
    ("real zg = minr(maxr(gx, gy), gp.w)", "real zg = maxr(gx, gy)", ["filt_sky_neighbours"]),
    ("sqrt((double)d2) * step", "sqrt((double)d2)", ["filt_depth_ramp"]),
    ("real zl = tG(t, qx - 1, qy).w, zr = tG(t, qx + 1, qy).w",
     "real zl = tG(t, qx - step, qy).w, zr = tG(t, qx + step, qy).w", ["filt_depth_ramp"]),
    ("tA(t, qx - 1, qy - 1).w + tA(t, qx + 1, qy - 1).w", "tA(t, qx + 1, qy + 1).w + tA(t, qx + 1, qy - 1).w",
     ["filt_corner"]),
    ("if (inside(qx + ox, qy + oy, t->w, t->h) && gq.w >= 0)", "if (gq.w >= 0)", ["filt_corner"]),
    ("pow128(maxr(dot3(xyz(gq), xyz(gp)), 0))", "pow128(SQRT(maxr(dot3(xyz(gq), xyz(gp)), 0)))", ["filt_normal_term"]),
    ("(RX(0.05) * z + RX(1e-3))", "(RX(0.05) * zq + RX(1e-3))", ["jbw_depth_step"]),
    ("return wb * EXP(-FABS(zq - z) / (RX(0.05) * z + RX(1e-3))) * dn4 * dn4",
     "return wb * EXP(-FABS(zq - z) / (RX(0.05) * z + RX(1e-3))) * dn4 * dn2", ["jbw_normal_and_fallback"]),
    ("ill.x = uB.x / swh + pr.w * (f.x - uA.x / swh)", "ill.x = uB.x / swh + pr.w * (f.x - uA.x)", ["jbw_depth_step"]),
    ("if (swh > RX(1e-4))", "if (swh >= RX(0))", ["jbw_normal_and_fallback"]),
    ("real cpx = ((real)px + RX(0.5)) / 4 - RX(0.5)", "real cpx = ((real)px + RX(0.5)) / 4", ["jbw_coarse"]),
    ("RX(0.5) + RX(7.5) * sigmoid(o[1])", "RX(0.5) + RX(7.0) * sigmoid(o[1])", ["infer_activations"]),
    ("(RX(31.0) / RX(32.0)) * sigmoid(o[0])", "sigmoid(o[0])", ["infer_activations"]),
    ("clampr(o[4 + k], -5, 5)", "clampr(o[4 + k], -4, 4)", ["infer_activations"]),
    ("qf16(s[1]), qf16(cc[1])", "qf16(s[1]), qf16(cc[2])", ["infer_activations"]),
    ("qf16(vr / N / N)", "qf16(vr / N)", ["down_counts"]),
    ("real w = F[27];", "real w = 1;", ["down_counts"]),
    ("real kx = RX(1.25) * sd.x", "real kx = RX(1.0) * sd.x", ["taa_clip"]),
    ("res = mk3(hr.x + (cu.x - hr.x) * RX(0.1)", "res = mk3(hr.x + (cu.x - hr.x) * RX(0.2)", ["taa_clip"]),
    ("ux <= 1", "ux < 1", ["taa_uv_bounds"]),
    ("ux >= 0 && uy >= 0", "ux > 0 && uy >= 0", ["taa_uv_bounds"]),
]


def _mutant_lib(tmpdir, i, old, new):
    src = open(os.path.join(ROOT, "oracle", "hush_oracle.c")).read()
    assert src.count(old) >= 1, f"mutation {i}: {old!r} not in the oracle"
    path_c = os.path.join(tmpdir, f"m{i}.c")
    with open(path_c, "w") as fh:
        fh.write(src.replace(old, new, 1))
    path_so = os.path.join(tmpdir, f"m{i}.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=gnu11", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                    "-fvisibility=hidden", "-w", "-DHO_DOUBLE", "-DHO_QUANT=0", "-DHO_PREFIX=ho64_", path_c, "-o",
                    path_so, "-lm"], check=True)
    return ctypes.CDLL(path_so)


def test_pins_catch_planted_mistakes(tmp_path):
    missed = []
    for i, (old, new, pins) in enumerate(MUTANTS):
        p = Pins(_mutant_lib(str(tmp_path), i, old, new))
        caught = False
        for name in pins:
            try:
                getattr(p, name)()
            except AssertionError:
                caught = True
        if not caught:
            missed.append((old, new))
    assert not missed, f"planted mistakes no pin caught: {missed}"
