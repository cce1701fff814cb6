"""Second, independent restatement of the Hush passes in vectorised numpy (float64).  TEST INFRASTRUCTURE
ONLY: tests/ use it to pin the C oracle (oracle/hush_oracle.c) — two restatements written separately from
the WGSL of src/denoise/hush.js must agree.

quant=False: no binary16 storage (compare with the C oracle's FP64-ideal build, ho64_*).
quant=True: every rgba16float store rounds toward zero and qf16 truncates (compare with ho64q_*).

Images are (H, W, C) arrays.  Each function cites the WGSL it follows.
"""
from __future__ import annotations

import numpy as np

LUMA = np.array([0.2126, 0.7152, 0.0722])
HMAX = 32.0


class Q:
    """binary16 storage model (hush.js:68, 98-101)"""

    def __init__(self, quant: bool):
        self.on = quant

    def qf16(self, x):
        if not self.on:
            return np.asarray(x, dtype=np.float64)
        x = np.clip(np.asarray(x, dtype=np.float64), -65504.0, 65504.0)
        m, e = np.frexp(x)
        return np.ldexp(np.trunc(m * 2048.0) / 2048.0, e)

    def st(self, x):
        """rgba16float store: round toward zero, subnormals kept"""
        if not self.on:
            return np.asarray(x, dtype=np.float64)
        x = np.clip(np.asarray(x, dtype=np.float64), -65504.0, 65504.0)
        small = np.abs(x) < 2.0 ** -14
        sub = np.trunc(x * 2.0 ** 24) / 2.0 ** 24
        return np.where(small, sub, self.qf16(x))


def lum(c):
    return c[..., 0] * LUMA[0] + c[..., 1] * LUMA[1] + c[..., 2] * LUMA[2]


def oct_enc(n):
    """octEnc (hush.js:34-39)"""
    v = n / np.maximum(np.abs(n[..., 0]) + np.abs(n[..., 1]) + np.abs(n[..., 2]), 1e-20)[..., None]
    s = np.stack([np.where(v[..., 0] >= 0, 1.0, -1.0), np.where(v[..., 1] >= 0, 1.0, -1.0)], -1)
    folded = (1.0 - np.abs(v[..., [1, 0]])) * s
    return np.where((v[..., 2] < 0)[..., None], folded, v[..., :2])


def oct_dec(e):
    """octDec (hush.js:40-46)"""
    x, y = e[..., 0], e[..., 1]
    z = 1.0 - np.abs(x) - np.abs(y)
    t = np.maximum(-z, 0.0)
    x = x + np.where(x >= 0, -t, t)
    y = y + np.where(y >= 0, -t, t)
    v = np.stack([x, y, z], -1)
    return v / np.maximum(np.linalg.norm(v, axis=-1), 1e-20)[..., None]


def shift(a, dy, dx, fill=None):
    """a[y + dy, x + dx]: out-of-image reads give `fill` (if given) or clamp to the edge"""
    H, W = a.shape[:2]
    ys = np.arange(H) + dy
    xs = np.arange(W) + dx
    yc, xc = np.clip(ys, 0, H - 1), np.clip(xs, 0, W - 1)
    out = a[yc][:, xc]
    if fill is not None:
        bad = ((ys < 0) | (ys >= H))[:, None] | ((xs < 0) | (xs >= W))[None, :]
        out = np.where(bad.reshape(bad.shape + (1,) * (a.ndim - 2)), fill, out)
    return out


def gather(a, yy, xx):
    """a at integer coordinates, clamped to the image"""
    H, W = a.shape[:2]
    return a[np.clip(yy, 0, H - 1), np.clip(xx, 0, W - 1)]


class State:
    """The Hush history and level surfaces (hush.js:780-793)."""

    def __init__(self, W, H):
        self.W, self.H = W, H
        self.W2, self.H2 = (W + 1) // 2, (H + 1) // 2
        self.W4, self.H4 = (self.W2 + 1) // 2, (self.H2 + 1) // 2
        z = lambda w, h, c=4: np.zeros((h, w, c))
        self.A0, self.X = z(W, H), z(W, H)
        self.mom = [z(W, H), z(W, H)]
        self.key = [z(W, H), z(W, H)]
        self.hist = [z(W, H), z(W, H)]
        self.taa = [z(W, H), z(W, H)]
        self.Bh = [z(self.W2, self.H2), z(self.W2, self.H2)]
        self.par0 = [z(self.W2, self.H2), z(self.W2, self.H2)]
        self.par1, self.par2 = z(self.W2, self.H2), z(self.W2, self.H2)
        self.F = z(self.W2, self.H2, 32)
        self.C = [z(self.W4, self.H4), z(self.W4, self.H4)]
        self.CG = z(self.W4, self.H4)
        self.mask = np.zeros((H, W), np.uint8)


def check(k, n, zprev, obj):
    """check (hush.js:103-105)"""
    d = np.sum(oct_dec(k[..., :2]) * n, axis=-1)
    return (k[..., 2] >= 0) & (k[..., 3] == obj) & (np.abs(k[..., 2] - zprev) <= 0.05 * zprev + 0.01) & (d >= 0.9)


def temporal(S: State, g: dict, cur: int, reset: bool, E: float, q: Q):
    """TEMPORAL (hush.js:79-308)"""
    prev = cur ^ 1
    H, W = S.H, S.W
    rad, nd, mv, mat = g["radiance"], g["normalDepth"], g["motion"], g["material"]
    valid = nd[..., 3] >= 0
    ill = np.where(valid[..., None], rad[..., :3] / np.maximum(mat[..., :3], 1e-3), 0.0)
    va = valid.astype(np.float64)
    # 3x3 stats over in-image valid neighbours (hush.js:124-142)
    offs = [(j // 3 - 1, j % 3 - 1) for j in range(9)]
    s1 = np.zeros((H, W, 3)); cnt = np.zeros((H, W))
    for dy, dx in offs:
        s1 += shift(ill, dy, dx, 0.0) * shift(va, dy, dx, 0.0)[..., None]
        cnt += shift(va, dy, dx, 0.0)
    cnt = np.maximum(cnt, 1.0)
    mu = s1 / cnt[..., None]
    slm = lum(mu)
    sv3 = np.zeros((H, W, 3)); svar = np.zeros((H, W))
    for dy, dx in offs:
        c = shift(ill, dy, dx, 0.0); a = shift(va, dy, dx, 0.0)
        sv3 += a[..., None] * (c - mu) ** 2
        svar += a * (lum(c) - slm) ** 2
    sd = np.sqrt(np.maximum(sv3 / cnt[..., None], 1e-12))
    svar = svar / cnt
    n = nd[..., :3]
    l = lum(ill)
    yy, xx = np.mgrid[0:H, 0:W]
    has = np.zeros((H, W), bool)
    hc = np.zeros((H, W, 3)); hm = np.zeros((H, W, 4)); v = np.zeros((H, W))
    alpha = np.ones((H, W)); gamma = np.ones((H, W))
    mask = np.zeros((H, W), np.uint8)
    if not reset:
        ppx, ppy = xx + mv[..., 0], yy + mv[..., 1]
        bx, by = np.floor(ppx), np.floor(ppy)
        fx, fy = ppx - bx, ppy - by
        bx, by = bx.astype(np.int64), by.astype(np.int64)
        wts = [(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy]
        sumW = np.zeros((H, W))
        for i in range(4):
            qx, qy = bx + (i & 1), by + (i >> 1)
            inside = (qx >= 0) & (qy >= 0) & (qx < W) & (qy < H)
            ok = (wts[i] > 0) & inside & check(gather(S.key[prev], qy, qx), n, mv[..., 3], mv[..., 2])
            mask |= (ok.astype(np.uint8) << i)
            w = np.where(ok, wts[i], 0.0)
            sumW += w
            hc += w[..., None] * gather(S.hist[prev], qy, qx)[..., :3]
            hm += w[..., None] * gather(S.mom[prev], qy, qx)
        good = sumW >= 0.01
        hc = np.where(good[..., None], hc / np.where(good, sumW, 1.0)[..., None], 0.0)
        hm = np.where(good[..., None], hm / np.where(good, sumW, 1.0)[..., None], 0.0)
        v = np.where(good, sumW, 0.0)
        has = good.copy()
        cx, cy = np.floor(ppx + 0.5).astype(np.int64), np.floor(ppy + 0.5).astype(np.int64)
        # 3x3 fallback of consistent texels where the bilinear weights failed
        n3 = np.zeros((H, W)); fc = np.zeros((H, W, 3)); fm = np.zeros((H, W, 4))
        for dy, dx in offs:
            qx, qy = cx + dx, cy + dy
            inside = (qx >= 0) & (qy >= 0) & (qx < W) & (qy < H)
            ok = inside & check(gather(S.key[prev], qy, qx), n, mv[..., 3], mv[..., 2])
            n3 += ok
            fc += np.where(ok[..., None], gather(S.hist[prev], qy, qx)[..., :3], 0.0)
            fm += np.where(ok[..., None], gather(S.mom[prev], qy, qx), 0.0)
        use_fb = ~good & (n3 > 0)
        nn = np.where(use_fb, n3, 1.0)[..., None]
        hc = np.where(use_fb[..., None], fc / nn, hc)
        hm = np.where(use_fb[..., None], fm / nn, hm)
        v = np.where(use_fb, 0.1, v)
        has = good | use_fb
        mask |= (use_fb.astype(np.uint8) << 4)
        pq = gather(S.par0[prev], np.clip(cy, 0, H - 1) >> 1, np.clip(cx, 0, W - 1) >> 1)
        alpha, gamma = pq[..., 0], pq[..., 1]
    has = has & (hm[..., 3] > 0)
    # blend (hush.js:189-211)
    len1 = np.where(has, np.minimum(hm[..., 3] + 1.0, HMAX), 1.0)
    a = np.maximum(alpha, 1.0 / len1)
    lo, hi = mu - gamma[..., None] * sd, mu + gamma[..., None] * sd
    hcl = np.minimum(np.maximum(hc, lo), hi)
    color = np.where(has[..., None], hcl + (ill - hcl) * a[..., None], ill)
    nmx = np.where(has, hm[..., 0] + (l - hm[..., 0]) * a, l)
    nmy = np.where(has, hm[..., 1] + (l * l - hm[..., 1]) * a, l * l)
    nfast = np.where(has, hm[..., 2] + (l - hm[..., 2]) * np.maximum(1.0 / 6.0, 1.0 / len1), l)
    lh = lum(hc)
    lag = np.where(has, np.abs(slm - lh) / (np.sqrt(np.maximum(svar / cnt, 1e-12)) + 0.02 * lh + 1e-3), 0.0)
    v = np.where(has, v, 0.0)
    tvar = np.maximum(nmy - nmx * nmx, 0.0)
    A = np.concatenate([color, (svar + (tvar - svar) * np.minimum(len1 / 4.0, 1.0))[..., None]], -1)
    oc = oct_enc(n)
    sky = ~valid
    S.A0 = np.where(sky[..., None], 0.0, q.st(A))
    S.mom[cur] = np.where(sky[..., None], 0.0, q.st(np.stack([nmx, nmy, nfast, len1], -1)))
    key = np.stack([oc[..., 0], oc[..., 1], nd[..., 3], mv[..., 2]], -1)
    key_sky = np.stack([np.zeros((H, W)), np.zeros((H, W)), -np.ones((H, W)), mv[..., 2]], -1)
    S.key[cur] = q.st(np.where(sky[..., None], key_sky, key))
    mask = np.where(sky, 0x80, mask | (has.astype(np.uint8) << 5)).astype(np.uint8)
    S.mask = mask
    # feature inputs (hush.js:212-222) from the stored values
    Aq = q.qf16(A)
    m1 = E * q.qf16(nmx)
    m2 = E * E * q.qf16(nmy)
    P = np.zeros((H, W, 20))
    P[..., 0] = E * lum(Aq); P[..., 1] = E * q.qf16(l); P[..., 2] = np.maximum(m2 - m1 * m1, 0.0); P[..., 3] = len1
    P[..., 4] = q.qf16(v); P[..., 5] = q.qf16(lag); P[..., 6] = E * q.qf16(nfast); P[..., 7] = m1
    P[..., 8] = q.qf16(oc[..., 0]); P[..., 9] = q.qf16(oc[..., 1]); P[..., 10] = q.qf16(nd[..., 3])
    P[..., 11] = q.qf16(np.hypot(mv[..., 0], mv[..., 1]))
    P[..., 12:16] = mat; P[..., 16:20] = Aq
    skyP = np.zeros(20); skyP[10] = -1.0
    P = np.where(sky[..., None], skyP, P)
    # block features (hush.js:226-282): pad to even size with sky
    H2, W2 = S.H2, S.W2
    Pp = np.broadcast_to(skyP, (2 * H2, 2 * W2, 20)).copy()
    Pp[:H, :W] = P
    blocks = [Pp[(j >> 1)::2, (j & 1)::2] for j in range(4)]
    ok = [b[..., 10] >= 0 for b in blocks]
    nv = sum(o.astype(np.float64) for o in ok)
    sums = sum(np.where(o[..., None], b, 0.0) for o, b in zip(ok, blocks))
    nsj = [oct_dec(b[..., 8:10]) for b in blocks]
    Nsum = sum(np.where(o[..., None], ns, 0.0) for o, ns in zip(ok, nsj))
    k = 1.0 / np.where(nv > 0, nv, 1.0)
    mean = sums * k[..., None]
    Nm = Nsum * k[..., None]
    Lacc = mean[..., 0]
    VA = sum(np.where(o, (b[..., 0] - Lacc) ** 2, 0.0) for o, b in zip(ok, blocks)) * k
    # 4x4 neighbourhood (block +-1) of the exposed truncated luminance over valid texels
    lq = np.where(valid, E * q.qf16(lum(ill)), 0.0)
    lq_p = np.zeros((2 * H2 + 2, 2 * W2 + 2)); lq_p[1:H + 1, 1:W + 1] = lq
    va_p = np.zeros((2 * H2 + 2, 2 * W2 + 2)); va_p[1:H + 1, 1:W + 1] = va
    s1b = np.zeros((H2, W2)); n4 = np.zeros((H2, W2))
    win = [(j >> 2, j & 3) for j in range(16)]
    for oy, ox in win:
        s1b += va_p[oy::2][:H2, ox::2][:, :W2] * lq_p[oy::2][:H2, ox::2][:, :W2]
        n4 += va_p[oy::2][:H2, ox::2][:, :W2]
    mu4 = s1b / np.maximum(n4, 1.0)
    var4 = np.zeros((H2, W2))
    for oy, ox in win:
        d = lq_p[oy::2][:H2, ox::2][:, :W2] - mu4
        var4 += va_p[oy::2][:H2, ox::2][:, :W2] * d * d
    var4 = var4 / np.maximum(n4, 1.0)
    Nn = Nm / np.maximum(np.linalg.norm(Nm, axis=-1), 1e-6)[..., None]
    zs = [b[..., 10] for b in blocks]
    zmax = np.max(np.stack([np.where(o, z, -1e30) for o, z in zip(ok, zs)]), 0)
    zmin = np.min(np.stack([np.where(o, z, 1e30) for o, z in zip(ok, zs)]), 0)
    mind = np.min(np.stack([np.where(o, np.sum(ns * Nn, -1), 1e30) for o, ns in zip(ok, nsj)]), 0)
    dz = np.clip((zmax - zmin) / np.maximum(zmin, 1e-6), 0.0, 1.0)
    dn = np.clip(1.0 - mind, 0.0, 2.0)
    ocb = oct_enc(Nn)
    lg = lambda x: np.log(1.0 + x)
    f = np.stack([lg(Lacc), lg(mean[..., 1]), lg(mean[..., 2]), lg(var4), mean[..., 3] / HMAX, mean[..., 4],
                  lg(mean[..., 5]), lg(np.abs(mean[..., 6] - mean[..., 7]) / (mean[..., 7] + 1e-3)), ocb[..., 0],
                  ocb[..., 1], dz, dn, mean[..., 12], mean[..., 13], mean[..., 14], mean[..., 15], 1.0 - nv / 4.0,
                  lg(mean[..., 11]), lg(mu4), lg(VA)], -1)
    F = np.zeros((H2, W2, 32))
    F[..., :20] = f
    F[..., 20:24] = mean[..., 16:20]
    F[..., 24], F[..., 25] = ocb[..., 0], ocb[..., 1]
    F[..., 26] = mean[..., 10]
    F[..., 27] = nv
    F = q.st(q.qf16(F))
    empty = nv == 0
    Fsky = np.zeros(32); Fsky[26] = -1.0
    S.F = np.where(empty[..., None], Fsky, F)
    S.F[..., 28:] = 0.0


def filt(A, G, R, step, prefilter, s, c):
    """filt (atrousFn, hush.js:431-476) over whole images; A: (rgb, var), G: (normal xyz, depth)"""
    H, W = A.shape[:2]
    h = [0.25, 0.5, 0.25] if R == 1 else [1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16]
    gp, center = G, A
    if prefilter:
        av = A[..., 3]
        gv = 0.0625 * (shift(av, -1, -1) + shift(av, -1, 1) + shift(av, 1, -1) + shift(av, 1, 1)) \
            + 0.125 * (shift(av, -1, 0) + shift(av, 0, -1) + shift(av, 0, 1) + shift(av, 1, 0)) + 0.25 * av
    else:
        gv = center[..., 3]
    lc = lum(center)
    sl = s / (np.sqrt(np.maximum(gv, 1e-12)) + 1e-6)
    z = G[..., 3]
    zl, zr, zu, zd = shift(z, 0, -1), shift(z, 0, 1), shift(z, -1, 0), shift(z, 1, 0)
    gx = np.minimum(np.where(zr > 0, np.abs(zr - z), 1e9), np.where(zl > 0, np.abs(zl - z), 1e9))
    gy = np.minimum(np.where(zd > 0, np.abs(zd - z), 1e9), np.where(zu > 0, np.abs(zu - z), 1e9))
    zg = np.minimum(np.maximum(gx, gy), z)
    cw = h[R] * h[R] * c
    sw = cw.copy()
    sc = center[..., :3] * cw[..., None]
    sv = center[..., 3] * cw * cw
    yy, xx = np.mgrid[0:H, 0:W]
    for dy in range(-R, R + 1):
        for dx in range(-R, R + 1):
            if dx == 0 and dy == 0:
                continue
            oy, ox = dy * step, dx * step
            gq = shift(G, oy, ox)
            inside = (yy + oy >= 0) & (yy + oy < H) & (xx + ox >= 0) & (xx + ox < W)
            ok = inside & (gq[..., 3] >= 0)
            cq = shift(A, oy, ox)
            iz = 1.0 / (zg * (np.sqrt(dx * dx + dy * dy) * step) + 1e-3)
            w = h[dx + R] * h[dy + R] * np.maximum(np.sum(gq[..., :3] * gp[..., :3], -1), 0.0) ** 128 \
                * np.exp(-np.abs(gq[..., 3] - z) * iz - sl * np.abs(lum(cq) - lc))
            w = np.where(ok, w, 0.0)
            sw += w
            sc += w[..., None] * cq[..., :3]
            sv += w * w * cq[..., 3]
    return np.concatenate([sc / sw[..., None], (sv / (sw * sw))[..., None]], -1)


def guide_F(F):
    return np.concatenate([oct_dec(F[..., 24:26]), F[..., 26:27]], -1)


def infer(S: State, cur: int, w: np.ndarray, q: Q):
    """INFER (hush.js:312-368): MLP, activations, mid1"""
    hdim = 16
    W1 = w[:hdim * 20].reshape(hdim, 20); b1 = w[hdim * 20:hdim * 21]
    o = hdim * 21
    W2 = w[o:o + hdim * hdim].reshape(hdim, hdim); b2 = w[o + hdim * hdim:o + hdim * hdim + hdim]
    o = o + hdim * hdim + hdim
    W3 = w[o:o + 12 * hdim].reshape(12, hdim); b3 = w[o + 12 * hdim:o + 12 * hdim + 12]
    f = S.F[..., :20]
    h1 = np.maximum(f @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    out = h2 @ W3.T + b3
    sig = lambda x: 1.0 / (1.0 + np.exp(-x))
    alpha = 1 / 32 + 31 / 32 * sig(out[..., 0]); gamma = 0.5 + 7.5 * sig(out[..., 1])
    beta, kappa = sig(out[..., 2]), sig(out[..., 3])
    s = 0.25 * np.exp(np.clip(out[..., 4:8], -5, 5))
    c = np.exp(np.clip(out[..., 8:12], -4, 4))
    G = guide_F(S.F)
    A = S.F[..., 20:24]
    sky = G[..., 3] < 0
    mid = filt(A, G, 1, 1, True, q.qf16(s[..., 1]), q.qf16(c[..., 1]))
    Z = lambda x: np.where(sky[..., None], 0.0, x)
    S.par0[cur] = Z(q.st(np.stack([alpha, gamma, beta, kappa], -1)))
    S.par1 = Z(q.st(s))
    S.par2 = Z(q.st(c))
    S.Bh[0] = Z(q.st(mid))


def mid(S: State, step: int, comp: int, src: int, dst: int, q: Q):
    G = guide_F(S.F)
    out = filt(S.Bh[src], G, 1, step, False, S.par1[..., comp], S.par2[..., comp])
    S.Bh[dst] = np.where((G[..., 3] < 0)[..., None], 0.0, q.st(out))


def coarse1(S: State, q: Q):
    """COARSE1 (hush.js:372-422) with down() (:385-402)"""
    H4, W4, H2, W2 = S.H4, S.W4, S.H2, S.W2
    Fp = np.zeros((2 * H4, 2 * W4, 32))
    Fp[:H2, :W2] = S.F
    inimg = np.zeros((2 * H4, 2 * W4), bool); inimg[:H2, :W2] = True
    N = np.zeros((H4, W4)); col = np.zeros((H4, W4, 3)); vr = np.zeros((H4, W4)); nrm = np.zeros((H4, W4, 3)); z = np.zeros((H4, W4))
    for j in range(4):
        b = Fp[(j >> 1)::2, (j & 1)::2]
        ok = inimg[(j >> 1)::2, (j & 1)::2] & (b[..., 27] > 0)
        wgt = np.where(ok, b[..., 27], 0.0)
        N += wgt
        col += wgt[..., None] * b[..., 20:23]
        vr += wgt * b[..., 23]
        nrm += wgt[..., None] * oct_dec(b[..., 24:26])
        z += wgt * b[..., 26]
    empty = N == 0
    Ns = np.where(empty, 1.0, N)
    o = oct_enc(nrm / np.maximum(np.linalg.norm(nrm, axis=-1), 1e-6)[..., None])
    c = q.qf16(np.concatenate([col / Ns[..., None], (vr / Ns / Ns)[..., None]], -1))
    g = np.stack([q.qf16(o[..., 0]), q.qf16(o[..., 1]), q.qf16(z / Ns), np.zeros((H4, W4))], -1)
    c = np.where(empty[..., None], 0.0, c)
    g = np.where(empty[..., None], np.array([0.0, 0.0, -1.0, 0.0]), g)
    S.CG = q.st(g)
    G = np.concatenate([oct_dec(g[..., :2]), g[..., 2:3]], -1)
    out = filt(c, G, 2, 1, True, np.full((H4, W4), 0.25), np.ones((H4, W4)))
    S.C[1] = np.where((G[..., 3] < 0)[..., None], 0.0, q.st(out))


def coarse2(S: State, q: Q):
    G = np.concatenate([oct_dec(S.CG[..., :2]), S.CG[..., 2:3]], -1)
    out = filt(S.C[1], G, 2, 2, True, np.full(G.shape[:2], 0.25), np.ones(G.shape[:2]))
    S.C[0] = np.where((G[..., 3] < 0)[..., None], 0.0, q.st(out))


def jbw(fx, fy, j, nq, zq, n, z):
    """jbw (hush.js:552-558)"""
    wb = (fx if (j & 1) else 1 - fx) * (fy if (j >> 1) else 1 - fy)
    dn = np.maximum(np.sum(nq * n, -1), 0.0)
    return wb * np.exp(-np.abs(zq - z) / (0.05 * z + 1e-3)) * dn ** 8


def fine_final(S: State, g: dict, cur: int, q: Q):
    """FINE_FINAL (hush.js:530-622)"""
    H, W = S.H, S.W
    key = S.key[cur]
    G = np.concatenate([oct_dec(key[..., :2]), key[..., 2:3]], -1)
    yy, xx = np.mgrid[0:H, 0:W]
    hy, hx = yy >> 1, xx >> 1
    p1, p2, pr = S.par1[hy, hx], S.par2[hy, hx], S.par0[cur][hy, hx]
    A1 = filt(S.A0, G, 1, 1, True, p1[..., 0], p2[..., 0])
    f = q.qf16(A1[..., :3])
    n, z = G[..., :3], G[..., 3]
    # half-res mid level
    hpx, hpy = (xx + 0.5) / 2 - 0.5, (yy + 0.5) / 2 - 0.5
    hbx, hby = np.floor(hpx), np.floor(hpy)
    uB = np.zeros((H, W, 3)); uA = np.zeros((H, W, 3)); swh = np.zeros((H, W))
    for j in range(4):
        qx, qy = hbx.astype(int) + (j & 1), hby.astype(int) + (j >> 1)
        ok = (qx >= 0) & (qy >= 0) & (qx < S.W2) & (qy < S.H2)
        Fq = gather(S.F, qy, qx)
        ok &= Fq[..., 26] >= 0
        w = np.where(ok, jbw(hpx - hbx, hpy - hby, j, oct_dec(Fq[..., 24:26]), Fq[..., 26], n, z), 0.0)
        swh += w
        uB += w[..., None] * gather(S.Bh[0], qy, qx)[..., :3]
        uA += w[..., None] * Fq[..., 20:23]
    sw_ = np.where(swh > 1e-4, swh, 1.0)[..., None]
    ill = np.where((swh > 1e-4)[..., None], uB / sw_ + pr[..., 3:4] * (f - uA / sw_), f)
    cpx, cpy = (xx + 0.5) / 4 - 0.5, (yy + 0.5) / 4 - 0.5
    cbx, cby = np.floor(cpx), np.floor(cpy)
    uC = np.zeros((H, W, 3)); swq = np.zeros((H, W))
    for j in range(4):
        qx, qy = cbx.astype(int) + (j & 1), cby.astype(int) + (j >> 1)
        ok = (qx >= 0) & (qy >= 0) & (qx < S.W4) & (qy < S.H4)
        cg = gather(S.CG, qy, qx)
        ok &= cg[..., 2] >= 0
        w = np.where(ok, jbw(cpx - cbx, cpy - cby, j, oct_dec(cg[..., :2]), cg[..., 2], n, z), 0.0)
        swq += w
        uC += w[..., None] * gather(S.C[0], qy, qx)[..., :3]
    up = np.where((swq > 1e-4)[..., None], uC / np.where(swq > 1e-4, swq, 1.0)[..., None], ill)
    ill = ill + (up - ill) * pr[..., 2:3]
    alb, emi = g["material"][..., :3], g["emission"][..., :3]
    sky = (z < 0)[..., None]
    S.hist[cur] = np.where(sky, 0.0, q.st(A1))
    X = np.concatenate([ill * alb + emi, np.zeros((H, W, 1))], -1)
    S.X = q.st(np.where(sky, np.concatenate([emi, np.zeros((H, W, 1))], -1), X))


def ycocg(c):
    return np.stack([c[..., 0] * 0.25 + c[..., 1] * 0.5 + c[..., 2] * 0.25, c[..., 0] * 0.5 - c[..., 2] * 0.5,
                     -c[..., 0] * 0.25 + c[..., 1] * 0.5 - c[..., 2] * 0.25], -1)


def from_ycocg(c):
    return np.stack([c[..., 0] + c[..., 1] - c[..., 2], c[..., 0] + c[..., 2], c[..., 0] - c[..., 1] - c[..., 2]], -1)


def tm(c):
    return c / (1.0 + lum(c))[..., None]


def itm(c):
    return c / np.maximum(1.0 - lum(c), 1e-4)[..., None]


def taa(S: State, g: dict, cur: int, reset: bool, q: Q):
    """TAA (hush.js:625-678)"""
    prev = cur ^ 1
    H, W = S.H, S.W
    cu = tm(S.X[..., :3])
    one = np.ones((H, W, 1))
    if reset:
        S.taa[cur] = q.st(np.concatenate([itm(cu), one], -1))
        return
    Y = ycocg(tm(S.X[..., :3]))
    offs = [(j // 3 - 1, j % 3 - 1) for j in range(9)]
    m1 = sum(shift(Y, dy, dx) for dy, dx in offs) / 9.0
    v3 = sum((shift(Y, dy, dx) - m1) ** 2 for dy, dx in offs)
    sd = np.sqrt(np.maximum(v3 / 9.0, 1e-12))
    mv = g["motion"]
    yy, xx = np.mgrid[0:H, 0:W]
    ux, uy = (xx + 0.5 + mv[..., 0]) / W, (yy + 0.5 + mv[..., 1]) / H
    inb = (ux >= 0) & (uy >= 0) & (ux <= 1) & (uy <= 1)
    ppx, ppy = ux * W - 0.5, uy * H - 0.5
    bx, by = np.floor(ppx), np.floor(ppy)
    fx, fy = ppx - bx, ppy - by
    hist = np.zeros((H, W, 3))
    for j in range(4):
        ox, oy = j & 1, j >> 1
        wb = (fx if ox else 1 - fx) * (fy if oy else 1 - fy)
        hist += wb[..., None] * gather(S.taa[prev], by.astype(int) + oy, bx.astype(int) + ox)[..., :3]
    hh = np.clip(ycocg(tm(hist)), m1 - 1.25 * sd, m1 + 1.25 * sd)
    hr = from_ycocg(hh)
    res = np.where(inb[..., None], hr + (cu - hr) * 0.1, cu)
    S.taa[cur] = q.st(np.concatenate([itm(res), one], -1))


class Hush:
    """State machine of the Hush class (hush.js:725-875) over numpy arrays."""

    def __init__(self, W, H, weights_f16: np.ndarray, exposure=1.0, quant=False):
        self.S = State(W, H)
        self.w = weights_f16.view(np.float16).astype(np.float64)
        assert self.w.size == 812
        self.E = exposure
        self.q = Q(quant)
        self.frame = 0
        self.reset_history = True

    def denoise(self, g16: dict) -> np.ndarray:
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):   # masked lanes (sky) only
            return self._denoise(g16)

    def _denoise(self, g16: dict) -> np.ndarray:
        g = {k: v.view(np.float16).astype(np.float64) for k, v in g16.items() if k != "clean"}
        cur = self.frame & 1
        S, q = self.S, self.q
        temporal(S, g, cur, self.reset_history, self.E, q)
        infer(S, cur, self.w, q)
        mid(S, 2, 2, 0, 1, q)
        mid(S, 4, 3, 1, 0, q)
        coarse1(S, q)
        coarse2(S, q)
        fine_final(S, g, cur, q)
        taa(S, g, cur, self.reset_history, q)
        self.reset_history = False
        self.frame += 1
        return S.taa[cur].copy()
