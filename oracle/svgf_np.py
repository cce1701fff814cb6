"""Second, independent restatement of the SVGF baseline (src/denoise/svgf.js) in vectorised numpy (float64).
TEST INFRASTRUCTURE ONLY: tests/ use it to pin the C oracle (oracle/svgf_oracle.c) — two restatements
written separately from the WGSL must agree.

quant=False: no binary16 storage (compare with the C oracle's FP64-ideal build, ho64_svgf_*).
quant=True: every rgba16float store rounds toward zero (compare with ho64q_svgf_*).
Images are (H, W, 4) float64 arrays; each function cites the WGSL it follows.
"""
from __future__ import annotations

import numpy as np

from .hush_np import Q, lum

DEFAULTS = dict(sigmaL=4.0, sigmaN=128.0, sigmaZ=1.0, alphaColor=0.2, alphaMoments=0.2, taaAlpha=0.1, taa=True,
                iterations=5)
KEYS = ("radiance", "normalDepth", "motion", "material", "emission", "prevNormalDepth", "prevMotion")


def h2f(u):
    return np.asarray(u, dtype=np.uint16).view(np.float16).astype(np.float64)


def gather(img, qx, qy):
    """img[qy, qx] with indices clipped into the image (callers mask what lies outside)"""
    H, W = img.shape[:2]
    return img[np.clip(qy, 0, H - 1), np.clip(qx, 0, W - 1)]


def inside(qx, qy, W, H):
    return (qx >= 0) & (qy >= 0) & (qx < W) & (qy < H)


class SVGF:
    """the SVGF class state machine (svgf.js:273-379)"""

    def __init__(self, W, H, quant=True, **opts):
        self.W, self.H = W, H
        self.o = dict(DEFAULTS, **opts)
        self.q = Q(quant)
        z = lambda: np.zeros((H, W, 4))
        self.histColor, self.histMoments, self.atrous, self.taa = [z(), z()], [z(), z()], [z(), z()], [z(), z()]
        self.tmpColor = z()
        self.frame, self.reset_history = 0, True
        self.ys, self.xs = np.mgrid[0:H, 0:W].astype(np.float64)
        self.iy, self.ix = np.mgrid[0:H, 0:W]

    def st(self, x):
        return self.q.st(x)

    # TEMPORAL (svgf.js:32-105)
    def temporal(self, g, cur, prev):
        nd, mv, alb, rad = g["normalDepth"], g["motion"], g["material"][..., :3], g["radiance"][..., :3]
        pnd, pmv = g["prevNormalDepth"], g["prevMotion"]
        W, H = self.W, self.H
        illum = rad / np.maximum(alb, 1e-3)
        l = lum(illum)
        ppx = self.xs + 0.5 + mv[..., 0] - 0.5
        ppy = self.ys + 0.5 + mv[..., 1] - 0.5
        bx, by = np.floor(ppx).astype(np.int64), np.floor(ppy).astype(np.int64)
        fx, fy = ppx - np.floor(ppx), ppy - np.floor(ppy)
        zexp, obj = mv[..., 3], mv[..., 2]

        def consistent(qx, qy):
            ok = inside(qx, qy, W, H)
            q = gather(pnd, qx, qy)
            ok &= q[..., 3] >= 0
            ok &= gather(pmv, qx, qy)[..., 2] == obj
            ok &= ~(np.abs(q[..., 3] - zexp) > 0.05 * zexp + 0.01)
            ok &= ~((q[..., :3] * nd[..., :3]).sum(-1) < 0.9)
            return ok

        pc, pm = self.histColor[prev], self.histMoments[prev]
        sumW = np.zeros((H, W))
        hc, hm = np.zeros((H, W, 3)), np.zeros((H, W, 3))
        if not self.reset_history:
            for (ox, oy), wt in (((0, 0), (1 - fx) * (1 - fy)), ((1, 0), fx * (1 - fy)), ((0, 1), (1 - fx) * fy),
                                 ((1, 1), fx * fy)):
                qx, qy = bx + ox, by + oy
                use = consistent(qx, qy) & (wt > 0)
                w = np.where(use, wt, 0.0)
                sumW = sumW + w
                hc = hc + w[..., None] * np.where(use[..., None], gather(pc, qx, qy)[..., :3], 0.0)
                hm = hm + w[..., None] * np.where(use[..., None], gather(pm, qx, qy)[..., :3], 0.0)
            fb = sumW < 0.01
            cx, cy = np.rint(ppx).astype(np.int64), np.rint(ppy).astype(np.int64)
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    qx, qy = cx + dx, cy + dy
                    use = fb & consistent(qx, qy)
                    sumW = sumW + use
                    hc = hc + np.where(use[..., None], gather(pc, qx, qy)[..., :3], 0.0)
                    hm = hm + np.where(use[..., None], gather(pm, qx, qy)[..., :3], 0.0)
        has = sumW >= 0.01
        sw = np.where(has, sumW, 1.0)[..., None]
        hc, hm = hc / sw, hm / sw
        length = np.where(has, np.minimum(hm[..., 2] + 1.0, 64.0), 1.0)
        ac = np.maximum(self.o["alphaColor"], 1.0 / length)[..., None]
        am = np.maximum(self.o["alphaMoments"], 1.0 / length)[..., None]
        mom = np.stack([l, l * l], -1)
        color = np.where(has[..., None], hc * (1 - ac) + illum * ac, illum)
        mom = np.where(has[..., None], hm[..., :2] * (1 - am) + mom * am, mom)
        bg = (nd[..., 3] < 0)[..., None]
        zero = np.zeros((H, W, 1))
        self.tmpColor = self.st(np.where(bg, 0.0, np.concatenate([color, zero], -1)))
        self.histMoments[cur] = self.st(np.where(bg, 0.0, np.concatenate([mom, length[..., None], zero], -1)))

    # VARIANCE (svgf.js:114-148)
    def variance(self, g, cur):
        nd = g["normalDepth"]
        c, m = self.tmpColor, self.histMoments[cur]
        H, W = self.H, self.W
        sw = np.zeros((H, W))
        sc, sm = np.zeros((H, W, 3)), np.zeros((H, W, 2))
        for dy in range(-3, 4):
            for dx in range(-3, 4):
                qx, qy = self.ix + dx, self.iy + dy
                ndq = gather(nd, qx, qy)
                use = inside(qx, qy, W, H) & (ndq[..., 3] >= 0)
                wz = np.exp(-np.abs(ndq[..., 3] - nd[..., 3]) / (0.02 * nd[..., 3] * np.hypot(dx, dy) + 1e-3))
                wn = np.maximum(0.0, (ndq[..., :3] * nd[..., :3]).sum(-1)) ** self.o["sigmaN"]
                w = np.where(use, wz * wn, 0.0)
                sw = sw + w
                sc = sc + w[..., None] * np.where(use[..., None], gather(c, qx, qy)[..., :3], 0.0)
                sm = sm + w[..., None] * np.where(use[..., None], gather(m, qx, qy)[..., :2], 0.0)
        d = np.maximum(sw, 1e-6)[..., None]
        sc, sm = sc / d, sm / d
        v_sp = np.maximum(0.0, sm[..., 1] - sm[..., 0] ** 2) * 4.0 / np.maximum(m[..., 2], 1.0)
        v_t = np.maximum(0.0, m[..., 1] - m[..., 0] ** 2)
        long = (m[..., 2] >= 4.0)[..., None]
        out = np.where(long, np.concatenate([c[..., :3], v_t[..., None]], -1),
                       np.concatenate([sc, v_sp[..., None]], -1))
        self.atrous[0] = self.st(np.where((nd[..., 3] < 0)[..., None], 0.0, out))

    # ATROUS (svgf.js:157-216)
    def atrous_pass(self, g, it, src, cur):
        nd = g["normalDepth"]
        inp = self.atrous[src]
        H, W = self.H, self.W
        step = 1 << it
        k3 = (0.25, 0.5, 0.25)
        gv = np.zeros((H, W))
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                gv = gv + k3[dx + 1] * k3[dy + 1] * gather(inp, self.ix + dx, self.iy + dy)[..., 3]
        lp = lum(inp[..., :3])
        denomL = self.o["sigmaL"] * np.sqrt(np.maximum(gv, 0.0)) + 1e-6
        z = nd[..., 3]
        zs = [gather(nd, self.ix + ox, self.iy + oy)[..., 3] for ox, oy in ((-1, 0), (1, 0), (0, -1), (0, 1))]
        d = [np.where(zz > 0, np.abs(zz - z), 1e9) for zz in zs]
        zg = np.minimum(np.maximum(np.minimum(d[1], d[0]), np.minimum(d[3], d[2])), z)
        hk = (1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16)
        sw = np.full((H, W), hk[2] * hk[2])
        sc = inp[..., :3] * sw[..., None]
        sv = inp[..., 3] * sw * sw
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                if dx == 0 and dy == 0:
                    continue
                qx, qy = self.ix + dx * step, self.iy + dy * step
                ndq, cq = gather(nd, qx, qy), gather(inp, qx, qy)
                use = inside(qx, qy, W, H) & (ndq[..., 3] >= 0)
                dist = np.hypot(dx, dy) * step
                wz = np.abs(ndq[..., 3] - z) / (self.o["sigmaZ"] * zg * dist + 1e-3)
                wn = np.maximum(0.0, (ndq[..., :3] * nd[..., :3]).sum(-1)) ** self.o["sigmaN"]
                wl = np.abs(lum(cq[..., :3]) - lp) / denomL
                w = np.where(use, hk[dx + 2] * hk[dy + 2] * wn * np.exp(-wz - wl), 0.0)
                sw = sw + w
                sc = sc + w[..., None] * np.where(use[..., None], cq[..., :3], 0.0)
                sv = sv + w * w * np.where(use, cq[..., 3], 0.0)
        o = np.concatenate([sc / sw[..., None], (sv / (sw * sw))[..., None]], -1)
        bg = (nd[..., 3] < 0)[..., None]
        self.atrous[src ^ 1] = self.st(np.where(bg, 0.0, o))
        if it == 0:
            fb = o.copy()
            fb[..., 3] = 0.0
            self.histColor[cur] = self.st(np.where(bg, 0.0, fb))

    # RESOLVE (svgf.js:228-270)
    def resolve(self, g, src, cur, prev):
        illum, mat, emi = self.atrous[src], g["material"], g["emission"]
        H, W = self.H, self.W

        def tm(c):
            return c / (1.0 + lum(c))[..., None]

        def itm(c):
            return c / np.maximum(1.0 - lum(c), 1e-4)[..., None]

        def ycocg(c):
            r, gg, b = c[..., 0], c[..., 1], c[..., 2]
            return np.stack([r * 0.25 + gg * 0.5 + b * 0.25, r * 0.5 - b * 0.5, -r * 0.25 + gg * 0.5 - b * 0.25], -1)

        def remod(ox, oy):
            qx, qy = self.ix + ox, self.iy + oy
            return gather(illum, qx, qy)[..., :3] * gather(mat, qx, qy)[..., :3] + gather(emi, qx, qy)[..., :3]

        cu = tm(remod(0, 0))
        one = np.ones((H, W, 1))
        if not self.o["taa"]:
            self.taa[cur] = self.st(np.concatenate([itm(cu), one], -1))
            return
        m1, m2 = np.zeros((H, W, 3)), np.zeros((H, W, 3))
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                c = ycocg(tm(remod(dx, dy)))
                m1, m2 = m1 + c, m2 + c * c
        m1, m2 = m1 / 9.0, m2 / 9.0
        sd = np.sqrt(np.maximum(m2 - m1 * m1, 0.0))
        mv = g["motion"]
        ux, uy = (self.xs + 0.5 + mv[..., 0]) / W, (self.ys + 0.5 + mv[..., 1]) / H
        ok = (ux >= 0) & (uy >= 0) & (ux <= 1) & (uy <= 1) & (not self.reset_history)
        # linear sampler with clamp-to-edge: bilinear weights at uv * size - 0.5
        px, py = ux * W - 0.5, uy * H - 0.5
        bx, by = np.floor(px), np.floor(py)
        fx, fy = (px - bx)[..., None], (py - by)[..., None]
        bxi, byi = bx.astype(np.int64), by.astype(np.int64)
        t = self.taa[prev]
        hist = ((1 - fx) * (1 - fy) * gather(t, bxi, byi)[..., :3] + fx * (1 - fy) * gather(t, bxi + 1, byi)[..., :3]
                + (1 - fx) * fy * gather(t, bxi, byi + 1)[..., :3] + fx * fy * gather(t, bxi + 1, byi + 1)[..., :3])
        h = np.clip(ycocg(tm(hist)), m1 - 1.25 * sd, m1 + 1.25 * sd)
        fr = np.stack([h[..., 0] + h[..., 1] - h[..., 2], h[..., 0] + h[..., 2], h[..., 0] - h[..., 1] - h[..., 2]], -1)
        a = self.o["taaAlpha"]
        res = np.where(ok[..., None], fr * (1 - a) + cu * a, cu)
        self.taa[cur] = self.st(np.concatenate([itm(res), one], -1))

    def encode(self, gbits: dict) -> np.ndarray:
        g = {k: h2f(gbits[k]) for k in KEYS}
        cur = self.frame & 1
        prev = cur ^ 1
        self.temporal(g, cur, prev)
        self.variance(g, cur)
        src = 0
        for i in range(int(self.o["iterations"])):
            self.atrous_pass(g, i, src, cur)
            src ^= 1
        self.resolve(g, src, cur, prev)
        self.reset_history = False
        self.frame += 1
        return self.taa[cur]
