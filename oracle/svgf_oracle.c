/*
 * CPU oracle of the SVGF baseline (src/denoise/svgf.js, Schied et al. 2017 + TAA resolve).  TEST
 * INFRASTRUCTURE ONLY: only tests/ and the smoke check call it, as the checker.
 *
 * Same build scheme as hush_oracle.c (FP32 mirror with binary16 round-toward-zero storage; FP64 with the
 * same storage; FP64 ideal).  Parity notes:
 *   - the resolve samples the previous output with a linear sampler (svgf.js:248); the sampler's
 *     sub-texel weight precision is implementation-defined in WebGPU, so this oracle uses exact
 *     bilinear weights at uv * size - 0.5 with clamp-to-edge (parity unpinned at weight rounding);
 *   - round() is WGSL's round-half-to-even (svgf.js:74);
 *   - pow(x, sigmaN) (WGSL bounds its error only, as that of exp2(y log2 x)): an integer exponent
 *     1 <= n < 65536 is binary exponentiation in the build's arithmetic (x^128 = seven squarings, the form
 *     of hush.js:70-72 pow128; within WGSL's bound), other exponents the C library's pow.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef HO_DOUBLE
typedef float real;
#define RX(x) x##f
#define EXP expf
#define SQRT sqrtf
#define FLOOR floorf
#define FABS fabsf
#define POW powf
#define RINT rintf
#else
typedef double real;
#define RX(x) x
#define EXP exp
#define SQRT sqrt
#define FLOOR floor
#define FABS fabs
#define POW pow
#define RINT rint
#endif
#ifndef HO_QUANT
#define HO_QUANT 1
#endif
#ifndef HO_PREFIX
#define HO_PREFIX ho32_
#endif
/* pow(x, y) as the passes evaluate it: binary exponentiation for an integer 1 <= y < 65536, else POW */
static real spow(real x, real y) {
  if (!(y >= 1 && y < 65536) || FLOOR(y) != y) return POW(x, y);
  real b = x, r = 1;
  for (int k = (int)y;;) {
    if (k & 1) r = r * b;
    k >>= 1;
    if (!k) break;
    b = b * b;
  }
  return r;
}
#define HO_CAT2(a, b) a##b
#define HO_CAT(a, b) HO_CAT2(a, b)
#define HO_NAME(x) HO_CAT(HO_PREFIX, x)
#define EXPORT __attribute__((visibility("default")))

static real h2r(uint16_t h) {
  int s = h >> 15, e = (h >> 10) & 31, m = h & 1023;
  double v;
  if (e == 0) v = ldexp((double)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = ldexp((double)(1024 + m), e - 25);
  return (real)(s ? -v : v);
}

/* rgba16float store: round toward zero, clamped to +-65504 (the rounding hush.js:98 states; NaN to -65504,
   the clamp with IEEE min / max) */
static real st(real x) {
#if HO_QUANT
  if (x != x) return -65504.0f;
  double a = fabs((double)x), r;
  if (a >= 65504.0) r = 65504.0;
  else if (a < ldexp(1.0, -14)) r = floor(a * 16777216.0) / 16777216.0;
  else { int e; double m = frexp(a, &e); r = ldexp(floor(m * 2048.0) / 2048.0, e); }
  return (real)(x < 0 ? -r : r);
#else
  return x;
#endif
}

static real minr(real a, real b) { return a < b ? a : b; }
static real maxr(real a, real b) { return a > b ? a : b; }
static real lum3(const real *c) { return c[0] * RX(0.2126) + c[1] * RX(0.7152) + c[2] * RX(0.0722); }
static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* f32 -> i32 toward zero, saturating (WGSL's conversion; positive side at 2^31 - 128), NaN -> 0, and i32
   addition with two's-complement wrap: the defined forms of the reprojection's index arithmetic (a motion
   vector that overflowed binary16 is +-inf) */
static int f2i(real x) {
  if (x != x) return 0;
  if (x < (real)-2147483648.0) return (int)(-2147483647 - 1);
  if (x > (real)2147483520.0) return 2147483520;
  return (int)x;
}
static int addw(int a, int b) { return (int)((unsigned)a + (unsigned)b); }

typedef struct {
  real sigma_l, sigma_n, sigma_z, alpha_color, alpha_moments, taa_alpha;
  int taa, iterations;
} sv_opts;

typedef struct {
  int W, H, frame, reset;
  sv_opts o;
  real *hist_color[2], *hist_mom[2], *tmp_color, *atrous[2], *taa[2];
  real *in[7];     /* radiance, normalDepth, motion, material, emission, prevNormalDepth, prevMotion */
} sv_ctx;

enum { I_RAD, I_ND, I_MV, I_MAT, I_EMI, I_PND, I_PMV };

static real *at(real *s, int W, int x, int y) { return s + ((size_t)y * W + x) * 4; }
static void store4(real *d, real a, real b, real c, real e) { d[0] = st(a); d[1] = st(b); d[2] = st(c); d[3] = st(e); }

EXPORT sv_ctx *HO_NAME(svgf_create)(int w, int h, double sigma_l, double sigma_n, double sigma_z, double alpha_color,
                                    double alpha_moments, double taa_alpha, int taa, int iterations) {
  if (w <= 0 || h <= 0) return NULL;
  sv_ctx *c = (sv_ctx *)calloc(1, sizeof(sv_ctx));
  c->W = w; c->H = h; c->reset = 1;
  c->o.sigma_l = (real)sigma_l; c->o.sigma_n = (real)sigma_n; c->o.sigma_z = (real)sigma_z;
  c->o.alpha_color = (real)alpha_color; c->o.alpha_moments = (real)alpha_moments; c->o.taa_alpha = (real)taa_alpha;
  c->o.taa = taa; c->o.iterations = iterations;
  size_t n = (size_t)w * h * 4;
  for (int i = 0; i < 2; i++) {
    c->hist_color[i] = (real *)calloc(n, sizeof(real)); c->hist_mom[i] = (real *)calloc(n, sizeof(real));
    c->atrous[i] = (real *)calloc(n, sizeof(real)); c->taa[i] = (real *)calloc(n, sizeof(real));
  }
  c->tmp_color = (real *)calloc(n, sizeof(real));
  for (int i = 0; i < 7; i++) c->in[i] = (real *)calloc(n, sizeof(real));
  return c;
}

EXPORT void HO_NAME(svgf_destroy)(sv_ctx *c) {
  if (!c) return;
  for (int i = 0; i < 2; i++) { free(c->hist_color[i]); free(c->hist_mom[i]); free(c->atrous[i]); free(c->taa[i]); }
  free(c->tmp_color);
  for (int i = 0; i < 7; i++) free(c->in[i]);
  free(c);
}

EXPORT void HO_NAME(svgf_reset)(sv_ctx *c) { c->reset = 1; }
EXPORT int HO_NAME(svgf_frame)(const sv_ctx *c) { return c->frame; }

/* 7 inputs as binary16 bit patterns (W*H*4 each) */
EXPORT void HO_NAME(svgf_set_inputs)(sv_ctx *c, const uint16_t *const *src) {
  size_t n = (size_t)c->W * c->H * 4;
  for (int k = 0; k < 7; k++)
    for (size_t i = 0; i < n; i++) c->in[k][i] = h2r(src[k][i]);
}

EXPORT real *HO_NAME(svgf_surface)(sv_ctx *c, int id) {
  switch (id) {
    case 0: return c->hist_color[0]; case 1: return c->hist_color[1];
    case 2: return c->hist_mom[0]; case 3: return c->hist_mom[1];
    case 4: return c->tmp_color; case 5: return c->atrous[0]; case 6: return c->atrous[1];
    case 7: return c->taa[0]; case 8: return c->taa[1];
    default: return NULL;
  }
}

/* consistent (svgf.js:34-41) */
static int consistent(sv_ctx *c, int qx, int qy, const real *n, real zexp, real obj) {
  if (!(qx >= 0 && qy >= 0 && qx < c->W && qy < c->H)) return 0;
  const real *nd = at(c->in[I_PND], c->W, qx, qy);
  if (nd[3] < 0) return 0;
  if (at(c->in[I_PMV], c->W, qx, qy)[2] != obj) return 0;
  real tol = RX(0.05) * zexp;
  tol = tol + RX(0.01);
  if (FABS(nd[3] - zexp) > tol) return 0;
  if (nd[0] * n[0] + nd[1] * n[1] + nd[2] * n[2] < RX(0.9)) return 0;
  return 1;
}

/* TEMPORAL (svgf.js:20-106) */
static void temporal(sv_ctx *c, int cur, int prev) {
  const int W = c->W, H = c->H;
  #pragma omp parallel for schedule(static)
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const real *nd = at(c->in[I_ND], W, x, y);
      real *oc = at(c->tmp_color, W, x, y), *om = at(c->hist_mom[cur], W, x, y);
      if (nd[3] < 0) { store4(oc, 0, 0, 0, 0); store4(om, 0, 0, 0, 0); continue; }
      const real *mv = at(c->in[I_MV], W, x, y), *alb = at(c->in[I_MAT], W, x, y), *rad = at(c->in[I_RAD], W, x, y);
      real il[3] = {rad[0] / maxr(alb[0], RX(1e-3)), rad[1] / maxr(alb[1], RX(1e-3)), rad[2] / maxr(alb[2], RX(1e-3))};
      real l = lum3(il);
      real ppx = ((real)x + RX(0.5) + mv[0]) - RX(0.5), ppy = ((real)y + RX(0.5) + mv[1]) - RX(0.5);
      int bx = f2i(FLOOR(ppx)), by = f2i(FLOOR(ppy));
      real fx = ppx - FLOOR(ppx), fy = ppy - FLOOR(ppy);
      real sumW = 0, hc[3] = {0, 0, 0}, hm[3] = {0, 0, 0};
      real wts[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
      int nohist = c->reset;
      for (int i = 0; i < 4 && !nohist; i++) {
        int qx = addw(bx, i & 1), qy = addw(by, i >> 1);
        if (consistent(c, qx, qy, nd, mv[3], mv[2]) && wts[i] > 0) {
          sumW += wts[i];
          const real *pc = at(c->hist_color[prev], W, qx, qy), *pm = at(c->hist_mom[prev], W, qx, qy);
          for (int k = 0; k < 3; k++) { hc[k] += wts[i] * pc[k]; hm[k] += wts[i] * pm[k]; }
        }
      }
      if (sumW < RX(0.01) && !nohist) {
        int cx = f2i(RINT(ppx)), cy = f2i(RINT(ppy));
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int qx = addw(cx, dx), qy = addw(cy, dy);
            if (consistent(c, qx, qy, nd, mv[3], mv[2])) {
              sumW += 1;
              const real *pc = at(c->hist_color[prev], W, qx, qy), *pm = at(c->hist_mom[prev], W, qx, qy);
              for (int k = 0; k < 3; k++) { hc[k] += pc[k]; hm[k] += pm[k]; }
            }
          }
      }
      real len, color[3] = {il[0], il[1], il[2]}, mom[2] = {l, l * l};
      if (sumW >= RX(0.01)) {
        for (int k = 0; k < 3; k++) { hc[k] /= sumW; hm[k] /= sumW; }
        len = minr(hm[2] + 1, RX(64.0));
        real ac = maxr(c->o.alpha_color, 1 / len), am = maxr(c->o.alpha_moments, 1 / len);
        for (int k = 0; k < 3; k++) color[k] = hc[k] * (1 - ac) + il[k] * ac;
        real m0 = hm[0] * (1 - am) + mom[0] * am, m1 = hm[1] * (1 - am) + mom[1] * am;
        mom[0] = m0; mom[1] = m1;
      } else {
        len = 1;
      }
      store4(oc, color[0], color[1], color[2], 0);
      store4(om, mom[0], mom[1], len, 0);
    }
}

/* VARIANCE (svgf.js:108-149) */
static void variance(sv_ctx *c, int cur) {
  const int W = c->W, H = c->H;
  #pragma omp parallel for schedule(static)
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const real *nd = at(c->in[I_ND], W, x, y);
      real *o = at(c->atrous[0], W, x, y);
      if (nd[3] < 0) { store4(o, 0, 0, 0, 0); continue; }
      const real *col = at(c->tmp_color, W, x, y), *m = at(c->hist_mom[cur], W, x, y);
      if (m[2] >= 4) { store4(o, col[0], col[1], col[2], maxr(0, m[1] - m[0] * m[0])); continue; }
      real sw = 0, sc[3] = {0, 0, 0}, sm[2] = {0, 0};
      for (int dy = -3; dy <= 3; dy++)
        for (int dx = -3; dx <= 3; dx++) {
          int qx = x + dx, qy = y + dy;
          if (!(qx >= 0 && qy >= 0 && qx < W && qy < H)) continue;
          const real *ndq = at(c->in[I_ND], W, qx, qy);
          if (ndq[3] < 0) continue;
          real len2 = SQRT((real)dx * (real)dx + (real)dy * (real)dy);
          real wz = EXP(-FABS(ndq[3] - nd[3]) / (RX(0.02) * nd[3] * len2 + RX(1e-3)));
          real wn = spow(maxr(0, ndq[0] * nd[0] + ndq[1] * nd[1] + ndq[2] * nd[2]), c->o.sigma_n);
          real w = wz * wn;
          sw += w;
          const real *cq = at(c->tmp_color, W, qx, qy), *mq = at(c->hist_mom[cur], W, qx, qy);
          for (int k = 0; k < 3; k++) sc[k] += w * cq[k];
          sm[0] += w * mq[0]; sm[1] += w * mq[1];
        }
      real d = maxr(sw, RX(1e-6));
      for (int k = 0; k < 3; k++) sc[k] /= d;
      sm[0] /= d; sm[1] /= d;
      real v = maxr(0, sm[1] - sm[0] * sm[0]) * 4 / maxr(m[2], 1);
      store4(o, sc[0], sc[1], sc[2], v);
    }
}

/* ATROUS (svgf.js:151-217), iteration `iter` with step 1 << iter */
static void atrous(sv_ctx *c, int iter, int src, int cur) {
  const int W = c->W, H = c->H, step = 1 << iter;
  static const double h[5] = {1.0 / 16, 1.0 / 4, 3.0 / 8, 1.0 / 4, 1.0 / 16};
  static const double k3[3] = {0.25, 0.5, 0.25};
  real *in = c->atrous[src], *out = c->atrous[src ^ 1];
  #pragma omp parallel for schedule(static)
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const real *nd = at(c->in[I_ND], W, x, y);
      real *o = at(out, W, x, y);
      if (nd[3] < 0) {
        store4(o, 0, 0, 0, 0);
        if (iter == 0) store4(at(c->hist_color[cur], W, x, y), 0, 0, 0, 0);
        continue;
      }
      const real *center = at(in, W, x, y);
      real gv = 0;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++)
          gv += (real)k3[dx + 1] * (real)k3[dy + 1] * at(in, W, clampi(x + dx, 0, W - 1), clampi(y + dy, 0, H - 1))[3];
      real lp = lum3(center);
      real denomL = c->o.sigma_l * SQRT(maxr(gv, 0)) + RX(1e-6);
      /* zGrad (svgf.js:159-167) */
      real z = nd[3];
      real zl = at(c->in[I_ND], W, clampi(x - 1, 0, W - 1), y)[3], zr = at(c->in[I_ND], W, clampi(x + 1, 0, W - 1), y)[3];
      real zu = at(c->in[I_ND], W, x, clampi(y - 1, 0, H - 1))[3], zd = at(c->in[I_ND], W, x, clampi(y + 1, 0, H - 1))[3];
      real gx = minr(zr > 0 ? FABS(zr - z) : RX(1e9), zl > 0 ? FABS(zl - z) : RX(1e9));
      real gy = minr(zd > 0 ? FABS(zd - z) : RX(1e9), zu > 0 ? FABS(zu - z) : RX(1e9));
      real zg = minr(maxr(gx, gy), z);
      real sw = (real)h[2] * (real)h[2];
      real sc[3] = {center[0] * sw, center[1] * sw, center[2] * sw};
      real sv = center[3] * sw * sw;
      for (int dy = -2; dy <= 2; dy++)
        for (int dx = -2; dx <= 2; dx++) {
          if (dx == 0 && dy == 0) continue;
          int qx = x + dx * step, qy = y + dy * step;
          if (!(qx >= 0 && qy >= 0 && qx < W && qy < H)) continue;
          const real *ndq = at(c->in[I_ND], W, qx, qy);
          if (ndq[3] < 0) continue;
          const real *cq = at(in, W, qx, qy);
          real dist = SQRT((real)dx * (real)dx + (real)dy * (real)dy) * (real)step;
          real wz = FABS(ndq[3] - z) / (c->o.sigma_z * zg * dist + RX(1e-3));
          real wn = spow(maxr(0, ndq[0] * nd[0] + ndq[1] * nd[1] + ndq[2] * nd[2]), c->o.sigma_n);
          real wl = FABS(lum3(cq) - lp) / denomL;
          real w = (real)h[dx + 2] * (real)h[dy + 2] * wn * EXP(-wz - wl);
          sw += w;
          for (int k = 0; k < 3; k++) sc[k] += w * cq[k];
          sv += w * w * cq[3];
        }
      real r0 = sc[0] / sw, r1 = sc[1] / sw, r2 = sc[2] / sw, r3 = sv / (sw * sw);
      store4(o, r0, r1, r2, r3);
      if (iter == 0) store4(at(c->hist_color[cur], W, x, y), r0, r1, r2, 0);
    }
}

static void ycocg(const real *c, real *o) {
  o[0] = c[0] * RX(0.25) + c[1] * RX(0.5) + c[2] * RX(0.25);
  o[1] = c[0] * RX(0.5) - c[2] * RX(0.5);
  o[2] = -c[0] * RX(0.25) + c[1] * RX(0.5) - c[2] * RX(0.25);
}
static void tmv(const real *c, real *o) { real d = 1 + lum3(c); o[0] = c[0] / d; o[1] = c[1] / d; o[2] = c[2] / d; }
static void itmv(const real *c, real *o) { real d = maxr(1 - lum3(c), RX(1e-4)); o[0] = c[0] / d; o[1] = c[1] / d; o[2] = c[2] / d; }

/* remod (svgf.js:229-232) */
static void remod(sv_ctx *c, const real *illum, int x, int y, real *o) {
  int qx = clampi(x, 0, c->W - 1), qy = clampi(y, 0, c->H - 1);
  const real *il = at((real *)illum, c->W, qx, qy), *m = at(c->in[I_MAT], c->W, qx, qy), *e = at(c->in[I_EMI], c->W, qx, qy);
  for (int k = 0; k < 3; k++) o[k] = il[k] * m[k] + e[k];
}

/* RESOLVE (svgf.js:219-271) */
static void resolve(sv_ctx *c, int src, int cur, int prev) {
  const int W = c->W, H = c->H;
  const real *illum = c->atrous[src], *ptaa = c->taa[prev];
  #pragma omp parallel for schedule(static)
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      real r[3], cu[3], out[3];
      remod(c, illum, x, y, r);
      tmv(r, cu);
      real *o = at(c->taa[cur], W, x, y);
      if (!c->o.taa) { itmv(cu, out); store4(o, out[0], out[1], out[2], 1); continue; }
      real m1[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          real q[3], t[3], yc[3];
          remod(c, illum, x + dx, y + dy, q);
          tmv(q, t);
          ycocg(t, yc);
          for (int k = 0; k < 3; k++) { m1[k] += yc[k]; m2[k] += yc[k] * yc[k]; }
        }
      real sd[3];
      for (int k = 0; k < 3; k++) { m1[k] /= 9; m2[k] /= 9; sd[k] = SQRT(maxr(m2[k] - m1[k] * m1[k], 0)); }
      const real *mv = at(c->in[I_MV], W, x, y);
      real ux = ((real)x + RX(0.5) + mv[0]) / (real)W, uy = ((real)y + RX(0.5) + mv[1]) / (real)H;
      real a = c->o.taa_alpha;
      real res[3] = {cu[0], cu[1], cu[2]};
      if (ux >= 0 && uy >= 0 && ux <= 1 && uy <= 1 && !c->reset) {
        /* linear sampler, clamp to edge: exact bilinear weights at uv * size - 0.5 */
        real px = ux * (real)W - RX(0.5), py = uy * (real)H - RX(0.5);
        real bx = FLOOR(px), by = FLOOR(py), fx = px - bx, fy = py - by;
        real hist[3] = {0, 0, 0};
        for (int j = 0; j < 4; j++) {
          int ox = j & 1, oy = j >> 1;
          real wb = (ox ? fx : 1 - fx) * (oy ? fy : 1 - fy);
          const real *t = at((real *)ptaa, W, clampi(addw(f2i(bx), ox), 0, W - 1), clampi(addw(f2i(by), oy), 0, H - 1));
          for (int k = 0; k < 3; k++) hist[k] += wb * t[k];
        }
        real th[3], hh[3];
        tmv(hist, th);
        ycocg(th, hh);
        for (int k = 0; k < 3; k++) {
          real lo = m1[k] - RX(1.25) * sd[k], hi = m1[k] + RX(1.25) * sd[k];
          hh[k] = minr(maxr(hh[k], lo), hi);
        }
        real fr[3] = {hh[0] + hh[1] - hh[2], hh[0] + hh[2], hh[0] - hh[1] - hh[2]};
        for (int k = 0; k < 3; k++) res[k] = fr[k] * (1 - a) + cu[k] * a;
      }
      itmv(res, out);
      store4(o, out[0], out[1], out[2], 1);
    }
}

/* encode (svgf.js:330-377): returns the id of taa[cur] (svgf_surface ids 7, 8) */
EXPORT int HO_NAME(svgf_encode)(sv_ctx *c) {
  int cur = c->frame & 1, prev = cur ^ 1;
  temporal(c, cur, prev);
  variance(c, cur);
  int src = 0;
  for (int i = 0; i < c->o.iterations; i++) { atrous(c, i, src, cur); src ^= 1; }
  resolve(c, src, cur, prev);
  c->reset = 0;
  c->frame++;
  return 7 + cur;
}

/* teacher-forcing hooks: set the frame counter / reset flag, run one pass (0 temporal, 1 variance,
   2 .. 1 + iterations a-trous, 2 + iterations resolve) without advancing */
EXPORT void HO_NAME(svgf_set_state)(sv_ctx *c, int frame, int reset) { c->frame = frame; c->reset = reset; }
EXPORT void HO_NAME(svgf_run_pass)(sv_ctx *c, int pass) {
  int cur = c->frame & 1, prev = cur ^ 1;
  if (pass == 0) temporal(c, cur, prev);
  else if (pass == 1) variance(c, cur);
  else if (pass < 2 + c->o.iterations) atrous(c, pass - 2, (pass - 2) & 1, cur);
  else if (pass == 2 + c->o.iterations) resolve(c, c->o.iterations & 1, cur, prev);
}
