/*
 * CPU oracle for the Hush denoiser hot path.  TEST INFRASTRUCTURE ONLY: nothing in the product links,
 * loads or calls this file; only tests/, the smoke check and the CPU-baseline leg of bench.py do, and
 * only as the checker.
 *
 * A plain-C restatement of the eight compute passes of src/denoise/hush.js (kanaworksai-hue/hush v0.1.0),
 * pixel by pixel over whole images.  Every function cites the WGSL it follows.  The same source is
 * compiled three ways (see oracle/Makefile):
 *   ho32_*   REAL=float,  QUANT=1  FP32 mirror: the reference's op order, no FMA contraction, f16 storage
 *                                  (round toward zero) at every rgba16float store and the explicit
 *                                  qf16 / q4 / pk truncations.  Used for parity with the HIP kernels.
 *   ho64q_*  REAL=double, QUANT=1  FP64 arithmetic, same f16 storage points (isolates arithmetic error).
 *   ho64_*   REAL=double, QUANT=0  FP64 ideal: no storage quantisation at all (PSNR reference).
 *
 * Storage rounding.  hush.js:98 states that an rgba16float storage write rounds toward zero ("f16 rounding
 * toward zero, like a rgba16float storage write"), and the feature inputs are built with qf16 to match
 * "the values as stored" (hush.js:212).  The oracle therefore stores every rgba16float value with
 * round-toward-zero to binary16 (clamped to +-65504, subnormals kept), and pack2x16float(qf16(a), qf16(b))
 * is the same RTZ conversion.  The inputs arrive as binary16 bit patterns and are widened exactly.
 *
 * Parity status of each function (pins are in tests/test_oracle_*.py; the closed-form pins of
 * tests/test_oracle_closedform.py run the FP64-ideal build of this same source, and each is shown to fail
 * on a one-token mistake planted in a rebuilt copy):
 *   f16 RTZ / qf16 / halfToFloat      pinned: exhaustive and closed-form checks against numpy float16.
 *   lum, octEnc, octDec               pinned: round trip, octant folds, closed forms.
 *   hiddenOf                          pinned: quadratic closed form.
 *   check                             pinned: hand-built cases either side of every threshold.
 *   pixel(): 3x3 stats, lag, blend    pinned: closed forms on stripes with a static history (the lag's
 *                                     second division by the count, the clamp, len, nfast, tvar).
 *   reprojection fallback             pinned: the 3x3 search around floor(p + mv + 0.5), v = 0.1.
 *   block(): the 20 features          pinned: constant plane, 1..4 valid pixels, depth step, normal
 *                                     crease, luminance statistics of a known 2x2 / 4x4 window.
 *   MLP                               pinned: numpy FP64 brute force, f = 0 closed form.
 *   activations + mid1 wiring         pinned: closed forms with the output bias alone.
 *   filt (a-trous)                    pinned: binomial-kernel moments (R = 1, 2; step 1, 2), the centre
 *                                     weight, sum w^2 var / (sum w)^2, the luminance edge-stopping slope,
 *                                     zg on a depth ramp and next to sky, prefilter clamping vs tap
 *                                     skipping at a corner, the pow128 normal weight.
 *   down, coarse                      pinned: count-weighted means and vr / N / N on a one-texel level.
 *   fine+final (jbw upsamples)        pinned: depth step, normal term, kappa, fallback, coarse level.
 *   TAA                               pinned: the 1.25 sigma clip of a known window, the 1/10 blend, the
 *                                     inclusive uv bounds.
 *   oracle/hush_np.py, a second restatement in numpy, remains a cross-check of the whole sequence.
 *   WebGPU store rounding, FMA contraction and transcendental ULPs of the browser stack cannot be run
 *   here (no WebGPU runtime): parity against the WGSL itself is unpinned at threshold ties.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef HO_DOUBLE
typedef float real;
#define RX(x) x##f
#define EXP expf
#define LOG logf
#define SQRT sqrtf
#define FLOOR floorf
#define FABS fabsf
#define FMA fmaf
#else
typedef double real;
#define RX(x) x
#define EXP exp
#define LOG log
#define SQRT sqrt
#define FLOOR floor
#define FABS fabs
#define FMA fma
#endif
#ifndef HO_QUANT
#define HO_QUANT 1
#endif
#ifndef HO_PREFIX
#define HO_PREFIX ho32_
#endif
#define HO_CAT2(a, b) a##b
#define HO_CAT(a, b) HO_CAT2(a, b)
#define HO_NAME(x) HO_CAT(HO_PREFIX, x)
#define EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------- binary16 helpers */

/* IEEE binary16 -> float, exact (hush.js:706-711 halfToFloat for the weights; unpack2x16float). */
static real h2r(uint16_t h) {
  int s = h >> 15, e = (h >> 10) & 31, m = h & 1023;
  double v;
  if (e == 0) v = ldexp((double)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = ldexp((double)(1024 + m), e - 25);
  return (real)(s ? -v : v);
}

/* value -> binary16 with round toward zero, clamped to +-65504 (the rgba16float storage write of
   hush.js:98, also pack2x16float(qf16(a), qf16(b)) of hush.js:101).  NaN stores as -65504: the clamp
   min(max(x, -65504), 65504) with IEEE maxNum / minNum, which return the non-NaN operand (WGSL leaves NaN
   at run time indeterminate; this is what its clamp gives on hardware with IEEE min / max). */
static uint16_t r2h_rtz(real x) {
  if (x != x) return 0xfbff;
  uint16_t s = 0;
  double a = (double)x;
  if (a < 0 || (a == 0 && signbit(a))) { s = 0x8000; a = -a; }
  if (a >= 65504.0) return s | 0x7bff;
  if (a < ldexp(1.0, -14)) {                       /* subnormal: k * 2^-24 */
    double k = floor(a * 16777216.0);
    return s | (uint16_t)k;
  }
  int e;
  double m = frexp(a, &e);                          /* a = m * 2^e, m in [0.5, 1) */
  double q = floor(m * 2048.0);                     /* 11 significant bits, truncated */
  int ee = e - 1 + 15;                              /* biased exponent of 1.xxx * 2^(e-1) */
  return s | (uint16_t)((ee << 10) | ((int)q - 1024));
}

/* the value an rgba16float store keeps */
static real st(real x) {
#if HO_QUANT
  return h2r(r2h_rtz(x));
#else
  return x;
#endif
}

/* qf16 (hush.js:68,99): clamp to +-65504, then drop the low 13 mantissa bits of the f32 (toward zero);
   NaN clamps to -65504 as in r2h_rtz. */
static real qf16(real x) {
#if HO_QUANT
  if (x != x) return -65504.0f;
#ifndef HO_DOUBLE
  float c = x < -65504.0f ? -65504.0f : (x > 65504.0f ? 65504.0f : x);
  uint32_t u;
  memcpy(&u, &c, 4);
  u &= 0xffffe000u;
  memcpy(&c, &u, 4);
  return c;
#else
  double c = x < -65504.0 ? -65504.0 : (x > 65504.0 ? 65504.0 : x);
  if (c == 0) return c;
  int e;
  double m = frexp(c, &e);
  return ldexp(trunc(m * 2048.0) / 2048.0, e);
#endif
#else
  return x;
#endif
}

/* ---------------------------------------------------------------- small vector types */

typedef struct { real x, y; } v2;
typedef struct { real x, y, z; } v3;
typedef struct { real x, y, z, w; } v4;

static v3 mk3(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static v4 mk4(real x, real y, real z, real w) { v4 r = {x, y, z, w}; return r; }
static v3 xyz(v4 a) { return mk3(a.x, a.y, a.z); }
static real minr(real a, real b) { return a < b ? a : b; }   /* WGSL min/max on non-NaN operands */
static real maxr(real a, real b) { return a > b ? a : b; }
static real clampr(real x, real lo, real hi) { return minr(maxr(x, lo), hi); }

/* lum (hush.js:31): dot with Rec.709 weights, summed left to right */
static real lum(v3 c) { return c.x * RX(0.2126) + c.y * RX(0.7152) + c.z * RX(0.0722); }
static real dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static real length3(v3 a) { return SQRT(a.x * a.x + a.y * a.y + a.z * a.z); }
static real length2(real x, real y) { return SQRT(x * x + y * y); }

/* octEnc (hush.js:34-39) */
static v2 oct_enc(v3 n) {
  real d = maxr(FABS(n.x) + FABS(n.y) + FABS(n.z), RX(1e-20));
  v3 v = mk3(n.x / d, n.y / d, n.z / d);
  real sx = v.x >= 0 ? RX(1.0) : RX(-1.0);
  real sy = v.y >= 0 ? RX(1.0) : RX(-1.0);
  v2 r;
  if (v.z < 0) { r.x = (RX(1.0) - FABS(v.y)) * sx; r.y = (RX(1.0) - FABS(v.x)) * sy; }
  else { r.x = v.x; r.y = v.y; }
  return r;
}

/* octDec (hush.js:40-46) */
static v3 oct_dec(real ex, real ey) {
  v3 v = mk3(ex, ey, RX(1.0) - FABS(ex) - FABS(ey));
  real t = maxr(-v.z, 0);
  v.x += v.x >= 0 ? -t : t;
  v.y += v.y >= 0 ? -t : t;
  real l = maxr(length3(v), RX(1e-20));
  return mk3(v.x / l, v.y / l, v.z / l);
}

/* log1p (hush.js:47) is literally log(1 + x) */
static real log1p_(real x) { return LOG(RX(1.0) + x); }

/* pow128 (hush.js:70-73): seven squarings */
static real pow128(real x) {
  real x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16, x64 = x32 * x32;
  return x64 * x64;
}

static real sigmoid(real x) { return RX(1.0) / (RX(1.0) + EXP(-x)); }

#define HMAX RX(32.0)
#define N_FEAT 20
#define N_OUT 12

/* ---------------------------------------------------------------- context */

enum { S_A0, S_X, S_BH0, S_BH1, S_MOM0, S_MOM1, S_KEY0, S_KEY1, S_HIST0, S_HIST1, S_TAA0, S_TAA1,
       S_PAR0_0, S_PAR0_1, S_PAR1, S_PAR2, S_F, S_C0, S_C1, S_CG, S_COUNT };

typedef struct {
  int W, H, W2, H2, W4, H4;
  int frame, reset, f16mlp, hidden;
  real exposure;
  real *w;                   /* MLP weights, widened from f16 */
  real *surf[S_COUNT];       /* rgba surfaces: 4 values per texel; S_F: 32 values (F0..F3 unpacked) */
  real *in[5];               /* this frame's inputs, widened */
  real *P;                   /* per-pixel feature inputs P0..P4 (20 values) */
  real *sI;                  /* per-pixel demodulated illumination + valid flag */
  uint8_t *mask;             /* reprojection validity bits per pixel */
} ho_ctx;

enum { IN_RAD, IN_ND, IN_MV, IN_MAT, IN_EMI };

static int surf_w(const ho_ctx *c, int s) {
  if (s == S_BH0 || s == S_BH1 || s == S_PAR0_0 || s == S_PAR0_1 || s == S_PAR1 || s == S_PAR2 || s == S_F) return c->W2;
  if (s == S_C0 || s == S_C1 || s == S_CG) return c->W4;
  return c->W;
}
static int surf_h(const ho_ctx *c, int s) {
  if (s == S_BH0 || s == S_BH1 || s == S_PAR0_0 || s == S_PAR0_1 || s == S_PAR1 || s == S_PAR2 || s == S_F) return c->H2;
  if (s == S_C0 || s == S_C1 || s == S_CG) return c->H4;
  return c->H;
}
static int surf_ch(int s) { return s == S_F ? 32 : 4; }

/* hiddenOf (hush.js:22-25): h*h + (N_FEAT + N_OUT + 2) h + N_OUT == n for h in [1, 256] */
EXPORT int HO_NAME(hidden_of)(long n) {
  for (long h = 1; h <= 256; h++)
    if (h * h + (N_FEAT + N_OUT + 2) * h + N_OUT == n) return (int)h;
  return 0;
}

static void free_surfaces(ho_ctx *c) {
  for (int s = 0; s < S_COUNT; s++) { free(c->surf[s]); c->surf[s] = NULL; }
  for (int i = 0; i < 5; i++) { free(c->in[i]); c->in[i] = NULL; }
  free(c->P); free(c->sI); free(c->mask);
  c->P = c->sI = NULL; c->mask = NULL;
}

/* resize (hush.js:764-794): half = ceil, quarter = ceil(half / 2); history reset, frame = 0 */
EXPORT int HO_NAME(resize)(ho_ctx *c, int w, int h) {
  if (!(w > 0 && h > 0)) return -1;
  if (w == c->W && h == c->H) return 0;
  free_surfaces(c);
  c->W = w; c->H = h;
  c->W2 = (w + 1) / 2; c->H2 = (h + 1) / 2;
  c->W4 = (c->W2 + 1) / 2; c->H4 = (c->H2 + 1) / 2;
  c->frame = 0; c->reset = 1;
  for (int s = 0; s < S_COUNT; s++)
    c->surf[s] = (real *)calloc((size_t)surf_w(c, s) * surf_h(c, s) * surf_ch(s), sizeof(real));
  for (int i = 0; i < 5; i++) c->in[i] = (real *)calloc((size_t)w * h * 4, sizeof(real));
  c->P = (real *)calloc((size_t)w * h * 20, sizeof(real));
  c->sI = (real *)calloc((size_t)w * h * 4, sizeof(real));
  c->mask = (uint8_t *)calloc((size_t)w * h, 1);
  return 0;
}

/* constructor (hush.js:731-761) minus the WebGPU pipeline set-up */
EXPORT ho_ctx *HO_NAME(create)(int w, int h, const uint16_t *weights, long count, real exposure, int f16mlp) {
  int hid = HO_NAME(hidden_of)(count);
  if (!hid) return NULL;
  ho_ctx *c = (ho_ctx *)calloc(1, sizeof(ho_ctx));
  c->hidden = hid;
  c->exposure = exposure;
  c->f16mlp = f16mlp;
  c->w = (real *)malloc(sizeof(real) * count);
  for (long i = 0; i < count; i++) c->w[i] = h2r(weights[i]);
  if (HO_NAME(resize)(c, w, h) != 0) { free(c->w); free(c); return NULL; }
  return c;
}

EXPORT void HO_NAME(destroy)(ho_ctx *c) {
  if (!c) return;
  free_surfaces(c);
  free(c->w);
  free(c);
}

EXPORT void HO_NAME(reset)(ho_ctx *c) { c->reset = 1; }
EXPORT void HO_NAME(set_exposure)(ho_ctx *c, double e) { c->exposure = (real)e; }
EXPORT void HO_NAME(set_state)(ho_ctx *c, int frame, int reset) { c->frame = frame; c->reset = reset; }
EXPORT int HO_NAME(frame)(const ho_ctx *c) { return c->frame; }
EXPORT int HO_NAME(hidden)(const ho_ctx *c) { return c->hidden; }
EXPORT real *HO_NAME(surface)(ho_ctx *c, int s, int *w, int *h, int *ch) {
  if (s < 0 || s >= S_COUNT) return NULL;
  if (w) *w = surf_w(c, s);
  if (h) *h = surf_h(c, s);
  if (ch) *ch = surf_ch(s);
  return c->surf[s];
}
EXPORT const uint8_t *HO_NAME(mask)(const ho_ctx *c) { return c->mask; }
EXPORT real *HO_NAME(features_in)(ho_ctx *c) { return c->P; }

/* inputs: five W*H*4 arrays of binary16 bit patterns (radiance, normalDepth, motion, material, emission) */
EXPORT void HO_NAME(set_inputs)(ho_ctx *c, const uint16_t *rad, const uint16_t *nd, const uint16_t *mv,
                                const uint16_t *mat, const uint16_t *emi) {
  const uint16_t *src[5] = {rad, nd, mv, mat, emi};
  size_t n = (size_t)c->W * c->H * 4;
  for (int k = 0; k < 5; k++)
    for (size_t i = 0; i < n; i++) c->in[k][i] = h2r(src[k][i]);
}

/* ---------------------------------------------------------------- texel access */

static v4 ld4(const real *s, int w, int x, int y) {
  const real *p = s + ((size_t)y * w + x) * 4;
  return mk4(p[0], p[1], p[2], p[3]);
}
static void st4(real *s, int w, int x, int y, v4 v) {
  real *p = s + ((size_t)y * w + x) * 4;
  p[0] = st(v.x); p[1] = st(v.y); p[2] = st(v.z); p[3] = st(v.w);
}
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
static int inside(int x, int y, int w, int h) { return x >= 0 && y >= 0 && x < w && y < h; }

/* F planes: the 32 binary16 values of F0..F3 per half-res texel (pk = qf16 + RTZ store == st) */
static real *fptr(ho_ctx *c, int x, int y) { return c->surf[S_F] + ((size_t)y * c->W2 + x) * 32; }

/* ---------------------------------------------------------------- pass 1: temporal (hush.js:79-308) */

/* check (hush.js:103-105) */
static int check(v4 k, v3 n, real zprev, real obj) {
  if (!(k.z >= 0)) return 0;
  if (!(k.w == obj)) return 0;
  real tol = RX(0.05) * zprev;
  tol = tol + RX(0.01);
  if (!(FABS(k.z - zprev) <= tol)) return 0;
  return dot3(oct_dec(k.x, k.y), n) >= RX(0.9);
}

/* tile staging of TEMPORAL main (hush.js:287-294): demodulated illumination, valid flag */
static v4 illum_at(ho_ctx *c, int x, int y) {
  if (!inside(x, y, c->W, c->H)) return mk4(0, 0, 0, 0);
  v4 nd = ld4(c->in[IN_ND], c->W, x, y);
  if (!(nd.w >= 0)) return mk4(0, 0, 0, 0);
  v4 r = ld4(c->in[IN_RAD], c->W, x, y);
  v4 m = ld4(c->in[IN_MAT], c->W, x, y);
  return mk4(r.x / maxr(m.x, RX(1e-3)), r.y / maxr(m.y, RX(1e-3)), r.z / maxr(m.z, RX(1e-3)), 1);
}
static v4 sI(ho_ctx *c, int x, int y) {
  if (!inside(x, y, c->W, c->H)) return mk4(0, 0, 0, 0);
  const real *p = c->sI + ((size_t)y * c->W + x) * 4;
  return mk4(p[0], p[1], p[2], p[3]);
}

/* pixel (hush.js:111-223) */
static void pixel(ho_ctx *c, int px, int py, int cur, int prev) {
  const int W = c->W, H = c->H;
  real *P = c->P + ((size_t)py * W + px) * 20;
  memset(P, 0, 20 * sizeof(real));
  P[10] = -1;                                             /* P[2].z = -1: sky */
  uint8_t *mask = c->mask + (size_t)py * W + px;
  *mask = 0;
  v4 nd = ld4(c->in[IN_ND], W, px, py);
  v4 mv = ld4(c->in[IN_MV], W, px, py);
  if (nd.w < 0) {
    st4(c->surf[S_A0], W, px, py, mk4(0, 0, 0, 0));
    st4(c->surf[cur ? S_MOM1 : S_MOM0], W, px, py, mk4(0, 0, 0, 0));
    st4(c->surf[cur ? S_KEY1 : S_KEY0], W, px, py, mk4(0, 0, -1, mv.z));
    *mask = 0x80;
    return;
  }
  v3 n = xyz(nd);
  v4 mr = ld4(c->in[IN_MAT], W, px, py);
  v4 ce = sI(c, px, py);
  v3 illum = xyz(ce);
  real l = lum(illum);
  /* 3x3 stats (hush.js:125-142) */
  v3 s1 = mk3(0, 0, 0);
  real cnt = 0;
  for (int j = 0; j < 9; j++) {
    v4 t = sI(c, px + j % 3 - 1, py + j / 3 - 1);
    s1.x += t.x * t.w; s1.y += t.y * t.w; s1.z += t.z * t.w;
    cnt += t.w;
  }
  cnt = maxr(cnt, 1);
  v3 mu = mk3(s1.x / cnt, s1.y / cnt, s1.z / cnt);
  real slm = lum(mu);
  v3 sv3 = mk3(0, 0, 0);
  real svar = 0;
  for (int j = 0; j < 9; j++) {
    v4 t = sI(c, px + j % 3 - 1, py + j / 3 - 1);
    v3 d = mk3(t.x - mu.x, t.y - mu.y, t.z - mu.z);
    real dl = lum(xyz(t)) - slm;
    sv3.x += t.w * d.x * d.x; sv3.y += t.w * d.y * d.y; sv3.z += t.w * d.z * d.z;
    svar += t.w * dl * dl;
  }
  v3 sd = mk3(SQRT(maxr(sv3.x / cnt, RX(1e-12))), SQRT(maxr(sv3.y / cnt, RX(1e-12))), SQRT(maxr(sv3.z / cnt, RX(1e-12))));
  svar /= cnt;
  /* reprojection (hush.js:144-187) */
  int has = 0;
  v3 hc = mk3(0, 0, 0);
  v4 hm = mk4(0, 0, 0, 0);
  real v = 0, alpha = 1, gamma = 1;
  if (!c->reset) {
    const real *pkey = c->surf[prev ? S_KEY1 : S_KEY0];
    const real *pcol = c->surf[prev ? S_HIST1 : S_HIST0];
    const real *pmom = c->surf[prev ? S_MOM1 : S_MOM0];
    const real *ppar = c->surf[prev ? S_PAR0_1 : S_PAR0_0];
    real ppx = (real)px + mv.x, ppy = (real)py + mv.y;
    real bx = FLOOR(ppx), by = FLOOR(ppy);
    real fx = ppx - bx, fy = ppy - by;
    int bix = f2i(bx), biy = f2i(by);
    real sumW = 0;
    real wts[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
    int cx = f2i(FLOOR(ppx + RX(0.5))), cy = f2i(FLOOR(ppy + RX(0.5)));
    v4 pq = ld4(ppar, c->W2, clampi(cx, 0, W - 1) >> 1, clampi(cy, 0, H - 1) >> 1);
    for (int i = 0; i < 4; i++) {
      int qx = addw(bix, i & 1), qy = addw(biy, i >> 1);
      int cqx = clampi(qx, 0, W - 1), cqy = clampi(qy, 0, H - 1);
      v4 k4 = ld4(pkey, W, cqx, cqy);
      v4 c4 = ld4(pcol, W, cqx, cqy);
      v4 m4 = ld4(pmom, W, cqx, cqy);
      if (wts[i] > 0 && inside(qx, qy, W, H) && check(k4, n, mv.w, mv.z)) {
        *mask |= (uint8_t)(1 << i);
        sumW += wts[i];
        hc.x += wts[i] * c4.x; hc.y += wts[i] * c4.y; hc.z += wts[i] * c4.z;
        hm.x += wts[i] * m4.x; hm.y += wts[i] * m4.y; hm.z += wts[i] * m4.z; hm.w += wts[i] * m4.w;
      }
    }
    if (sumW >= RX(0.01)) {
      hc.x /= sumW; hc.y /= sumW; hc.z /= sumW;
      hm.x /= sumW; hm.y /= sumW; hm.z /= sumW; hm.w /= sumW;
      v = sumW; has = 1;
    } else {
      real n3 = 0;
      hc = mk3(0, 0, 0); hm = mk4(0, 0, 0, 0);
      for (int j = 0; j < 9; j++) {
        int qx = addw(cx, j % 3 - 1), qy = addw(cy, j / 3 - 1);
        if (inside(qx, qy, W, H) && check(ld4(pkey, W, qx, qy), n, mv.w, mv.z)) {
          n3 += 1;
          v4 cc = ld4(pcol, W, qx, qy), mm = ld4(pmom, W, qx, qy);
          hc.x += cc.x; hc.y += cc.y; hc.z += cc.z;
          hm.x += mm.x; hm.y += mm.y; hm.z += mm.z; hm.w += mm.w;
        }
      }
      if (n3 > 0) {
        hc.x /= n3; hc.y /= n3; hc.z /= n3;
        hm.x /= n3; hm.y /= n3; hm.z /= n3; hm.w /= n3;
        v = RX(0.1); has = 1;
        *mask |= 0x10;
      }
    }
    alpha = pq.x; gamma = pq.y;
  }
  has = has && hm.w > 0;
  if (has) *mask |= 0x20;
  /* blend (hush.js:189-211) */
  v3 color = illum;
  real nmx = l, nmy = l * l, nfast = l, len1 = 1, lag = 0;
  if (has) {
    real gx = gamma * sd.x, gy = gamma * sd.y, gz = gamma * sd.z;
    v3 hcl = mk3(minr(maxr(hc.x, mu.x - gx), mu.x + gx), minr(maxr(hc.y, mu.y - gy), mu.y + gy),
                 minr(maxr(hc.z, mu.z - gz), mu.z + gz));
    len1 = minr(hm.w + 1, HMAX);
    real a = maxr(alpha, 1 / len1);
    color = mk3(hcl.x + (illum.x - hcl.x) * a, hcl.y + (illum.y - hcl.y) * a, hcl.z + (illum.z - hcl.z) * a);
    nmx = hm.x + (l - hm.x) * a;
    nmy = hm.y + (l * l - hm.y) * a;
    nfast = hm.z + (l - hm.z) * maxr(RX(1.0) / RX(6.0), 1 / len1);
    real lh = lum(hc);
    lag = FABS(slm - lh) / (SQRT(maxr(svar / cnt, RX(1e-12))) + RX(0.02) * lh + RX(1e-3));
  } else {
    v = 0;
  }
  real tvar = maxr(nmy - nmx * nmx, 0);
  v4 A = mk4(color.x, color.y, color.z, svar + (tvar - svar) * minr(len1 / 4, 1));
  v2 oc = oct_enc(n);
  st4(c->surf[S_A0], W, px, py, A);
  st4(c->surf[cur ? S_MOM1 : S_MOM0], W, px, py, mk4(nmx, nmy, nfast, len1));
  st4(c->surf[cur ? S_KEY1 : S_KEY0], W, px, py, mk4(oc.x, oc.y, nd.w, mv.z));
  /* feature inputs from the stored values (hush.js:212-222) */
  real E = c->exposure;
  v4 Aq = mk4(qf16(A.x), qf16(A.y), qf16(A.z), qf16(A.w));
  real m1 = E * qf16(nmx);
  real m2 = E * E * qf16(nmy);
  P[0] = E * lum(xyz(Aq)); P[1] = E * qf16(l); P[2] = maxr(m2 - m1 * m1, 0); P[3] = len1;
  P[4] = qf16(v); P[5] = qf16(lag); P[6] = E * qf16(nfast); P[7] = m1;
  P[8] = qf16(oc.x); P[9] = qf16(oc.y); P[10] = qf16(nd.w); P[11] = qf16(length2(mv.x, mv.y));
  P[12] = mr.x; P[13] = mr.y; P[14] = mr.z; P[15] = mr.w;
  P[16] = Aq.x; P[17] = Aq.y; P[18] = Aq.z; P[19] = Aq.w;
}

static const real P_SKY[20] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0};

/* block (hush.js:226-282): the 20 features of half-res texel (X, Y) */
static void block(ho_ctx *c, int X, int Y) {
  const int W = c->W, H = c->H;
  real E = c->exposure;
  real nv = 0;
  v4 S0 = {0, 0, 0, 0}, S1 = {0, 0, 0, 0}, S3 = {0, 0, 0, 0}, AC = {0, 0, 0, 0};
  v3 N = {0, 0, 0};
  real MV = 0, Z = 0;
  real la[4] = {0, 0, 0, 0}, zs[4] = {0, 0, 0, 0};
  v3 ns[4];
  int ok[4];
  for (int j = 0; j < 4; j++) {
    int x = 2 * X + (j & 1), y = 2 * Y + (j >> 1);
    const real *P = inside(x, y, W, H) ? c->P + ((size_t)y * W + x) * 20 : P_SKY;
    ok[j] = P[10] >= 0;
    if (!ok[j]) continue;
    nv += 1;
    la[j] = P[0]; ns[j] = oct_dec(P[8], P[9]); zs[j] = P[10];
    S0.x += P[0]; S0.y += P[1]; S0.z += P[2]; S0.w += P[3];
    S1.x += P[4]; S1.y += P[5]; S1.z += P[6]; S1.w += P[7];
    N.x += ns[j].x; N.y += ns[j].y; N.z += ns[j].z;
    S3.x += P[12]; S3.y += P[13]; S3.z += P[14]; S3.w += P[15];
    AC.x += P[16]; AC.y += P[17]; AC.z += P[18]; AC.w += P[19];
    MV += P[11]; Z += P[10];
  }
  real *F = fptr(c, X, Y);
  if (nv == 0) {
    memset(F, 0, 32 * sizeof(real));
    F[24] = st(qf16(0)); F[25] = st(qf16(0)); F[26] = st(qf16(-1)); F[27] = st(qf16(0));
    return;
  }
  real k = 1 / nv;
  S0.x *= k; S0.y *= k; S0.z *= k; S0.w *= k;
  S1.x *= k; S1.y *= k; S1.z *= k; S1.w *= k;
  N.x *= k; N.y *= k; N.z *= k;
  S3.x *= k; S3.y *= k; S3.z *= k; S3.w *= k;
  AC.x *= k; AC.y *= k; AC.z *= k; AC.w *= k;
  MV *= k; Z *= k;
  real Lacc = S0.x;
  real VA = 0;
  for (int j = 0; j < 4; j++) if (ok[j]) VA += (la[j] - Lacc) * (la[j] - Lacc);
  VA *= k;
  /* 4x4 neighbourhood (hush.js:253-265) */
  real s1 = 0, n4 = 0;
  for (int j = 0; j < 16; j++) {
    v4 t = sI(c, 2 * X + (j & 3) - 1, 2 * Y + (j >> 2) - 1);
    s1 += t.w * E * qf16(lum(xyz(t)));
    n4 += t.w;
  }
  real mu4 = s1 / maxr(n4, 1);
  real var4 = 0;
  for (int j = 0; j < 16; j++) {
    v4 t = sI(c, 2 * X + (j & 3) - 1, 2 * Y + (j >> 2) - 1);
    real d = E * qf16(lum(xyz(t))) - mu4;
    var4 += t.w * d * d;
  }
  var4 /= maxr(n4, 1);
  real ln = maxr(length3(N), RX(1e-6));
  v3 Nn = mk3(N.x / ln, N.y / ln, N.z / ln);
  real zmax = RX(-1e30), zmin = RX(1e30), mind = RX(1e30);
  for (int j = 0; j < 4; j++)
    if (ok[j]) { zmax = maxr(zmax, zs[j]); zmin = minr(zmin, zs[j]); mind = minr(mind, dot3(ns[j], Nn)); }
  real dz = clampr((zmax - zmin) / maxr(zmin, RX(1e-6)), 0, 1);
  real dn = clampr(1 - mind, 0, 2);
  v2 oc = oct_enc(Nn);
  real f[20] = {log1p_(Lacc), log1p_(S0.y), log1p_(S0.z), log1p_(var4), S0.w / HMAX, S1.x, log1p_(S1.y),
                log1p_(FABS(S1.z - S1.w) / (S1.w + RX(1e-3))), oc.x, oc.y, dz, dn, S3.x, S3.y, S3.z, S3.w,
                1 - nv / 4, log1p_(MV), log1p_(mu4), log1p_(VA)};
  for (int i = 0; i < 20; i++) F[i] = st(qf16(f[i]));
  F[20] = st(qf16(AC.x)); F[21] = st(qf16(AC.y)); F[22] = st(qf16(AC.z)); F[23] = st(qf16(AC.w));
  F[24] = st(qf16(oc.x)); F[25] = st(qf16(oc.y)); F[26] = st(qf16(Z)); F[27] = st(qf16(nv));
  F[28] = F[29] = F[30] = F[31] = 0;
}

static void pass_temporal(ho_ctx *c) {
  int cur = c->frame & 1, prev = cur ^ 1;
  #pragma omp parallel for schedule(static)
  for (int y = 0; y < c->H; y++)
    for (int x = 0; x < c->W; x++) {
      v4 e = illum_at(c, x, y);
      real *p = c->sI + ((size_t)y * c->W + x) * 4;
      p[0] = e.x; p[1] = e.y; p[2] = e.z; p[3] = e.w;
    }
  #pragma omp parallel for schedule(static)
  for (int y = 0; y < c->H; y++)
    for (int x = 0; x < c->W; x++) pixel(c, x, y, cur, prev);
  #pragma omp parallel for schedule(static)
  for (int Y = 0; Y < c->H2; Y++)
    for (int X = 0; X < c->W2; X++) block(c, X, Y);
}

/* ---------------------------------------------------------------- a-trous (atrousFn, hush.js:431-476) */

typedef struct {
  const real *A;    /* source: 4 values per texel (rgb, var) */
  const real *G;    /* guide: decoded normal xyz + depth per texel */
  int w, h;
} tile;

static v4 tA(const tile *t, int x, int y) { return ld4(t->A, t->w, clampi(x, 0, t->w - 1), clampi(y, 0, t->h - 1)); }
static v4 tG(const tile *t, int x, int y) { return ld4(t->G, t->w, clampi(x, 0, t->w - 1), clampi(y, 0, t->h - 1)); }

/* one edge-aware a-trous iteration of radius R (1: 3x3, 2: 5x5) and tap step at texel (qx, qy) */
static v4 filt(const tile *t, int qx, int qy, int R, int step, int prefilter, real s, real cc) {
  static const double h1[3] = {0.25, 0.5, 0.25};
  static const double h2[5] = {1.0 / 16, 1.0 / 4, 3.0 / 8, 1.0 / 4, 1.0 / 16};
  const double *hk = R == 1 ? h1 : h2;
  v4 gp = tG(t, qx, qy);
  v4 center = tA(t, qx, qy);
  real gv;
  if (prefilter) {
    gv = RX(0.0625) * (tA(t, qx - 1, qy - 1).w + tA(t, qx + 1, qy - 1).w + tA(t, qx - 1, qy + 1).w + tA(t, qx + 1, qy + 1).w)
       + RX(0.125) * (tA(t, qx, qy - 1).w + tA(t, qx - 1, qy).w + tA(t, qx + 1, qy).w + tA(t, qx, qy + 1).w)
       + RX(0.25) * center.w;
  } else {
    gv = center.w;
  }
  real lc = lum(xyz(center));
  real sl = s / (SQRT(maxr(gv, RX(1e-12))) + RX(1e-6));
  real zl = tG(t, qx - 1, qy).w, zr = tG(t, qx + 1, qy).w, zu = tG(t, qx, qy - 1).w, zd = tG(t, qx, qy + 1).w;
  real gx = minr(zr > 0 ? FABS(zr - gp.w) : RX(1e9), zl > 0 ? FABS(zl - gp.w) : RX(1e9));
  real gy = minr(zd > 0 ? FABS(zd - gp.w) : RX(1e9), zu > 0 ? FABS(zu - gp.w) : RX(1e9));
  real zg = minr(maxr(gx, gy), gp.w);
  /* 1 / (zg * |o| + 1e-3) per distinct tap distance |o| = sqrt(dx^2 + dy^2) * step */
  real iz[9];
  for (int d2 = 1; d2 <= 8; d2++) iz[d2] = 1 / (zg * (real)(sqrt((double)d2) * step) + RX(1e-3));
  real cw = (real)(hk[R] * hk[R]) * cc;
  real sw = cw;
  v3 sc = mk3(center.x * cw, center.y * cw, center.z * cw);
  real sv = center.w * cw * cw;
  for (int dy = -R; dy <= R; dy++)
    for (int dx = -R; dx <= R; dx++) {
      if (dx == 0 && dy == 0) continue;
      int ox = dx * step, oy = dy * step;
      v4 gq = tG(t, qx + ox, qy + oy);
      if (inside(qx + ox, qy + oy, t->w, t->h) && gq.w >= 0) {
        v4 cq = tA(t, qx + ox, qy + oy);
        real w = (real)(hk[dx + R] * hk[dy + R]) * pow128(maxr(dot3(xyz(gq), xyz(gp)), 0))
               * EXP(-FABS(gq.w - gp.w) * iz[dx * dx + dy * dy] - sl * FABS(lum(xyz(cq)) - lc));
        sw += w;
        sc.x += w * cq.x; sc.y += w * cq.y; sc.z += w * cq.z;
        sv += w * w * cq.w;
      }
    }
  return mk4(sc.x / sw, sc.y / sw, sc.z / sw, sv / (sw * sw));
}

/* guide from the F3 plane (half res): decoded oct normal, block depth (hush.js:337, 496) */
static void guide_from_F(ho_ctx *c, real *G) {
  for (int i = 0; i < c->W2 * c->H2; i++) {
    const real *F = c->surf[S_F] + (size_t)i * 32;
    v3 n = oct_dec(F[24], F[25]);
    G[i * 4 + 0] = n.x; G[i * 4 + 1] = n.y; G[i * 4 + 2] = n.z; G[i * 4 + 3] = F[26];
  }
}

/* ---------------------------------------------------------------- MLP (mlpCode, hush.js:683-704)
   Per output: bias, then the inputs in order (hush.js:691-692).  Each multiply-add is one fused FMA, a
   contraction WGSL permits for `b + f0 * w0 + f1 * w1 + ...`; it is the one fused operation of the FP32
   mirror (everything else is evaluated unfused). */

#if HO_QUANT && !defined(HO_DOUBLE)
/* f16 arithmetic of the shader-f16 MLP variant: every product and sum rounded to binary16 (RNE) */
static float h16(float x) {
  /* round to nearest even onto the binary16 grid (overflow to inf) */
  double a = fabs((double)x);
  if (x != x) return x;
  if (a >= 65520.0) return x < 0 ? -INFINITY : INFINITY;
  double q;
  if (a < ldexp(1.0, -14)) q = nearbyint(a * 16777216.0) / 16777216.0;
  else {
    int e;
    double m = frexp(a, &e);
    q = ldexp(nearbyint(m * 2048.0) / 2048.0, e);
  }
  return (float)(x < 0 ? -q : q);
}
#else
static real h16(real x) { return x; }
#endif

static void mlp(const ho_ctx *c, const real *f, real *o) {
  const int h = c->hidden;
  const real *W1 = c->w, *B1 = W1 + h * N_FEAT, *W2 = B1 + h, *B2 = W2 + h * h, *W3 = B2 + h, *B3 = W3 + N_OUT * h;
  real h1[256], h2[256];
  if (!c->f16mlp) {
    for (int j = 0; j < h; j++) {
      real a = B1[j];
      for (int i = 0; i < N_FEAT; i++) a = FMA(f[i], W1[j * N_FEAT + i], a);
      h1[j] = maxr(a, 0);
    }
    for (int j = 0; j < h; j++) {
      real a = B2[j];
      for (int i = 0; i < h; i++) a = FMA(h1[i], W2[j * h + i], a);
      h2[j] = maxr(a, 0);
    }
    for (int j = 0; j < N_OUT; j++) {
      real a = B3[j];
      for (int i = 0; i < h; i++) a = FMA(h2[i], W3[j * h + i], a);
      o[j] = a;
    }
  } else {
    /* vec4h arithmetic: f16(f[i]) inputs, f16 products and sums (hush.js:690,699-702) */
    for (int j = 0; j < h; j++) {
      real a = h16(B1[j]);
      for (int i = 0; i < N_FEAT; i++) a = h16(a + h16(h16(f[i]) * h16(W1[j * N_FEAT + i])));
      h1[j] = maxr(a, 0);
    }
    for (int j = 0; j < h; j++) {
      real a = h16(B2[j]);
      for (int i = 0; i < h; i++) a = h16(a + h16(h1[i] * h16(W2[j * h + i])));
      h2[j] = maxr(a, 0);
    }
    for (int j = 0; j < N_OUT; j++) {
      real a = h16(B3[j]);
      for (int i = 0; i < h; i++) a = h16(a + h16(h2[i] * h16(W3[j * h + i])));
      o[j] = a;
    }
  }
}

/* MLP forward for one feature vector (exported for the MLP unit checks) */
EXPORT void HO_NAME(mlp)(const ho_ctx *c, const real *f, real *o) { mlp(c, f, o); }

/* ---------------------------------------------------------------- pass 2: infer + mid1 (hush.js:312-368) */

static void pass_infer(ho_ctx *c, real *G, real *Acc) {
  const int W2 = c->W2, H2 = c->H2;
  int cur = c->frame & 1;
  guide_from_F(c, G);
  for (int i = 0; i < W2 * H2; i++) {
    const real *F = c->surf[S_F] + (size_t)i * 32;
    Acc[i * 4 + 0] = F[20]; Acc[i * 4 + 1] = F[21]; Acc[i * 4 + 2] = F[22]; Acc[i * 4 + 3] = F[23];
  }
  tile t = {Acc, G, W2, H2};
  real *o0 = c->surf[cur ? S_PAR0_1 : S_PAR0_0], *o1 = c->surf[S_PAR1], *o2 = c->surf[S_PAR2], *om = c->surf[S_BH0];
  #pragma omp parallel for schedule(static)
  for (int Y = 0; Y < H2; Y++)
    for (int X = 0; X < W2; X++) {
      if (G[((size_t)Y * W2 + X) * 4 + 3] < 0) {
        v4 z = mk4(0, 0, 0, 0);
        st4(o0, W2, X, Y, z); st4(o1, W2, X, Y, z); st4(o2, W2, X, Y, z); st4(om, W2, X, Y, z);
        continue;
      }
      const real *F = fptr(c, X, Y);
      real o[N_OUT];
      mlp(c, F, o);
      real alpha = RX(1.0) / RX(32.0) + (RX(31.0) / RX(32.0)) * sigmoid(o[0]);
      real gamma = RX(0.5) + RX(7.5) * sigmoid(o[1]);
      real beta = sigmoid(o[2]);
      real kappa = sigmoid(o[3]);
      real s[4], cc[4];
      for (int k = 0; k < 4; k++) {
        s[k] = RX(0.25) * EXP(clampr(o[4 + k], -5, 5));
        cc[k] = EXP(clampr(o[8 + k], -4, 4));
      }
      st4(o0, W2, X, Y, mk4(alpha, gamma, beta, kappa));
      st4(o1, W2, X, Y, mk4(s[0], s[1], s[2], s[3]));
      st4(o2, W2, X, Y, mk4(cc[0], cc[1], cc[2], cc[3]));
      st4(om, W2, X, Y, filt(&t, X, Y, 1, 1, 1, qf16(s[1]), qf16(cc[1])));
    }
}

/* ---------------------------------------------------------------- passes 3-4: mid2, mid3 (atrousShader) */

static void pass_mid(ho_ctx *c, int step, int comp, int src, int dst, real *G) {
  const int W2 = c->W2, H2 = c->H2;
  guide_from_F(c, G);
  tile t = {c->surf[src], G, W2, H2};
  const real *p1 = c->surf[S_PAR1], *p2 = c->surf[S_PAR2];
  real *out = c->surf[dst];
  #pragma omp parallel for schedule(static)
  for (int Y = 0; Y < H2; Y++)
    for (int X = 0; X < W2; X++) {
      size_t i = (size_t)Y * W2 + X;
      if (G[i * 4 + 3] < 0) { st4(out, W2, X, Y, mk4(0, 0, 0, 0)); continue; }
      st4(out, W2, X, Y, filt(&t, X, Y, 1, step, 0, p1[i * 4 + comp], p2[i * 4 + comp]));
    }
}

/* ---------------------------------------------------------------- passes 5-6: coarse (hush.js:372-422) */

/* down (hush.js:385-402): count-weighted 2x2 half -> quarter */
static void down(ho_ctx *c, int QX, int QY, v4 *col, v4 *g) {
  real N = 0, vr = 0, z = 0;
  v3 cl = {0, 0, 0}, nrm = {0, 0, 0};
  for (int j = 0; j < 4; j++) {
    int hx = QX * 2 + (j & 1), hy = QY * 2 + (j >> 1);
    if (hx >= c->W2 || hy >= c->H2) continue;
    const real *F = fptr(c, hx, hy);
    if (F[27] <= 0) continue;
    real w = F[27];
    N += w;
    cl.x += w * F[20]; cl.y += w * F[21]; cl.z += w * F[22];
    vr += w * F[23];
    v3 d = oct_dec(F[24], F[25]);
    nrm.x += w * d.x; nrm.y += w * d.y; nrm.z += w * d.z;
    z += w * F[26];
  }
  if (N == 0) { *col = mk4(0, 0, 0, 0); *g = mk4(0, 0, -1, 0); return; }
  real ln = maxr(length3(nrm), RX(1e-6));
  v2 o = oct_enc(mk3(nrm.x / ln, nrm.y / ln, nrm.z / ln));
  *col = mk4(qf16(cl.x / N), qf16(cl.y / N), qf16(cl.z / N), qf16(vr / N / N));
  *g = mk4(qf16(o.x), qf16(o.y), qf16(z / N), 0);
}

static void pass_coarse1(ho_ctx *c, real *G, real *Acc) {
  const int W4 = c->W4, H4 = c->H4;
  for (int Y = 0; Y < H4; Y++)
    for (int X = 0; X < W4; X++) {
      v4 col, g;
      down(c, X, Y, &col, &g);
      size_t i = (size_t)Y * W4 + X;
      Acc[i * 4 + 0] = col.x; Acc[i * 4 + 1] = col.y; Acc[i * 4 + 2] = col.z; Acc[i * 4 + 3] = col.w;
      v3 n = oct_dec(g.x, g.y);
      G[i * 4 + 0] = n.x; G[i * 4 + 1] = n.y; G[i * 4 + 2] = n.z; G[i * 4 + 3] = g.z;
      st4(c->surf[S_CG], W4, X, Y, g);
    }
  tile t = {Acc, G, W4, H4};
  for (int Y = 0; Y < H4; Y++)
    for (int X = 0; X < W4; X++) {
      size_t i = (size_t)Y * W4 + X;
      if (G[i * 4 + 3] < 0) { st4(c->surf[S_C1], W4, X, Y, mk4(0, 0, 0, 0)); continue; }
      st4(c->surf[S_C1], W4, X, Y, filt(&t, X, Y, 2, 1, 1, RX(0.25), 1));
    }
}

static void pass_coarse2(ho_ctx *c, real *G) {
  const int W4 = c->W4, H4 = c->H4;
  const real *cg = c->surf[S_CG];
  for (int i = 0; i < W4 * H4; i++) {
    v3 n = oct_dec(cg[i * 4 + 0], cg[i * 4 + 1]);
    G[i * 4 + 0] = n.x; G[i * 4 + 1] = n.y; G[i * 4 + 2] = n.z; G[i * 4 + 3] = cg[i * 4 + 2];
  }
  tile t = {c->surf[S_C1], G, W4, H4};
  for (int Y = 0; Y < H4; Y++)
    for (int X = 0; X < W4; X++) {
      size_t i = (size_t)Y * W4 + X;
      if (G[i * 4 + 3] < 0) { st4(c->surf[S_C0], W4, X, Y, mk4(0, 0, 0, 0)); continue; }
      st4(c->surf[S_C0], W4, X, Y, filt(&t, X, Y, 2, 2, 1, RX(0.25), 1));
    }
}

/* ---------------------------------------------------------------- pass 7: fine + final (hush.js:530-622) */

/* jbw (hush.js:552-558) */
static real jbw(real frx, real fry, int j, v3 nq, real zq, v3 n, real z) {
  real wb = ((j & 1) == 1 ? frx : 1 - frx) * ((j >> 1) == 1 ? fry : 1 - fry);
  real dn = maxr(dot3(nq, n), 0);
  real dn2 = dn * dn;
  real dn4 = dn2 * dn2;
  return wb * EXP(-FABS(zq - z) / (RX(0.05) * z + RX(1e-3))) * dn4 * dn4;
}

static void pass_fine_final(ho_ctx *c, real *G) {
  const int W = c->W, H = c->H, W2 = c->W2, H2 = c->H2, W4 = c->W4, H4 = c->H4;
  int cur = c->frame & 1;
  const real *key = c->surf[cur ? S_KEY1 : S_KEY0];
  for (int i = 0; i < W * H; i++) {
    v3 n = oct_dec(key[i * 4 + 0], key[i * 4 + 1]);
    G[i * 4 + 0] = n.x; G[i * 4 + 1] = n.y; G[i * 4 + 2] = n.z; G[i * 4 + 3] = key[i * 4 + 2];
  }
  tile t = {c->surf[S_A0], G, W, H};
  real *hist = c->surf[cur ? S_HIST1 : S_HIST0], *X = c->surf[S_X];
  const real *par0 = c->surf[cur ? S_PAR0_1 : S_PAR0_0], *par1 = c->surf[S_PAR1], *par2 = c->surf[S_PAR2];
  const real *mid = c->surf[S_BH0], *coarse = c->surf[S_C0], *cg = c->surf[S_CG];
  #pragma omp parallel for schedule(static)
  for (int py = 0; py < H; py++)
    for (int px = 0; px < W; px++) {
      size_t i = (size_t)py * W + px;
      v4 alb = ld4(c->in[IN_MAT], W, px, py), emi = ld4(c->in[IN_EMI], W, px, py);
      v4 gp = ld4(G, W, px, py);
      if (gp.w < 0) {
        st4(hist, W, px, py, mk4(0, 0, 0, 0));
        st4(X, W, px, py, mk4(emi.x, emi.y, emi.z, 0));
        continue;
      }
      int hx = px >> 1, hy = py >> 1;
      size_t hi = (size_t)hy * W2 + hx;
      v4 A1 = filt(&t, px, py, 1, 1, 1, par1[hi * 4 + 0], par2[hi * 4 + 0]);
      st4(hist, W, px, py, A1);
      v3 f = mk3(qf16(A1.x), qf16(A1.y), qf16(A1.z));
      v3 n = xyz(gp);
      real z = gp.w;
      v4 pr = ld4(par0, W2, hx, hy);
      /* half-res mid level (hush.js:585-602) */
      real hpx = ((real)px + RX(0.5)) / 2 - RX(0.5), hpy = ((real)py + RX(0.5)) / 2 - RX(0.5);
      real hbx = FLOOR(hpx), hby = FLOOR(hpy);
      v3 uB = {0, 0, 0}, uA = {0, 0, 0};
      real swh = 0;
      for (int j = 0; j < 4; j++) {
        int qx = (int)hbx + (j & 1), qy = (int)hby + (j >> 1);
        if (qx < 0 || qy < 0 || qx >= W2 || qy >= H2) continue;
        const real *F = fptr(c, qx, qy);
        if (F[26] < 0) continue;
        real w = jbw(hpx - hbx, hpy - hby, j, oct_dec(F[24], F[25]), F[26], n, z);
        swh += w;
        v4 m = ld4(mid, W2, qx, qy);
        uB.x += w * m.x; uB.y += w * m.y; uB.z += w * m.z;
        uA.x += w * F[20]; uA.y += w * F[21]; uA.z += w * F[22];
      }
      v3 ill = f;
      if (swh > RX(1e-4)) {
        ill.x = uB.x / swh + pr.w * (f.x - uA.x / swh);
        ill.y = uB.y / swh + pr.w * (f.y - uA.y / swh);
        ill.z = uB.z / swh + pr.w * (f.z - uA.z / swh);
      }
      /* quarter-res coarse level (hush.js:604-619) */
      real cpx = ((real)px + RX(0.5)) / 4 - RX(0.5), cpy = ((real)py + RX(0.5)) / 4 - RX(0.5);
      real cbx = FLOOR(cpx), cby = FLOOR(cpy);
      v3 uC = {0, 0, 0};
      real swq = 0;
      for (int j = 0; j < 4; j++) {
        int qx = (int)cbx + (j & 1), qy = (int)cby + (j >> 1);
        if (qx < 0 || qy < 0 || qx >= W4 || qy >= H4) continue;
        v4 g = ld4(cg, W4, qx, qy);
        if (g.z < 0) continue;
        real w = jbw(cpx - cbx, cpy - cby, j, oct_dec(g.x, g.y), g.z, n, z);
        swq += w;
        v4 cc = ld4(coarse, W4, qx, qy);
        uC.x += w * cc.x; uC.y += w * cc.y; uC.z += w * cc.z;
      }
      v3 up = ill;
      if (swq > RX(1e-4)) up = mk3(uC.x / swq, uC.y / swq, uC.z / swq);
      ill.x = ill.x + (up.x - ill.x) * pr.z;
      ill.y = ill.y + (up.y - ill.y) * pr.z;
      ill.z = ill.z + (up.z - ill.z) * pr.z;
      st4(X, W, px, py, mk4(ill.x * alb.x + emi.x, ill.y * alb.y + emi.y, ill.z * alb.z + emi.z, 0));
      (void)i;
    }
}

/* ---------------------------------------------------------------- pass 8: TAA (hush.js:625-678) */

static v3 to_ycocg(v3 c) {
  return mk3(c.x * RX(0.25) + c.y * RX(0.5) + c.z * RX(0.25), c.x * RX(0.5) - c.z * RX(0.5),
             -c.x * RX(0.25) + c.y * RX(0.5) - c.z * RX(0.25));
}
static v3 from_ycocg(v3 c) { return mk3(c.x + c.y - c.z, c.x + c.z, c.x - c.y - c.z); }
static v3 tm(v3 c) { real d = 1 + lum(c); return mk3(c.x / d, c.y / d, c.z / d); }
static v3 itm(v3 c) { real d = maxr(1 - lum(c), RX(1e-4)); return mk3(c.x / d, c.y / d, c.z / d); }

static void pass_taa(ho_ctx *c, real *YC) {
  const int W = c->W, H = c->H;
  int cur = c->frame & 1, prev = cur ^ 1;
  const real *X = c->surf[S_X], *ptaa = c->surf[prev ? S_TAA1 : S_TAA0];
  real *out = c->surf[cur ? S_TAA1 : S_TAA0];
  for (int i = 0; i < W * H; i++) {
    v3 y = to_ycocg(tm(mk3(X[i * 4], X[i * 4 + 1], X[i * 4 + 2])));
    YC[i * 4] = y.x; YC[i * 4 + 1] = y.y; YC[i * 4 + 2] = y.z; YC[i * 4 + 3] = 0;
  }
  #pragma omp parallel for schedule(static)
  for (int py = 0; py < H; py++)
    for (int px = 0; px < W; px++) {
      v3 cu = tm(xyz(ld4(X, W, px, py)));
      if (c->reset) { v3 r = itm(cu); st4(out, W, px, py, mk4(r.x, r.y, r.z, 1)); continue; }
      v3 m1 = {0, 0, 0};
      for (int j = 0; j < 9; j++) {
        v4 s = ld4(YC, W, clampi(px + j % 3 - 1, 0, W - 1), clampi(py + j / 3 - 1, 0, H - 1));
        m1.x += s.x; m1.y += s.y; m1.z += s.z;
      }
      m1.x /= 9; m1.y /= 9; m1.z /= 9;
      v3 v3s = {0, 0, 0};
      for (int j = 0; j < 9; j++) {
        v4 s = ld4(YC, W, clampi(px + j % 3 - 1, 0, W - 1), clampi(py + j / 3 - 1, 0, H - 1));
        real dx = s.x - m1.x, dy = s.y - m1.y, dz = s.z - m1.z;
        v3s.x += dx * dx; v3s.y += dy * dy; v3s.z += dz * dz;
      }
      v3 sd = mk3(SQRT(maxr(v3s.x / 9, RX(1e-12))), SQRT(maxr(v3s.y / 9, RX(1e-12))), SQRT(maxr(v3s.z / 9, RX(1e-12))));
      v4 mv = ld4(c->in[IN_MV], W, px, py);
      real szx = (real)W, szy = (real)H;
      real ux = ((real)px + RX(0.5) + mv.x) / szx, uy = ((real)py + RX(0.5) + mv.y) / szy;
      v3 res = cu;
      if (ux >= 0 && uy >= 0 && ux <= 1 && uy <= 1) {
        real ppx = ux * szx - RX(0.5), ppy = uy * szy - RX(0.5);
        real bx = FLOOR(ppx), by = FLOOR(ppy);
        real fx = ppx - bx, fy = ppy - by;
        v3 hist = {0, 0, 0};
        for (int j = 0; j < 4; j++) {
          int ox = j & 1, oy = j >> 1;
          real wb = (ox == 1 ? fx : 1 - fx) * (oy == 1 ? fy : 1 - fy);
          v4 t = ld4(ptaa, W, clampi(addw(f2i(bx), ox), 0, W - 1), clampi(addw(f2i(by), oy), 0, H - 1));
          hist.x += wb * t.x; hist.y += wb * t.y; hist.z += wb * t.z;
        }
        v3 hh = to_ycocg(tm(hist));
        real kx = RX(1.25) * sd.x, ky = RX(1.25) * sd.y, kz = RX(1.25) * sd.z;
        hh.x = clampr(hh.x, m1.x - kx, m1.x + kx);
        hh.y = clampr(hh.y, m1.y - ky, m1.y + ky);
        hh.z = clampr(hh.z, m1.z - kz, m1.z + kz);
        v3 hr = from_ycocg(hh);
        res = mk3(hr.x + (cu.x - hr.x) * RX(0.1), hr.y + (cu.y - hr.y) * RX(0.1), hr.z + (cu.z - hr.z) * RX(0.1));
      }
      v3 r = itm(res);
      st4(out, W, px, py, mk4(r.x, r.y, r.z, 1));
    }
}

/* ---------------------------------------------------------------- sequencing (hush.js:835-874) */

enum { P_TEMPORAL, P_INFER, P_MID2, P_MID3, P_COARSE1, P_COARSE2, P_FINE, P_TAA };

/* one pass at the current frame / reset state (no advance); used for teacher-forced checks */
EXPORT int HO_NAME(run_pass)(ho_ctx *c, int pass) {
  size_t n = (size_t)c->W * c->H * 4;
  real *G = (real *)malloc(n * sizeof(real));
  real *T = (real *)malloc(n * sizeof(real));
  switch (pass) {
    case P_TEMPORAL: pass_temporal(c); break;
    case P_INFER: pass_infer(c, G, T); break;
    case P_MID2: pass_mid(c, 2, 2, S_BH0, S_BH1, G); break;
    case P_MID3: pass_mid(c, 4, 3, S_BH1, S_BH0, G); break;
    case P_COARSE1: pass_coarse1(c, G, T); break;
    case P_COARSE2: pass_coarse2(c, G); break;
    case P_FINE: pass_fine_final(c, G); break;
    case P_TAA: pass_taa(c, T); break;
    default: free(G); free(T); return -1;
  }
  free(G); free(T);
  return 0;
}

/* denoise (hush.js:813-822) over already-set inputs: 8 passes, then clear reset and advance the frame */
EXPORT int HO_NAME(denoise)(ho_ctx *c) {
  for (int p = P_TEMPORAL; p <= P_TAA; p++) HO_NAME(run_pass)(c, p);
  c->reset = 0;
  c->frame++;
  return (c->frame + 1) & 1 ? S_TAA1 : S_TAA0;   /* surface id of the output */
}

/* output surface id (hush.js:798): taa[(frame + 1) & 1] */
EXPORT int HO_NAME(output_id)(const ho_ctx *c) { return ((c->frame + 1) & 1) ? S_TAA1 : S_TAA0; }

/* binary16 helpers exported for the oracle's own pins */
EXPORT int HO_NAME(threads)(void) { return omp_get_max_threads(); }   /* OpenMP threads of the passes */
EXPORT uint16_t HO_NAME(f2h_rtz)(real x) { return r2h_rtz(x); }
EXPORT real HO_NAME(h2f)(uint16_t h) { return h2r(h); }
EXPORT real HO_NAME(qf16)(real x) { return qf16(x); }
EXPORT void HO_NAME(oct_enc)(real x, real y, real z, real *out) { v2 r = oct_enc(mk3(x, y, z)); out[0] = r.x; out[1] = r.y; }
EXPORT void HO_NAME(oct_dec)(real x, real y, real *out) { v3 r = oct_dec(x, y); out[0] = r.x; out[1] = r.y; out[2] = r.z; }
EXPORT int HO_NAME(check)(const real *k, const real *n, real zprev, real obj) {
  return check(mk4(k[0], k[1], k[2], k[3]), mk3(n[0], n[1], n[2]), zprev, obj);
}
