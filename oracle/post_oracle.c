/*
 * CPU oracle of the display post (src/demo/post.js).  TEST INFRASTRUCTURE ONLY: only tests/ call it, as the
 * checker.
 *
 * FP32 restatement of the four render passes (post.js:21-127) and their sequencing (post.js:179-209), with
 * the conventions the GPU implementation documents:
 *   - a fragment at pixel (x, y) of a w x h target has uv = ((x + 0.5) / w, (y + 0.5) / h);
 *   - the linear clamp-to-edge sampler is exact bilinear weights at uv * size - 0.5 (the hardware's
 *     sub-texel weight precision is implementation-defined: parity unpinned there);
 *   - mat3x3f(c0, c1, c2) * v sums the columns in order; mix(x, y, a) = x (1 - a) + y a;
 *   - pow(x, y) = exp2(y log2(x)), WGSL's definition of pow;
 *   - every rgba16float store (render target or additive blend, the blend computed in FP32 on the stored
 *     destination) rounds toward zero with a clamp at +-65504; the 8-bit target is
 *     rint(clamp(x, 0, 1) * 255) (round half to even);
 *   - log2f / exp2f are the C library's; the GPU uses the hardware log2 / exp2 (WGSL bounds their error
 *     only: parity unpinned at the last ulp, the 8-bit target within one step).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))
#define LEVELS 6

static float h2f(uint16_t h) {
  int s = h >> 15, e = (h >> 10) & 31, m = h & 1023;
  double v;
  if (e == 0) v = ldexp((double)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = ldexp((double)(1024 + m), e - 25);
  return (float)(s ? -v : v);
}

/* rgba16float store: round toward zero, clamped to +-65504 (NaN to -65504, the clamp's IEEE min / max) */
static float st(float x) {
  if (x != x) return -65504.0f;
  double a = fabs((double)x), r;
  if (a >= 65504.0) r = 65504.0;
  else if (a < ldexp(1.0, -14)) r = floor(a * 16777216.0) / 16777216.0;
  else { int e; double m = frexp(a, &e); r = ldexp(floor(m * 2048.0) / 2048.0, e); }
  return (float)(x < 0 ? -r : r);
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int f2i(float x) {
  if (x != x) return 0;
  if (x < -2147483648.0f) return (int)(-2147483647 - 1);
  if (x > 2147483520.0f) return 2147483520;
  return (int)x;
}
static float lum(float r, float g, float b) { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }

typedef struct { float *p; int w, h; } surf;   /* rgba, row-major, stored (quantised) values */

static surf alloc_surf(int w, int h) {
  surf s = {(float *)calloc((size_t)w * h * 4, sizeof(float)), w, h};
  return s;
}
static float *px(const surf *s, int x, int y) { return s->p + ((size_t)y * s->w + x) * 4; }

/* textureSampleLevel with a linear clamp-to-edge sampler (post.js:166) */
static void sample(const surf *s, float u, float v, float *o) {
  const float x = u * (float)s->w - 0.5f, y = v * (float)s->h - 0.5f;
  const float bx = floorf(x), by = floorf(y);
  const float fx = x - bx, fy = y - by;
  const int ix = f2i(bx), iy = f2i(by);
  const int x0 = clampi(ix, 0, s->w - 1), x1 = clampi((int)((unsigned)ix + 1u), 0, s->w - 1);
  const int y0 = clampi(iy, 0, s->h - 1), y1 = clampi((int)((unsigned)iy + 1u), 0, s->h - 1);
  const float *t00 = px(s, x0, y0), *t10 = px(s, x1, y0), *t01 = px(s, x0, y1), *t11 = px(s, x1, y1);
  const float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
  for (int c = 0; c < 3; c++) o[c] = w00 * t00[c] + w10 * t10[c] + w01 * t01[c] + w11 * t11[c];
}

/* COMPOSITE (post.js:21-40) */
static void composite(const surf *left, const surf *right, const surf *mid, surf *out, float split, float split2,
                      int same, float exposure) {
  const int W = out->w, H = out->h;
  for (int y = 0; y < H; y++)
    for (int x0 = 0; x0 < W; x0++) {
      int x = x0;
      if (same) {
        int k = x * 3 / W;
        x = x - ((k < 2 ? k : 2) - 1) * (W / 3);
      }
      const float u = ((float)x0 + 0.5f) / (float)W;
      const surf *src = right;
      if (u < split2) src = mid;
      if (u < split) src = left;
      const float *c = px(src, x, y);
      float *d = px(out, x0, y);
      for (int k = 0; k < 3; k++) d[k] = st(fmaxf(c[k], 0.0f) * exposure);
      d[3] = 1.0f;
    }
}

static float karis_w(const float *c) { return 1.0f / (1.0f + lum(c[0], c[1], c[2])); }

/* DOWN (post.js:42-66) */
static void down(const surf *src, surf *out, int karis) {
  const float tx = (float)(1.0 / src->w), ty = (float)(1.0 / src->h);
  static const float off[13][2] = {{-2, -2}, {0, -2}, {2, -2}, {-2, 0}, {0, 0}, {2, 0}, {-2, 2}, {0, 2}, {2, 2},
                                   {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
  enum { A, B, C, D, E, F, G, H, K, J, L, M, N };
  for (int y = 0; y < out->h; y++)
    for (int x = 0; x < out->w; x++) {
      const float u = ((float)x + 0.5f) / (float)out->w, v = ((float)y + 0.5f) / (float)out->h;
      float s[13][3];
      for (int i = 0; i < 13; i++) sample(src, u + off[i][0] * tx, v + off[i][1] * ty, s[i]);
      float o[3];
      if (karis) {
        float g[5][3];
        static const int grp[5][4] = {{A, B, D, E}, {B, C, E, F}, {D, E, G, H}, {E, F, H, K}, {J, L, M, N}};
        for (int q = 0; q < 5; q++)
          for (int c = 0; c < 3; c++)
            g[q][c] = (((s[grp[q][0]][c] + s[grp[q][1]][c]) + s[grp[q][2]][c]) + s[grp[q][3]][c]) * 0.25f;
        float w[5];
        for (int q = 0; q < 4; q++) w[q] = karis_w(g[q]) * 0.125f;
        w[4] = karis_w(g[4]) * 0.5f;
        const float d = (((w[0] + w[1]) + w[2]) + w[3]) + w[4];
        for (int c = 0; c < 3; c++)
          o[c] = ((((g[0][c] * w[0] + g[1][c] * w[1]) + g[2][c] * w[2]) + g[3][c] * w[3]) + g[4][c] * w[4]) / d;
      } else {
        for (int c = 0; c < 3; c++)
          o[c] = ((s[E][c] * 0.125f + (((s[A][c] + s[C][c]) + s[G][c]) + s[K][c]) * 0.03125f) +
                  (((s[B][c] + s[D][c]) + s[F][c]) + s[H][c]) * 0.0625f) +
                 (((s[J][c] + s[L][c]) + s[M][c]) + s[N][c]) * 0.125f;
      }
      float *d = px(out, x, y);
      for (int c = 0; c < 3; c++) d[c] = st(o[c]);
      d[3] = 1.0f;
    }
}

/* UP (post.js:68-82) with the additive blend (post.js:141) */
static void up(const surf *src, surf *dst) {
  const float tx = (float)(1.0 / src->w), ty = (float)(1.0 / src->h);
  for (int y = 0; y < dst->h; y++)
    for (int x = 0; x < dst->w; x++) {
      const float u = ((float)x + 0.5f) / (float)dst->w, v = ((float)y + 0.5f) / (float)dst->h;
      float c0[3], l[3], r[3], t[3], b[3], d0[3], d1[3], d2[3], d3[3];
      sample(src, u, v, c0);
      sample(src, u + -1.0f * tx, v + 0.0f * ty, l); sample(src, u + 1.0f * tx, v + 0.0f * ty, r);
      sample(src, u + 0.0f * tx, v + -1.0f * ty, t); sample(src, u + 0.0f * tx, v + 1.0f * ty, b);
      sample(src, u + -1.0f * tx, v + -1.0f * ty, d0); sample(src, u + 1.0f * tx, v + -1.0f * ty, d1);
      sample(src, u + -1.0f * tx, v + 1.0f * ty, d2); sample(src, u + 1.0f * tx, v + 1.0f * ty, d3);
      float *d = px(dst, x, y);
      for (int c = 0; c < 3; c++) {
        float o = c0[c] * 4.0f;
        o = o + (((l[c] + r[c]) + t[c]) + b[c]) * 2.0f;
        o = o + (((d0[c] + d1[c]) + d2[c]) + d3[c]);
        d[c] = st(d[c] + o / 16.0f);
      }
      d[3] = st(d[3] + 1.0f);
    }
}

static float powx(float x, float y) { return exp2f(y * log2f(x)); }

static void mat_mul(const float *m, const float *v, float *o) {
  float r[3];
  for (int k = 0; k < 3; k++) r[k] = m[k] * v[0] + m[3 + k] * v[1] + m[6 + k] * v[2];
  memcpy(o, r, sizeof(r));
}
static float agx_contrast(float x) {   /* post.js:94-98 */
  const float x2 = x * x, x4 = x2 * x2;
  return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x - 0.00232f;
}
/* AgX with the look of post.js:99-119 (matrices column by column as written there) */
EXPORT void po32_agx(const float *c0, float *out) {
  static const float toRec2020[9] = {0.6274f, 0.0691f, 0.0164f, 0.3293f, 0.9195f, 0.0880f, 0.0433f, 0.0113f, 0.8956f};
  static const float toSrgb[9] = {1.6605f, -0.1246f, -0.0182f, -0.5876f, 1.1329f, -0.1006f, -0.0728f, -0.0083f, 1.1187f};
  static const float inset[9] = {0.856627153315983f, 0.137318972929847f, 0.11189821299995f,
                                 0.0951212405381588f, 0.761241990602591f, 0.0767994186031903f,
                                 0.0482516061458583f, 0.101439036467562f, 0.811302368396859f};
  static const float outset[9] = {1.1271005818144368f, -0.1413297634984383f, -0.14132976349843826f,
                                  -0.11060664309660323f, 1.157823702216272f, -0.11060664309660294f,
                                  -0.016493938717834573f, -0.016493938717834257f, 1.2519364065950405f};
  const float minEv = -12.47393f, maxEv = 4.026069f;
  float c[3];
  mat_mul(toRec2020, c0, c);
  mat_mul(inset, c, c);
  const float r = maxEv - minEv;
  for (int k = 0; k < 3; k++) c[k] = fminf(fmaxf((log2f(fmaxf(c[k], 1e-10f)) - minEv) / r, 0.0f), 1.0f);
  for (int k = 0; k < 3; k++) c[k] = agx_contrast(c[k]);
  const float l = lum(c[0], c[1], c[2]);
  for (int k = 0; k < 3; k++) c[k] = powx(fmaxf(c[k], 0.0f), 1.22f);
  for (int k = 0; k < 3; k++) c[k] = l + (c[k] - l) * 1.15f;
  mat_mul(outset, c, c);
  for (int k = 0; k < 3; k++) c[k] = powx(fmaxf(0.0f, c[k]), 2.2f);
  mat_mul(toSrgb, c, c);
  for (int k = 0; k < 3; k++) out[k] = fminf(fmaxf(c[k], 0.0f), 1.0f);
}
EXPORT float po32_srgb(float c) {     /* post.js:120 */
  return c <= 0.0031308f ? c * 12.92f : 1.055f * powx(c, (float)(1.0 / 2.4)) - 0.055f;
}
static uint32_t unorm8(float x) { return (uint32_t)rintf(fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f); }

/* FINAL (post.js:84-127) */
static void final_pass(const surf *hdr, const surf *bloom, uint8_t *target, int W, int H, float split, float split2,
                       float bloom_mix, float vignette, float line, float width, int bgra) {
  const float tx = (float)(1.0 / bloom->w), ty = (float)(1.0 / bloom->h);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
      float c[3], s0[3], a[3], b[3], cc[3], d[3], e0[3], e1[3], e2[3], e3[3];
      sample(hdr, u, v, c);
      sample(bloom, u, v, s0);
      sample(bloom, u + -tx, v + 0.0f, a); sample(bloom, u + tx, v + 0.0f, b);
      sample(bloom, u + 0.0f, v + -ty, cc); sample(bloom, u + 0.0f, v + ty, d);
      sample(bloom, u - tx, v - ty, e0); sample(bloom, u + tx, v + ty, e1);
      sample(bloom, u + tx, v + -ty, e2); sample(bloom, u + -tx, v + ty, e3);
      for (int k = 0; k < 3; k++) {
        float bb = s0[k] * 4.0f;
        bb = bb + (((a[k] + b[k]) + cc[k]) + d[k]) * 2.0f;
        bb = bb + (((e0[k] + e1[k]) + e2[k]) + e3[k]);
        c[k] = c[k] * (1.0f - bloom_mix) + (bb / 16.0f) * bloom_mix;
      }
      const float qx = (u - 0.5f) * 1.0f, qy = (v - 0.5f) * 0.75f;
      const float vg = 1.0f * (1.0f - vignette) + (1.0f - (qx * qx + qy * qy) * 1.6f) * vignette;
      for (int k = 0; k < 3; k++) c[k] = c[k] * vg;
      float t[3], o[3];
      po32_agx(c, t);
      for (int k = 0; k < 3; k++) o[k] = po32_srgb(t[k]);
      if (line > 0.0f && fminf(fabsf(u - split), fabsf(u - split2)) * width < line) {
        const float kk = 1.0f - 0.85f;
        for (int k = 0; k < 3; k++) o[k] = o[k] * kk + 1.0f * 0.85f;
      }
      uint8_t *p = target + ((size_t)y * W + x) * 4;
      const uint32_t r = unorm8(o[0]), g = unorm8(o[1]), bl = unorm8(o[2]);
      p[0] = (uint8_t)(bgra ? bl : r); p[1] = (uint8_t)g; p[2] = (uint8_t)(bgra ? r : bl); p[3] = 255;
    }
}

/* the sampler alone, for its pins: img w x h x 4 floats */
EXPORT void po32_sample(const float *img, int w, int h, float u, float v, float *out) {
  surf s = {(float *)img, w, h};
  sample(&s, u, v, out);
}
/* one pass alone: 0 down (karis), 1 down, 2 up (dst in / out) */
EXPORT void po32_pass(int which, const float *src, int sw, int sh, float *dst, int w, int h) {
  surf a = {(float *)src, sw, sh}, d = {dst, w, h};
  if (which == 2) up(&a, &d);
  else down(&a, &d, which == 0);
}

static surf load_image(const uint16_t *h, int w, int ht) {
  surf s = alloc_surf(w, ht);
  for (size_t i = 0; i < (size_t)w * ht * 4; i++) s.p[i] = h2f(h[i]);
  return s;
}

/* Post.encode (post.js:179-209).  Images: binary16 rgba, w x h, row-major; mid NULL = right.  Outputs:
   `target` tw x th x 4 bytes; `hdr` (w x h x 4 floats) and `levels` (levels 1..6 concatenated, each
   max(1, w >> i) x max(1, h >> i) x 4 floats) when not NULL. */
EXPORT int po32_encode(const uint16_t *left, const uint16_t *right, const uint16_t *mid, int w, int h, float split,
                       float split2, int same, float exposure, float bloom_mix, float vignette, float line, int tw,
                       int th, int bgra, uint8_t *target, float *hdr_out, float *levels_out) {
  if (w <= 0 || h <= 0 || tw <= 0 || th <= 0) return -1;
  surf L = load_image(left, w, h), R = load_image(right, w, h), M = mid ? load_image(mid, w, h) : R;
  surf lv[1 + LEVELS];
  lv[0] = alloc_surf(w, h);
  for (int i = 1; i <= LEVELS; i++) lv[i] = alloc_surf(w >> i > 0 ? w >> i : 1, h >> i > 0 ? h >> i : 1);
  composite(&L, &R, &M, &lv[0], split, split2, same, exposure);
  for (int i = 1; i <= LEVELS; i++) down(&lv[i - 1], &lv[i], i == 1);
  for (int i = LEVELS - 1; i >= 1; i--) up(&lv[i + 1], &lv[i]);
  final_pass(&lv[0], &lv[1], target, tw, th, split, split2, bloom_mix, vignette, line, (float)tw, bgra);
  if (hdr_out) memcpy(hdr_out, lv[0].p, (size_t)w * h * 4 * sizeof(float));
  if (levels_out) {
    size_t o = 0;
    for (int i = 1; i <= LEVELS; i++) {
      memcpy(levels_out + o, lv[i].p, (size_t)lv[i].w * lv[i].h * 4 * sizeof(float));
      o += (size_t)lv[i].w * lv[i].h * 4;
    }
  }
  free(L.p); free(R.p);
  if (mid) free(M.p);
  for (int i = 0; i <= LEVELS; i++) free(lv[i].p);
  return 0;
}
