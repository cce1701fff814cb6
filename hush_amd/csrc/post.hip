// Display post on gfx950: the four render passes of src/demo/post.js (composite, bloom downsample, bloom
// upsample, final AgX pass) as compute kernels, and the C ABI of include/post.h.
//
// A fragment of the reference's full-screen triangle at pixel (x, y) of a w x h target sees
// uv = ((x + 0.5) / w, (y + 0.5) / h); the kernels evaluate exactly that, one thread per target pixel.
// The linear clamp-to-edge sampler is evaluated as exact bilinear weights at uv * size - 0.5 (a hardware
// sampler's sub-texel weight precision is implementation-defined).  Matrix-vector products sum the
// columns in order, mix(x, y, a) is x (1 - a) + y a and pow(x, y) is exp2(y log2(x)), WGSL's definitions
// (evaluated with the hardware v_log_f32 / v_exp_f32, ~1 ulp; a display pass).  Every rgba16float store,
// the additive blend's included (computed in FP32 on the stored destination), rounds toward zero like
// the denoisers' surfaces; the 8-bit target takes clamp(x, 0, 1) * 255 rounded to nearest even.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "hush_device.h"
#include "post.h"

using namespace hush;

namespace {

constexpr int TW = 32, TH = 8, NT = 256;

__device__ __forceinline__ int2 pixel_of(int W, int H) {
  const int2 o = tile_origin<TW, TH>(W, H);
  return make_int2(o.x + (int)(threadIdx.x & (TW - 1)), o.y + (int)(threadIdx.x / TW));
}

__device__ __forceinline__ float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 mul3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }

// One axis of a linear clamp-to-edge sampler at normalised coordinate u of an n-texel extent: the two
// clamped texel indices and the weights 1 - f, f at u * n - 0.5.  A filter whose taps share a row or column
// offset computes it once for all of them.
struct Axis {
  int i0, i1;
  float g, f;                    // 1 - f, f
};
__device__ __forceinline__ Axis axis(float u, int n) {
  const float x = u * (float)n - 0.5f;
  const float b = floorf(x);
  const int i = f2i(b);
  Axis a;
  a.f = x - b;
  a.g = 1.0f - a.f;
  a.i0 = clampi(i, 0, n - 1);
  a.i1 = clampi(addw(i, 1), 0, n - 1);
  return a;
}
// the two texel rows of a vertical axis
struct Rows {
  const uint2 *r0, *r1;
  float g, f;
};
__device__ __forceinline__ Rows rows(const Tex &t, const Axis &ay) {
  return Rows{t.p + (size_t)ay.i0 * t.pitch, t.p + (size_t)ay.i1 * t.pitch, ay.g, ay.f};
}
// textureSampleLevel(src, linear clamp-to-edge sampler, uv).rgb of an rgba16float surface
__device__ __forceinline__ float3 bilerp(const Rows &ry, const Axis &ax) {
  const float4 t00 = unpack4(ry.r0[ax.i0]), t10 = unpack4(ry.r0[ax.i1]);
  const float4 t01 = unpack4(ry.r1[ax.i0]), t11 = unpack4(ry.r1[ax.i1]);
  const float w00 = ax.g * ry.g, w10 = ax.f * ry.g, w01 = ax.g * ry.f, w11 = ax.f * ry.f;
  return make_float3(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
                     w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
                     w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z);
}
__device__ __forceinline__ float3 sample(const Tex &t, int w, int h, float u, float v) {
  return bilerp(rows(t, axis(v, h)), axis(u, w));
}

struct CompositeArgs {
  Tex left, right, mid, out;
  int W, H, same;
  float split, split2, exposure;
};
struct DownArgs {
  Tex src, out;
  int sw, sh, w, h, karis;
  float tx, ty;                  // RN(1 / sw), RN(1 / sh)
};
struct UpArgs {
  Tex src, dst;
  int sw, sh, w, h;
  float tx, ty;
};
struct FinalArgs {
  Tex hdr, bloom;
  int hw, hh, bw, bh;
  uint32_t *target;
  int pitch;                     // texels per row
  int W, H, bgra;
  float split, split2, bloom_mix, vignette, line, width;
  float tx, ty;                  // RN(1 / bw), RN(1 / bh)
};

// COMPOSITE (post.js:21-40): left | mid | right at split / split2, clamped at 0, times exposure
__global__ __launch_bounds__(NT) void k_post_composite(CompositeArgs a) {
  const int2 p = pixel_of(a.W, a.H);
  if (!inside(p.x, p.y, a.W, a.H)) return;
  int x = p.x;
  if (a.same) x = x - (min(x * 3 / a.W, 2) - 1) * (a.W / 3);
  const float u = ((float)p.x + 0.5f) / (float)a.W;
  const Tex &src = u < a.split ? a.left : (u < a.split2 ? a.mid : a.right);
  const float4 c = ld(src, x, p.y);
  st(a.out, p.x, p.y, make_float4(fmaxf(c.x, 0.0f) * a.exposure, fmaxf(c.y, 0.0f) * a.exposure,
                                  fmaxf(c.z, 0.0f) * a.exposure, 1.0f));
}

// 1 / (1 + lum(c)) (post.js:51)
__device__ __forceinline__ float karis_w(float3 c) { return 1.0f / (1.0f + lum(c.x, c.y, c.z)); }

// DOWN (post.js:42-66): 13-tap downsample; Karis-weighted average of its five 2x2 groups on the first level
__global__ __launch_bounds__(NT) void k_post_down(DownArgs a) {
  const int2 p = pixel_of(a.w, a.h);
  if (!inside(p.x, p.y, a.w, a.h)) return;
  const float u = ((float)p.x + 0.5f) / (float)a.w, v = ((float)p.y + 0.5f) / (float)a.h;
  Axis ax[5];                       // offsets -2 .. 2 texels
  Rows ry[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    ax[k] = axis(u + (float)(k - 2) * a.tx, a.sw);
    ry[k] = rows(a.src, axis(v + (float)(k - 2) * a.ty, a.sh));
  }
  auto s = [&](int ox, int oy) { return bilerp(ry[oy + 2], ax[ox + 2]); };
  const float3 A = s(-2, -2), B = s(0, -2), C = s(2, -2);
  const float3 D = s(-2, 0), E = s(0, 0), F = s(2, 0);
  const float3 G = s(-2, 2), H = s(0, 2), K = s(2, 2);
  const float3 J = s(-1, -1), L = s(1, -1), M = s(-1, 1), N = s(1, 1);
  float3 o;
  if (a.karis) {
    const float3 g0 = mul3(add3(add3(add3(A, B), D), E), 0.25f), g1 = mul3(add3(add3(add3(B, C), E), F), 0.25f);
    const float3 g2 = mul3(add3(add3(add3(D, E), G), H), 0.25f), g3 = mul3(add3(add3(add3(E, F), H), K), 0.25f);
    const float3 g4 = mul3(add3(add3(add3(J, L), M), N), 0.25f);
    const float w0 = karis_w(g0) * 0.125f, w1 = karis_w(g1) * 0.125f, w2 = karis_w(g2) * 0.125f;
    const float w3 = karis_w(g3) * 0.125f, w4 = karis_w(g4) * 0.5f;
    const float3 n = add3(add3(add3(add3(mul3(g0, w0), mul3(g1, w1)), mul3(g2, w2)), mul3(g3, w3)), mul3(g4, w4));
    const float d = (((w0 + w1) + w2) + w3) + w4;
    o = make_float3(n.x / d, n.y / d, n.z / d);
  } else {
    o = add3(add3(add3(mul3(E, 0.125f), mul3(add3(add3(add3(A, C), G), K), 0.03125f)),
                  mul3(add3(add3(add3(B, D), F), H), 0.0625f)),
             mul3(add3(add3(add3(J, L), M), N), 0.125f));
  }
  st(a.out, p.x, p.y, make_float4(o.x, o.y, o.z, 1.0f));
}

// UP (post.js:68-82) with the additive blend of post.js:141: dst += 3x3 tent of the smaller level
__global__ __launch_bounds__(NT) void k_post_up(UpArgs a) {
  const int2 p = pixel_of(a.w, a.h);
  if (!inside(p.x, p.y, a.w, a.h)) return;
  const float u = ((float)p.x + 0.5f) / (float)a.w, v = ((float)p.y + 0.5f) / (float)a.h;
  Axis ax[3];                       // offsets -1, 0, 1 texel
  Rows ry[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ax[k] = axis(u + (float)(k - 1) * a.tx, a.sw);
    ry[k] = rows(a.src, axis(v + (float)(k - 1) * a.ty, a.sh));
  }
  auto s = [&](int ox, int oy) { return bilerp(ry[oy + 1], ax[ox + 1]); };
  float3 o = mul3(s(0, 0), 4.0f);
  o = add3(o, mul3(add3(add3(add3(s(-1, 0), s(1, 0)), s(0, -1)), s(0, 1)), 2.0f));
  o = add3(o, add3(add3(add3(s(-1, -1), s(1, -1)), s(-1, 1)), s(1, 1)));
  const float4 d = ld(a.dst, p.x, p.y);
  st(a.dst, p.x, p.y, make_float4(d.x + o.x / 16.0f, d.y + o.y / 16.0f, d.z + o.z / 16.0f, d.w + 1.0f));
}

// pow(x, y) = exp2(y log2(x)) for x >= 0 (0 -> 0 for y > 0)
__device__ __forceinline__ float powx(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// mat3x3f(c0, c1, c2) * v: the columns in order
__device__ __forceinline__ float3 mat_mul(const float (&m)[9], float3 v) {
  return make_float3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
                     m[2] * v.x + m[5] * v.y + m[8] * v.z);
}
__device__ __forceinline__ float agx_contrast(float x) {     // post.js:94-98
  const float x2 = x * x, x4 = x2 * x2;
  return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x - 0.00232f;
}
// AgX with the "punchy" look (post.js:99-119); matrices column by column as written there
__device__ float3 agx(float3 c0) {
  const float toRec2020[9] = {0.6274f, 0.0691f, 0.0164f, 0.3293f, 0.9195f, 0.0880f, 0.0433f, 0.0113f, 0.8956f};
  const float toSrgb[9] = {1.6605f, -0.1246f, -0.0182f, -0.5876f, 1.1329f, -0.1006f, -0.0728f, -0.0083f, 1.1187f};
  const float inset[9] = {0.856627153315983f, 0.137318972929847f, 0.11189821299995f,
                          0.0951212405381588f, 0.761241990602591f, 0.0767994186031903f,
                          0.0482516061458583f, 0.101439036467562f, 0.811302368396859f};
  const float outset[9] = {1.1271005818144368f, -0.1413297634984383f, -0.14132976349843826f,
                           -0.11060664309660323f, 1.157823702216272f, -0.11060664309660294f,
                           -0.016493938717834573f, -0.016493938717834257f, 1.2519364065950405f};
  const float minEv = -12.47393f, maxEv = 4.026069f;
  float3 c = mat_mul(inset, mat_mul(toRec2020, c0));
  // (log2(c) - minEv) / (maxEv - minEv): a shared-reciprocal quotient, the IEEE one here (the numerator is 0
  // or at least ~1e-6 in magnitude, so the quotient is normal or zero)
  const float r = maxEv - minEv, rr = 1.0f / r;
  using M = Math<false>;
  c.x = fminf(fmaxf(M::divs(__builtin_amdgcn_logf(fmaxf(c.x, 1e-10f)) - minEv, r, rr), 0.0f), 1.0f);
  c.y = fminf(fmaxf(M::divs(__builtin_amdgcn_logf(fmaxf(c.y, 1e-10f)) - minEv, r, rr), 0.0f), 1.0f);
  c.z = fminf(fmaxf(M::divs(__builtin_amdgcn_logf(fmaxf(c.z, 1e-10f)) - minEv, r, rr), 0.0f), 1.0f);
  c = make_float3(agx_contrast(c.x), agx_contrast(c.y), agx_contrast(c.z));
  const float l = lum(c.x, c.y, c.z);
  c = make_float3(powx(fmaxf(c.x, 0.0f), 1.22f), powx(fmaxf(c.y, 0.0f), 1.22f), powx(fmaxf(c.z, 0.0f), 1.22f));
  c = make_float3(l + (c.x - l) * 1.15f, l + (c.y - l) * 1.15f, l + (c.z - l) * 1.15f);
  c = mat_mul(outset, c);
  c = make_float3(powx(fmaxf(0.0f, c.x), 2.2f), powx(fmaxf(0.0f, c.y), 2.2f), powx(fmaxf(0.0f, c.z), 2.2f));
  c = mat_mul(toSrgb, c);
  return make_float3(fminf(fmaxf(c.x, 0.0f), 1.0f), fminf(fmaxf(c.y, 0.0f), 1.0f), fminf(fmaxf(c.z, 0.0f), 1.0f));
}
__device__ __forceinline__ float srgb(float c) {               // post.js:120
  return c <= 0.0031308f ? c * 12.92f : 1.055f * powx(c, (float)(1.0 / 2.4)) - 0.055f;
}
__device__ __forceinline__ uint32_t unorm8(float x) {
  return (uint32_t)rintf(fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f);
}

// FINAL (post.js:84-127): bloom mix, vignette, AgX, sRGB, divider lines -> the 8-bit target
__global__ __launch_bounds__(NT) void k_post_final(FinalArgs a) {
  const int2 p = pixel_of(a.W, a.H);
  if (!inside(p.x, p.y, a.W, a.H)) return;
  const float u = ((float)p.x + 0.5f) / (float)a.W, v = ((float)p.y + 0.5f) / (float)a.H;
  float3 c = sample(a.hdr, a.hw, a.hh, u, v);
  // the tent's taps at uv + (-t.x, 0), uv - t, ...: three positions per axis (u - t equals u + (-t))
  const Axis xm = axis(u - a.tx, a.bw), x0 = axis(u + 0.0f, a.bw), xp = axis(u + a.tx, a.bw);
  const Rows ym = rows(a.bloom, axis(v - a.ty, a.bh)), y0 = rows(a.bloom, axis(v + 0.0f, a.bh));
  const Rows yp = rows(a.bloom, axis(v + a.ty, a.bh));
  float3 b = mul3(bilerp(y0, x0), 4.0f);
  b = add3(b, mul3(add3(add3(add3(bilerp(y0, xm), bilerp(y0, xp)), bilerp(ym, x0)), bilerp(yp, x0)), 2.0f));
  b = add3(b, add3(add3(add3(bilerp(ym, xm), bilerp(yp, xp)), bilerp(ym, xp)), bilerp(yp, xm)));
  const float m = a.bloom_mix;
  c = make_float3(c.x * (1.0f - m) + (b.x / 16.0f) * m, c.y * (1.0f - m) + (b.y / 16.0f) * m,
                  c.z * (1.0f - m) + (b.z / 16.0f) * m);
  const float qx = (u - 0.5f) * 1.0f, qy = (v - 0.5f) * 0.75f;
  const float vg = 1.0f * (1.0f - a.vignette) + (1.0f - (qx * qx + qy * qy) * 1.6f) * a.vignette;
  c = mul3(c, vg);
  const float3 t = agx(c);
  float3 o = make_float3(srgb(t.x), srgb(t.y), srgb(t.z));
  if (a.line > 0.0f && fminf(fabsf(u - a.split), fabsf(u - a.split2)) * a.width < a.line) {
    const float k = 1.0f - 0.85f;
    o = make_float3(o.x * k + 1.0f * 0.85f, o.y * k + 1.0f * 0.85f, o.z * k + 1.0f * 0.85f);
  }
  const uint32_t r = unorm8(o.x), g = unorm8(o.y), bb = unorm8(o.z);
  a.target[(size_t)p.y * a.pitch + p.x] = (a.bgra ? (bb | g << 8 | r << 16) : (r | g << 8 | bb << 16)) | 0xff000000u;
}

dim3 grid_for(int w, int h) { return dim3((unsigned)(((w + TW - 1) / TW) * ((h + TH - 1) / TH))); }

}  // namespace

// ============================================================================ C ABI

struct post_ctx {
  int device = 0;
  post_options o{};
  int W = 0, H = 0;
  void *arena = nullptr;
  Tex tex[1 + POST_LEVELS] = {};     // the composited HDR image, then the levels 1..6
  int lw[1 + POST_LEVELS] = {}, lh[1 + POST_LEVELS] = {};
  char err[512] = {0};
};

static thread_local char g_err[512];

static int fail(post_t *p, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
static int fail(post_t *p, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(p ? p->err : g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}
static int hip_fail(post_t *p, hipError_t e, const char *what) {
  return fail(p, HUSH_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

extern "C" void post_default_options(post_options *o) {
  if (!o) return;
  o->bloom = 0.03f;
  o->vignette = 0.35f;
}

extern "C" int32_t post_create(int32_t device, const post_options *opts, post_t **out) {
  if (!out) return fail(nullptr, HUSH_ERR_ARGUMENT, "post_create: null argument");
  *out = nullptr;
  post_t *p = new (std::nothrow) post_ctx;
  if (!p) return fail(nullptr, HUSH_ERR_DEVICE, "post_create: out of host memory");
  p->device = device;
  post_default_options(&p->o);
  if (opts) p->o = *opts;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) { delete p; return fail(nullptr, HUSH_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e)); }
  *out = p;
  return HUSH_OK;
}

extern "C" void post_destroy(post_t *p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  if (p->arena) {
    (void)hipDeviceSynchronize();      // work still in flight on the surfaces
    (void)hipFree(p->arena);
  }
  delete p;
}

extern "C" int32_t post_resize(post_t *p, int32_t w, int32_t h) {
  if (!p) return fail(nullptr, HUSH_ERR_ARGUMENT, "post_resize: null context");
  if (w == p->W && h == p->H && p->arena) return HUSH_OK;
  if (w <= 0 || h <= 0 || w > HUSH_MAX_SIDE || h > HUSH_MAX_SIDE) return fail(p, HUSH_ERR_SIZE, "Post: bad size %dx%d", w, h);
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return hip_fail(p, e, "hipSetDevice");
  int lw[1 + POST_LEVELS], lh[1 + POST_LEVELS];
  size_t off[1 + POST_LEVELS], total = 0;
  for (int i = 0; i <= POST_LEVELS; i++) {
    lw[i] = w >> i > 0 ? w >> i : 1;
    lh[i] = h >> i > 0 ? h >> i : 1;
    off[i] = total;
    total += ((size_t)((lw[i] + 63) / 64 * 64) * lh[i] * 8 + 4095) / 4096 * 4096;
  }
  void *arena = nullptr;
  e = hipMalloc(&arena, total);
  if (e != hipSuccess) return hip_fail(p, e, "hipMalloc");
  if (p->arena) {
    (void)hipDeviceSynchronize();
    (void)hipFree(p->arena);
  }
  p->arena = arena;
  for (int i = 0; i <= POST_LEVELS; i++) {
    p->tex[i] = Tex{(uint2 *)((char *)arena + off[i]), (lw[i] + 63) / 64 * 64};
    p->lw[i] = lw[i];
    p->lh[i] = lh[i];
  }
  p->W = w;
  p->H = h;
  return HUSH_OK;
}

extern "C" const char *post_last_error(const post_t *p) { return p ? p->err : g_err; }

static Tex in_tex(const hush_surface &s) { return Tex{(uint2 *)s.ptr, (int)(s.pitch / 8)}; }

static int check_image(post_t *p, const hush_surface &s, const char *name, int w, int h) {
  if (!s.ptr) return fail(p, HUSH_ERR_INPUT, "Post.encode: missing image '%s'", name);
  if (s.format != HUSH_RGBA16F) return fail(p, HUSH_ERR_INPUT, "Post.encode: '%s' must be rgba16float", name);
  if (s.width != w || s.height != h)
    return fail(p, HUSH_ERR_INPUT, "Post.encode: '%s' is %dx%d, 'left' %dx%d (the images share one size)", name,
                s.width, s.height, w, h);
  if (s.pitch % 8 != 0 || s.pitch < (size_t)w * 8 || ((uintptr_t)s.ptr & 7) != 0 || !hush_pitch_ok(s.pitch / 8, h))
    return fail(p, HUSH_ERR_INPUT, "Post.encode: '%s' has pitch %zu bytes", name, s.pitch);
  return HUSH_OK;
}

extern "C" int32_t post_encode(post_t *p, const post_args *a, void *stream) {
  if (!p || !a) return fail(p, HUSH_ERR_ARGUMENT, "post_encode: null argument");
  const int w = a->left.width, h = a->left.height;
  if (!a->left.ptr) return fail(p, HUSH_ERR_INPUT, "Post.encode: missing image 'left'");
  if (w <= 0 || h <= 0 || w > HUSH_MAX_SIDE || h > HUSH_MAX_SIDE) return fail(p, HUSH_ERR_SIZE, "Post: bad size %dx%d", w, h);
  int r = check_image(p, a->left, "left", w, h);
  if (r == HUSH_OK) r = check_image(p, a->right, "right", w, h);
  const hush_surface &mid = a->mid.ptr ? a->mid : a->right;
  if (r == HUSH_OK) r = check_image(p, mid, "mid", w, h);
  if (r != HUSH_OK) return r;
  if (!a->target) return fail(p, HUSH_ERR_INPUT, "Post.encode: missing target");
  if (a->width <= 0 || a->height <= 0 || a->width > HUSH_MAX_SIDE || a->height > HUSH_MAX_SIDE)
    return fail(p, HUSH_ERR_SIZE, "Post.encode: bad target size %dx%d", a->width, a->height);
  if (a->target_pitch % 4 != 0 || a->target_pitch < (size_t)a->width * 4 || ((uintptr_t)a->target & 3) != 0 ||
      !hush_pitch_ok(a->target_pitch / 4, a->height))
    return fail(p, HUSH_ERR_INPUT, "Post.encode: target pitch %zu bytes; rows need 4-byte alignment and >= %d bytes",
                a->target_pitch, a->width * 4);
  if (a->format != POST_RGBA8_UNORM && a->format != POST_BGRA8_UNORM)
    return fail(p, HUSH_ERR_ARGUMENT, "Post.encode: unknown target format %d", a->format);
  r = post_resize(p, w, h);
  if (r != HUSH_OK) return r;
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return hip_fail(p, e, "hipSetDevice");
  hipStream_t s = (hipStream_t)stream;

  CompositeArgs ca;
  ca.left = in_tex(a->left); ca.right = in_tex(a->right); ca.mid = in_tex(mid); ca.out = p->tex[0];
  ca.W = w; ca.H = h; ca.same = a->same ? 1 : 0; ca.split = a->split; ca.split2 = a->split2; ca.exposure = a->exposure;
  hipLaunchKernelGGL(k_post_composite, grid_for(w, h), dim3(NT), 0, s, ca);
  for (int i = 1; i <= POST_LEVELS; i++) {      // levels 1..6 from the next larger one (post.js:185-188)
    DownArgs da;
    da.src = p->tex[i - 1]; da.out = p->tex[i];
    da.sw = p->lw[i - 1]; da.sh = p->lh[i - 1]; da.w = p->lw[i]; da.h = p->lh[i]; da.karis = i == 1;
    da.tx = (float)(1.0 / da.sw); da.ty = (float)(1.0 / da.sh);
    hipLaunchKernelGGL(k_post_down, grid_for(da.w, da.h), dim3(NT), 0, s, da);
  }
  for (int i = POST_LEVELS - 1; i >= 1; i--) {  // level i += up(level i + 1), 5 .. 1 (post.js:189-192)
    UpArgs ua;
    ua.src = p->tex[i + 1]; ua.dst = p->tex[i];
    ua.sw = p->lw[i + 1]; ua.sh = p->lh[i + 1]; ua.w = p->lw[i]; ua.h = p->lh[i];
    ua.tx = (float)(1.0 / ua.sw); ua.ty = (float)(1.0 / ua.sh);
    hipLaunchKernelGGL(k_post_up, grid_for(ua.w, ua.h), dim3(NT), 0, s, ua);
  }
  FinalArgs fa;
  fa.hdr = p->tex[0]; fa.bloom = p->tex[1];
  fa.hw = w; fa.hh = h; fa.bw = p->lw[1]; fa.bh = p->lh[1];
  fa.target = (uint32_t *)a->target; fa.pitch = (int)(a->target_pitch / 4);
  fa.W = a->width; fa.H = a->height; fa.bgra = a->format == POST_BGRA8_UNORM;
  fa.split = a->split; fa.split2 = a->split2; fa.bloom_mix = p->o.bloom; fa.vignette = p->o.vignette;
  fa.line = a->line; fa.width = (float)a->width;
  fa.tx = (float)(1.0 / fa.bw); fa.ty = (float)(1.0 / fa.bh);
  hipLaunchKernelGGL(k_post_final, grid_for(a->width, a->height), dim3(NT), 0, s, fa);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(p, e, "post_encode");
  return HUSH_OK;
}

extern "C" int32_t post_debug_surface(post_t *p, int32_t id, hush_surface *out) {
  if (!p || !out) return fail(p, HUSH_ERR_ARGUMENT, "post_debug_surface: null argument");
  if (id < 0 || id > POST_LEVELS) return fail(p, HUSH_ERR_ARGUMENT, "post_debug_surface: no surface %d", id);
  if (!p->arena) return fail(p, HUSH_ERR_ARGUMENT, "post_debug_surface: nothing encoded yet");
  hush_surface r{};
  r.ptr = p->tex[id].p; r.pitch = (size_t)p->tex[id].pitch * 8; r.width = p->lw[id]; r.height = p->lh[id];
  r.format = HUSH_RGBA16F;
  *out = r;
  return HUSH_OK;
}
