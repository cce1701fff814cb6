// GPU render of the seeded procedural scenes (include/hush_scene.h): one workgroup per 16x16 tile of the
// crop window, one thread per pixel.  Every workgroup walks its tile's primitive list (uniform across the
// workgroup, so primitive data comes through scalar loads), intersects in FP64 — whose vector rate on
// CDNA4 equals non-packed FP32 — and writes the five binary16 G-buffer surfaces and the clean target.
//
// The arithmetic is the CPU renderer's (hush_amd/csrc/scene.c: hs_render) operation for operation with
// FMA contraction off, so every value agrees bit for bit except where the noise's log / cos / exp (the
// device library vs the C library, both within an ulp in FP64) lands a binary16 rounding differently.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "hush_device.h"
#include "hush_scene.h"

using namespace hush;

namespace {

constexpr int TS = 16, NPD = 20;
constexpr double PI_D = 3.14159265358979323846;

struct D3 { double x, y, z; };
__device__ __forceinline__ D3 v3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 add(D3 a, D3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 mul(D3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 norm(D3 a) {
  const double l = sqrt(dot(a, a));
  return l > 0 ? mul(a, 1.0 / l) : a;
}

struct Cam { D3 pos, fwd, right, up; double tx, ty; };
__device__ __forceinline__ Cam cam_of(const double *c) {
  return Cam{v3(c[0], c[1], c[2]), v3(c[3], c[4], c[5]), v3(c[6], c[7], c[8]), v3(c[9], c[10], c[11]), c[12], c[13]};
}

// counter-based random numbers (scene.c: mix64, u01, hrand)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double hrand(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  return (double)(mix64(seed ^ mix64(a * 0x100000001b3ull ^ mix64(b ^ mix64(c)))) >> 11) * (1.0 / 9007199254740992.0);
}

// projectPix (pathtracer.wgsl.js:505-509): w = 1 point or w = 0 direction
__device__ __forceinline__ bool project(const Cam &c, D3 p, bool point, int W, int H, double &px, double &py) {
  const D3 rel = point ? sub(p, c.pos) : p;
  const double z = dot(rel, c.fwd);
  if (z <= 1e-9) { px = py = 0; return false; }
  const double x = dot(rel, c.right) / (z * c.tx), y = dot(rel, c.up) / (z * c.ty);
  px = (x * 0.5 + 0.5) * W;
  py = (0.5 - y * 0.5) * H;
  return true;
}

__device__ __forceinline__ D3 sky(D3 d) {
  const double t = sqrt(d.y > 0 ? d.y : 0);
  return v3(1.1 * (1 - t) + 0.35 * t, 1.05 * (1 - t) + 0.55 * t, 0.95 * (1 - t) + 1.0 * t);
}
__device__ __forceinline__ D3 irradiance(D3 n) {
  const D3 L = v3(0.4523, 0.8040, 0.3854);
  double ndl = dot(n, L);
  if (ndl < 0) ndl = 0;
  const double hemi = 0.5 + 0.5 * n.y;
  return v3(0.22 * hemi + 2.9 * ndl, 0.27 * hemi + 2.7 * ndl, 0.36 * hemi + 2.4 * ndl);
}
__device__ __forceinline__ bool ray_sphere(D3 o, D3 d, D3 c, double r, double &t) {
  const D3 oc = sub(o, c);
  const double b = dot(oc, d), cc = dot(oc, oc) - r * r, disc = b * b - cc;
  if (disc < 0) return false;
  const double sq = sqrt(disc), t0 = -b - sq;
  if (t0 > 1e-4) { t = t0; return true; }
  return false;
}
__device__ __forceinline__ bool ray_box(D3 o, D3 d, D3 c, D3 h, double &t, D3 &n) {
  double tmin = -1e30, tmax = 1e30, sg = 1;
  int axis = 0;
  const double oo[3] = {o.x - c.x, o.y - c.y, o.z - c.z}, dd[3] = {d.x, d.y, d.z}, hh[3] = {h.x, h.y, h.z};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (fabs(dd[a]) < 1e-12) {
      if (oo[a] < -hh[a] || oo[a] > hh[a]) return false;
      continue;
    }
    const double inv = 1.0 / dd[a];
    double t1 = (-hh[a] - oo[a]) * inv, t2 = (hh[a] - oo[a]) * inv, s = -1;
    if (t1 > t2) { const double tt = t1; t1 = t2; t2 = tt; s = 1; }
    if (t1 > tmin) { tmin = t1; axis = a; sg = s; }
    if (t2 < tmax) tmax = t2;
  }
  if (tmin > tmax || tmin <= 1e-4) return false;
  t = tmin;
  n = v3(axis == 0 ? sg : 0, axis == 1 ? sg : 0, axis == 2 ? sg : 0);
  return true;
}

// (float) then binary16, round to nearest even (scene.c: put / f2h)
__device__ __forceinline__ uint2 pack_rne(double a, double b, double c, double d) {
  const _Float16 h[4] = {(_Float16)(float)a, (_Float16)(float)b, (_Float16)(float)c, (_Float16)(float)d};
  uint2 r;
  __builtin_memcpy(&r, h, 8);
  return r;
}

struct RenderArgs {
  const double *cams, *prims;
  const int *off, *lst;
  uint64_t seed;
  int frame, W, H, x0, y0, w, h, tw;
  Tex rad, nd, mv, mat, emi;
  float *clean;
  size_t clean_pitch;
};

__global__ __launch_bounds__(TS * TS) void k_scene(RenderArgs a) {
  const int tile = blockIdx.x, tx = tile % a.tw, ty = tile / a.tw;
  const int xx = tx * TS + (int)(threadIdx.x % TS), yy = ty * TS + (int)(threadIdx.x / TS);
  if (xx >= a.w || yy >= a.h) return;
  const int X = a.x0 + xx, Y = a.y0 + yy;
  const Cam cam = cam_of(a.cams), pcam = cam_of(a.cams + 14);
  const double ndx = ((X + 0.5) / a.W) * 2 - 1, ndy = ((Y + 0.5) / a.H) * 2 - 1;
  const D3 d = norm(add(cam.fwd, sub(mul(cam.right, ndx * cam.tx), mul(cam.up, ndy * cam.ty))));
  double bt = 1e30;
  int bk = -2;
  D3 bn = v3(0, 0, 0);
  if (d.y < -1e-9) {
    const double t = -cam.pos.y / d.y;
    if (t > 1e-4 && t * dot(d, cam.fwd) < 300.0) { bt = t; bk = -1; bn = v3(0, 1, 0); }
  }
  for (int li = a.off[tile]; li < a.off[tile + 1]; li++) {
    const int k = a.lst[li];
    const double *q = a.prims + NPD * k;
    const D3 pos = v3(q[2], q[3], q[4]);
    double tt;
    D3 nn;
    if (q[0] == 0.0) {
      if (ray_sphere(cam.pos, d, pos, q[8], tt) && tt < bt) { bt = tt; bk = k; bn = norm(sub(add(cam.pos, mul(d, tt)), pos)); }
    } else if (ray_box(cam.pos, d, pos, v3(q[8], q[9], q[10]), tt, nn) && tt < bt) {
      bt = tt; bk = k; bn = nn;
    }
  }
  double rr, rg, rb, ar, ag, ab, ro, er = 0, eg = 0, eb = 0, id, mvx, mvy, pdepth, depth;
  double c0 = 0, c1 = 0, c2 = 0;
  D3 nrm;
  if (bk == -2) {                                   // sky
    const D3 e = sky(d);
    er = e.x; eg = e.y; eb = e.z;
    rr = rg = rb = 0;
    ar = ag = ab = 1; ro = 1;
    nrm = mul(d, -1);
    depth = -1; pdepth = -1; id = 0;
    double cx, cy, px, py;
    if (project(cam, d, false, a.W, a.H, cx, cy) && project(pcam, d, false, a.W, a.H, px, py)) { mvx = px - cx; mvy = py - cy; }
    else { mvx = mvy = 0; }
    c0 = er; c1 = eg; c2 = eb;
  } else {
    const D3 p = add(cam.pos, mul(d, bt));
    D3 prevP = p;
    if (bk >= 0) {
      const double *q = a.prims + NPD * bk;
      ar = q[11]; ag = q[12]; ab = q[13]; ro = q[14]; id = q[1];
      er = q[15]; eg = q[16]; eb = q[17];
      prevP = sub(p, v3(q[5], q[6], q[7]));
    } else {                                        // checkered ground
      const int ch = ((int)floor(p.x) + (int)floor(p.z)) & 1;
      ar = ch ? 0.72 : 0.32; ag = ch ? 0.70 : 0.34; ab = ch ? 0.64 : 0.38; ro = 0.8; id = 1;
    }
    nrm = bn;
    depth = dot(sub(p, cam.pos), cam.fwd);
    pdepth = dot(sub(prevP, pcam.pos), pcam.fwd);
    double cx, cy, px, py;
    if (project(cam, p, true, a.W, a.H, cx, cy) && project(pcam, prevP, true, a.W, a.H, px, py)) { mvx = px - cx; mvy = py - cy; }
    else { mvx = mvy = 0; }
    const bool emitter = er > 0 || eg > 0 || eb > 0;
    const D3 E = irradiance(nrm);
    if (emitter) {
      rr = rg = rb = 0;
    } else {
      const uint64_t pix = (uint64_t)Y * 65536u + X;
      const double z = sqrt(-2.0 * log(1.0 - hrand(a.seed, 101, a.frame, pix))) * cos(2 * PI_D * hrand(a.seed, 102, a.frame, pix));
      double nz = exp(z);
      if (hrand(a.seed, 103, a.frame, pix) < 0.02) nz *= 50.0;
      rr = ar * E.x * nz; rg = ag * E.y * nz; rb = ab * E.z * nz;
    }
    const double noise_mean = 1.6487212707001282 * (1.0 + 0.02 * 49.0);     // exp(0.5) E[firefly factor]
    c0 = emitter ? er : ar * E.x * noise_mean + er;
    c1 = emitter ? eg : ag * E.y * noise_mean + eg;
    c2 = emitter ? eb : ab * E.z * noise_mean + eb;
  }
  a.rad.p[(size_t)yy * a.rad.pitch + xx] = pack_rne(rr, rg, rb, 1);
  a.nd.p[(size_t)yy * a.nd.pitch + xx] = pack_rne(nrm.x, nrm.y, nrm.z, depth);
  a.mv.p[(size_t)yy * a.mv.pitch + xx] = pack_rne(mvx, mvy, id, pdepth);
  a.mat.p[(size_t)yy * a.mat.pitch + xx] = pack_rne(ar, ag, ab, ro);
  a.emi.p[(size_t)yy * a.emi.pitch + xx] = pack_rne(er, eg, eb, 0);
  if (a.clean) {
    float *c = (float *)((char *)a.clean + (size_t)yy * a.clean_pitch) + 4 * xx;
    *(float4 *)c = make_float4((float)c0, (float)c1, (float)c2, 0.0f);
  }
}

thread_local char g_err[512];

}  // namespace

struct hush_scene_ctx {
  int device = 0;
  void *buf = nullptr;
  size_t cap = 0;
  char err[512] = {0};
};

static int sfail(hush_scene_t *g, int code, const char *msg, hipError_t e = hipSuccess) {
  if (e != hipSuccess) std::snprintf(g ? g->err : g_err, 512, "%s: %s", msg, hipGetErrorString(e));
  else std::snprintf(g ? g->err : g_err, 512, "%s", msg);
  return code;
}

extern "C" const char *hush_scene_last_error(const hush_scene_t *g) { return g ? g->err : g_err; }

extern "C" int32_t hush_scene_create(int32_t device, hush_scene_t **out) {
  if (!out) return sfail(nullptr, HUSH_ERR_ARGUMENT, "hush_scene_create: null argument");
  *out = nullptr;
  hush_scene_t *g = new (std::nothrow) hush_scene_ctx;
  if (!g) return sfail(nullptr, HUSH_ERR_DEVICE, "hush_scene_create: out of host memory");
  g->device = device;
  *out = g;
  return HUSH_OK;
}

extern "C" void hush_scene_destroy(hush_scene_t *g) {
  if (!g) return;
  if (g->buf) {
    (void)hipSetDevice(g->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(g->buf);
  }
  delete g;
}

static bool surf_ok(const hush_surface &s, int w, int h) {
  return s.ptr && s.width == w && s.height == h && s.format == HUSH_RGBA16F && s.pitch % 8 == 0 && s.pitch >= (size_t)w * 8 &&
         ((uintptr_t)s.ptr & 7) == 0;
}

extern "C" int32_t hush_scene_render(hush_scene_t *g, const hush_scene_frame *fr, const hush_inputs *out, float *clean,
                                     size_t clean_pitch, void *stream) {
  if (!g || !fr || !out || !fr->cams || !fr->prims || !fr->off || (fr->n_list > 0 && !fr->lst))
    return sfail(g, HUSH_ERR_ARGUMENT, "hush_scene_render: null argument");
  if (fr->w <= 0 || fr->h <= 0 || fr->x0 < 0 || fr->y0 < 0 || fr->x0 + fr->w > fr->width || fr->y0 + fr->h > fr->height ||
      fr->n_prims < 0 || fr->n_list < 0)
    return sfail(g, HUSH_ERR_SIZE, "hush_scene_render: crop window outside the image");
  const hush_surface *s[5] = {&out->radiance, &out->normal_depth, &out->motion, &out->material, &out->emission};
  for (auto *p : s)
    if (!surf_ok(*p, fr->w, fr->h)) return sfail(g, HUSH_ERR_INPUT, "hush_scene_render: outputs must be w x h rgba16float surfaces");
  if (clean && (clean_pitch < (size_t)fr->w * 16 || clean_pitch % 16 != 0 || ((uintptr_t)clean & 15) != 0))
    return sfail(g, HUSH_ERR_INPUT, "hush_scene_render: clean needs 16-byte aligned rows of >= w * 16 bytes");
  const int tw = (fr->w + TS - 1) / TS, th = (fr->h + TS - 1) / TS, nt = tw * th;
  if (fr->off[nt] != fr->n_list || fr->off[0] != 0) return sfail(g, HUSH_ERR_ARGUMENT, "hush_scene_render: tile offsets do not match the list");
  for (int i = 0; i < fr->n_list; i++)
    if (fr->lst[i] < 0 || fr->lst[i] >= fr->n_prims) return sfail(g, HUSH_ERR_ARGUMENT, "hush_scene_render: primitive index out of range");
  for (int t = 0; t < nt; t++)
    if (fr->off[t + 1] < fr->off[t]) return sfail(g, HUSH_ERR_ARGUMENT, "hush_scene_render: tile offsets must not decrease");
  // stage the set-up: cams | prims | off | lst
  const size_t b_cam = 28 * sizeof(double), b_prim = (size_t)fr->n_prims * NPD * sizeof(double);
  const size_t b_off = ((size_t)nt + 1) * sizeof(int), b_lst = (size_t)fr->n_list * sizeof(int);
  const size_t o_prim = 256, o_off = o_prim + (b_prim + 255) / 256 * 256, o_lst = o_off + (b_off + 255) / 256 * 256;
  const size_t need = o_lst + b_lst + 256;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipSetDevice(g->device);
  if (e == hipSuccess && need > g->cap) {
    if (g->buf) { (void)hipStreamSynchronize(st); (void)hipFree(g->buf); g->buf = nullptr; g->cap = 0; }
    e = hipMalloc(&g->buf, need * 2);
    if (e == hipSuccess) g->cap = need * 2;
  }
  char *base = (char *)g->buf;
  if (e == hipSuccess) e = hipMemcpyAsync(base, fr->cams, b_cam, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && b_prim) e = hipMemcpyAsync(base + o_prim, fr->prims, b_prim, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(base + o_off, fr->off, b_off, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && b_lst) e = hipMemcpyAsync(base + o_lst, fr->lst, b_lst, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return sfail(g, HUSH_ERR_DEVICE, "hush_scene_render: staging", e);
  RenderArgs a;
  a.cams = (const double *)base; a.prims = (const double *)(base + o_prim);
  a.off = (const int *)(base + o_off); a.lst = (const int *)(base + o_lst);
  a.seed = fr->seed; a.frame = fr->frame; a.W = fr->width; a.H = fr->height;
  a.x0 = fr->x0; a.y0 = fr->y0; a.w = fr->w; a.h = fr->h; a.tw = tw;
  Tex *t[5] = {&a.rad, &a.nd, &a.mv, &a.mat, &a.emi};
  for (int k = 0; k < 5; k++) *t[k] = Tex{(uint2 *)s[k]->ptr, (int)(s[k]->pitch / 8)};
  a.clean = clean; a.clean_pitch = clean_pitch;
  hipLaunchKernelGGL(k_scene, dim3(nt), dim3(TS * TS), 0, st, a);
  e = hipGetLastError();
  if (e != hipSuccess) return sfail(g, HUSH_ERR_DEVICE, "hush_scene_render: launch", e);
  return HUSH_OK;
}
