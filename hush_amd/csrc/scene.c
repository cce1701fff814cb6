/*
 * Seeded procedural G-buffer generator: the synthetic input of the denoiser's workloads.
 *
 * It writes the five rgba16float inputs of the denoiser (layout of src/index.d.ts:2-18) following the
 * path tracer's G-buffer conventions (src/pt/pathtracer.wgsl.js:505-577, 657-661):
 *   radiance     1 spp radiance without the primary hit's emission (0 on emitters, which end the path)
 *   normalDepth  world normal, linear view depth dot(p - camPos, camFwd); sky: normal = -d, depth = -1
 *   motion       xy = projectPix(prevVP, prevP) - projectPix(VP, p) in pixels (y down), z = object id,
 *                w = linear depth of prevP in the previous frame's camera (-1 on sky)
 *   material     albedo (1 on emitters and sky), roughness (1 on sky)
 *   emission     primary emission; sky radiance on sky pixels
 * Shading is analytic Lambert (sun + hemispherical ambient, no shadows).  The 1 spp noise is
 * clean * lognormal(0, 1) per pixel and frame, with 2% firefly pixels scaled by 50.  Every random number
 * comes from a counter hash of (seed, frame, pixel), so any crop or thread split gives identical bytes.
 *
 * Scene kinds:
 * The ground plane ends at view depth 300 (sky beyond), which keeps every depth finite in binary16.
 *   HS_SPHERES  16 spheres on a checkered ground plane, camera orbiting 1 degree per frame.
 *   HS_BOXES    200 boxes, 40 of them moving independently, ~5% emissive, camera panning 3 px per frame
 *               (yaw), horizon near the top tenth of the frame (~10% sky).
 *   HS_TRAIN    300 spheres and boxes, 30% moving, new random camera every 48 frames (a camera cut).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

enum { HS_SPHERES = 0, HS_BOXES = 1, HS_TRAIN = 2 };
enum { PRIM_SPHERE = 0, PRIM_BOX = 1 };

typedef struct { double x, y, z; } d3;
static d3 v(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static d3 add(d3 a, d3 b) { return v(a.x + b.x, a.y + b.y, a.z + b.z); }
static d3 sub(d3 a, d3 b) { return v(a.x - b.x, a.y - b.y, a.z - b.z); }
static d3 mul(d3 a, double s) { return v(a.x * s, a.y * s, a.z * s); }
static double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 cross(d3 a, d3 b) { return v(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static d3 norm(d3 a) { double l = sqrt(dot(a, a)); return l > 0 ? mul(a, 1.0 / l) : a; }

typedef struct {
  int type, id, moving;
  d3 c, h;                 /* centre at frame 0; half extents (sphere: radius in h.x) */
  d3 dir;                  /* oscillation direction */
  double amp, omega, phase;
  d3 albedo, emit;
  double rough;
} prim;

typedef struct { d3 pos, fwd, right, up; double tx, ty; } camera;

typedef struct hs_scene {
  int kind, W, H;
  uint64_t seed;
  int n;
  prim *p;
} hs_scene;

/* ---------------------------------------------------------------- counter-based random numbers */

static uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }
static double hrand(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  return u01(mix64(seed ^ mix64(a * 0x100000001b3ull ^ mix64(b ^ mix64(c)))));
}

/* ---------------------------------------------------------------- binary16 (round to nearest even) */

static uint16_t f2h(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  uint32_t s = (u >> 16) & 0x8000u, a = u & 0x7fffffffu;
  if (a >= 0x7f800000u) return (uint16_t)(s | (a > 0x7f800000u ? 0x7e00u : 0x7c00u));
  if (a >= 0x477ff000u) return (uint16_t)(s | 0x7c00u);               /* rounds to inf */
  if (a >= 0x38800000u) {                                              /* normal */
    uint32_t r = a - 0x38000000u;
    uint32_t lsb = (r >> 13) & 1u;
    r += 0x0fffu + lsb;
    return (uint16_t)(s | (r >> 13));
  }
  if (a < 0x33000000u) return (uint16_t)s;                             /* < 2^-25 -> 0 */
  uint32_t e = a >> 23, m = (a & 0x7fffffu) | 0x800000u;
  uint32_t sh = 126u - e;                                              /* 14 .. 24 */
  uint32_t q = m >> sh, rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
  if (rem > half || (rem == half && (q & 1u))) q++;
  return (uint16_t)(s | q);
}

/* ---------------------------------------------------------------- scene set-up */

static void place(hs_scene *s, prim *q, int idx, double xmin, double xmax, double zmin, double zmax,
                  double smin, double smax, int type) {
  uint64_t sd = s->seed;
  q->type = type;
  q->id = 2 + idx;
  double r0 = hrand(sd, 11, idx, 0), r1 = hrand(sd, 11, idx, 1), r2 = hrand(sd, 11, idx, 2);
  double r3 = hrand(sd, 11, idx, 3), r4 = hrand(sd, 11, idx, 4);
  if (type == PRIM_SPHERE) {
    double r = smin + (smax - smin) * r2;
    q->h = v(r, r, r);
    q->c = v(xmin + (xmax - xmin) * r0, r, zmin + (zmax - zmin) * r1);
  } else {
    d3 h = v(0.5 * (smin + (smax - smin) * r2), 0.5 * (smin + (smax - smin) * r3), 0.5 * (smin + (smax - smin) * r4));
    q->h = h;
    q->c = v(xmin + (xmax - xmin) * r0, h.y, zmin + (zmax - zmin) * r1);
  }
  q->albedo = v(0.1 + 0.8 * hrand(sd, 12, idx, 0), 0.1 + 0.8 * hrand(sd, 12, idx, 1), 0.1 + 0.8 * hrand(sd, 12, idx, 2));
  q->rough = hrand(sd, 12, idx, 3);
  q->emit = v(0, 0, 0);
  q->moving = 0;
  double a = 2 * M_PI * hrand(sd, 13, idx, 0);
  q->dir = v(cos(a), 0, sin(a));
  q->amp = 0.5 + 2.0 * hrand(sd, 13, idx, 1);
  q->omega = 0.03 + 0.09 * hrand(sd, 13, idx, 2);
  q->phase = 2 * M_PI * hrand(sd, 13, idx, 3);
}

EXPORT hs_scene *hs_create(int kind, int width, int height, unsigned long long seed) {
  if (width <= 0 || height <= 0 || kind < 0 || kind > 2) return NULL;
  hs_scene *s = (hs_scene *)calloc(1, sizeof(hs_scene));
  s->kind = kind; s->W = width; s->H = height; s->seed = mix64(seed + 0x51ed27ull);
  if (kind == HS_SPHERES) {
    s->n = 16;
    s->p = (prim *)calloc(s->n, sizeof(prim));
    for (int i = 0; i < s->n; i++) place(s, &s->p[i], i, -6, 6, -6, 6, 0.35, 1.1, PRIM_SPHERE);
  } else if (kind == HS_BOXES) {
    s->n = 200;
    s->p = (prim *)calloc(s->n, sizeof(prim));
    for (int i = 0; i < s->n; i++) {
      place(s, &s->p[i], i, -30, 30, 2, 55, 0.4, 2.2, PRIM_BOX);
      if (i < 40) s->p[i].moving = 1;
      if (hrand(s->seed, 14, i, 0) < 0.05) {
        double e = 1 + 19 * hrand(s->seed, 14, i, 1);
        s->p[i].emit = v(e, e * (0.7 + 0.3 * hrand(s->seed, 14, i, 2)), e * (0.5 + 0.5 * hrand(s->seed, 14, i, 3)));
        s->p[i].albedo = v(1, 1, 1);
        s->p[i].h = mul(s->p[i].h, 0.5);
        s->p[i].c.y = s->p[i].h.y + 0.2;
      }
    }
  } else {
    s->n = 300;
    s->p = (prim *)calloc(s->n, sizeof(prim));
    for (int i = 0; i < s->n; i++) {
      int type = hrand(s->seed, 15, i, 0) < 0.5 ? PRIM_SPHERE : PRIM_BOX;
      place(s, &s->p[i], i, -20, 20, -20, 20, 0.3, 2.0, type);
      s->p[i].moving = hrand(s->seed, 15, i, 1) < 0.3;
      if (hrand(s->seed, 14, i, 0) < 0.04) {
        double e = 1 + 19 * hrand(s->seed, 14, i, 1);
        s->p[i].emit = v(e, e * 0.85, e * 0.7);
        s->p[i].albedo = v(1, 1, 1);
      }
    }
  }
  return s;
}

EXPORT void hs_destroy(hs_scene *s) {
  if (!s) return;
  free(s->p);
  free(s);
}

EXPORT int hs_num_prims(const hs_scene *s) { return s->n; }
/* the per-scene seed every random number of the scene hashes (what a renderer elsewhere needs) */
EXPORT unsigned long long hs_seed(const hs_scene *s) { return s->seed; }
EXPORT int hs_cut_interval(const hs_scene *s) { return s->kind == HS_TRAIN ? 48 : 0; }
/* 1 when `frame` starts a new camera segment (the caller resets the denoiser's history) */
EXPORT int hs_is_cut(const hs_scene *s, int frame) { return s->kind == HS_TRAIN && frame % 48 == 0; }

static d3 prim_pos(const prim *q, double f) {
  if (!q->moving) return q->c;
  return add(q->c, mul(q->dir, q->amp * sin(q->omega * f + q->phase)));
}

static camera make_camera(const hs_scene *s, double f) {
  camera c;
  double fovy = 50.0 * M_PI / 180.0;
  c.ty = tan(0.5 * fovy);
  c.tx = c.ty * (double)s->W / (double)s->H;
  d3 target;
  if (s->kind == HS_SPHERES) {
    double a = (20.0 + f) * M_PI / 180.0;
    c.pos = v(11.0 * sin(a), 4.5, -11.0 * cos(a));
    target = v(0, 0.8, 0);
    c.fwd = norm(sub(target, c.pos));
  } else if (s->kind == HS_BOXES) {
    /* yaw pan of 3 px per frame at the image centre; pitch puts the horizon 10% below the top edge */
    double focal = 0.5 * s->H / c.ty;
    double yaw = -0.25 + f * atan(3.0 / focal);
    double pitch = atan(0.8 * c.ty);
    c.pos = v(0, 3.0, -6.0);
    c.fwd = v(sin(yaw) * cos(pitch), -sin(pitch), cos(yaw) * cos(pitch));
  } else {
    int seg = (int)floor(f / 48.0);
    double t = f - seg * 48.0;
    double a0 = 2 * M_PI * hrand(s->seed, 21, seg, 0);
    double rad = 14 + 10 * hrand(s->seed, 21, seg, 1);
    double hgt = 2 + 6 * hrand(s->seed, 21, seg, 2);
    double spd = (hrand(s->seed, 21, seg, 3) - 0.5) * 0.04;
    double a = a0 + spd * t;
    c.pos = v(rad * sin(a), hgt, -rad * cos(a));
    target = v(6 * (hrand(s->seed, 21, seg, 4) - 0.5), 0.5, 6 * (hrand(s->seed, 21, seg, 5) - 0.5));
    c.fwd = norm(sub(target, c.pos));
  }
  c.right = norm(cross(v(0, 1, 0), c.fwd));
  c.up = cross(c.fwd, c.right);
  return c;
}

/* projectPix (pathtracer.wgsl.js:505-509): w = 1 point or w = 0 direction */
static int project(const hs_scene *s, const camera *c, d3 p, int point, double *px, double *py) {
  d3 rel = point ? sub(p, c->pos) : p;
  double z = dot(rel, c->fwd);
  if (z <= 1e-9) { *px = *py = 0; return 0; }
  double x = dot(rel, c->right) / (z * c->tx), y = dot(rel, c->up) / (z * c->ty);
  *px = (x * 0.5 + 0.5) * s->W;
  *py = (0.5 - y * 0.5) * s->H;
  return 1;
}

static d3 sky(d3 d) {
  double t = sqrt(d.y > 0 ? d.y : 0);
  return v(1.1 * (1 - t) + 0.35 * t, 1.05 * (1 - t) + 0.55 * t, 0.95 * (1 - t) + 1.0 * t);
}

static d3 irradiance(d3 n) {
  const d3 L = {0.4523, 0.8040, 0.3854};
  double ndl = dot(n, L);
  if (ndl < 0) ndl = 0;
  double hemi = 0.5 + 0.5 * n.y;
  return v(0.22 * hemi + 2.9 * ndl, 0.27 * hemi + 2.7 * ndl, 0.36 * hemi + 2.4 * ndl);
}

typedef struct { double t; int k; d3 n; } hit;

static int ray_sphere(d3 o, d3 d, d3 c, double r, double *t) {
  d3 oc = sub(o, c);
  double b = dot(oc, d), cc = dot(oc, oc) - r * r, disc = b * b - cc;
  if (disc < 0) return 0;
  double sq = sqrt(disc), t0 = -b - sq;
  if (t0 > 1e-4) { *t = t0; return 1; }
  return 0;
}

static int ray_box(d3 o, d3 d, d3 c, d3 h, double *t, d3 *n) {
  double tmin = -1e30, tmax = 1e30;
  int axis = 0;
  double sgn = 1;
  double oo[3] = {o.x - c.x, o.y - c.y, o.z - c.z}, dd[3] = {d.x, d.y, d.z}, hh[3] = {h.x, h.y, h.z};
  for (int a = 0; a < 3; a++) {
    if (fabs(dd[a]) < 1e-12) {
      if (oo[a] < -hh[a] || oo[a] > hh[a]) return 0;
      continue;
    }
    double inv = 1.0 / dd[a];
    double t1 = (-hh[a] - oo[a]) * inv, t2 = (hh[a] - oo[a]) * inv;
    double s = -1;
    if (t1 > t2) { double tt = t1; t1 = t2; t2 = tt; s = 1; }
    if (t1 > tmin) { tmin = t1; axis = a; sgn = s; }
    if (t2 < tmax) tmax = t2;
  }
  if (tmin > tmax || tmin <= 1e-4) return 0;
  *t = tmin;
  double nn[3] = {0, 0, 0};
  nn[axis] = sgn;
  *n = v(nn[0], nn[1], nn[2]);
  return 1;
}

#define TILE 16

static void put(uint16_t *dst, size_t i, double a, double b, double c, double d) {
  dst[i * 4 + 0] = f2h((float)a); dst[i * 4 + 1] = f2h((float)b);
  dst[i * 4 + 2] = f2h((float)c); dst[i * 4 + 3] = f2h((float)d);
}

/*
 * Renders frame `frame` into the crop window [x0, x0+w) x [y0, y0+h) of the scene's full W x H image.
 * Outputs are w*h*4 binary16 arrays (row-major, tight pitch); `clean` (optional) receives w*h*4 floats:
 * the expected radiance albedo * irradiance * E[noise] plus emission (the noise-free target).
 */
/* per-frame set-up shared by the CPU renderer and the GPU one (hush_amd/csrc/scene_gpu.hip): cameras,
   primitive positions and motions, and per-16x16-tile primitive lists of the crop window (conservative
   projected bounds) */
typedef struct {
  camera cam, pcam;
  d3 *pos, *dpos;
  int tw, th;
  int *off, *lst;
} frame_setup;

static void setup_free(frame_setup *fs) {
  free(fs->pos); free(fs->dpos); free(fs->off); free(fs->lst);
}

static void setup_frame(const hs_scene *s, int frame, int x0, int y0, int w, int h, frame_setup *fs) {
  const double f = frame;
  fs->cam = make_camera(s, f);
  fs->pcam = make_camera(s, f - 1);
  const camera cam = fs->cam;
  const int n = s->n;
  d3 *pos = (d3 *)malloc(sizeof(d3) * (n > 0 ? n : 1));
  d3 *dpos = (d3 *)malloc(sizeof(d3) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) {
    pos[i] = prim_pos(&s->p[i], f);
    dpos[i] = sub(pos[i], prim_pos(&s->p[i], f - 1));
  }
  int tw = (w + TILE - 1) / TILE, th = (h + TILE - 1) / TILE;
  int *cnt = (int *)calloc((size_t)tw * th + 1, sizeof(int));
  int *bx = (int *)malloc(sizeof(int) * 4 * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) {
    d3 hh = s->p[i].h;
    double mnx = 1e30, mny = 1e30, mxx = -1e30, mxy = -1e30;
    int all = 1;
    for (int k = 0; k < 8; k++) {
      d3 q = add(pos[i], v(k & 1 ? hh.x : -hh.x, k & 2 ? hh.y : -hh.y, k & 4 ? hh.z : -hh.z));
      double px, py;
      if (dot(sub(q, cam.pos), cam.fwd) < 0.05 || !project(s, &cam, q, 1, &px, &py)) { all = 0; break; }
      if (px < mnx) mnx = px;
      if (py < mny) mny = py;
      if (px > mxx) mxx = px;
      if (py > mxy) mxy = py;
    }
    int a0, a1, b0, b1;
    if (!all) { a0 = 0; a1 = tw - 1; b0 = 0; b1 = th - 1; }
    else {
      a0 = (int)floor((mnx - 1 - x0) / TILE); a1 = (int)floor((mxx + 1 - x0) / TILE);
      b0 = (int)floor((mny - 1 - y0) / TILE); b1 = (int)floor((mxy + 1 - y0) / TILE);
      if (a0 < 0) a0 = 0;
      if (b0 < 0) b0 = 0;
      if (a1 > tw - 1) a1 = tw - 1;
      if (b1 > th - 1) b1 = th - 1;
    }
    bx[i * 4 + 0] = a0; bx[i * 4 + 1] = a1; bx[i * 4 + 2] = b0; bx[i * 4 + 3] = b1;
    for (int ty = b0; ty <= b1; ty++)
      for (int tx = a0; tx <= a1; tx++) cnt[ty * tw + tx]++;
  }
  int *off = (int *)malloc(sizeof(int) * ((size_t)tw * th + 1));
  off[0] = 0;
  for (int t = 0; t < tw * th; t++) off[t + 1] = off[t] + cnt[t];
  int *lst = (int *)malloc(sizeof(int) * (off[tw * th] > 0 ? off[tw * th] : 1));
  memset(cnt, 0, sizeof(int) * (size_t)tw * th);
  for (int i = 0; i < n; i++)
    for (int ty = bx[i * 4 + 2]; ty <= bx[i * 4 + 3]; ty++)
      for (int tx = bx[i * 4 + 0]; tx <= bx[i * 4 + 1]; tx++) {
        int t = ty * tw + tx;
        lst[off[t] + cnt[t]++] = i;
      }
  free(cnt); free(bx);
  fs->pos = pos; fs->dpos = dpos; fs->tw = tw; fs->th = th; fs->off = off; fs->lst = lst;
}

/* The set-up of frame `frame` for a renderer elsewhere (the GPU one).  Layouts, all doubles unless noted:
     cams   2 x 14: current then previous camera, each pos.xyz fwd.xyz right.xyz up.xyz tx ty
     prims  n x HS_PRIM_DOUBLES: type, id, pos.xyz, dpos.xyz, h.xyz, albedo.xyz, roughness, emit.xyz
     off    (ntiles + 1) ints, lst off[ntiles] ints (ntiles = ceil(w/16) * ceil(h/16), row-major tiles)
   Returns the list length; when it exceeds lst_cap nothing is written to lst (call again with room). */
#define HS_PRIM_DOUBLES 20
EXPORT int hs_frame_setup(const hs_scene *s, int frame, int x0, int y0, int w, int h, double *cams, double *prims,
                          int *off, int *lst, int lst_cap) {
  if (!s || w <= 0 || h <= 0 || x0 < 0 || y0 < 0 || x0 + w > s->W || y0 + h > s->H) return -1;
  frame_setup fs;
  setup_frame(s, frame, x0, y0, w, h, &fs);
  const camera *cs[2] = {&fs.cam, &fs.pcam};
  for (int k = 0; k < 2; k++) {
    const camera *c = cs[k];
    double *o = cams + 14 * k;
    o[0] = c->pos.x; o[1] = c->pos.y; o[2] = c->pos.z; o[3] = c->fwd.x; o[4] = c->fwd.y; o[5] = c->fwd.z;
    o[6] = c->right.x; o[7] = c->right.y; o[8] = c->right.z; o[9] = c->up.x; o[10] = c->up.y; o[11] = c->up.z;
    o[12] = c->tx; o[13] = c->ty;
  }
  for (int i = 0; i < s->n; i++) {
    const prim *q = &s->p[i];
    double *o = prims + HS_PRIM_DOUBLES * i;
    o[0] = q->type; o[1] = q->id;
    o[2] = fs.pos[i].x; o[3] = fs.pos[i].y; o[4] = fs.pos[i].z;
    o[5] = fs.dpos[i].x; o[6] = fs.dpos[i].y; o[7] = fs.dpos[i].z;
    o[8] = q->h.x; o[9] = q->h.y; o[10] = q->h.z;
    o[11] = q->albedo.x; o[12] = q->albedo.y; o[13] = q->albedo.z; o[14] = q->rough;
    o[15] = q->emit.x; o[16] = q->emit.y; o[17] = q->emit.z; o[18] = 0; o[19] = 0;
  }
  const int nt = fs.tw * fs.th, len = fs.off[nt];
  memcpy(off, fs.off, sizeof(int) * ((size_t)nt + 1));
  if (len <= lst_cap) memcpy(lst, fs.lst, sizeof(int) * (size_t)(len > 0 ? len : 0));
  setup_free(&fs);
  return len;
}

/*
 * Renders frame `frame` into the crop window [x0, x0+w) x [y0, y0+h) of the scene's full W x H image.
 * Outputs are w*h*4 binary16 arrays (row-major, tight pitch); `clean` (optional) receives w*h*4 floats:
 * the expected radiance albedo * irradiance * E[noise] plus emission (the noise-free target).
 */
EXPORT int hs_render(const hs_scene *s, int frame, int x0, int y0, int w, int h, uint16_t *rad, uint16_t *nd,
                     uint16_t *mvo, uint16_t *mat, uint16_t *emi, float *clean) {
  if (!s || w <= 0 || h <= 0 || x0 < 0 || y0 < 0 || x0 + w > s->W || y0 + h > s->H) return -1;
  frame_setup fs;
  setup_frame(s, frame, x0, y0, w, h, &fs);
  const camera cam = fs.cam, pcam = fs.pcam;
  const d3 *pos = fs.pos, *dpos = fs.dpos;
  const int tw = fs.tw;
  const int *off = fs.off, *lst = fs.lst;
  const double noise_mean = exp(0.5) * (1.0 + 0.02 * 49.0);
#pragma omp parallel for schedule(dynamic, 4)
  for (int yy = 0; yy < h; yy++) {
    for (int xx = 0; xx < w; xx++) {
      int X = x0 + xx, Y = y0 + yy;
      size_t idx = (size_t)yy * w + xx;
      double ndx = ((X + 0.5) / s->W) * 2 - 1, ndy = ((Y + 0.5) / s->H) * 2 - 1;
      d3 d = norm(add(cam.fwd, sub(mul(cam.right, ndx * cam.tx), mul(cam.up, ndy * cam.ty))));
      hit best = {1e30, -2, {0, 0, 0}};
      if (d.y < -1e-9) {
        double t = -cam.pos.y / d.y;
        if (t > 1e-4 && t * dot(d, cam.fwd) < 300.0) { best.t = t; best.k = -1; best.n = v(0, 1, 0); }
      }
      int t = (yy / TILE) * tw + (xx / TILE);
      for (int li = off[t]; li < off[t + 1]; li++) {
        int k = lst[li];
        const prim *q = &s->p[k];
        double tt;
        d3 nn;
        if (q->type == PRIM_SPHERE) {
          if (ray_sphere(cam.pos, d, pos[k], q->h.x, &tt) && tt < best.t) {
            best.t = tt; best.k = k; best.n = norm(sub(add(cam.pos, mul(d, tt)), pos[k]));
          }
        } else if (ray_box(cam.pos, d, pos[k], q->h, &tt, &nn) && tt < best.t) {
          best.t = tt; best.k = k; best.n = nn;
        }
      }
      double rr, rg, rb, ar, ag, ab, ro, er = 0, eg = 0, eb = 0, id;
      double mvx, mvy, pdepth, depth;
      d3 nrm;
      if (best.k == -2) {                                  /* sky */
        d3 e = sky(d);
        er = e.x; eg = e.y; eb = e.z;
        rr = rg = rb = 0;
        ar = ag = ab = 1; ro = 1;
        nrm = mul(d, -1);
        depth = -1; pdepth = -1; id = 0;
        double cx, cy, px, py;
        if (project(s, &cam, d, 0, &cx, &cy) && project(s, &pcam, d, 0, &px, &py)) { mvx = px - cx; mvy = py - cy; }
        else { mvx = mvy = 0; }
      } else {
        d3 p = add(cam.pos, mul(d, best.t));
        d3 prevP = p;
        if (best.k >= 0) {
          const prim *q = &s->p[best.k];
          ar = q->albedo.x; ag = q->albedo.y; ab = q->albedo.z; ro = q->rough; id = q->id;
          er = q->emit.x; eg = q->emit.y; eb = q->emit.z;
          prevP = sub(p, dpos[best.k]);
        } else {                                           /* checkered ground */
          int ch = ((int)floor(p.x) + (int)floor(p.z)) & 1;
          ar = ch ? 0.72 : 0.32; ag = ch ? 0.70 : 0.34; ab = ch ? 0.64 : 0.38; ro = 0.8; id = 1;
        }
        nrm = best.n;
        depth = dot(sub(p, cam.pos), cam.fwd);
        pdepth = dot(sub(prevP, pcam.pos), pcam.fwd);
        double cx, cy, px, py;
        if (project(s, &cam, p, 1, &cx, &cy) && project(s, &pcam, prevP, 1, &px, &py)) { mvx = px - cx; mvy = py - cy; }
        else { mvx = mvy = 0; }
        int emitter = er > 0 || eg > 0 || eb > 0;
        if (emitter) { rr = rg = rb = 0; }
        else {
          d3 E = irradiance(nrm);
          double z = sqrt(-2.0 * log(1.0 - hrand(s->seed, 101, frame, (uint64_t)Y * 65536u + X)))
                   * cos(2 * M_PI * hrand(s->seed, 102, frame, (uint64_t)Y * 65536u + X));
          double nz = exp(z);
          if (hrand(s->seed, 103, frame, (uint64_t)Y * 65536u + X) < 0.02) nz *= 50.0;
          rr = ar * E.x * nz; rg = ag * E.y * nz; rb = ab * E.z * nz;
        }
        if (clean) {
          d3 E = irradiance(nrm);
          clean[idx * 4 + 0] = (float)(emitter ? er : ar * E.x * noise_mean + er);
          clean[idx * 4 + 1] = (float)(emitter ? eg : ag * E.y * noise_mean + eg);
          clean[idx * 4 + 2] = (float)(emitter ? eb : ab * E.z * noise_mean + eb);
          clean[idx * 4 + 3] = 0;
        }
      }
      if (best.k == -2 && clean) {
        clean[idx * 4 + 0] = (float)er; clean[idx * 4 + 1] = (float)eg; clean[idx * 4 + 2] = (float)eb; clean[idx * 4 + 3] = 0;
      }
      put(rad, idx, rr, rg, rb, 1);
      put(nd, idx, nrm.x, nrm.y, nrm.z, depth);
      put(mvo, idx, mvx, mvy, id, pdepth);
      put(mat, idx, ar, ag, ab, ro);
      put(emi, idx, er, eg, eb, 0);
    }
  }
  setup_free(&fs);
  return 0;
}
