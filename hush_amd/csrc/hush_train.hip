// Differentiable frame of the Hush denoiser (include/hush_train.h): forward and backward kernels templated on
// the element type T (float for training, double for finite-difference checks).
//
// Forward (no binary16 rounding between passes): MLP -> mid1 / mid2 / mid3 a-trous -> fine a-trous ->
// recombination -> TAA, following hush.js:348-366, 431-476, 560-620, 625-677.  Backward: one kernel per
// pass and direction; input gradients of the a-trous are gathered per source texel (every output that
// reads it recomputes its tap weight), so no kernel scatters and every result is deterministic.  The MLP
// weight gradient is reduced in FP64: a partial sum per workgroup, then one fixed-order sum per weight.
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include "hush.h"
#include "hush_ctx.h"
#include "hush_device.h"
#include "hush_internal.h"
#include "hush_train.h"

using namespace hush;

namespace {

constexpr int NTH = 256;
constexpr int NW = 812, HID = 16;
constexpr int OW1 = 0, OB1 = 320, OW2 = 336, OB2 = 592, OW3 = 608, OB3 = 800;

template <class T> __device__ __forceinline__ T texp(T x);
template <> __device__ __forceinline__ float texp(float x) { return expf(x); }
template <> __device__ __forceinline__ double texp(double x) { return exp(x); }
template <class T> __device__ __forceinline__ T tsqrt(T x);
template <> __device__ __forceinline__ float tsqrt(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double tsqrt(double x) { return sqrt(x); }

template <class T> __device__ __forceinline__ T tabs(T x) { return x < T(0) ? -x : x; }
template <class T> __device__ __forceinline__ T tfloor(T x);
template <> __device__ __forceinline__ float tfloor(float x) { return floorf(x); }
template <> __device__ __forceinline__ double tfloor(double x) { return floor(x); }

template <class T> struct V4 { T x, y, z, w; };
template <class T> __device__ __forceinline__ V4<T> ld4(const T *p, int i) { return V4<T>{p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; }
template <class T> __device__ __forceinline__ void st4(T *p, int i, V4<T> v) { p[4 * i] = v.x; p[4 * i + 1] = v.y; p[4 * i + 2] = v.z; p[4 * i + 3] = v.w; }
template <class T> __device__ __forceinline__ T lumT(T r, T g, T b) { return r * T(0.2126) + g * T(0.7152) + b * T(0.0722); }
template <class T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <class T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <class T> __device__ __forceinline__ T sgn(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }
template <class T> __device__ __forceinline__ T sigm(T x) { return T(1) / (T(1) + texp(-x)); }
template <class T> __device__ __forceinline__ T p128(T x) {
  T x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16, x64 = x32 * x32;
  return x64 * x64;
}

// guide texel: decoded normal and depth (w of GH / KD holds the binary16 pair (depth, id | count))
template <class T> struct G { T nx, ny, nz, z; };
template <class T> __device__ __forceinline__ G<T> guide(const TexF &g, int x, int y) {
  const float4 v = ldf(g, x, y);
  return G<T>{(T)v.x, (T)v.y, (T)v.z, (T)pair_of(v.w).x};
}
template <class T> __device__ __forceinline__ T ndot(const G<T> &a, const G<T> &b) { return a.nx * b.nx + a.ny * b.ny + a.nz * b.nz; }

__device__ __forceinline__ int tid1() { return blockIdx.x * NTH + threadIdx.x; }
inline dim3 grid1(long n) { return dim3((unsigned)((n + NTH - 1) / NTH)); }

// ------------------------------------------------------------------------------------------- MLP forward
template <class T>
__device__ __forceinline__ void features(const TexU &F0, const TexU &F1, const TexU &F2, int x, int y, T *f) {
  const uint4 a = ldu(F0, x, y), b = ldu(F1, x, y), c = ldu(F2, x, y);
  const uint32_t u[10] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y};
#pragma unroll
  for (int i = 0; i < 10; i++) {
    const float2 v = unpack2(u[i]);
    f[2 * i] = (T)v.x;
    f[2 * i + 1] = (T)v.y;
  }
}

template <class T>
__device__ __forceinline__ void mlp_fwd(const T *w, const T *f, T *h1, T *h2, T *o) {
#pragma unroll
  for (int j = 0; j < HID; j++) {
    T acc = w[OB1 + j];
#pragma unroll
    for (int i = 0; i < 20; i++) acc += w[OW1 + j * 20 + i] * f[i];
    h1[j] = tmax(acc, T(0));
  }
#pragma unroll
  for (int j = 0; j < HID; j++) {
    T acc = w[OB2 + j];
#pragma unroll
    for (int i = 0; i < HID; i++) acc += w[OW2 + j * HID + i] * h1[i];
    h2[j] = tmax(acc, T(0));
  }
#pragma unroll
  for (int j = 0; j < 12; j++) {
    T acc = w[OB3 + j];
#pragma unroll
    for (int i = 0; i < HID; i++) acc += w[OW3 + j * HID + i] * h2[i];
    o[j] = acc;
  }
}

// activations (hush.js:357-366): par = (alpha, gamma, beta, kappa, s0..s3, c0..c3) from the 12 outputs
template <class T>
__device__ __forceinline__ void acts(const T *r, T *pp) {
  pp[0] = T(1.0 / 32.0) + T(31.0 / 32.0) * sigm(r[0]);
  pp[1] = T(0.5) + T(7.5) * sigm(r[1]);
  pp[2] = sigm(r[2]);
  pp[3] = sigm(r[3]);
  for (int k = 0; k < 4; k++) {
    pp[4 + k] = T(0.25) * texp(tmin(tmax(r[4 + k], T(-5)), T(5)));
    pp[8 + k] = texp(tmin(tmax(r[8 + k], T(-4)), T(4)));
  }
}

// MLP and activations per half-res texel (hush.js:348-366): o (12) and par = (alpha, gamma, beta, kappa,
// s0..s3, c0..c3); sky blocks give zeros
template <class T>
__global__ __launch_bounds__(NTH) void tk_infer(TexU F0, TexU F1, TexU F2, TexF GH, int W2, int H2, const T *w, T *o, T *par) {
  const int i = tid1();
  if (i >= W2 * H2) return;
  const int x = i % W2, y = i / W2;
  T *oo = o + 12 * i, *pp = par + 12 * i;
  if (guide<T>(GH, x, y).z < T(0)) {
    for (int k = 0; k < 12; k++) { oo[k] = T(0); pp[k] = T(0); }
    return;
  }
  T f[20], h1[HID], h2[HID], r[12];
  features(F0, F1, F2, x, y, f);
  mlp_fwd(w, f, h1, h2, r);
  for (int k = 0; k < 12; k++) oo[k] = r[k];
  acts(r, pp);
}

// the activations alone, on MLP outputs already in o (the MFMA forward option); sky blocks give zeros
template <class T>
__global__ __launch_bounds__(NTH) void tk_act(TexF GH, int W2, int H2, T *o, T *par) {
  const int i = tid1();
  if (i >= W2 * H2) return;
  T *oo = o + 12 * i, *pp = par + 12 * i;
  if (guide<T>(GH, i % W2, i / W2).z < T(0)) {
    for (int k = 0; k < 12; k++) { oo[k] = T(0); pp[k] = T(0); }
    return;
  }
  T r[12];
  for (int k = 0; k < 12; k++) r[k] = oo[k];
  acts(r, pp);
}

// ------------------------------------------------------------------------------------------- a-trous
__host__ __device__ constexpr double tdist(int d2) {
  return d2 == 1 ? 1.0 : d2 == 2 ? 1.4142135623730951 : d2 == 4 ? 2.0 : d2 == 5 ? 2.23606797749979 : d2 == 8 ? 2.8284271247461903 : 0.0;
}
template <int R> __host__ __device__ constexpr double kh(int i) {
  return R == 1 ? (i == 1 ? 0.5 : 0.25) : (i == 2 ? 0.375 : (i == 1 || i == 3) ? 0.25 : 0.0625);
}

// one a-trous level: source (rgb, var) from the packed block means (SRC 0, F2.zw) or a T buffer (SRC 1);
// guide GH or KD; s / c from the par buffer (12 per texel, at (x, y) or (x >> 1, y >> 1) when PH)
template <class T> struct FiltArgs {
  TexU F2;
  const T *src;
  TexF guide;
  const T *par;
  int pw, ks, kc;      // par row width, component of s and c
  int w, h;
  T *out;              // forward output (4 per texel)
  const T *gout;       // backward: dL/dout
  T *gp;               // backward: dL/ds, dL/dc per output texel (stride gps at offsets js, jc)
  int gps, js, jc;
  T *coef;             // backward: 12 per output texel
  T *gsrc;             // backward gather: dL/dsrc (4 per texel)
};

template <class T, int SRC>
__device__ __forceinline__ V4<T> fsrc(const FiltArgs<T> &a, int x, int y) {
  if (SRC == 0) {
    const uint4 v = ldu(a.F2, x, y);
    const float2 p = unpack2(v.z), q = unpack2(v.w);
    return V4<T>{(T)p.x, (T)p.y, (T)q.x, (T)q.y};
  }
  return ld4(a.src, y * a.w + x);
}

// prefilter (hush.js:460-462): 3x3 [1/16 1/8 1/4] of the variance over clamp-to-edge texels
template <class T, int SRC>
__device__ __forceinline__ T prefilter(const FiltArgs<T> &a, int x, int y, T cvar) {
  T gv = T(0);
  const int xm = clampi(x - 1, 0, a.w - 1), xp = clampi(x + 1, 0, a.w - 1), ym = clampi(y - 1, 0, a.h - 1), yp = clampi(y + 1, 0, a.h - 1);
  gv = T(0.0625) * (fsrc<T, SRC>(a, xm, ym).w + fsrc<T, SRC>(a, xp, ym).w + fsrc<T, SRC>(a, xm, yp).w + fsrc<T, SRC>(a, xp, yp).w)
     + T(0.125) * (fsrc<T, SRC>(a, x, ym).w + fsrc<T, SRC>(a, xm, y).w + fsrc<T, SRC>(a, xp, y).w + fsrc<T, SRC>(a, x, yp).w)
     + T(0.25) * cvar;
  return gv;
}

template <class T>
__device__ __forceinline__ T zgrad(const TexF &g, int x, int y, int w, int h, T z) {
  const T zl = guide<T>(g, clampi(x - 1, 0, w - 1), y).z, zr = guide<T>(g, clampi(x + 1, 0, w - 1), y).z;
  const T zu = guide<T>(g, x, clampi(y - 1, 0, h - 1)).z, zd = guide<T>(g, x, clampi(y + 1, 0, h - 1)).z;
  const T gx = tmin(zr > T(0) ? tabs(zr - z) : T(1e9), zl > T(0) ? tabs(zl - z) : T(1e9));
  const T gy = tmin(zd > T(0) ? tabs(zd - z) : T(1e9), zu > T(0) ? tabs(zu - z) : T(1e9));
  return tmin(tmax(gx, gy), z);
}

// tap weight of source texel q for output p (hush.js:449-451)
template <class T, int R, int STEP>
__device__ __forceinline__ T tapw(int dx, int dy, const G<T> &gp, const G<T> &gq, T zg, T sl, T lc, T lq) {
  const T iz = T(1) / (zg * T(tdist(dx * dx + dy * dy) * STEP) + T(1e-3));
  return T(kh<R>(dx + R) * kh<R>(dy + R)) * p128(tmax(ndot(gq, gp), T(0))) * texp(-tabs(gq.z - gp.z) * iz - sl * tabs(lq - lc));
}

// forward (BWD = false) or backward first half (BWD = true: parameter gradients and per-output coefficients)
template <class T, int R, int STEP, bool PRE, int SRC, bool PH, bool BWD>
__global__ __launch_bounds__(NTH) void tk_filt(FiltArgs<T> a) {
  const int i = tid1();
  if (i >= a.w * a.h) return;
  const int x = i % a.w, y = i / a.w;
  const G<T> gp = guide<T>(a.guide, x, y);
  if (gp.z < T(0)) {
    if (!BWD) st4(a.out, i, V4<T>{T(0), T(0), T(0), T(0)});
    else {
      a.gp[i * a.gps + a.js] = T(0);
      a.gp[i * a.gps + a.jc] = T(0);
      for (int k = 0; k < 12; k++) a.coef[12 * i + k] = T(0);
    }
    return;
  }
  const int pidx = PH ? (y >> 1) * a.pw + (x >> 1) : y * a.pw + x;
  const T s = a.par[12 * pidx + a.ks], c = a.par[12 * pidx + a.kc];
  const V4<T> ce = fsrc<T, SRC>(a, x, y);
  const T gv = PRE ? prefilter<T, SRC>(a, x, y, ce.w) : ce.w;
  const T lc = lumT(ce.x, ce.y, ce.z);
  const T den = tsqrt(tmax(gv, T(1e-12))) + T(1e-6);
  const T sl = s / den;
  const T zg = zgrad<T>(a.guide, x, y, a.w, a.h, gp.z);
  const T H = T(kh<R>(R) * kh<R>(R));
  const T cw = H * c;
  T S = cw, C0 = ce.x * cw, C1 = ce.y * cw, C2 = ce.z * cw, V = ce.w * cw * cw;
  for (int dy = -R; dy <= R; dy++)
    for (int dx = -R; dx <= R; dx++) {
      if (dx == 0 && dy == 0) continue;
      const int qx = x + dx * STEP, qy = y + dy * STEP;
      if (!inside(qx, qy, a.w, a.h)) continue;
      const G<T> gq = guide<T>(a.guide, qx, qy);
      if (gq.z < T(0)) continue;
      const V4<T> cq = fsrc<T, SRC>(a, qx, qy);
      const T ww = tapw<T, R, STEP>(dx, dy, gp, gq, zg, sl, lc, lumT(cq.x, cq.y, cq.z));
      S += ww; C0 += ww * cq.x; C1 += ww * cq.y; C2 += ww * cq.z; V += ww * ww * cq.w;
    }
  const V4<T> o{C0 / S, C1 / S, C2 / S, V / (S * S)};
  if (!BWD) { st4(a.out, i, o); return; }
  const V4<T> g = ld4(a.gout, i);
  const T ax = g.x / S, ay = g.y / S, az = g.z / S, b = g.w / (S * S);
  const T dS = -(g.x * o.x + g.y * o.y + g.z * o.z + T(2) * g.w * o.w) / S;
  T dsl = T(0), dlc = T(0);
  for (int dy = -R; dy <= R; dy++)
    for (int dx = -R; dx <= R; dx++) {
      if (dx == 0 && dy == 0) continue;
      const int qx = x + dx * STEP, qy = y + dy * STEP;
      if (!inside(qx, qy, a.w, a.h)) continue;
      const G<T> gq = guide<T>(a.guide, qx, qy);
      if (gq.z < T(0)) continue;
      const V4<T> cq = fsrc<T, SRC>(a, qx, qy);
      const T lq = lumT(cq.x, cq.y, cq.z);
      const T ww = tapw<T, R, STEP>(dx, dy, gp, gq, zg, sl, lc, lq);
      const T t = (ax * cq.x + ay * cq.y + az * cq.z + T(2) * b * ww * cq.w + dS) * ww;
      dsl -= t * tabs(lq - lc);
      dlc += t * sl * sgn(lq - lc);
    }
  const T dwc = ax * ce.x + ay * ce.y + az * ce.z + T(2) * b * cw * ce.w + dS;
  a.gp[i * a.gps + a.js] = dsl / den;
  a.gp[i * a.gps + a.jc] = H * dwc;
  const T dgv = gv > T(1e-12) ? dsl * (-s / (den * den)) * (T(0.5) / tsqrt(gv)) : T(0);
  T *cf = a.coef + 12 * i;
  cf[0] = ax; cf[1] = ay; cf[2] = az; cf[3] = b; cf[4] = dS; cf[5] = sl; cf[6] = lc; cf[7] = zg;
  cf[8] = dgv; cf[9] = dlc; cf[10] = cw; cf[11] = T(1);
}

// backward second half: dL/dsrc per source texel r, gathered over the outputs that read it (as tap, as
// centre, through the variance prefilter)
template <class T, int R, int STEP, bool PRE, int SRC>
__global__ __launch_bounds__(NTH) void tk_filt_gather(FiltArgs<T> a) {
  const int i = tid1();
  if (i >= a.w * a.h) return;
  const int x = i % a.w, y = i / a.w;
  const V4<T> cr = fsrc<T, SRC>(a, x, y);
  const T lr = lumT(cr.x, cr.y, cr.z);
  T g0 = T(0), g1 = T(0), g2 = T(0), gv = T(0), gl = T(0);
  const T *cf = a.coef + 12 * i;
  if (cf[11] != T(0)) {                       // centre of its own output
    g0 += cf[10] * cf[0]; g1 += cf[10] * cf[1]; g2 += cf[10] * cf[2];
    gl += cf[9];
    gv += cf[10] * cf[10] * cf[3];
    if (!PRE) gv += cf[8];
  }
  const G<T> gr = guide<T>(a.guide, x, y);
  if (gr.z >= T(0)) {                         // tap of the outputs around it
    for (int dy = -R; dy <= R; dy++)
      for (int dx = -R; dx <= R; dx++) {
        if (dx == 0 && dy == 0) continue;
        const int px = x - dx * STEP, py = y - dy * STEP;
        if (!inside(px, py, a.w, a.h)) continue;
        const T *cp = a.coef + 12 * (py * a.w + px);
        if (cp[11] == T(0)) continue;
        const G<T> gp = guide<T>(a.guide, px, py);
        const T ww = tapw<T, R, STEP>(dx, dy, gp, gr, cp[7], cp[5], cp[6], lr);
        const T t = (cp[0] * cr.x + cp[1] * cr.y + cp[2] * cr.z + T(2) * cp[3] * ww * cr.w + cp[4]) * ww;
        g0 += ww * cp[0]; g1 += ww * cp[1]; g2 += ww * cp[2];
        gl -= t * cp[5] * sgn(lr - cp[6]);
        gv += ww * ww * cp[3];
      }
  }
  if (PRE) {                                  // variance prefilter over clamp-to-edge texels
    for (int py = y - 1; py <= y + 1; py++)
      for (int px = x - 1; px <= x + 1; px++) {
        if (!inside(px, py, a.w, a.h)) continue;
        const T *cp = a.coef + 12 * (py * a.w + px);
        if (cp[11] == T(0) || cp[8] == T(0)) continue;
        for (int ky = -1; ky <= 1; ky++)
          for (int kx = -1; kx <= 1; kx++) {
            if (clampi(px + kx, 0, a.w - 1) != x || clampi(py + ky, 0, a.h - 1) != y) continue;
            const T wk = (kx == 0 && ky == 0) ? T(0.25) : (kx == 0 || ky == 0) ? T(0.125) : T(0.0625);
            gv += wk * cp[8];
          }
      }
  }
  st4(a.gsrc, i, V4<T>{g0 + gl * T(0.2126), g1 + gl * T(0.7152), g2 + gl * T(0.0722), gv});
}

// ------------------------------------------------------------------------------------------- recombination
template <class T> struct FinalArgs {
  const T *fine, *par, *mid;       // fine level (full), par (half, 12), mid3 output (half)
  TexF kd, GH, cgd;
  TexU F2;
  Tex coarse, mat, emi;
  int W, H, W2, H2, W4, H4;
  T *X;
  const T *gX;                     // backward
  T *gfine, *gill, *gpf;           // backward outputs (full res, 4 per pixel)
  T *gmid, *gpar;                  // backward gather outputs (half res)
};

// jbw (hush.js:552-558): bilinear x depth x normal^8 weight of tap j
template <class T>
__device__ __forceinline__ T jbwT(T fx, T fy, int j, const G<T> &q, const G<T> &p) {
  const T wb = ((j & 1) ? fx : T(1) - fx) * ((j >> 1) ? fy : T(1) - fy);
  const T dn = tmax(ndot(q, p), T(0));
  const T dn2 = dn * dn, dn4 = dn2 * dn2;
  return wb * texp(-tabs(q.z - p.z) / (T(0.05) * p.z + T(1e-3))) * dn4 * dn4;
}

// the half-res mid upsample of pixel (x, y): uB, uA and the weight sum (hush.js:577-593)
template <class T>
__device__ __forceinline__ void mid_up(const FinalArgs<T> &a, int x, int y, const G<T> &gp, T *uB, T *uA, T &sw) {
  const T hx = (T(x) + T(0.5)) / T(2) - T(0.5), hy = (T(y) + T(0.5)) / T(2) - T(0.5);
  const T bx = tfloor(hx), by = tfloor(hy);
  sw = T(0);
  for (int k = 0; k < 3; k++) { uB[k] = T(0); uA[k] = T(0); }
  for (int j = 0; j < 4; j++) {
    const int qx = (int)bx + (j & 1), qy = (int)by + (j >> 1);
    if (qx < 0 || qy < 0 || qx >= a.W2 || qy >= a.H2) continue;
    const G<T> gq = guide<T>(a.GH, qx, qy);
    if (gq.z < T(0)) continue;
    const T w = jbwT(hx - bx, hy - by, j, gq, gp);
    const uint4 f2 = ldu(a.F2, qx, qy);
    const float2 az = unpack2(f2.z), aw = unpack2(f2.w);
    const V4<T> m = ld4(a.mid, qy * a.W2 + qx);
    sw += w;
    uB[0] += w * m.x; uB[1] += w * m.y; uB[2] += w * m.z;
    uA[0] += w * (T)az.x; uA[1] += w * (T)az.y; uA[2] += w * (T)aw.x;
  }
}

// the quarter-res coarse upsample (hush.js:596-610); constant with respect to the weights
template <class T>
__device__ __forceinline__ void coarse_up(const FinalArgs<T> &a, int x, int y, const G<T> &gp, T *uC, T &sw) {
  const T cx = (T(x) + T(0.5)) / T(4) - T(0.5), cy = (T(y) + T(0.5)) / T(4) - T(0.5);
  const T bx = tfloor(cx), by = tfloor(cy);
  sw = T(0);
  uC[0] = uC[1] = uC[2] = T(0);
  for (int j = 0; j < 4; j++) {
    const int qx = (int)bx + (j & 1), qy = (int)by + (j >> 1);
    if (qx < 0 || qy < 0 || qx >= a.W4 || qy >= a.H4) continue;
    const float4 cg = ldf(a.cgd, qx, qy);
    if (cg.w < 0.0f) continue;
    const G<T> gq{(T)cg.x, (T)cg.y, (T)cg.z, (T)cg.w};
    const T w = jbwT(cx - bx, cy - by, j, gq, gp);
    const float4 c = ld(a.coarse, qx, qy);
    sw += w;
    uC[0] += w * (T)c.x; uC[1] += w * (T)c.y; uC[2] += w * (T)c.z;
  }
}

template <class T, bool BWD>
__global__ __launch_bounds__(NTH) void tk_final(FinalArgs<T> a) {
  const int i = tid1();
  if (i >= a.W * a.H) return;
  const int x = i % a.W, y = i / a.W;
  const float4 emi = ld(a.emi, x, y), alb = ld(a.mat, x, y);
  const G<T> gp = guide<T>(a.kd, x, y);
  if (gp.z < T(0)) {
    if (!BWD) st4(a.X, i, V4<T>{(T)emi.x, (T)emi.y, (T)emi.z, T(0)});
    else {
      st4(a.gfine, i, V4<T>{T(0), T(0), T(0), T(0)});
      st4(a.gill, i, V4<T>{T(0), T(0), T(0), T(0)});
      st4(a.gpf, i, V4<T>{T(0), T(0), T(0), T(0)});
    }
    return;
  }
  const V4<T> fv = ld4(a.fine, i);
  const T f[3] = {fv.x, fv.y, fv.z};
  const T *pr = a.par + 12 * ((y >> 1) * a.W2 + (x >> 1));
  const T beta = pr[2], kappa = pr[3];
  T uB[3], uA[3], swh, uC[3], swq;
  mid_up(a, x, y, gp, uB, uA, swh);
  coarse_up(a, x, y, gp, uC, swq);
  const bool hm = swh > T(1e-4), hc = swq > T(1e-4);
  T ill[3], up[3];
  for (int k = 0; k < 3; k++) {
    ill[k] = hm ? uB[k] / swh + kappa * (f[k] - uA[k] / swh) : f[k];
    up[k] = hc ? uC[k] / swq : ill[k];
  }
  const T al[3] = {(T)alb.x, (T)alb.y, (T)alb.z};
  if (!BWD) {
    T o[3];
    for (int k = 0; k < 3; k++) o[k] = (ill[k] + (up[k] - ill[k]) * beta) * al[k] + (k == 0 ? (T)emi.x : k == 1 ? (T)emi.y : (T)emi.z);
    st4(a.X, i, V4<T>{o[0], o[1], o[2], T(0)});
    return;
  }
  const V4<T> gx = ld4(a.gX, i);
  const T g2[3] = {gx.x * al[0], gx.y * al[1], gx.z * al[2]};
  T dbeta = T(0), dkappa = T(0), gi[3], gf[3], gc[3];
  for (int k = 0; k < 3; k++) {
    dbeta += g2[k] * (up[k] - ill[k]);
    gi[k] = hc ? g2[k] * (T(1) - beta) : g2[k];
  }
  for (int k = 0; k < 3; k++) {
    if (hm) {
      dkappa += gi[k] * (f[k] - uA[k] / swh);
      gf[k] = kappa * gi[k];
      gc[k] = gi[k] / swh;       // dL/duB: spread over the mid taps by their weights
    } else {
      gf[k] = gi[k];
      gc[k] = T(0);
    }
  }
  if (!hc) dbeta = T(0);
  st4(a.gfine, i, V4<T>{gf[0], gf[1], gf[2], T(0)});
  st4(a.gill, i, V4<T>{gc[0], gc[1], gc[2], T(0)});
  st4(a.gpf, i, V4<T>{dbeta, dkappa, T(0), T(0)});     // .z, .w: ds0, dc0 from the fine level's backward
}

// per half-res texel q: dL/dmid3(q) gathered over the full-res pixels whose upsample taps q, and the
// gradients of (beta, kappa, s0, c0) summed over the four pixels that read q's parameters
template <class T>
__global__ __launch_bounds__(NTH) void tk_final_gather(FinalArgs<T> a) {
  const int i = tid1();
  if (i >= a.W2 * a.H2) return;
  const int qx = i % a.W2, qy = i / a.W2;
  T *gq = a.gpar + 12 * i;
  T gb = T(0), gk = T(0), gs = T(0), gc = T(0);
  for (int j = 0; j < 4; j++) {
    const int x = 2 * qx + (j & 1), y = 2 * qy + (j >> 1);
    if (x >= a.W || y >= a.H) continue;
    const V4<T> v = ld4(a.gpf, y * a.W + x);
    gb += v.x; gk += v.y; gs += v.z; gc += v.w;
  }
  for (int k = 0; k < 12; k++) gq[k] = T(0);
  gq[2] = gb; gq[3] = gk; gq[4] = gs; gq[8] = gc;
  T m0 = T(0), m1 = T(0), m2 = T(0);
  const G<T> gqg = guide<T>(a.GH, qx, qy);
  if (gqg.z >= T(0)) {
    for (int y = 2 * qy - 1; y <= 2 * qy + 2; y++)
      for (int x = 2 * qx - 1; x <= 2 * qx + 2; x++) {
        if (x < 0 || y < 0 || x >= a.W || y >= a.H) continue;
        const T hx = (T(x) + T(0.5)) / T(2) - T(0.5), hy = (T(y) + T(0.5)) / T(2) - T(0.5);
        const T bx = tfloor(hx), by = tfloor(hy);
        const int jx = qx - (int)bx, jy = qy - (int)by;
        if (jx < 0 || jx > 1 || jy < 0 || jy > 1) continue;
        const V4<T> gc4 = ld4(a.gill, y * a.W + x);
        if (gc4.x == T(0) && gc4.y == T(0) && gc4.z == T(0)) continue;
        const G<T> gp = guide<T>(a.kd, x, y);
        if (gp.z < T(0)) continue;
        const T w = jbwT(hx - bx, hy - by, jx + 2 * jy, gqg, gp);
        m0 += w * gc4.x; m1 += w * gc4.y; m2 += w * gc4.z;
      }
  }
  st4(a.gmid, i, V4<T>{m0, m1, m2, T(0)});
}

// ------------------------------------------------------------------------------------------- TAA
template <class T> struct TaaTArgs {
  const T *X, *prev;
  Tex mv;
  int W, H, reset;
  T *out;
  const T *gout;
  T *coef, *gX;
};

template <class T> __device__ __forceinline__ void tmT(const T *c, T *o, T &e) {
  e = T(1) + lumT(c[0], c[1], c[2]);
  o[0] = c[0] / e; o[1] = c[1] / e; o[2] = c[2] / e;
}
template <class T> __device__ __forceinline__ void ycT(const T *c, T *o) {
  o[0] = c[0] * T(0.25) + c[1] * T(0.5) + c[2] * T(0.25);
  o[1] = c[0] * T(0.5) - c[2] * T(0.5);
  o[2] = -c[0] * T(0.25) + c[1] * T(0.5) - c[2] * T(0.25);
}
// J_tm(c)^T g for tm(c) = c / e, e = 1 + lum(c)
template <class T> __device__ __forceinline__ void tm_bwd(const T *c, T e, const T *g, T *o) {
  const T gc = (g[0] * c[0] + g[1] * c[1] + g[2] * c[2]) / (e * e);
  o[0] = g[0] / e - gc * T(0.2126); o[1] = g[1] / e - gc * T(0.7152); o[2] = g[2] / e - gc * T(0.0722);
}
template <class T> __device__ __forceinline__ void yc_bwd(const T *g, T *o) {   // YCoCg^T g
  o[0] = T(0.25) * g[0] + T(0.5) * g[1] - T(0.25) * g[2];
  o[1] = T(0.5) * g[0] + T(0.5) * g[2];
  o[2] = T(0.25) * g[0] - T(0.5) * g[1] - T(0.25) * g[2];
}

template <class T, bool BWD>
__global__ __launch_bounds__(NTH) void tk_taa(TaaTArgs<T> a) {
  const int i = tid1();
  if (i >= a.W * a.H) return;
  const int x = i % a.W, y = i / a.W;
  const V4<T> xv = ld4(a.X, i);
  const T xc[3] = {xv.x, xv.y, xv.z};
  T cur[3], e0;
  tmT(xc, cur, e0);
  T res[3] = {cur[0], cur[1], cur[2]};
  bool inr = false;
  T m1[3] = {0, 0, 0}, sd[3], v3[3] = {0, 0, 0}, hh0[3], lo[3], hi[3], hh[3];
  if (!a.reset) {
    T sx[9][3];
    for (int j = 0; j < 9; j++) {
      const V4<T> q = ld4(a.X, clampi(y + j / 3 - 1, 0, a.H - 1) * a.W + clampi(x + j % 3 - 1, 0, a.W - 1));
      const T qc[3] = {q.x, q.y, q.z};
      T t[3], e;
      tmT(qc, t, e);
      ycT(t, sx[j]);
      for (int k = 0; k < 3; k++) m1[k] += sx[j][k];
    }
    for (int k = 0; k < 3; k++) m1[k] /= T(9);
    for (int j = 0; j < 9; j++)
      for (int k = 0; k < 3; k++) { const T d = sx[j][k] - m1[k]; v3[k] += d * d; }
    for (int k = 0; k < 3; k++) sd[k] = tsqrt(tmax(v3[k] / T(9), T(1e-12)));
    const float4 mv = ld(a.mv, x, y);
    const T ux = (T(x) + T(0.5) + (T)mv.x) / T(a.W), uy = (T(y) + T(0.5) + (T)mv.y) / T(a.H);
    inr = ux >= T(0) && uy >= T(0) && ux <= T(1) && uy <= T(1);
    if (inr) {
      const T px = ux * T(a.W) - T(0.5), py = uy * T(a.H) - T(0.5);
      const T bx = tfloor(px), by = tfloor(py), fx = px - bx, fy = py - by;
      T hist[3] = {0, 0, 0};
      for (int j = 0; j < 4; j++) {
        const int ox = j & 1, oy = j >> 1;
        const T wb = (ox ? fx : T(1) - fx) * (oy ? fy : T(1) - fy);
        const V4<T> t = ld4(a.prev, clampi(addw(f2i((float)by), oy), 0, a.H - 1) * a.W + clampi(addw(f2i((float)bx), ox), 0, a.W - 1));
        hist[0] += wb * t.x; hist[1] += wb * t.y; hist[2] += wb * t.z;
      }
      T th[3], eh;
      tmT(hist, th, eh);
      ycT(th, hh0);
      for (int k = 0; k < 3; k++) {
        lo[k] = m1[k] - T(1.25) * sd[k];
        hi[k] = m1[k] + T(1.25) * sd[k];
        hh[k] = tmin(tmax(hh0[k], lo[k]), hi[k]);
      }
      const T hr[3] = {hh[0] + hh[1] - hh[2], hh[0] + hh[2], hh[0] - hh[1] - hh[2]};
      for (int k = 0; k < 3; k++) res[k] = hr[k] + (cur[k] - hr[k]) * T(0.1);
    }
  }
  const T d = tmax(T(1) - lumT(res[0], res[1], res[2]), T(1e-4));
  if (!BWD) {
    st4(a.out, i, V4<T>{res[0] / d, res[1] / d, res[2] / d, T(1)});
    return;
  }
  // backward: itm, then the blend, the clip box and the tone map of the centre
  const V4<T> go = ld4(a.gout, i);
  const T g[3] = {go.x, go.y, go.z};
  T gr[3];
  {
    const bool act = T(1) - lumT(res[0], res[1], res[2]) > T(1e-4);
    const T gc = act ? (g[0] * res[0] + g[1] * res[1] + g[2] * res[2]) / (d * d) : T(0);
    gr[0] = g[0] / d + gc * T(0.2126); gr[1] = g[1] / d + gc * T(0.7152); gr[2] = g[2] / d + gc * T(0.0722);
  }
  T gcur[3], cA[3] = {0, 0, 0}, cB[3] = {0, 0, 0};
  if (inr) {
    const T gh[3] = {T(0.9) * gr[0], T(0.9) * gr[1], T(0.9) * gr[2]};
    const T ghh[3] = {gh[0] + gh[1] + gh[2], gh[0] - gh[2], -gh[0] + gh[1] - gh[2]};
    for (int k = 0; k < 3; k++) {
      gcur[k] = T(0.1) * gr[k];
      const T mxv = tmax(hh0[k], lo[k]);
      T glo = T(0), ghi = T(0);
      if (mxv > hi[k]) ghi = ghh[k];
      else if (!(hh0[k] > lo[k])) glo = ghh[k];
      const T gm1 = glo + ghi, gsd = T(1.25) * (ghi - glo);
      const T gv3 = v3[k] / T(9) > T(1e-12) ? gsd * (T(0.5) / sd[k]) / T(9) : T(0);
      cA[k] = gm1 / T(9);
      cB[k] = T(2) * gv3;
    }
  } else {
    for (int k = 0; k < 3; k++) gcur[k] = gr[k];
  }
  T gx[3];
  tm_bwd(xc, e0, gcur, gx);
  T *cf = a.coef + 12 * i;
  for (int k = 0; k < 3; k++) { cf[k] = cA[k]; cf[3 + k] = cB[k]; cf[6 + k] = m1[k]; }
  cf[9] = inr ? T(1) : T(0);
  st4(a.gX, i, V4<T>{gx[0], gx[1], gx[2], T(0)});
}

// dL/dX(r) += the clip-box statistics of the (up to nine) pixels whose 3x3 window reads r
template <class T>
__global__ __launch_bounds__(NTH) void tk_taa_gather(TaaTArgs<T> a) {
  const int i = tid1();
  if (i >= a.W * a.H || a.reset) return;
  const int x = i % a.W, y = i / a.W;
  const V4<T> xv = ld4(a.X, i);
  const T xc[3] = {xv.x, xv.y, xv.z};
  T t[3], e, sx[3];
  tmT(xc, t, e);
  ycT(t, sx);
  T gs[3] = {0, 0, 0};
  for (int py = y - 1; py <= y + 1; py++)
    for (int px = x - 1; px <= x + 1; px++) {
      if (!inside(px, py, a.W, a.H)) continue;
      const T *cf = a.coef + 12 * (py * a.W + px);
      if (cf[9] == T(0)) continue;
      int mult = 0;
      for (int ky = -1; ky <= 1; ky++)
        for (int kx = -1; kx <= 1; kx++)
          if (clampi(px + kx, 0, a.W - 1) == x && clampi(py + ky, 0, a.H - 1) == y) mult++;
      for (int k = 0; k < 3; k++) gs[k] += T(mult) * (cf[k] + cf[3 + k] * (sx[k] - cf[6 + k]));
    }
  T gy[3], gx[3];
  yc_bwd(gs, gy);
  tm_bwd(xc, e, gy, gx);
  const V4<T> g = ld4(a.gX, i);
  st4(a.gX, i, V4<T>{g.x + gx[0], g.y + gx[1], g.z + gx[2], T(0)});
}

// ------------------------------------------------------------------------------------------- loss
template <class T> __device__ __forceinline__ void tl(const T *c, T *o, T &e, bool &act) {
  const T l = lumT(c[0], c[1], c[2]);
  act = l > T(0);
  e = T(1) + (act ? l : T(0));
  o[0] = c[0] / e; o[1] = c[1] / e; o[2] = c[2] / e;
}
template <class T> __device__ __forceinline__ void tl_bwd(const T *c, T e, bool act, const T *g, T *o) {
  const T gc = act ? (g[0] * c[0] + g[1] * c[1] + g[2] * c[2]) / (e * e) : T(0);
  o[0] = g[0] / e - gc * T(0.2126); o[1] = g[1] / e - gc * T(0.7152); o[2] = g[2] / e - gc * T(0.0722);
}

template <class T>
__global__ __launch_bounds__(NTH) void tk_loss(const T *out, const T *tgt, const T *pout, const T *ptgt, T lambda, int n,
                                                 T *gout, double *part) {
  __shared__ double red[NTH];
  const int i = tid1();
  double l = 0.0;
  if (i < n) {
    const T o[3] = {out[4 * i], out[4 * i + 1], out[4 * i + 2]}, r[3] = {tgt[4 * i], tgt[4 * i + 1], tgt[4 * i + 2]};
    T to[3], tr[3], eo, er;
    bool ao, ar;
    tl(o, to, eo, ao);
    tl(r, tr, er, ar);
    T g[3];
    for (int k = 0; k < 3; k++) {
      const T d = to[k] - tr[k];
      l += tabs((double)d);
      g[k] = sgn(d);
    }
    if (pout) {
      const T po[3] = {pout[4 * i], pout[4 * i + 1], pout[4 * i + 2]}, pr[3] = {ptgt[4 * i], ptgt[4 * i + 1], ptgt[4 * i + 2]};
      T tpo[3], tpr[3], e1, e2;
      bool a1, a2;
      tl(po, tpo, e1, a1);
      tl(pr, tpr, e2, a2);
      for (int k = 0; k < 3; k++) {
        const T d = (to[k] - tpo[k]) - (tr[k] - tpr[k]);
        l += (double)lambda * tabs((double)d);
        g[k] += lambda * sgn(d);
      }
    }
    const T inv = T(1) / T(n);
    for (int k = 0; k < 3; k++) g[k] *= inv;
    T go[3];
    tl_bwd(o, eo, ao, g, go);
    gout[4 * i] = go[0]; gout[4 * i + 1] = go[1]; gout[4 * i + 2] = go[2]; gout[4 * i + 3] = T(0);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = NTH / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// fixed-order FP64 sum of `n` partials, scaled (one workgroup)
__global__ __launch_bounds__(NTH) void tk_sum(const double *part, int n, double scale, double *res) {
  __shared__ double red[NTH];
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += NTH) s += part[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int t = NTH / 2; t > 0; t >>= 1) {
    if ((int)threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) *res = red[0] * scale;
}

// ------------------------------------------------------------------------------------------- MLP backward
constexpr int MB = 128;                 // texels per workgroup of the weight-gradient reduction
constexpr int NV = 20 + 4 * HID + 12;   // f, h1, h2, gh1, gh2, go

// One texel of the MLP backward: features f, hidden activations h1, h2 (recomputed forward), the output
// gradient go through the activations (hush.js:357-366) and the hidden gradients gh2, gh1 through the ReLUs;
// all zero for a sky or out-of-range texel.
template <class T>
__device__ __forceinline__ void mlp_bwd_texel(TexU F0, TexU F1, TexU F2, TexF GH, int W2, int H2, const T *w, const T *o,
                                              const T *par, const T *gpar, int i, T *f, T *h1, T *h2, T *go, T *gh1,
                                              T *gh2) {
  bool live = false;
  if (i < W2 * H2) {
    const int x = i % W2, y = i / W2;
    live = guide<T>(GH, x, y).z >= T(0);
    if (live) {
      features(F0, F1, F2, x, y, f);
      T r[12];
      mlp_fwd(w, f, h1, h2, r);
      const T *p = par + 12 * i, *g = gpar + 12 * i, *oo = o + 12 * i;
      const T s0 = sigm(oo[0]), s1 = sigm(oo[1]);
      go[0] = g[0] * T(31.0 / 32.0) * s0 * (T(1) - s0);
      go[1] = g[1] * T(7.5) * s1 * (T(1) - s1);
      go[2] = g[2] * p[2] * (T(1) - p[2]);
      go[3] = g[3] * p[3] * (T(1) - p[3]);
      for (int k = 0; k < 4; k++) {
        go[4 + k] = (oo[4 + k] > T(-5) && oo[4 + k] < T(5)) ? g[4 + k] * p[4 + k] : T(0);
        go[8 + k] = (oo[8 + k] > T(-4) && oo[8 + k] < T(4)) ? g[8 + k] * p[8 + k] : T(0);
      }
      for (int j = 0; j < HID; j++) {
        T acc = T(0);
        for (int k = 0; k < 12; k++) acc += w[OW3 + k * HID + j] * go[k];
        gh2[j] = h2[j] > T(0) ? acc : T(0);
      }
      for (int j = 0; j < HID; j++) {
        T acc = T(0);
        for (int k = 0; k < HID; k++) acc += w[OW2 + k * HID + j] * gh2[k];
        gh1[j] = h1[j] > T(0) ? acc : T(0);
      }
    }
  }
  if (!live) {
    for (int k = 0; k < 20; k++) f[k] = T(0);
    for (int k = 0; k < HID; k++) { h1[k] = h2[k] = gh1[k] = gh2[k] = T(0); }
    for (int k = 0; k < 12; k++) go[k] = T(0);
  }
}

// rows of the per-texel value table: f, h1, h2, gh1, gh2, go; one column per texel (padded by one against
// LDS bank conflicts between rows)
constexpr int RF = 0, RH1 = 20, RH2 = 20 + HID, RG1 = 20 + 2 * HID, RG2 = 20 + 3 * HID, RGO = 20 + 4 * HID;
constexpr int SVP = MB + 1;

template <class T>
__device__ __forceinline__ void store_texel(T (*sv)[SVP], int t, const T *f, const T *h1, const T *h2, const T *go,
                                            const T *gh1, const T *gh2) {
  for (int k = 0; k < 20; k++) sv[RF + k][t] = f[k];
  for (int k = 0; k < HID; k++) {
    sv[RH1 + k][t] = h1[k]; sv[RH2 + k][t] = h2[k]; sv[RG1 + k][t] = gh1[k]; sv[RG2 + k][t] = gh2[k];
  }
  for (int k = 0; k < 12; k++) sv[RGO + k][t] = go[k];
}

// The MLP weight gradient, FP64 (the gradient-check contexts): a partial per 128-texel workgroup, each weight
// the FP64 sum of its products over the workgroup's texels in texel order
__global__ __launch_bounds__(MB) void tk_mlp_bwd64(TexU F0, TexU F1, TexU F2, TexF GH, int W2, int H2, const double *w,
                                                  const double *o, const double *par, const double *gpar, double *part) {
  __shared__ double sv[NV][SVP];
  const int t = threadIdx.x;
  double f[20], h1[HID], h2[HID], go[12], gh1[HID], gh2[HID];
  mlp_bwd_texel<double>(F0, F1, F2, GH, W2, H2, w, o, par, gpar, blockIdx.x * MB + t, f, h1, h2, go, gh1, gh2);
  store_texel<double>(sv, t, f, h1, h2, go, gh1, gh2);
  __syncthreads();
  for (int k = t; k < NW; k += MB) {
    int ra, rb;          // rows of sv whose product sums to weight k (rb < 0: a bias, the sum of row ra)
    if (k < OB1) { ra = RG1 + k / 20; rb = RF + k % 20; }
    else if (k < OW2) { ra = RG1 + (k - OB1); rb = -1; }
    else if (k < OB2) { ra = RG2 + (k - OW2) / HID; rb = RH1 + (k - OW2) % HID; }
    else if (k < OW3) { ra = RG2 + (k - OB2); rb = -1; }
    else if (k < OB3) { ra = RGO + (k - OW3) / HID; rb = RH2 + (k - OW3) % HID; }
    else { ra = RGO + (k - OB3); rb = -1; }
    double acc = 0.0;
    if (rb >= 0)
      for (int j = 0; j < MB; j++) acc += sv[ra][j] * sv[rb][j];
    else
      for (int j = 0; j < MB; j++) acc += sv[ra][j];
    part[(size_t)blockIdx.x * NW + k] = acc;
  }
}

// The MLP weight gradient, FP32 (training contexts), on the matrix cores.  Per texel tile the gradients are
// outer-product sums over the texels: dW1 = sum gh1 f^T, dW2 = sum gh2 h1^T, dW3 = sum go h2^T and the biases
// sum gh1, sum gh2, sum go.  Each is a 16 x 16 tile D = A B with the texels as the inner dimension, so wave v
// of the workgroup feeds its 64 texels of each 128-texel tile, four at a time, to v_mfma_f32_16x16x4f32 (lane
// l: A[l % 16][l / 16], B[l / 16][l % 16]; D[4 (l / 16) + r][l % 16] in accumulator r).  Six tiles:
//   T0  A = gh1, B = f[0..15]                -> dW1[:, 0..15]
//   T1  A = gh1, B = (f[16..19], 1, 0 ...)   -> dW1[:, 16..19], db1
//   T2  A = gh2, B = h1                      -> dW2
//   T3  A = (go, 0), B = h2                  -> dW3
//   T4  A = gh2, B = (1, 0 ...)              -> db2
//   T5  A = (go, 0), B = (1, 0 ...)          -> db3
// The accumulators run over MT tiles per workgroup (FP32 fused multiply-add chains in texel order); the two
// waves' sums add in wave order and the workgroup writes one partial per weight, which the FP64 stage
// (tk_wsum1 / tk_wsum2) reduces.
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int MT = 4;                   // 128-texel tiles per workgroup of the FP32 form
__global__ __launch_bounds__(MB) void tk_mlp_bwd_mfma(TexU F0, TexU F1, TexU F2, TexF GH, int W2, int H2, const float *w,
                                                      const float *o, const float *par, const float *gpar, double *part) {
  __shared__ float sv[NV][SVP];
  __shared__ f32x4 red[6][64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, li = lane & 15, lk = lane >> 4;
  f32x4 acc[6];
  for (int q = 0; q < 6; q++) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int it = 0; it < MT; it++) {
    float f[20], h1[HID], h2[HID], go[12], gh1[HID], gh2[HID];
    mlp_bwd_texel<float>(F0, F1, F2, GH, W2, H2, w, o, par, gpar, (blockIdx.x * MT + it) * MB + t, f, h1, h2, go, gh1, gh2);
    if (it) __syncthreads();                     // the previous tile's operands have been read
    store_texel<float>(sv, t, f, h1, h2, go, gh1, gh2);
    __syncthreads();
#pragma unroll 4
    for (int st = 0; st < 16; st++) {
      const int j = wv * 64 + st * 4 + lk;        // texel of this lane's k
      const float a1 = sv[RG1 + li][j], a2 = sv[RG2 + li][j], ao = li < 12 ? sv[RGO + li][j] : 0.0f;
      const float bf0 = sv[RF + li][j], bf1 = li < 4 ? sv[RF + 16 + li][j] : (li == 4 ? 1.0f : 0.0f);
      const float bh1 = sv[RH1 + li][j], bh2 = sv[RH2 + li][j], b1 = li == 0 ? 1.0f : 0.0f;
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, bf0, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, bf1, acc[1], 0, 0, 0);
      acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, bh1, acc[2], 0, 0, 0);
      acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(ao, bh2, acc[3], 0, 0, 0);
      acc[4] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b1, acc[4], 0, 0, 0);
      acc[5] = __builtin_amdgcn_mfma_f32_16x16x4f32(ao, b1, acc[5], 0, 0, 0);
    }
  }
  if (wv == 1)
    for (int q = 0; q < 6; q++) red[q][lane] = acc[q];
  __syncthreads();
  if (wv == 0) {
    double *pb = part + (size_t)blockIdx.x * NW;
    for (int r = 0; r < 4; r++) {
      const int i = 4 * lk + r, j = li;           // D[i][j] in accumulator r of this lane
      float d[6];
      for (int q = 0; q < 6; q++) d[q] = acc[q][r] + red[q][lane][r];
      pb[OW1 + i * 20 + j] = d[0];
      if (j < 4) pb[OW1 + i * 20 + 16 + j] = d[1];
      else if (j == 4) pb[OB1 + i] = d[1];
      pb[OW2 + i * HID + j] = d[2];
      if (i < 12) pb[OW3 + i * HID + j] = d[3];
      if (j == 0) pb[OB2 + i] = d[4];
      if (j == 0 && i < 12) pb[OB3 + i] = d[5];
    }
  }
}

// Per weight: the FP64 sum of the nb workgroup partials part[b * NW + k], in a fixed order (the same bits on
// every run), in two stages so that it spreads over the GPU.  Stage 1: workgroup (x, y) takes 64 consecutive
// weights (one per lane, so each wave reads 512 contiguous bytes per partial row) and the y-th of WSL slices
// of the rows; wave v sums rows v, v + 4, ... of the slice, then the four wave sums add in wave order.
// Stage 2: the WSL slice sums of each weight, the same way.  A single thread per weight walking all rows
// was a dependent chain of nb loads (2.6 ms at 1920x1080).
constexpr int WSL = 64;
__global__ __launch_bounds__(NTH) void tk_wsum1(const double *part, int nb, double *mid) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, k = blockIdx.x * 64 + lane;
  const int per = (nb + WSL - 1) / WSL, b0 = blockIdx.y * per, b1 = min(nb, b0 + per);
  double acc = 0.0;
  if (k < NW) {
#pragma unroll 8
    for (int b = b0 + wv; b < b1; b += 4) acc += part[(size_t)b * NW + k];
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && k < NW) mid[(size_t)blockIdx.y * NW + k] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}
template <class T>
__global__ __launch_bounds__(NTH) void tk_wsum2(const double *mid, T *gw) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, k = blockIdx.x * 64 + lane;
  double acc = 0.0;
  if (k < NW) {
#pragma unroll
    for (int q = 0; q < WSL / 4; q++) acc += mid[(size_t)(wv + 4 * q) * NW + k];
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && k < NW) gw[k] = (T)(((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane]);
}


// ------------------------------------------------------------------------------------------- temporal blend
// A0 from the temporal pass's blend inputs (aux, TEMPORAL_AUX floats per pixel) and the previous frame's
// alpha / gamma (hush.js:189-211): colour = lerp(clamp(hc, mu +- gamma sd), illum, a), a = max(alpha, 1/len1);
// variance = svar + (tvar - svar) min(len1 / 4, 1), tvar from the blended moments
template <class T>
__global__ __launch_bounds__(NTH) void tk_blend(const float *aux, const T *pprev, int n, T *a0) {
  const int i = tid1();
  if (i >= n) return;
  const float *x = aux + (size_t)i * TEMPORAL_AUX;
  if (x[0] < 0.0f) { st4(a0, i, V4<T>{T(0), T(0), T(0), T(0)}); return; }
  const T ce[3] = {x[1], x[2], x[3]}, l = x[16], svar = x[17];
  T col[3], nmx = l, nmy = l * l, len1 = T(1);
  if (x[0] > 0.0f) {
    const int q = (int)x[18];
    const T al = pprev[12 * q], ga = pprev[12 * q + 1];
    len1 = tmin((T)x[15] + T(1), T(32));
    const T aa = tmax(al, T(1) / len1);
    for (int k = 0; k < 3; k++) {
      const T g = ga * (T)x[7 + k], mu = x[4 + k];
      const T hl = tmin(tmax((T)x[10 + k], mu - g), mu + g);
      col[k] = hl + (ce[k] - hl) * aa;
    }
    nmx = (T)x[13] + (l - (T)x[13]) * aa;
    nmy = (T)x[14] + (l * l - (T)x[14]) * aa;
  } else {
    for (int k = 0; k < 3; k++) col[k] = ce[k];
  }
  const T tvar = tmax(nmy - nmx * nmx, T(0));
  st4(a0, i, V4<T>{col[0], col[1], col[2], svar + (tvar - svar) * tmin(len1 / T(4), T(1))});
}

// dL/dalpha, dL/dgamma of the previous frame's half-res texels from dL/dA0.  A pixel's reprojected texel
// follows its motion vector, so several pixels may feed one texel in any order; the sums are made
// order-independent (and so bit-reproducible) with 64-bit fixed point: pass 1 stores every pixel's
// contributions and finds their largest magnitude (an exact maximum over the float bit patterns), pass 2
// adds every contribution as an integer, scaled by a power of two that keeps the sum of all of them below
// 2^62, and a per-texel pass converts the two sums back.  The fixed point keeps 2^-62 of (pixels x the
// largest contribution).
template <class T>
__device__ __forceinline__ bool blend_bwd_terms(const float *aux, const T *pprev, const T *ga0, int i, int &q, T &daa,
                                                T &dga) {
  const float *x = aux + (size_t)i * TEMPORAL_AUX;
  daa = dga = T(0);
  if (!(x[0] > 0.0f)) return false;
  const V4<T> g = ld4(ga0, i);
  const T gv[3] = {g.x, g.y, g.z};
  q = (int)x[18];
  const T al = pprev[12 * q], ga = pprev[12 * q + 1];
  const T l = x[16], len1 = tmin((T)x[15] + T(1), T(32));
  const T aa = tmax(al, T(1) / len1);
  for (int k = 0; k < 3; k++) {
    const T sd = x[7 + k], mu = x[4 + k], hc = x[10 + k];
    const T lo = mu - ga * sd, hi = mu + ga * sd;
    const T mv = tmax(hc, lo);
    const T hl = tmin(mv, hi);
    daa += gv[k] * ((T)x[1 + k] - hl);
    const T dhl = gv[k] * (T(1) - aa);
    if (!(mv < hi)) dga += dhl * sd;
    else if (!(hc > lo)) dga -= dhl * sd;
  }
  const T nmx = (T)x[13] + (l - (T)x[13]) * aa, nmy = (T)x[14] + (l * l - (T)x[14]) * aa;
  if (nmy - nmx * nmx > T(0)) daa += g.w * tmin(len1 / T(4), T(1)) * ((l * l - (T)x[14]) - T(2) * nmx * (l - (T)x[13]));
  if (!(al > T(1) / len1)) daa = T(0);          // alpha below 1 / len1 does not reach the blend
  return true;
}
__device__ __forceinline__ unsigned long long mag_bits(double v) { return (unsigned long long)__double_as_longlong(fabs(v)); }
// pass 1: each pixel's two contributions and target texel (q = -1: none), and each workgroup's largest
// |contribution| into mx[blockIdx] (non-negative doubles order like their bit patterns)
template <class T>
__global__ __launch_bounds__(NTH) void tk_blend_bwd_max(const float *aux, const T *pprev, const T *ga0, int n, T *terms,
                                                        int *tq, unsigned long long *mx) {
  const int i = tid1();
  unsigned long long m = 0;
  if (i < n) {
    int q = -1;
    T daa, dga;
    if (!blend_bwd_terms(aux, pprev, ga0, i, q, daa, dga)) q = -1;
    terms[2 * (size_t)i] = daa; terms[2 * (size_t)i + 1] = dga; tq[i] = q;
    m = max(mag_bits((double)daa), mag_bits((double)dga));
  }
  // the workgroup's maximum (waves by lane shuffles, every lane taking part, then across the four waves)
  // into its partial; tk_max_bits reduces the partials
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned lo = __shfl_xor((unsigned)m, off), hi = __shfl_xor((unsigned)(m >> 32), off);
    m = max(m, ((unsigned long long)hi << 32) | lo);
  }
  __shared__ unsigned long long wm[NTH / 64];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < NTH / 64; k++) m = max(m, wm[k]);
    mx[blockIdx.x] = m;
  }
}
// the largest of n partial maxima, by one workgroup: mx[n] = max(mx[0 .. n-1])
__global__ __launch_bounds__(NTH) void tk_max_bits(unsigned long long *mx, int n) {
  unsigned long long m = 0;
  for (int k = threadIdx.x; k < n; k += NTH) m = max(m, mx[k]);
  __shared__ unsigned long long sm[NTH];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = NTH / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) mx[n] = sm[0];
}
// the power of two s with n x max|contribution| x s < 2^62 (at most 2^1000); 0 when every contribution is
// zero, -1 when one is not finite
__device__ __forceinline__ double fixed_scale(unsigned long long mxb, int n) {
  if (!mxb) return 0.0;
  if (mxb >= 0x7ff0000000000000ull) return -1.0;
  int e;
  (void)frexp(__longlong_as_double((long long)mxb), &e);     // max < 2^e
  const int L = 32 - __clz(n);                                // n < 2^L
  return ldexp(1.0, min(62 - e - L, 1000));
}
// pass 2: the contributions as integers, summed exactly
template <class T>
__global__ __launch_bounds__(NTH) void tk_blend_bwd_add(const T *terms, const int *tq, int n, const unsigned long long *mx,
                                                        long long *acc) {
  const int i = tid1();
  if (i >= n) return;
  const int q = tq[i];
  if (q < 0) return;
  const T daa = terms[2 * (size_t)i], dga = terms[2 * (size_t)i + 1];
  const double sc = fixed_scale(*mx, n);
  if (!(sc > 0.0)) return;
  const long long ia = (long long)rint((double)daa * sc), ig = (long long)rint((double)dga * sc);
  if (ia) atomicAdd((unsigned long long *)&acc[2 * q], (unsigned long long)ia);
  if (ig) atomicAdd((unsigned long long *)&acc[2 * q + 1], (unsigned long long)ig);
}
// pass 3: the sums into the previous frame's parameter gradient (the other ten entries zero)
template <class T>
__global__ __launch_bounds__(NTH) void tk_blend_bwd_store(const long long *acc, const unsigned long long *mx, int n, int n2,
                                                          T *gprev) {
  const int q = tid1();
  if (q >= n2) return;
  const double sc = fixed_scale(*mx, n);
  T *g = gprev + 12 * (size_t)q;
  const T bad = sc < 0.0 ? (T)NAN : T(0);                    // a non-finite contribution poisons the gradient
  g[0] = sc > 0.0 ? (T)((double)acc[2 * q] / sc) : bad;
  g[1] = sc > 0.0 ? (T)((double)acc[2 * q + 1] / sc) : bad;
  for (int k = 2; k < 12; k++) g[k] = T(0);
}

// (alpha, gamma, beta, kappa) of a stored binary16 par0 texel into the 12-wide parameter layout
template <class T>
__global__ __launch_bounds__(NTH) void tk_par_from_tex(Tex src, int W2, int H2, T *par) {
  const int i = tid1();
  if (i >= W2 * H2) return;
  const float4 v = ld(src, i % W2, i / W2);
  T *p = par + 12 * i;
  p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w;
  for (int k = 4; k < 12; k++) p[k] = T(0);
}

template <class T>
__global__ __launch_bounds__(NTH) void tk_zero(T *p, long n) {
  const long i = (long)blockIdx.x * NTH + threadIdx.x;
  if (i < n) p[i] = T(0);
}

__global__ __launch_bounds__(NTH) void tk_copy_u4(const TexU src, int W2, int H2, uint4 *dst) {
  const int i = tid1();
  if (i < W2 * H2) dst[i] = ldu(src, i % W2, i / W2);
}
__global__ __launch_bounds__(NTH) void tk_copy_f4(const TexF src, int W2, int H2, float4 *dst) {
  const int i = tid1();
  if (i < W2 * H2) dst[i] = ldf(src, i % W2, i / W2);
}

// ------------------------------------------------------------------------------------------- state
template <class T>
__global__ __launch_bounds__(NTH) void tk_from_tex(Tex src, int W, int H, T *dst) {
  const int i = tid1();
  if (i >= W * H) return;
  const float4 v = ld(src, i % W, i / W);
  st4(dst, i, V4<T>{(T)v.x, (T)v.y, (T)v.z, (T)v.w});
}
template <class T>
__global__ __launch_bounds__(NTH) void tk_to_tex(const T *src, int stride, int W, int H, Tex dst) {
  const int i = tid1();
  if (i >= W * H) return;
  const T *p = src + (size_t)stride * i;
  st(dst, i % W, i / W, make_float4((float)p[0], (float)p[1], (float)p[2], (float)p[3]));
}
template <class T>
__global__ __launch_bounds__(NTH) void tk_copy(const T *src, T *dst, long n) {
  const long i = (long)blockIdx.x * NTH + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

struct hush_train_ctx {
  hush_t *h = nullptr;
  bool fp64 = false;
  size_t tsz = 4;
  int W = 0, H = 0, W2 = 0, H2 = 0, W4 = 0, H4 = 0;
  void *arena = nullptr;
  // element-type buffers (T = float or double)
  void *o = nullptr, *par = nullptr, *mid[3] = {}, *gpar = nullptr, *gmid[3] = {}, *coefH = nullptr;
  void *a0 = nullptr, *fine = nullptr, *X = nullptr, *prev = nullptr, *gX = nullptr, *taac = nullptr;
  void *gfine = nullptr, *gill = nullptr, *gpf = nullptr, *coefF = nullptr;
  double *wpart = nullptr, *lpart = nullptr;
  int nwb = 0, nlb = 0;
  float *aux = nullptr;                 // temporal-blend inputs of the prepared frame
  uint4 *Fprev[3] = {};                 // the previous frame's features F0..F2 (half res)
  float4 *GHprev = nullptr;             // and its half-res guide (sky blocks)
  void *oprev = nullptr, *pprev = nullptr, *gpprev = nullptr, *ga0 = nullptr;
  long long *bacc = nullptr;            // fixed-point blend-gradient sums (2 per half-res texel)
  unsigned long long *bmax = nullptr;   // per-workgroup maxima of the blend-gradient contributions, then the maximum
  void *bterm = nullptr;                // per pixel: the two blend-gradient contributions (element type)
  int *bq = nullptr;                    // per pixel: their half-res texel, -1 for none
  bool has_prev = false;                // Fprev holds the committed previous frame
  int prev_mode = 0;                    // this frame: 0 no history, 1 alpha/gamma from the stored par0, 2 recomputed
  bool a0_fixed = false;
  int fwd_prec = HUSH_MLP_FP32;         // forward MLP: FP32, or an MFMA form over packed weights (wq)
  void *wq = nullptr;
  hush_inputs in{};
  bool prepared = false, forwarded = false, reset = true;
  char err[512] = {0};
};

static thread_local char g_terr[512];

static int tfail(hush_train_t *t, int code, const char *msg, hipError_t e = hipSuccess) {
  if (e != hipSuccess) std::snprintf(t ? t->err : g_terr, 512, "%s: %s", msg, hipGetErrorString(e));
  else std::snprintf(t ? t->err : g_terr, 512, "%s", msg);
  return code;
}

extern "C" const char *hush_train_last_error(const hush_train_t *t) { return t ? t->err : g_terr; }

extern "C" int32_t hush_train_create(hush_t *h, int32_t fp64, hush_train_t **out) {
  if (!h || !out) return tfail(nullptr, HUSH_ERR_ARGUMENT, "hush_train_create: null argument");
  *out = nullptr;
  if (h->hidden != HID) return tfail(nullptr, HUSH_ERR_UNSUPPORTED, "hush_train_create: the training path is built for hidden width 16");
  hipError_t e = hipSetDevice(h->device);
  if (e != hipSuccess) return tfail(nullptr, HUSH_ERR_DEVICE, "hipSetDevice", e);
  hush_train_t *t = new (std::nothrow) hush_train_ctx;
  if (!t) return tfail(nullptr, HUSH_ERR_DEVICE, "hush_train_create: out of host memory");
  t->h = h; t->fp64 = fp64 != 0; t->tsz = fp64 ? 8 : 4;
  t->W = h->W; t->H = h->H; t->W2 = h->W2; t->H2 = h->H2; t->W4 = h->W4; t->H4 = h->H4;
  const size_t N = (size_t)t->W * t->H, N2 = (size_t)t->W2 * t->H2, ts = t->tsz;
  t->nwb = (int)((N2 + (size_t)MB * (fp64 ? 1 : MT) - 1) / ((size_t)MB * (fp64 ? 1 : MT)));
  t->nlb = (int)((N + NTH - 1) / NTH);
  // half: o, par, gpar, coefH (12 each), mid x3, gmid x3 (4 each), the fixed-point blend sums; full: a0, fine,
  // X, prev, gX, gfine, gill, gpf (4 each), taac, coefF (12 each); FP64 partials
  const size_t half = N2 * ts * (7 * 12 + 6 * 4) + N2 * 16 * 4 + N2 * 2 * 8,
               full = N * ts * (9 * 4 + 2 * 12 + 2) + N * 4 * TEMPORAL_AUX + N * 4 + ((N + NTH - 1) / NTH + 1) * 8;
  const size_t parts = ((size_t)2 * t->nwb * NW + (size_t)WSL * NW + t->nlb + 64) * sizeof(double);
  const size_t total = half + full + parts + 64 * 256;
  e = hipMalloc(&t->arena, total);
  if (e == hipSuccess) e = hipMemset(t->arena, 0, total);
  if (e != hipSuccess) {
    int r = tfail(nullptr, HUSH_ERR_DEVICE, "hipMalloc of the training buffers", e);
    if (t->arena) (void)hipFree(t->arena);
    delete t;
    return r;
  }
  char *p = (char *)t->arena;
  auto take = [&](size_t bytes) { void *r = p; p += (bytes + 255) / 256 * 256; return r; };
  t->o = take(N2 * ts * 12); t->par = take(N2 * ts * 12); t->gpar = take(N2 * ts * 12); t->coefH = take(N2 * ts * 12);
  for (int k = 0; k < 3; k++) { t->mid[k] = take(N2 * ts * 4); t->gmid[k] = take(N2 * ts * 4); }
  t->a0 = take(N * ts * 4); t->fine = take(N * ts * 4); t->X = take(N * ts * 4); t->prev = take(N * ts * 4);
  t->gX = take(N * ts * 4); t->gfine = take(N * ts * 4); t->gill = take(N * ts * 4); t->gpf = take(N * ts * 4);
  t->taac = take(N * ts * 12); t->coefF = take(N * ts * 12);
  t->oprev = take(N2 * ts * 12); t->pprev = take(N2 * ts * 12); t->gpprev = take(N2 * ts * 12); t->ga0 = take(N * ts * 4);
  t->aux = (float *)take(N * 4 * TEMPORAL_AUX);
  t->bacc = (long long *)take(N2 * 2 * 8);
  t->bmax = (unsigned long long *)take(((N + NTH - 1) / NTH + 1) * 8);
  t->bterm = take(N * ts * 2); t->bq = (int *)take(N * 4);
  for (int k = 0; k < 3; k++) t->Fprev[k] = (uint4 *)take(N2 * 16);
  t->GHprev = (float4 *)take(N2 * 16);
  t->wpart = (double *)take(((size_t)2 * t->nwb + WSL) * NW * sizeof(double));   // partials, then the slice sums
  t->lpart = (double *)take((size_t)t->nlb * sizeof(double));
  *out = t;
  return HUSH_OK;
}

extern "C" void hush_train_destroy(hush_train_t *t) {
  if (!t) return;
  (void)hipSetDevice(t->h->device);
  if (t->arena) (void)hipFree(t->arena);
  if (t->wq) (void)hipFree(t->wq);
  delete t;
}

static int check_size(hush_train_t *t) {
  if (t->h->W != t->W || t->h->H != t->H) return tfail(t, HUSH_ERR_SIZE, "hush_train: the denoiser was resized; create a new training context");
  return HUSH_OK;
}

extern "C" int32_t hush_train_prepare(hush_train_t *t, const hush_inputs *in, void *stream) {
  if (!t || !in) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_prepare: null argument");
  int r = check_size(t);
  if (r != HUSH_OK) return r;
  hush_t *h = t->h;
  if (hush_check_inputs(h, in) != HUSH_OK) return tfail(t, HUSH_ERR_INPUT, hush_last_error(h));
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(h->device);
  const int passes[3] = {0, 4, 5};
  h->train_aux = t->aux;
  for (int k = 0; k < 3 && e == hipSuccess; k++) e = hush_run_pass(h, passes[k], in, s);
  h->train_aux = nullptr;
  t->prev_mode = h->reset ? 0 : (t->has_prev ? 2 : 1);
  if (e == hipSuccess && t->prev_mode == 1) {     // history from outside this context: alpha / gamma as stored
    const Tex p0 = h->tex[S_PAR0_0 + (int)((h->frame & 1) ^ 1)];
    if (t->fp64) hipLaunchKernelGGL(tk_par_from_tex<double>, grid1((long)t->W2 * t->H2), dim3(NTH), 0, s, p0, t->W2, t->H2, (double *)t->pprev);
    else hipLaunchKernelGGL(tk_par_from_tex<float>, grid1((long)t->W2 * t->H2), dim3(NTH), 0, s, p0, t->W2, t->H2, (float *)t->pprev);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    if (t->fp64) hipLaunchKernelGGL(tk_from_tex<double>, grid1((long)t->W * t->H), dim3(NTH), 0, s, h->tex[S_A0], t->W, t->H, (double *)t->a0);
    else hipLaunchKernelGGL(tk_from_tex<float>, grid1((long)t->W * t->H), dim3(NTH), 0, s, h->tex[S_A0], t->W, t->H, (float *)t->a0);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return tfail(t, HUSH_ERR_DEVICE, "hush_train_prepare", e);
  t->in = *in;
  t->reset = h->reset;
  t->prepared = true;
  t->forwarded = false;
  return HUSH_OK;
}

template <class T>
static FiltArgs<T> filt_args(hush_train_t *t, int level, const T *gout, T *gsrc) {
  // level 0: fine (full res, source A0, guide KD, params at p >> 1); 1..3: mid levels (half res, guide GH)
  hush_t *h = t->h;
  FiltArgs<T> a{};
  const int cur = (int)(h->frame & 1);
  a.par = (const T *)t->par; a.pw = t->W2;
  a.ks = 4 + level; a.kc = 8 + level;
  if (level == 0) {
    a.src = (const T *)t->a0; a.guide = h->kd[cur]; a.w = t->W; a.h = t->H; a.out = (T *)t->fine;
    a.gp = (T *)t->gpf; a.gps = 4; a.js = 2; a.jc = 3; a.coef = (T *)t->coefF;
  } else {
    a.F2 = h->F[2];
    a.src = level == 1 ? nullptr : (const T *)t->mid[level - 2];
    a.guide = h->gh; a.w = t->W2; a.h = t->H2; a.out = (T *)t->mid[level - 1];
    a.gp = (T *)t->gpar; a.gps = 12; a.js = 4 + level; a.jc = 8 + level; a.coef = (T *)t->coefH;
  }
  a.gout = gout;
  a.gsrc = gsrc;
  return a;
}

static TexU fprev_tex(hush_train_t *t, int k) { return TexU{t->Fprev[k], t->W2}; }

template <class T>
static hipError_t forward_t(hush_train_t *t, const T *w, T *out, hipStream_t s) {
  hush_t *h = t->h;
  const long N = (long)t->W * t->H, N2 = (long)t->W2 * t->H2;
  // MFMA forward option (FP32 contexts only): the weights are packed for the matrix cores on every forward,
  // the MLP runs over the feature surfaces in place, the activations follow; the backward stays FP32
  const bool mf = sizeof(T) == sizeof(float) && t->fwd_prec != HUSH_MLP_FP32;
  if (mf) {
    const float one[3] = {1.0f, 1.0f, 1.0f};      // the binary16 / bfloat16 forms use no scales
    hipError_t e = launch_pack_mlp((const float *)w, (MfmaPack *)t->wq, one, one, s);
    if (e != hipSuccess) return e;
  }
  auto mlp = [&](TexU F0, TexU F1, TexU F2, TexF GH, void *o, void *par) -> hipError_t {
    if (!mf) {
      hipLaunchKernelGGL(tk_infer<T>, grid1(N2), dim3(NTH), 0, s, F0, F1, F2, GH, t->W2, t->H2, w, (T *)o, (T *)par);
      return hipGetLastError();
    }
    hipError_t e = launch_mlp_tex(t->fwd_prec, F0, F1, F2, t->W2, t->H2, (float *)o, t->wq, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(tk_act<T>, grid1(N2), dim3(NTH), 0, s, GH, t->W2, t->H2, (T *)o, (T *)par);
    return hipGetLastError();
  };
  if (!t->a0_fixed) {
    // the previous frame's alpha / gamma with these weights (one step back through the temporal blend)
    if (t->prev_mode == 2) {
      hipError_t e = mlp(fprev_tex(t, 0), fprev_tex(t, 1), fprev_tex(t, 2), TexF{t->GHprev, t->W2}, t->oprev, t->pprev);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(tk_blend<T>, grid1(N), dim3(NTH), 0, s, (const float *)t->aux, (const T *)t->pprev, (int)N, (T *)t->a0);
  }
  {
    hipError_t e = mlp(h->F[0], h->F[1], h->F[2], h->gh, t->o, t->par);
    if (e != hipSuccess) return e;
  }
  // mid1 (R 1, step 1, prefilter, source = block means), mid2 (step 2), mid3 (step 4) (hush.js:366, 753-754)
  hipLaunchKernelGGL((tk_filt<T, 1, 1, true, 0, false, false>), grid1(N2), dim3(NTH), 0, s, filt_args<T>(t, 1, nullptr, nullptr));
  hipLaunchKernelGGL((tk_filt<T, 1, 2, false, 1, false, false>), grid1(N2), dim3(NTH), 0, s, filt_args<T>(t, 2, nullptr, nullptr));
  hipLaunchKernelGGL((tk_filt<T, 1, 4, false, 1, false, false>), grid1(N2), dim3(NTH), 0, s, filt_args<T>(t, 3, nullptr, nullptr));
  // fine (R 1, step 1, prefilter, source A0, params at p >> 1) (hush.js:569-571)
  hipLaunchKernelGGL((tk_filt<T, 1, 1, true, 1, true, false>), grid1(N), dim3(NTH), 0, s, filt_args<T>(t, 0, nullptr, nullptr));
  FinalArgs<T> fa{};
  fa.fine = (const T *)t->fine; fa.par = (const T *)t->par; fa.mid = (const T *)t->mid[2];
  fa.kd = h->kd[(int)(h->frame & 1)]; fa.GH = h->gh; fa.cgd = h->cgd; fa.F2 = h->F[2];
  fa.coarse = h->tex[S_C0];
  fa.mat = Tex{(uint2 *)t->in.material.ptr, (int)(t->in.material.pitch / 8)};
  fa.emi = Tex{(uint2 *)t->in.emission.ptr, (int)(t->in.emission.pitch / 8)};
  fa.W = t->W; fa.H = t->H; fa.W2 = t->W2; fa.H2 = t->H2; fa.W4 = t->W4; fa.H4 = t->H4;
  fa.X = (T *)t->X;
  hipLaunchKernelGGL((tk_final<T, false>), grid1(N), dim3(NTH), 0, s, fa);
  TaaTArgs<T> ta{};
  ta.X = (const T *)t->X; ta.prev = (const T *)t->prev;
  ta.mv = Tex{(uint2 *)t->in.motion.ptr, (int)(t->in.motion.pitch / 8)};
  ta.W = t->W; ta.H = t->H; ta.reset = t->reset ? 1 : 0; ta.out = out;
  hipLaunchKernelGGL((tk_taa<T, false>), grid1(N), dim3(NTH), 0, s, ta);
  return hipGetLastError();
}

extern "C" int32_t hush_train_set_forward_precision(hush_train_t *t, int32_t precision) {
  if (!t) return tfail(nullptr, HUSH_ERR_ARGUMENT, "hush_train_set_forward_precision: null context");
  if (precision != HUSH_MLP_FP32 && precision != HUSH_MLP_FP16_MFMA && precision != HUSH_MLP_BF16_MFMA)
    return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_set_forward_precision: the training forward takes fp32, fp16-mfma or bf16-mfma");
  if (precision != HUSH_MLP_FP32 && t->fp64)
    return tfail(t, HUSH_ERR_UNSUPPORTED, "hush_train_set_forward_precision: the MFMA forward needs an FP32 context");
  if (precision != HUSH_MLP_FP32 && !t->wq) {
    hipError_t e = hipSetDevice(t->h->device);
    if (e == hipSuccess) e = hipMalloc(&t->wq, sizeof(MfmaPack));
    if (e != hipSuccess) { t->wq = nullptr; return tfail(t, HUSH_ERR_DEVICE, "hush_train_set_forward_precision: hipMalloc", e); }
  }
  t->fwd_prec = precision;
  return HUSH_OK;
}

extern "C" int32_t hush_train_forward(hush_train_t *t, const void *weights, void *out, void *stream) {
  if (!t || !weights || !out) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_forward: null argument");
  if (!t->prepared) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_forward: call hush_train_prepare first");
  int r = check_size(t);
  if (r != HUSH_OK) return r;
  hipError_t e = hipSetDevice(t->h->device);
  if (e == hipSuccess)
    e = t->fp64 ? forward_t<double>(t, (const double *)weights, (double *)out, (hipStream_t)stream)
                : forward_t<float>(t, (const float *)weights, (float *)out, (hipStream_t)stream);
  if (e != hipSuccess) return tfail(t, HUSH_ERR_DEVICE, "hush_train_forward", e);
  t->forwarded = true;
  return HUSH_OK;
}

template <class T>
static hipError_t loss_t(hush_train_t *t, const T *out, const T *tgt, const T *pout, const T *ptgt, double lambda, T *gout,
                         double *loss, hipStream_t s) {
  const int n = t->W * t->H;
  hipLaunchKernelGGL(tk_loss<T>, dim3(t->nlb), dim3(NTH), 0, s, out, tgt, pout, ptgt, (T)lambda, n, gout, t->lpart);
  hipLaunchKernelGGL(tk_sum, dim3(1), dim3(NTH), 0, s, (const double *)t->lpart, t->nlb, 1.0 / n, loss);
  return hipGetLastError();
}

extern "C" int32_t hush_train_loss(hush_train_t *t, const void *out, const void *target, const void *prev_out,
                                   const void *prev_target, double lambda, void *grad_out, double *loss_dev, void *stream) {
  if (!t || !out || !target || !grad_out || !loss_dev) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_loss: null argument");
  if ((prev_out == nullptr) != (prev_target == nullptr)) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_loss: give both prev_out and prev_target, or neither");
  int r = check_size(t);
  if (r != HUSH_OK) return r;
  hipError_t e = hipSetDevice(t->h->device);
  if (e == hipSuccess)
    e = t->fp64 ? loss_t<double>(t, (const double *)out, (const double *)target, (const double *)prev_out, (const double *)prev_target,
                                 lambda, (double *)grad_out, loss_dev, (hipStream_t)stream)
                : loss_t<float>(t, (const float *)out, (const float *)target, (const float *)prev_out, (const float *)prev_target,
                                lambda, (float *)grad_out, loss_dev, (hipStream_t)stream);
  if (e != hipSuccess) return tfail(t, HUSH_ERR_DEVICE, "hush_train_loss", e);
  return HUSH_OK;
}

// ---- the loss without a context and without allocation (hush_loss_noalloc)
// The per-workgroup partial sums of tk_loss live in this file-scope array; each GPU holds its own copy, whose
// device address hipGetSymbolAddress gives once per device.  tk_loss writes every partial it launches and
// tk_sum reads exactly those, so nothing needs clearing between calls; calls that overlap on one GPU share
// the array and race.
constexpr int NOALLOC_PARTS = 65536;                     // frames up to 65536 * 256 = 16,777,216 pixels
static __device__ __align__(256) double scratch[NOALLOC_PARTS];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[64];                         // per device ordinal

static int scratch_ptr(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return tfail(nullptr, HUSH_ERR_DEVICE, "hush_loss_noalloc: hipGetDevice", e);
  if (dev < 0 || dev >= 64) return tfail(nullptr, HUSH_ERR_UNSUPPORTED, "hush_loss_noalloc: device ordinals above 63 are not supported");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));   // with `dev` current: this GPU's copy
    if (e != hipSuccess || !p) return tfail(nullptr, HUSH_ERR_DEVICE, "hush_loss_noalloc: hipGetSymbolAddress of the scratch array", e);
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return HUSH_OK;
}

extern "C" int32_t hush_loss_noalloc(int32_t fp64, int32_t width, int32_t height, const void *out, const void *target,
                                     const void *prev_out, const void *prev_target, double lambda, void *grad_out,
                                     double *loss_dev, void *stream) {
  if (!out || !target || !grad_out || !loss_dev) return tfail(nullptr, HUSH_ERR_ARGUMENT, "hush_loss_noalloc: null argument");
  if ((prev_out == nullptr) != (prev_target == nullptr)) return tfail(nullptr, HUSH_ERR_ARGUMENT, "hush_loss_noalloc: give both prev_out and prev_target, or neither");
  if (!(width > 0 && height > 0)) return tfail(nullptr, HUSH_ERR_SIZE, "hush_loss_noalloc: non-positive size");
  const long n = (long)width * height, nb = (n + NTH - 1) / NTH;
  if (nb > NOALLOC_PARTS) {
    std::snprintf(g_terr, sizeof g_terr, "hush_loss_noalloc: %dx%d exceeds the static scratch (at most %ld pixels)",
                  width, height, (long)NOALLOC_PARTS * NTH);
    return HUSH_ERR_SIZE;
  }
  double *part = nullptr;
  int r = scratch_ptr(&part);
  if (r != HUSH_OK) return r;
  hipStream_t s = (hipStream_t)stream;
  if (fp64)
    hipLaunchKernelGGL(tk_loss<double>, dim3((unsigned)nb), dim3(NTH), 0, s, (const double *)out, (const double *)target,
                       (const double *)prev_out, (const double *)prev_target, lambda, (int)n, (double *)grad_out, part);
  else
    hipLaunchKernelGGL(tk_loss<float>, dim3((unsigned)nb), dim3(NTH), 0, s, (const float *)out, (const float *)target,
                       (const float *)prev_out, (const float *)prev_target, (float)lambda, (int)n, (float *)grad_out, part);
  hipLaunchKernelGGL(tk_sum, dim3(1), dim3(NTH), 0, s, (const double *)part, (int)nb, 1.0 / (double)n, loss_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tfail(nullptr, HUSH_ERR_DEVICE, "hush_loss_noalloc", e);
  return HUSH_OK;
}

// the MLP weight gradient's workgroup partials: FP64 contexts one per 128 texels, FP32 contexts one per
// MT x 128 texels on the matrix cores (t->nwb of them either way)
template <class T>
static void launch_mlp_bwd(hush_train_t *t, TexU F0, TexU F1, TexU F2, TexF GH, const T *w, const T *o, const T *par,
                           const T *gpar, double *part, hipStream_t s) {
  if constexpr (sizeof(T) == 8)
    hipLaunchKernelGGL(tk_mlp_bwd64, dim3(t->nwb), dim3(MB), 0, s, F0, F1, F2, GH, t->W2, t->H2, w, o, par, gpar, part);
  else
    hipLaunchKernelGGL(tk_mlp_bwd_mfma, dim3(t->nwb), dim3(MB), 0, s, F0, F1, F2, GH, t->W2, t->H2, w, o, par, gpar, part);
}

template <class T>
static hipError_t backward_t(hush_train_t *t, const T *w, const T *gout, T *gw, T *ga0, hipStream_t s) {
  hush_t *h = t->h;
  const long N = (long)t->W * t->H, N2 = (long)t->W2 * t->H2;
  TaaTArgs<T> ta{};
  ta.X = (const T *)t->X; ta.prev = (const T *)t->prev;
  ta.mv = Tex{(uint2 *)t->in.motion.ptr, (int)(t->in.motion.pitch / 8)};
  ta.W = t->W; ta.H = t->H; ta.reset = t->reset ? 1 : 0;
  ta.gout = gout; ta.coef = (T *)t->taac; ta.gX = (T *)t->gX;
  hipLaunchKernelGGL((tk_taa<T, true>), grid1(N), dim3(NTH), 0, s, ta);
  hipLaunchKernelGGL(tk_taa_gather<T>, grid1(N), dim3(NTH), 0, s, ta);
  FinalArgs<T> fa{};
  fa.fine = (const T *)t->fine; fa.par = (const T *)t->par; fa.mid = (const T *)t->mid[2];
  fa.kd = h->kd[(int)(h->frame & 1)]; fa.GH = h->gh; fa.cgd = h->cgd; fa.F2 = h->F[2];
  fa.coarse = h->tex[S_C0];
  fa.mat = Tex{(uint2 *)t->in.material.ptr, (int)(t->in.material.pitch / 8)};
  fa.emi = Tex{(uint2 *)t->in.emission.ptr, (int)(t->in.emission.pitch / 8)};
  fa.W = t->W; fa.H = t->H; fa.W2 = t->W2; fa.H2 = t->H2; fa.W4 = t->W4; fa.H4 = t->H4;
  fa.gX = (const T *)t->gX; fa.gfine = (T *)t->gfine; fa.gill = (T *)t->gill; fa.gpf = (T *)t->gpf;
  fa.gmid = (T *)t->gmid[2]; fa.gpar = (T *)t->gpar;
  hipLaunchKernelGGL((tk_final<T, true>), grid1(N), dim3(NTH), 0, s, fa);
  // fine level: dL/ds0, dL/dc0 per pixel (into gpf.zw), optionally dL/dA0
  const bool chain = !t->a0_fixed && t->prev_mode == 2;
  T *ga = ga0 ? ga0 : (T *)t->ga0;
  hipLaunchKernelGGL((tk_filt<T, 1, 1, true, 1, true, true>), grid1(N), dim3(NTH), 0, s, filt_args<T>(t, 0, (const T *)t->gfine, ga));
  if (ga0 || chain) hipLaunchKernelGGL((tk_filt_gather<T, 1, 1, true, 1>), grid1(N), dim3(NTH), 0, s, filt_args<T>(t, 0, (const T *)t->gfine, ga));
  hipLaunchKernelGGL(tk_final_gather<T>, grid1(N2), dim3(NTH), 0, s, fa);
  // mid3 -> mid2 -> mid1
  hipLaunchKernelGGL((tk_filt<T, 1, 4, false, 1, false, true>), grid1(N2), dim3(NTH), 0, s, filt_args<T>(t, 3, (const T *)t->gmid[2], (T *)t->gmid[1]));
  hipLaunchKernelGGL((tk_filt_gather<T, 1, 4, false, 1>), grid1(N2), dim3(NTH), 0, s, filt_args<T>(t, 3, (const T *)t->gmid[2], (T *)t->gmid[1]));
  hipLaunchKernelGGL((tk_filt<T, 1, 2, false, 1, false, true>), grid1(N2), dim3(NTH), 0, s, filt_args<T>(t, 2, (const T *)t->gmid[1], (T *)t->gmid[0]));
  hipLaunchKernelGGL((tk_filt_gather<T, 1, 2, false, 1>), grid1(N2), dim3(NTH), 0, s, filt_args<T>(t, 2, (const T *)t->gmid[1], (T *)t->gmid[0]));
  hipLaunchKernelGGL((tk_filt<T, 1, 1, true, 0, false, true>), grid1(N2), dim3(NTH), 0, s, filt_args<T>(t, 1, (const T *)t->gmid[0], nullptr));
  launch_mlp_bwd<T>(t, h->F[0], h->F[1], h->F[2], h->gh, w, (const T *)t->o, (const T *)t->par, (const T *)t->gpar, t->wpart, s);
  int nb = t->nwb;
  if (chain) {
    // temporal blend -> the previous frame's alpha / gamma -> its MLP (features of that frame)
    hipLaunchKernelGGL(tk_zero<long long>, grid1(N2 * 2), dim3(NTH), 0, s, t->bacc, (long)(N2 * 2));
    const int nblk = (int)grid1(N).x;
    unsigned long long *mxp = t->bmax + nblk;          // the maximum, after the nblk workgroup partials
    hipLaunchKernelGGL(tk_blend_bwd_max<T>, grid1(N), dim3(NTH), 0, s, (const float *)t->aux, (const T *)t->pprev, (const T *)ga,
                       (int)N, (T *)t->bterm, t->bq, t->bmax);
    hipLaunchKernelGGL(tk_max_bits, dim3(1), dim3(NTH), 0, s, t->bmax, nblk);
    hipLaunchKernelGGL(tk_blend_bwd_add<T>, grid1(N), dim3(NTH), 0, s, (const T *)t->bterm, (const int *)t->bq, (int)N,
                       (const unsigned long long *)mxp, t->bacc);
    hipLaunchKernelGGL(tk_blend_bwd_store<T>, grid1(N2), dim3(NTH), 0, s, (const long long *)t->bacc,
                       (const unsigned long long *)mxp, (int)N, (int)N2, (T *)t->gpprev);
    launch_mlp_bwd<T>(t, fprev_tex(t, 0), fprev_tex(t, 1), fprev_tex(t, 2), TexF{t->GHprev, t->W2}, w, (const T *)t->oprev,
                      (const T *)t->pprev, (const T *)t->gpprev, t->wpart + (size_t)t->nwb * NW, s);
    nb *= 2;
  }
  double *mid = t->wpart + (size_t)2 * t->nwb * NW;
  hipLaunchKernelGGL(tk_wsum1, dim3((NW + 63) / 64, WSL), dim3(NTH), 0, s, (const double *)t->wpart, nb, mid);
  hipLaunchKernelGGL(tk_wsum2<T>, dim3((NW + 63) / 64), dim3(NTH), 0, s, (const double *)mid, gw);
  return hipGetLastError();
}

extern "C" int32_t hush_train_backward(hush_train_t *t, const void *weights, const void *grad_out, void *grad_w, void *grad_a0,
                                       void *stream) {
  if (!t || !weights || !grad_out || !grad_w) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_backward: null argument");
  if (!t->forwarded) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_backward: no forward to differentiate");
  int r = check_size(t);
  if (r != HUSH_OK) return r;
  hipError_t e = hipSetDevice(t->h->device);
  if (e == hipSuccess)
    e = t->fp64 ? backward_t<double>(t, (const double *)weights, (const double *)grad_out, (double *)grad_w, (double *)grad_a0, (hipStream_t)stream)
                : backward_t<float>(t, (const float *)weights, (const float *)grad_out, (float *)grad_w, (float *)grad_a0, (hipStream_t)stream);
  if (e != hipSuccess) return tfail(t, HUSH_ERR_DEVICE, "hush_train_backward", e);
  return HUSH_OK;
}

template <class T>
static hipError_t commit_t(hush_train_t *t, const T *out, hipStream_t s) {
  hush_t *h = t->h;
  const int cur = (int)(h->frame & 1);
  const long N = (long)t->W * t->H, N2 = (long)t->W2 * t->H2;
  hipLaunchKernelGGL(tk_to_tex<T>, grid1(N), dim3(NTH), 0, s, (const T *)t->fine, 4, t->W, t->H, h->tex[S_HIST0 + cur]);
  hipLaunchKernelGGL(tk_to_tex<T>, grid1(N2), dim3(NTH), 0, s, (const T *)t->par, 12, t->W2, t->H2, h->tex[S_PAR0_0 + cur]);
  hipLaunchKernelGGL(tk_to_tex<T>, grid1(N), dim3(NTH), 0, s, out, 4, t->W, t->H, h->tex[S_TAA0 + cur]);
  hipLaunchKernelGGL(tk_copy<T>, grid1(N * 4), dim3(NTH), 0, s, out, (T *)t->prev, N * 4);
  for (int k = 0; k < 3; k++) hipLaunchKernelGGL(tk_copy_u4, grid1(N2), dim3(NTH), 0, s, h->F[k], t->W2, t->H2, t->Fprev[k]);
  hipLaunchKernelGGL(tk_copy_f4, grid1(N2), dim3(NTH), 0, s, h->gh, t->W2, t->H2, t->GHprev);
  return hipGetLastError();
}

extern "C" int32_t hush_train_commit(hush_train_t *t, const void *out, void *stream) {
  if (!t || !out) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_commit: null argument");
  if (!t->forwarded) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_commit: no forward to commit");
  // a resize between forward and commit leaves the denoiser's surfaces at another size (and pitch), and its
  // pending reset must stand: refuse before any launch
  int r = check_size(t);
  if (r != HUSH_OK) return r;
  hipError_t e = hipSetDevice(t->h->device);
  if (e == hipSuccess)
    e = t->fp64 ? commit_t<double>(t, (const double *)out, (hipStream_t)stream) : commit_t<float>(t, (const float *)out, (hipStream_t)stream);
  if (e != hipSuccess) return tfail(t, HUSH_ERR_DEVICE, "hush_train_commit", e);
  t->h->reset = false;
  t->h->frame++;
  t->prepared = t->forwarded = false;
  t->has_prev = true;
  return HUSH_OK;
}

extern "C" int32_t hush_train_set_a0_fixed(hush_train_t *t, int32_t fixed) {
  if (!t) return tfail(nullptr, HUSH_ERR_ARGUMENT, "hush_train_set_a0_fixed: null context");
  t->a0_fixed = fixed != 0;
  return HUSH_OK;
}

extern "C" int32_t hush_train_a0(hush_train_t *t, void **ptr) {
  if (!t || !ptr) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_a0: null argument");
  *ptr = t->a0;
  return HUSH_OK;
}

extern "C" int32_t hush_train_mlp_out(hush_train_t *t, void **ptr) {
  if (!t || !ptr) return tfail(t, HUSH_ERR_ARGUMENT, "hush_train_mlp_out: null argument");
  *ptr = t->o;
  return HUSH_OK;
}
