// SVGF baseline on gfx950: the four compute passes of src/denoise/svgf.js (temporal accumulation,
// variance estimate, edge-avoiding a-trous wavelet, TAA resolve) and the C ABI of include/svgf.h.
//
// Arithmetic mirrors the WGSL operation by operation (FMA contraction off for the whole library), with
// correctly rounded division and square root.  Two choices go beyond the reference:
//   - pow(x, sigmaN) with a positive integer exponent (the default 128) is evaluated by binary
//     exponentiation in FP32 (x^128 is seven squarings, hush.js's own pow128), the products the oracle's
//     FP32 mirror forms in the same order; other exponents use the library powf;
//   - the resolve's linear sampler (svgf.js:248, 261) is evaluated as exact bilinear weights at
//     uv * size - 0.5 with clamp-to-edge, the same formula as the oracle; a hardware sampler's sub-texel
//     weight precision is implementation-defined.
// Surfaces are rgba16float device arrays (8-byte texels); every store rounds toward zero (pack_rtz),
// the storage rounding the reference's textures are taken to have.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <type_traits>

#include "hush_device.h"
#include "svgf.h"

using namespace hush;

namespace {

constexpr int TW = 32, TH = 8, NT = 256;

struct SvTemporalArgs {
  Tex rad, nd, mv, mat, pnd, pmv, pcol, pmom;
  Tex outC, outM;
  int W, H, reset;
  float alpha_color, alpha_moments;
};

struct SvVarianceArgs {
  Tex col, mom, nd, out;
  int W, H, n_int;
  float sigma_n;
};

struct SvAtrousArgs {
  Tex in, nd, out, fb;
  int W, H, lg_step, n_int;
  int n_items;                       // row-phase tiling: regions (of TX x TY s pixels) x s row phases
  float sigma_l, sigma_n, sigma_z;
};

struct SvResolveArgs {
  Tex illum, mat, emi, mv, prev, out;
  int W, H, reset;
  float taa_alpha;
};

__device__ __forceinline__ int2 pixel_of(int W, int H) {
  const int2 o = tile_origin<TW, TH>(W, H);
  return make_int2(o.x + (int)(threadIdx.x & (TW - 1)), o.y + (int)(threadIdx.x / TW));
}

// pow(x, sigmaN) (svgf.js:137, 205); an integer exponent 1 <= n < 65536 runs as binary exponentiation in
// FP32 (x^128, the default, is seven squarings, the form of hush.js's own pow128), other exponents as powf.
// WGSL bounds pow's error only (that of exp2(y log2 x)); the squarings' error, at most ~(2n - 1) half-ulps,
// lies within it, and the oracle's FP32 mirror evaluates the same products in the same order.
__device__ __forceinline__ float sv_pow128(float x) {
#pragma unroll
  for (int k = 0; k < 7; k++) x = x * x;
  return x;
}
__device__ __forceinline__ float sv_pow(float x, int n, float y) {
  if (n > 0) {
    float b = x, r = 1.0f;
    for (int k = n;;) {
      if (k & 1) r = r * b;
      k >>= 1;
      if (!k) break;
      b = b * b;
    }
    return r;
  }
  return powf(x, y);
}

// a / b for a b shared by many quotients (rb = sv_rcp(b)): the shared-reciprocal quotient (Math<false>::divs),
// bit-identical to the IEEE one for a = 0, +-inf, NaN or 2^-102 <= |a| < inf, b = 0, +-inf, NaN or
// 2^-126 <= |b| < 2^126, and a normal quotient when both are finite and non-zero; sv_rcp takes the IEEE
// reciprocal outside the reduced-range reciprocal's domain.  The filters' finite numerators are differences
// of binary16 depths (0 or >= 2^-24) and of luminances of binary16 colours (0 or >= 2^-51); the variance
// pass's finite divisors lie in [1e-3, 5600] and the a-trous divisors in [1e-6, 2^82] while sigmaL and
// sigmaZ lie in [0, 2^60] (the host launches the IEEE quotient otherwise), so every such quotient is normal
// or zero.
__device__ __forceinline__ float sv_rcp(float b) {
  return __builtin_fabsf(b) < 0x1p126f ? Math<false>::rcpn(b) : 1.0f / b;
}
__device__ __forceinline__ float sv_div(float a, float b, float rb) { return Math<false>::divs(a, b, rb); }

// consistent (svgf.js:32-40) of an in-image texel from its previous normal / depth and object id
__device__ __forceinline__ bool consistent_at(float4 nd, float pobj, float3 n, float zexp, float obj) {
  if (nd.w < 0.0f) return false;
  if (pobj != obj) return false;
  float tol = 0.05f * zexp;
  tol = tol + 0.01f;
  if (fabsf(nd.w - zexp) > tol) return false;
  if (nd.x * n.x + nd.y * n.y + nd.z * n.z < 0.9f) return false;
  return true;
}
__device__ __forceinline__ bool consistent(const SvTemporalArgs &a, int qx, int qy, float3 n, float zexp, float obj) {
  if (!inside(qx, qy, a.W, a.H)) return false;
  return consistent_at(ld(a.pnd, qx, qy), ld(a.pmv, qx, qy).z, n, zexp, obj);
}

// TEMPORAL (svgf.js:42-105): bilinear reprojection with per-tap consistency, 3x3 fallback search,
// exponential moving average of colour and the first two luminance moments.  The pixel's own texels and the
// four bilinear taps' history texels (previous normal / depth and motion, colour, moments) are loaded at
// clamped addresses before any test, so their latencies overlap; a tap that fails a test, or lies outside
// the image, adds exact zeros (its texels are finite stored values), which leave every running sum unchanged
// bit for bit (each starts at +0).
__global__ __launch_bounds__(NT) void k_sv_temporal(SvTemporalArgs a) {
  const int2 p = pixel_of(a.W, a.H);
  if (p.x >= a.W || p.y >= a.H) return;
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint2 rnd = a.nd.p[(size_t)p.y * a.nd.pitch + p.x], rmv = a.mv.p[(size_t)p.y * a.mv.pitch + p.x];
  const uint2 ralb = a.mat.p[(size_t)p.y * a.mat.pitch + p.x], rrad = a.rad.p[(size_t)p.y * a.rad.pitch + p.x];
  const float4 nd = unpack4(rnd);
  if (nd.w < 0.0f) {
    st(a.outC, p.x, p.y, zero);
    st(a.outM, p.x, p.y, zero);
    return;
  }
  const float4 mv = unpack4(rmv), alb = unpack4(ralb), rad = unpack4(rrad);
  const float ppx = ((float)p.x + 0.5f + mv.x) - 0.5f, ppy = ((float)p.y + 0.5f + mv.y) - 0.5f;
  const int bx = f2i(floorf(ppx)), by = f2i(floorf(ppy));
  const bool nohist = a.reset != 0;
  uint2 tnd[4], tmv[4], tpc[4], tpm[4];
  if (!nohist) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int qx = clampi(addw(bx, i & 1), 0, a.W - 1), qy = clampi(addw(by, i >> 1), 0, a.H - 1);
      tnd[i] = a.pnd.p[(size_t)qy * a.pnd.pitch + qx];
      tmv[i] = a.pmv.p[(size_t)qy * a.pmv.pitch + qx];
      tpc[i] = a.pcol.p[(size_t)qy * a.pcol.pitch + qx];
      tpm[i] = a.pmom.p[(size_t)qy * a.pmom.pitch + qx];
    }
  }
  const float ir = rad.x / mx(alb.x, 1e-3f), ig = rad.y / mx(alb.y, 1e-3f), ib = rad.z / mx(alb.z, 1e-3f);
  const float l = lum(ir, ig, ib);
  const float fx = ppx - floorf(ppx), fy = ppy - floorf(ppy);
  const float3 n = make_float3(nd.x, nd.y, nd.z);
  float wts[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
  float sumW = 0.0f, hc0 = 0.0f, hc1 = 0.0f, hc2 = 0.0f, hm0 = 0.0f, hm1 = 0.0f, hm2 = 0.0f;
  if (!nohist) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int qx = addw(bx, i & 1), qy = addw(by, i >> 1);
      const bool ok = inside(qx, qy, a.W, a.H) && consistent_at(unpack4(tnd[i]), unpack4(tmv[i]).z, n, mv.w, mv.z) &&
                      wts[i] > 0.0f;
      const float w = ok ? wts[i] : 0.0f;
      const float4 pc = unpack4(tpc[i]), pm = unpack4(tpm[i]);
      sumW += w;
      hc0 += w * pc.x; hc1 += w * pc.y; hc2 += w * pc.z;
      hm0 += w * pm.x; hm1 += w * pm.y; hm2 += w * pm.z;
    }
    if (sumW < 0.01f) {
      const int cx = f2i(rintf(ppx)), cy = f2i(rintf(ppy));   // WGSL round(): half to even
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int qx = addw(cx, dx), qy = addw(cy, dy);
          if (consistent(a, qx, qy, n, mv.w, mv.z)) {
            sumW += 1.0f;
            const float4 pc = ld(a.pcol, qx, qy), pm = ld(a.pmom, qx, qy);
            hc0 += pc.x; hc1 += pc.y; hc2 += pc.z;
            hm0 += pm.x; hm1 += pm.y; hm2 += pm.z;
          }
        }
    }
  }
  float c0 = ir, c1 = ig, c2 = ib, m0 = l, m1 = l * l, len;
  if (sumW >= 0.01f) {
    hc0 /= sumW; hc1 /= sumW; hc2 /= sumW;
    hm0 /= sumW; hm1 /= sumW; hm2 /= sumW;
    len = mn(hm2 + 1.0f, 64.0f);
    const float ac = mx(a.alpha_color, 1.0f / len), am = mx(a.alpha_moments, 1.0f / len);
    c0 = hc0 * (1.0f - ac) + ir * ac;                     // mix(hc, illum, ac)
    c1 = hc1 * (1.0f - ac) + ig * ac;
    c2 = hc2 * (1.0f - ac) + ib * ac;
    const float n0 = hm0 * (1.0f - am) + m0 * am, n1 = hm1 * (1.0f - am) + m1 * am;
    m0 = n0; m1 = n1;
  } else {
    len = 1.0f;
  }
  st(a.outC, p.x, p.y, make_float4(c0, c1, c2, 0.0f));
  st(a.outM, p.x, p.y, make_float4(m0, m1, len, 0.0f));
}

// VARIANCE (svgf.js:114-148): temporal variance for long histories, a 7x7 depth / normal weighted
// spatial estimate (scaled by 4 / history length) for histories shorter than 4 frames.  A tile with any such
// pixel stages its 38x14 window once in LDS (normal and depth, colour, moments; unpacked, clamp-to-edge, the
// taps outside the image skipped); a tile without one skips the staging.  Each tap's arithmetic and the sum
// order are those of the WGSL loop.
template <bool P128>                 // sigmaN = 128
__global__ __launch_bounds__(NT) void k_sv_variance(SvVarianceArgs a) {
  constexpr int AP = 3, SW = TW + 2 * AP, SH = TH + 2 * AP, SN = SW * SH, NI = (SN + NT - 1) / NT;
  __shared__ float4 sN[SN], sC[SN];
  __shared__ float2 sM[SN];
  const int tid = (int)threadIdx.x;
  const int2 o = tile_origin<TW, TH>(a.W, a.H);
  const int lx = tid & (TW - 1), ly = tid / TW, px = o.x + lx, py = o.y + ly;
  const bool own = px < a.W && py < a.H;
  const int cx = mn(px, a.W - 1), cy = mn(py, a.H - 1);
  const float4 nd = ld(a.nd, cx, cy), c = ld(a.col, cx, cy), m = ld(a.mom, cx, cy);
  const bool spatial = own && !(nd.w < 0.0f) && !(m.z >= 4.0f);
  if (!__syncthreads_or(spatial)) {
    if (!own) return;
    if (nd.w < 0.0f) st(a.out, px, py, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    else st(a.out, px, py, make_float4(c.x, c.y, c.z, mx(0.0f, m.y - m.x * m.x)));
    return;
  }
  uint2 nr[NI], cr[NI], mr[NI];
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = mn(tid + k * NT, SN - 1);
    const int qx = clampi(o.x - AP + i % SW, 0, a.W - 1), qy = clampi(o.y - AP + i / SW, 0, a.H - 1);
    nr[k] = a.nd.p[(size_t)qy * a.nd.pitch + qx];
    cr[k] = a.col.p[(size_t)qy * a.col.pitch + qx];
    mr[k] = a.mom.p[(size_t)qy * a.mom.pitch + qx];
  }
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = tid + k * NT;
    if (i < SN) {
      sN[i] = unpack4(nr[k]);
      sC[i] = unpack4(cr[k]);
      const float4 mq = unpack4(mr[k]);
      sM[i] = make_float2(mq.x, mq.y);
    }
  }
  __syncthreads();
  if (!own) return;
  if (nd.w < 0.0f) { st(a.out, px, py, make_float4(0.0f, 0.0f, 0.0f, 0.0f)); return; }
  if (m.z >= 4.0f) {
    st(a.out, px, py, make_float4(c.x, c.y, c.z, mx(0.0f, m.y - m.x * m.x)));
    return;
  }
  // the depth term's divisor 0.02 z len + 1e-3, one per squared tap distance len^2 = dx^2 + dy^2 (0 .. 18)
  const float zs = 0.02f * nd.w;
  float zdiv[19], zrcp[19];
#pragma unroll
  for (int d2 = 0; d2 <= 18; d2++) {
    zdiv[d2] = zs * sqrtf((float)d2) + 1e-3f;
    zrcp[d2] = sv_rcp(zdiv[d2]);
  }
  float sw = 0.0f, s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sm0 = 0.0f, sm1 = 0.0f;
  const int cc = (ly + AP) * SW + lx + AP;
  // interior tiles (every tap of every pixel inside the image) skip the per-tap bounds tests
  const bool interior = o.x >= AP && o.y >= AP && o.x + TW + AP <= a.W && o.y + TH + AP <= a.H;
  auto taps = [&](auto in_tile) {
    constexpr bool CHK = !decltype(in_tile)::value;
#pragma unroll
    for (int dy = -3; dy <= 3; dy++) {
#pragma unroll
      for (int dx = -3; dx <= 3; dx++) {
        if (CHK && !inside(px + dx, py + dy, a.W, a.H)) continue;
        const int q = cc + dy * SW + dx;
        const float4 ndq = sN[q];
        if (ndq.w < 0.0f) continue;
        const int d2 = dx * dx + dy * dy;
        const float wz = Math<false>::expn(sv_div(-fabsf(ndq.w - nd.w), zdiv[d2], zrcp[d2]));
        const float dn = mx(0.0f, ndq.x * nd.x + ndq.y * nd.y + ndq.z * nd.z);
        const float wn = P128 ? sv_pow128(dn) : sv_pow(dn, a.n_int, a.sigma_n);
        const float w = wz * wn;
        sw += w;
        const float4 cq = sC[q];
        const float2 mq = sM[q];
        s0 += w * cq.x; s1 += w * cq.y; s2 += w * cq.z;
        sm0 += w * mq.x; sm1 += w * mq.y;
      }
    }
  };
  if (interior) taps(std::true_type{});
  else taps(std::false_type{});
  const float d = mx(sw, 1e-6f);
  s0 /= d; s1 /= d; s2 /= d; sm0 /= d; sm1 /= d;
  const float v = mx(0.0f, sm1 - sm0 * sm0) * 4.0f / mx(m.z, 1.0f);
  st(a.out, px, py, make_float4(s0, s1, s2, v));
}

// ATROUS (svgf.js:157-216): one 5x5 B3-spline wavelet iteration with taps `step` = 2^LG texels apart, depth /
// normal / luminance edge-stopping; iteration 0 also writes the colour feedback (the next frame's history).
// Row-phase tiling: the rows y = ry + s v of one phase ry in [0, s) hold every vertical tap of their pixels,
// so a workgroup filters TX contiguous columns x TY rows of one phase from a window of (TX + 4s) columns x
// (TY + 4) such rows staged once in LDS (colour and variance, luminance, normal and depth, unpacked; clamp-to-edge,
// the taps outside the image are skipped), in place of loading and unpacking every tap per pixel.  Window
// rows and the pixels' own +-1 neighbourhoods are contiguous runs of a row, so every load is coalesced.
// Workgroups walk regions of TX x TY s pixels x their s row phases, phases fastest, in one contiguous band per
// XCD (tile_origin's mapping), so the phases of one region share an L2.  Each tap's arithmetic and the sum
// order are those of the WGSL loop.
// SHARED: the edge-stopping quotients as shared-reciprocal ones (sigmaL and sigmaZ in [0, 2^60]), else IEEE
// tile of sv_tile_x(lg) x 256 / sv_tile_x(lg) pixels: wider and lower at the largest step, where the x apron
// dominates the window
constexpr int sv_tile_x(int lg) { return lg >= 4 ? 64 : 32; }
template <int LG> struct SvWin {
  static constexpr int S = 1 << LG, TX = sv_tile_x(LG), TY = NT / TX, AX = 2 * S, SW = TX + 4 * S, SH = TY + 4,
                       SN = SW * SH, NI = (SN + NT - 1) / NT;
};
template <bool FIRST, int LG, bool P128, bool SHARED>     // P128: sigmaN = 128
__global__ __launch_bounds__(NT) void k_sv_atrous(SvAtrousArgs a) {
  using Wn = SvWin<LG>;
  constexpr int S = Wn::S, TX = Wn::TX, TY = Wn::TY, AX = Wn::AX, SW = Wn::SW, SN = Wn::SN, NI = Wn::NI;
  __shared__ float4 sC[SN], sN[SN];
  __shared__ float sL[SN];
  auto quot = [](float x, float d, float rd) { return SHARED ? sv_div(x, d, rd) : x / d; };
  const int tid = (int)threadIdx.x;
  // workgroup -> (region, row phase)
  const int T = a.n_items, q = T >> 3, r = T & 7, xcd = (int)blockIdx.x & 7, n = (int)blockIdx.x >> 3;
  const int L = xcd * q + (xcd < r ? xcd : r) + n;
  const int region = L >> LG, ry = L & (S - 1);
  const int regions_x = (a.W + TX - 1) / TX;
  const int x0 = (region % regions_x) * TX, y0 = (region / regions_x) * (TY * S) + ry;
  // window texel (wx, wy) is image texel (x0 - 2s + wx, y0 + s (wy - 2))
  uint2 cr[NI], nr[NI];
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = mn(tid + k * NT, SN - 1);
    const int qx = clampi(x0 - AX + i % SW, 0, a.W - 1), qy = clampi(y0 + S * (i / SW - 2), 0, a.H - 1);
    cr[k] = a.in.p[(size_t)qy * a.in.pitch + qx];
    nr[k] = a.nd.p[(size_t)qy * a.nd.pitch + qx];
  }
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = tid + k * NT;
    if (i < SN) {
      const float4 c = unpack4(cr[k]);
      sC[i] = c;
      sL[i] = lum(c.x, c.y, c.z);
      sN[i] = unpack4(nr[k]);
    }
  }
  __syncthreads();
  const int lx = tid % TX, ly = tid / TX;
  const int px = x0 + lx, py = y0 + S * ly;
  if (px >= a.W || py >= a.H) return;
  const int c = (ly + 2) * SW + lx + AX;
  const float4 nd = sN[c];
  if (nd.w < 0.0f) {
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    st(a.out, px, py, zero);
    if (FIRST) st(a.fb, px, py, zero);
    return;
  }
  const float4 center = sC[c];
  const float k3[3] = {0.25f, 0.5f, 0.25f};
  float gv = 0.0f;
#pragma unroll
  for (int dy = -1; dy <= 1; dy++)
#pragma unroll
    for (int dx = -1; dx <= 1; dx++)
      gv += k3[dx + 1] * k3[dy + 1] * ld(a.in, clampi(px + dx, 0, a.W - 1), clampi(py + dy, 0, a.H - 1)).w;
  const float lp = sL[c];
  const float denomL = a.sigma_l * sqrtf(mx(gv, 0.0f)) + 1e-6f, rL = SHARED ? sv_rcp(denomL) : 0.0f;
  // zGrad (svgf.js:157-165): the pixel's own +-1 neighbours
  const float z = nd.w;
  const float zl = ld(a.nd, clampi(px - 1, 0, a.W - 1), py).w, zr = ld(a.nd, clampi(px + 1, 0, a.W - 1), py).w;
  const float zu = ld(a.nd, px, clampi(py - 1, 0, a.H - 1)).w, zd = ld(a.nd, px, clampi(py + 1, 0, a.H - 1)).w;
  const float gx = mn(zr > 0.0f ? fabsf(zr - z) : 1e9f, zl > 0.0f ? fabsf(zl - z) : 1e9f);
  const float gy = mn(zd > 0.0f ? fabsf(zd - z) : 1e9f, zu > 0.0f ? fabsf(zu - z) : 1e9f);
  const float zg = mn(mx(gx, gy), z);
  const float h[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
  const float zden = a.sigma_z * zg;
  // the depth term's divisor zden * dist + 1e-3, one per tap distance (dx^2 + dy^2 = 1, 2, 4, 5, 8)
  float zdiv[9], zrcp[9];
#pragma unroll
  for (int d2 = 1; d2 <= 8; d2++) {
    zdiv[d2] = zden * (sqrtf((float)d2) * (float)S) + 1e-3f;
    zrcp[d2] = SHARED ? sv_rcp(zdiv[d2]) : 0.0f;
  }
  float sw = h[2] * h[2];
  float s0 = center.x * sw, s1 = center.y * sw, s2 = center.z * sw;
  float sv = center.w * sw * sw;
  // interior tiles (every tap of every pixel inside the image) skip the per-tap bounds tests
  const bool interior = x0 >= AX && y0 >= 2 * S && x0 + TX + AX <= a.W && y0 + S * (TY + 1) < a.H;
  auto taps = [&](auto in_tile) {
    constexpr bool CHK = !decltype(in_tile)::value;
#pragma unroll
    for (int dy = -2; dy <= 2; dy++) {
#pragma unroll
      for (int dx = -2; dx <= 2; dx++) {
        if (dx == 0 && dy == 0) continue;
        if (CHK && !inside(px + dx * S, py + dy * S, a.W, a.H)) continue;
        const int o = c + dy * SW + dx * S;
        const float4 ndq = sN[o];
        if (ndq.w < 0.0f) continue;
        const float4 cq = sC[o];
        const int d2 = dx * dx + dy * dy;
        const float wz = quot(fabsf(ndq.w - z), zdiv[d2], zrcp[d2]);
        const float dn = mx(0.0f, ndq.x * nd.x + ndq.y * nd.y + ndq.z * nd.z);
        const float wn = P128 ? sv_pow128(dn) : sv_pow(dn, a.n_int, a.sigma_n);
        const float wl = quot(fabsf(sL[o] - lp), denomL, rL);
        // (P128 and SHARED: the exponent is finite and <= 0 and h h wn <= 3/32, so expw's product is expn's)
        const float x = -wz - wl;
        const float w = h[dx + 2] * h[dy + 2] * wn * (P128 && SHARED ? Math<false>::expw(x) : Math<false>::expn(x));
        sw += w;
        s0 += w * cq.x; s1 += w * cq.y; s2 += w * cq.z;
        sv += w * w * cq.w;
      }
    }
  };
  if (interior) taps(std::true_type{});
  else taps(std::false_type{});
  const float r0 = s0 / sw, r1 = s1 / sw, r2 = s2 / sw, r3 = sv / (sw * sw);
  st(a.out, px, py, make_float4(r0, r1, r2, r3));
  if (FIRST) st(a.fb, px, py, make_float4(r0, r1, r2, 0.0f));
}

// tm / itm (svgf.js:237-238) and the YCoCg transform (svgf.js:232-235).  The three quotients share one
// reciprocal (sv_div): tm's divisor 1 + lum(c) is 0 or normal (near lum = -1 the sum is exact, at least
// 2^-25 when not 0) or +-inf, itm's is at least 1e-4 or +inf; the numerators are 0, at least 2^-102 (the
// remodulated colour is a sum of products of binary16 values, a multiple of 2^-48, and its tone-mapped
// value a quotient of it by at most ~2^33) or non-finite, so every quotient is the IEEE one.
__device__ __forceinline__ float3 tm3(float3 c) {
  const float d = 1.0f + lum(c.x, c.y, c.z), r = sv_rcp(d);
  return make_float3(sv_div(c.x, d, r), sv_div(c.y, d, r), sv_div(c.z, d, r));
}
__device__ __forceinline__ float3 itm3(float3 c) {
  const float d = mx(1.0f - lum(c.x, c.y, c.z), 1e-4f), r = sv_rcp(d);
  return make_float3(sv_div(c.x, d, r), sv_div(c.y, d, r), sv_div(c.z, d, r));
}
__device__ __forceinline__ float3 ycocg(float3 c) {
  return make_float3(c.x * 0.25f + c.y * 0.5f + c.z * 0.25f, c.x * 0.5f - c.z * 0.5f, -c.x * 0.25f + c.y * 0.5f - c.z * 0.25f);
}
// remod (svgf.js:228-231): demodulated illumination back to radiance, clamp-to-edge
__device__ __forceinline__ float3 remod(const SvResolveArgs &a, int x, int y) {
  const int qx = clampi(x, 0, a.W - 1), qy = clampi(y, 0, a.H - 1);
  const float4 il = ld(a.illum, qx, qy), m = ld(a.mat, qx, qy), e = ld(a.emi, qx, qy);
  return make_float3(il.x * m.x + e.x, il.y * m.y + e.y, il.z * m.z + e.z);
}

// RESOLVE (svgf.js:240-270): remodulation, TAA with YCoCg variance clipping (gamma 1.25) in the
// tone-mapped domain.  tm(remod(q)) of the tile plus a one-texel apron is computed once into LDS.
template <bool TAA>
__global__ __launch_bounds__(NT) void k_sv_resolve(SvResolveArgs a) {
  constexpr int LW = TW + 2, LH = TH + 2;
  __shared__ float4 s_t[LH][LW];     // tm(remod(q)), w unused
  __shared__ float4 s_y[LH][LW];     // its YCoCg
  const int2 o = tile_origin<TW, TH>(a.W, a.H);
  const int tx = (int)(threadIdx.x & (TW - 1)), ty = (int)(threadIdx.x / TW);
  const int x = o.x + tx, y = o.y + ty;
  if (!TAA) {
    if (x >= a.W || y >= a.H) return;
    const float3 out = itm3(tm3(remod(a, x, y)));
    st(a.out, x, y, make_float4(out.x, out.y, out.z, 1.0f));
    return;
  }
  // the motion vector, then the four history taps it addresses (clamped, loaded whether or not they are
  // used), issued before the tile staging so that their latency overlaps it
  const int cx = mn(x, a.W - 1), cy = mn(y, a.H - 1);
  const float4 mv = ld(a.mv, cx, cy);
  const float ux = ((float)x + 0.5f + mv.x) / (float)a.W, uy = ((float)y + 0.5f + mv.y) / (float)a.H;
  const float hpx = ux * (float)a.W - 0.5f, hpy = uy * (float)a.H - 0.5f;
  const float hbx = floorf(hpx), hby = floorf(hpy);
  uint2 hr[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int qx = clampi(addw(f2i(hbx), j & 1), 0, a.W - 1), qy = clampi(addw(f2i(hby), j >> 1), 0, a.H - 1);
    hr[j] = a.prev.p[(size_t)qy * a.prev.pitch + qx];
  }
  constexpr int NI = (LW * LH + NT - 1) / NT;
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = threadIdx.x + k * NT;
    if (i < LW * LH) {
      const int ly = i / LW, lx = i - ly * LW;
      const float3 t = tm3(remod(a, o.x + lx - 1, o.y + ly - 1));
      const float3 c = ycocg(t);
      s_t[ly][lx] = make_float4(t.x, t.y, t.z, 0.0f);
      s_y[ly][lx] = make_float4(c.x, c.y, c.z, 0.0f);
    }
  }
  __syncthreads();
  if (x >= a.W || y >= a.H) return;
  const float4 cu4 = s_t[ty + 1][tx + 1];
  const float3 cur = make_float3(cu4.x, cu4.y, cu4.z);
  float m10 = 0.0f, m11 = 0.0f, m12 = 0.0f, m20 = 0.0f, m21 = 0.0f, m22 = 0.0f;
#pragma unroll
  for (int dy = -1; dy <= 1; dy++)
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
      const float4 c = s_y[ty + 1 + dy][tx + 1 + dx];
      m10 += c.x; m11 += c.y; m12 += c.z;
      m20 += c.x * c.x; m21 += c.y * c.y; m22 += c.z * c.z;
    }
  m10 /= 9.0f; m11 /= 9.0f; m12 /= 9.0f; m20 /= 9.0f; m21 /= 9.0f; m22 /= 9.0f;
  const float sd0 = sqrtf(mx(m20 - m10 * m10, 0.0f)), sd1 = sqrtf(mx(m21 - m11 * m11, 0.0f)), sd2 = sqrtf(mx(m22 - m12 * m12, 0.0f));
  float3 res = cur;
  if (ux >= 0.0f && uy >= 0.0f && ux <= 1.0f && uy <= 1.0f && !a.reset) {
    // linear sampler, clamp to edge
    const float fx = hpx - hbx, fy = hpy - hby;
    float h0 = 0.0f, h1 = 0.0f, h2 = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int ox = j & 1, oy = j >> 1;
      const float wb = (ox ? fx : 1.0f - fx) * (oy ? fy : 1.0f - fy);
      const float4 t = unpack4(hr[j]);
      h0 += wb * t.x; h1 += wb * t.y; h2 += wb * t.z;
    }
    float3 hh = ycocg(tm3(make_float3(h0, h1, h2)));
    hh.x = mn(mx(hh.x, m10 - 1.25f * sd0), m10 + 1.25f * sd0);
    hh.y = mn(mx(hh.y, m11 - 1.25f * sd1), m11 + 1.25f * sd1);
    hh.z = mn(mx(hh.z, m12 - 1.25f * sd2), m12 + 1.25f * sd2);
    const float f0 = hh.x + hh.y - hh.z, f1 = hh.x + hh.z, f2 = hh.x - hh.y - hh.z;
    const float al = a.taa_alpha;
    res = make_float3(f0 * (1.0f - al) + cur.x * al, f1 * (1.0f - al) + cur.y * al, f2 * (1.0f - al) + cur.z * al);
  }
  const float3 out = itm3(res);
  st(a.out, x, y, make_float4(out.x, out.y, out.z, 1.0f));
}

dim3 grid_of(int W, int H) { return dim3((unsigned)(((W + TW - 1) / TW) * ((H + TH - 1) / TH))); }

thread_local char g_err[512];

enum { SV_HC0, SV_HC1, SV_HM0, SV_HM1, SV_TMP, SV_AT0, SV_AT1, SV_TAA0, SV_TAA1, SV_COUNT };

const char *kInputName[7] = {"radiance", "normalDepth", "motion", "material", "emission", "prevNormalDepth", "prevMotion"};

// sigmaN as a positive integer below 2^16, else 0 (powf)
int int_exponent(float y) {
  if (!(y >= 1.0f && y < 65536.0f) || std::floor(y) != y) return 0;
  return (int)y;
}

}  // namespace

struct svgf_ctx {
  int device = 0;
  int W = 0, H = 0;
  int64_t frame = 0;
  bool reset = true;
  svgf_options o{};
  void *arena = nullptr;
  Tex tex[SV_COUNT] = {};
  bool timing = false, ev_ready = false;
  int ev_passes = 0;
  hipEvent_t ev[9] = {};
  char err[512] = {0};
};

static int fail(svgf_t *s, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
static int fail(svgf_t *s, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(s ? s->err : g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}
static int hip_fail(svgf_t *s, hipError_t e, const char *what) { return fail(s, HUSH_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e)); }

extern "C" void svgf_default_options(svgf_options *o) {
  if (!o) return;
  o->sigma_l = 4.0f; o->sigma_n = 128.0f; o->sigma_z = 1.0f; o->alpha_color = 0.2f; o->alpha_moments = 0.2f;
  o->taa_alpha = 0.1f; o->taa = 1; o->iterations = 5;
}

static int alloc_surfaces(svgf_t *s, int W, int H) {
  const int pitch = (W + 63) / 64 * 64;
  const size_t bytes = (size_t)pitch * H * 8;
  void *arena = nullptr;
  hipError_t e = hipMalloc(&arena, bytes * SV_COUNT);
  if (e != hipSuccess) return hip_fail(s, e, "hipMalloc");
  e = hipMemset(arena, 0, bytes * SV_COUNT);
  if (e != hipSuccess) { (void)hipFree(arena); return hip_fail(s, e, "hipMemset"); }
  if (s->arena) (void)hipFree(s->arena);
  s->arena = arena;
  for (int i = 0; i < SV_COUNT; i++) s->tex[i] = Tex{(uint2 *)((char *)arena + bytes * i), pitch};
  s->W = W; s->H = H;
  return HUSH_OK;
}

extern "C" int32_t svgf_resize(svgf_t *s, int32_t w, int32_t h) {
  if (!s) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_resize: null context");
  if (w == s->W && h == s->H) return HUSH_OK;
  if (w <= 0 || h <= 0 || w > 16384 || h > 16384) return fail(s, HUSH_ERR_SIZE, "SVGF: bad size %dx%d", w, h);
  hipError_t e = hipSetDevice(s->device);
  if (e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
  int r = alloc_surfaces(s, w, h);
  if (r != HUSH_OK) return r;
  s->frame = 0;
  s->reset = true;
  return HUSH_OK;
}

extern "C" int32_t svgf_create(int32_t device, int32_t width, int32_t height, const svgf_options *opts, svgf_t **out) {
  if (!out) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_create: null argument");
  *out = nullptr;
  svgf_options o;
  svgf_default_options(&o);
  if (opts) o = *opts;
  if (o.iterations < 0 || o.iterations > 5)
    return fail(nullptr, HUSH_ERR_UNSUPPORTED, "SVGF: %d a-trous iterations (the reference has uniform buffers for 0..5)", o.iterations);
  if (width <= 0 || height <= 0 || width > 16384 || height > 16384) return fail(nullptr, HUSH_ERR_SIZE, "SVGF: bad size %dx%d", width, height);
  svgf_t *s = new (std::nothrow) svgf_ctx;
  if (!s) return fail(nullptr, HUSH_ERR_DEVICE, "svgf_create: out of host memory");
  s->device = device;
  s->o = o;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) { delete s; return fail(nullptr, HUSH_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e)); }
  int r = alloc_surfaces(s, width, height);
  if (r != HUSH_OK) { std::snprintf(g_err, sizeof(g_err), "%s", s->err); delete s; return r; }
  *out = s;
  return HUSH_OK;
}

extern "C" void svgf_destroy(svgf_t *s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->arena) (void)hipFree(s->arena);
  for (auto &ev : s->ev)
    if (ev) (void)hipEventDestroy(ev);
  delete s;
}

extern "C" int32_t svgf_reset(svgf_t *s) {
  if (!s) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_reset: null context");
  s->reset = true;
  return HUSH_OK;
}

extern "C" int32_t svgf_get_size(const svgf_t *s, int32_t *w, int32_t *h) {
  if (!s) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_get_size: null context");
  if (w) *w = s->W;
  if (h) *h = s->H;
  return HUSH_OK;
}
extern "C" int64_t svgf_frame(const svgf_t *s) { return s ? s->frame : -1; }
extern "C" int32_t svgf_get_options(const svgf_t *s, svgf_options *o) {
  if (!s || !o) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_get_options: null argument");
  *o = s->o;
  return HUSH_OK;
}
extern "C" const char *svgf_last_error(const svgf_t *s) { return s ? s->err : g_err; }

static hush_surface surf(const svgf_t *s, int id) {
  hush_surface r{};
  r.ptr = s->tex[id].p; r.pitch = (size_t)s->tex[id].pitch * 8; r.width = s->W; r.height = s->H; r.format = HUSH_RGBA16F;
  return r;
}
static Tex in_tex(const hush_surface &s) { return Tex{(uint2 *)s.ptr, (int)(s.pitch / 8)}; }

static int check_inputs(svgf_t *s, const svgf_inputs *in) {
  const hush_surface *v[7] = {&in->radiance, &in->normal_depth, &in->motion, &in->material, &in->emission,
                              &in->prev_normal_depth, &in->prev_motion};
  for (int k = 0; k < 7; k++) {
    if (!v[k]->ptr) return fail(s, HUSH_ERR_INPUT, "SVGF.encode: missing input '%s'", kInputName[k]);
    if (v[k]->width != s->W || v[k]->height != s->H)
      return fail(s, HUSH_ERR_INPUT, "SVGF.encode: '%s' is %dx%d, the denoiser %dx%d (call resize)", kInputName[k], v[k]->width,
                  v[k]->height, s->W, s->H);
    if (v[k]->format != HUSH_RGBA16F) return fail(s, HUSH_ERR_INPUT, "SVGF.encode: '%s' must be rgba16float", kInputName[k]);
    if (v[k]->pitch % 8 != 0 || v[k]->pitch < (size_t)s->W * 8 || ((uintptr_t)v[k]->ptr & 7) != 0)
      return fail(s, HUSH_ERR_INPUT, "SVGF.encode: '%s' has pitch %zu bytes; rows need 8-byte alignment and >= %d bytes", kInputName[k],
                  v[k]->pitch, s->W * 8);
    if (!hush_pitch_ok(v[k]->pitch / 8, s->H))
      return fail(s, HUSH_ERR_INPUT, "SVGF.encode: '%s' has pitch %zu bytes; pitch x height must stay below 2^31 texels",
                  kInputName[k], v[k]->pitch);
  }
  return HUSH_OK;
}

// one pass at the current frame / reset state (svgf.js:359-374)
template <bool FIRST, int LG>
static void launch_atrous_lg(bool p128, bool shared, dim3 g, dim3 b, hipStream_t st, const SvAtrousArgs &a) {
  if (p128) {
    if (shared) hipLaunchKernelGGL((k_sv_atrous<FIRST, LG, true, true>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_sv_atrous<FIRST, LG, true, false>), g, b, 0, st, a);
  } else {
    if (shared) hipLaunchKernelGGL((k_sv_atrous<FIRST, LG, false, true>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_sv_atrous<FIRST, LG, false, false>), g, b, 0, st, a);
  }
}
// iteration i runs at step 2^i; only iteration 0 writes the colour feedback
static void launch_atrous(int i, bool p128, bool shared, dim3 g, dim3 b, hipStream_t st, const SvAtrousArgs &a) {
  switch (i) {
    case 0: launch_atrous_lg<true, 0>(p128, shared, g, b, st, a); break;
    case 1: launch_atrous_lg<false, 1>(p128, shared, g, b, st, a); break;
    case 2: launch_atrous_lg<false, 2>(p128, shared, g, b, st, a); break;
    case 3: launch_atrous_lg<false, 3>(p128, shared, g, b, st, a); break;
    default: launch_atrous_lg<false, 4>(p128, shared, g, b, st, a); break;
  }
}

static hipError_t run_pass(svgf_t *s, int pass, const svgf_inputs *in, hipStream_t st) {
  const int cur = (int)(s->frame & 1), prev = cur ^ 1, W = s->W, H = s->H;
  const Tex *t = s->tex;
  const dim3 g = grid_of(W, H), b(NT);
  const int n_int = int_exponent(s->o.sigma_n);
  if (pass == 0) {
    SvTemporalArgs a;
    a.rad = in_tex(in->radiance); a.nd = in_tex(in->normal_depth); a.mv = in_tex(in->motion); a.mat = in_tex(in->material);
    a.pnd = in_tex(in->prev_normal_depth); a.pmv = in_tex(in->prev_motion);
    a.pcol = t[SV_HC0 + prev]; a.pmom = t[SV_HM0 + prev]; a.outC = t[SV_TMP]; a.outM = t[SV_HM0 + cur];
    a.W = W; a.H = H; a.reset = s->reset ? 1 : 0; a.alpha_color = s->o.alpha_color; a.alpha_moments = s->o.alpha_moments;
    hipLaunchKernelGGL(k_sv_temporal, g, b, 0, st, a);
  } else if (pass == 1) {
    SvVarianceArgs a;
    a.col = t[SV_TMP]; a.mom = t[SV_HM0 + cur]; a.nd = in_tex(in->normal_depth); a.out = t[SV_AT0];
    a.W = W; a.H = H; a.n_int = n_int; a.sigma_n = s->o.sigma_n;
    if (a.n_int == 128) hipLaunchKernelGGL(k_sv_variance<true>, g, b, 0, st, a);
    else hipLaunchKernelGGL(k_sv_variance<false>, g, b, 0, st, a);
  } else if (pass < 2 + s->o.iterations) {
    const int i = pass - 2, src = i & 1;
    SvAtrousArgs a;
    a.in = t[SV_AT0 + src]; a.nd = in_tex(in->normal_depth); a.out = t[SV_AT0 + (src ^ 1)]; a.fb = t[SV_HC0 + cur];
    a.W = W; a.H = H; a.lg_step = i; a.n_int = n_int;
    const int tx = sv_tile_x(i), ty = NT / tx;
    a.n_items = (((W + tx - 1) / tx) * ((H + (ty << i) - 1) / (ty << i))) << i;
    a.sigma_l = s->o.sigma_l; a.sigma_n = s->o.sigma_n; a.sigma_z = s->o.sigma_z;
    const bool p128 = a.n_int == 128;
    // the shared-reciprocal quotients' domain holds for sigmaL, sigmaZ in [0, 2^60] (NaN fails the test)
    const bool shared = a.sigma_l >= 0.0f && a.sigma_l <= 0x1p60f && a.sigma_z >= 0.0f && a.sigma_z <= 0x1p60f;
    launch_atrous(i, p128, shared, dim3((unsigned)a.n_items), b, st, a);
  } else if (pass == 2 + s->o.iterations) {
    SvResolveArgs a;
    a.illum = t[SV_AT0 + (s->o.iterations & 1)]; a.mat = in_tex(in->material); a.emi = in_tex(in->emission);
    a.mv = in_tex(in->motion); a.prev = t[SV_TAA0 + prev]; a.out = t[SV_TAA0 + cur];
    a.W = W; a.H = H; a.reset = s->reset ? 1 : 0; a.taa_alpha = s->o.taa_alpha;
    if (s->o.taa) hipLaunchKernelGGL(k_sv_resolve<true>, g, b, 0, st, a);
    else hipLaunchKernelGGL(k_sv_resolve<false>, g, b, 0, st, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static const char *pass_label(const svgf_t *s, int p) {
  return p == 0 ? "temporal" : p == 1 ? "variance" : p < 2 + s->o.iterations ? "atrous" : "resolve";
}

extern "C" int32_t svgf_encode(svgf_t *s, const svgf_inputs *in, void *stream, hush_surface *out) {
  if (!s) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_encode: null context");
  if (!in) return fail(s, HUSH_ERR_INPUT, "SVGF.encode: missing input '%s'", kInputName[0]);
  int r = check_inputs(s, in);
  if (r != HUSH_OK) return r;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipSetDevice(s->device);
  if (e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
  const int np = 3 + s->o.iterations;
  if (s->timing && !s->ev[0])
    for (auto &ev : s->ev)
      if ((e = hipEventCreate(&ev)) != hipSuccess) return hip_fail(s, e, "hipEventCreate");
  for (int p = 0; p < np; p++) {
    if (s->timing) (void)hipEventRecord(s->ev[p], st);
    e = run_pass(s, p, in, st);
    if (e != hipSuccess) return hip_fail(s, e, pass_label(s, p));
  }
  if (s->timing) { (void)hipEventRecord(s->ev[np], st); s->ev_ready = true; s->ev_passes = np; }
  const int cur = (int)(s->frame & 1);
  s->reset = false;
  s->frame++;
  if (out) *out = surf(s, SV_TAA0 + cur);
  return HUSH_OK;
}

extern "C" int32_t svgf_output(const svgf_t *s, hush_surface *out) {
  if (!s || !out) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_output: null argument");
  *out = surf(s, SV_TAA0 + (int)(s->frame & 1));
  return HUSH_OK;
}

extern "C" int32_t svgf_set_timing(svgf_t *s, int32_t enabled) {
  if (!s) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_set_timing: null context");
  s->timing = enabled != 0;
  s->ev_ready = false;
  return HUSH_OK;
}

extern "C" int32_t svgf_pass_times(svgf_t *s, float *ms, const char **labels, int32_t *count) {
  if (!s) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_pass_times: null context");
  const int np = 3 + s->o.iterations;
  if (count) *count = np;
  if (labels)
    for (int i = 0; i < np; i++) labels[i] = pass_label(s, i);
  if (!ms) return HUSH_OK;
  if (!s->ev_ready) return fail(s, HUSH_ERR_ARGUMENT, "svgf_pass_times: no timed frame (enable timing first)");
  hipError_t e = hipEventSynchronize(s->ev[s->ev_passes]);
  if (e != hipSuccess) return hip_fail(s, e, "hipEventSynchronize");
  for (int i = 0; i < s->ev_passes; i++)
    if ((e = hipEventElapsedTime(&ms[i], s->ev[i], s->ev[i + 1])) != hipSuccess) return hip_fail(s, e, "hipEventElapsedTime");
  return HUSH_OK;
}

extern "C" int32_t svgf_debug_surface(svgf_t *s, int32_t id, hush_surface *out) {
  if (!s || !out) return fail(s, HUSH_ERR_ARGUMENT, "svgf_debug_surface: null argument");
  if (id < 0 || id >= SV_COUNT) return fail(s, HUSH_ERR_ARGUMENT, "svgf_debug_surface: unknown surface %d", id);
  *out = surf(s, id);
  return HUSH_OK;
}

extern "C" int32_t svgf_debug_set_state(svgf_t *s, int64_t frame, int32_t reset) {
  if (!s) return fail(nullptr, HUSH_ERR_ARGUMENT, "svgf_debug_set_state: null context");
  s->frame = frame;
  s->reset = reset != 0;
  return HUSH_OK;
}

extern "C" int32_t svgf_debug_run_pass(svgf_t *s, int32_t pass, const svgf_inputs *in, void *stream) {
  if (!s || !in) return fail(s, HUSH_ERR_ARGUMENT, "svgf_debug_run_pass: null argument");
  if (pass < 0 || pass > 2 + s->o.iterations) return fail(s, HUSH_ERR_ARGUMENT, "svgf_debug_run_pass: unknown pass %d", pass);
  int r = check_inputs(s, in);
  if (r != HUSH_OK) return r;
  hipError_t e = hipSetDevice(s->device);
  if (e == hipSuccess) e = run_pass(s, pass, in, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(s, e, pass_label(s, pass));
  return HUSH_OK;
}
