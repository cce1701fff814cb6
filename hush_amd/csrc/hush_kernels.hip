// Hush denoiser hot path: the eight per-frame passes of src/denoise/hush.js as gfx950 HIP kernels.
//
// Launch shape: every pass runs 256-thread workgroups (four wave64s) over 32x8 output texels; each wave
// covers a 32x2 strip, so a row segment is 32 texels = 256 contiguous bytes of an rgba16float surface.
// Neighbourhood passes stage a clamp-to-edge tile plus apron in LDS (the semantics of the reference's
// clampP-staged workgroup tiles, hush.js:287-295, 332-339, 407-415, 512-517, 563-568, 642-645).
//
// The arithmetic follows the reference's operation order with no FMA contraction (build flag
// -ffp-contract=off); the MLP's multiply-adds are fused explicitly (__builtin_fmaf).
#include "hush.h"
#include "hush_device.h"
#include <hip/hip_ext.h>
#include <type_traits>

#include "hush_internal.h"

namespace hush {

constexpr int TW = 32, TH = 8, NT = 256;

// The half- and quarter-resolution surfaces are the context's own, and at the largest frame side they span
// at most (HUSH_MAX_SIDE / 2)^2 texels of 16 B, 1 GiB: the passes that touch only them (infer + mid1, mid2,
// mid3, coarse1, coarse2) address them with 32-bit byte offsets from the surface base at every size.
static_assert((uint64_t)((HUSH_MAX_SIDE + 1) / 2 + 63) / 64 * 64 * ((HUSH_MAX_SIDE + 1) / 2) * 16 < ((uint64_t)1 << 31),
              "half-resolution surfaces must stay below 2 GiB for 32-bit byte offsets");

// ============================================================================ pass 1: temporal
// TEMPORAL (hush.js:79-308): demodulate, 3x3 stats, reproject (bilinear + 3x3 fallback), clamp and blend,
// store A0 / mom / key; then per 2x2 block the 20 MLP features and the half-res guide (F0..F3).

// check (hush.js:103-105) on a decoded key texel: xyz = octDec(k.xy) (exact), w = binary16 (k.z, k.w)
// Every term is evaluated and the results combined without short-circuit: the test is branch-free, so a
// caller's per-tap accumulation needs no nested divergent blocks.
__device__ __forceinline__ bool check(float4 kd, float3 n, float zprev, float obj) {
  const float2 zi = pair_of(kd.w);
  float tol = 0.05f * zprev;
  tol = tol + 0.01f;
  return (zi.x >= 0.0f) & (zi.y == obj) & (fabsf(zi.x - zprev) <= tol) & (dot3(make_float3(kd.x, kd.y, kd.z), n) >= 0.9f);
}

// per-pixel feature inputs kept in LDS for the block phase (structure-of-arrays, conflict-free).  P0.y
// (E qf16(lum)) is the staged sQ of the pixel and is read from there; the material P3 is kept as its two
// packed binary16 words (its exact input bits): 18 rows instead of 21, fewer LDS writes and reads.
enum { P0X, P0Z, P0W, P1X, P1Y, P1Z, P1W, P2Z, P2W, NSX, NSY, NSZ, P3A, P3B, P4X, P4Y, P4Z, P4W, NP };
// LDS slot of pixel t (t = 32 ly + lx) in the sP rows: bit 0 flipped by bit 6 of t and bit 5 by bit 7 (a
// permutation of 0 .. 255).  The block phase's lane (bxl, byl) reads pixel 64 byl + 32 jy + 2 bxl + jx: the
// 64 lanes then hit 64 distinct banks (4 lanes per bank without the flips); a wave's own writes stay
// conflict free.
__device__ __forceinline__ int pslot(int t) { return t ^ (((t >> 6) & 1) | (((t >> 7) & 1) << 5)); }

// Temporal statistics (moments, variances, cancellation-prone differences such as m2 - m1^2) are
// ill-conditioned, so this pass is evaluated with IEEE arithmetic under both math policies.
// AUX (training only): also writes the per-pixel inputs of the temporal blend (TEMPORAL_AUX floats each:
// history flag, illumination, 3x3 mean and deviation, reprojected colour and moments, luminance, spatial
// variance, the half-res texel of alpha / gamma), from which the differentiable frame re-blends A0.
// Running sum; KAHAN: Neumaier-compensated (the rounding error of every addition is carried separately and
// added back at the end), the compensated-moment form of the temporal statistics.
template <bool KAHAN> struct Sum {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float x) {
    if (!KAHAN) { s += x; return; }
    const float t = s + x;
    c += fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ float get() const { return KAHAN ? s + c : s; }
};

// BUF: the history gathers through buffer descriptors with 32-bit byte offsets (every history surface is
// under 2 GiB, checked by the host), in place of 64-bit address arithmetic per tap.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t hist_rsrc(const void *p, int pitch, int rows, int texel) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0,
                                           (int)((uint32_t)pitch * (uint32_t)rows * (uint32_t)texel), 0x00020000);
}

// 4x4 neighbourhood (block +-1) of the exposed, truncated luminance of a half-res texel's block
// (hush.js:253-265) over the staged tile (row pitch TW + 2): its mean, and its variance about that mean.
// sI[.].w is exactly 1 or 0 and an invalid texel's sQ is E * qf16(0) = 0, so w * E * qf16(lum) is sQ.
// The count divides through rcpn's exact reciprocal (a count of 1 .. 16); a sum below 2^-102, where the
// shared-reciprocal quotient could miss by an ulp, only occurs at exposures below 2^-58, and its quotient
// then vanishes in the features' log(1 + x) and in the variance's squares.
template <bool KAHAN>
__device__ __forceinline__ float mean4x4(const float4 *sI, const float *sQ, int byl, int bxl) {
  constexpr int SW = TW + 2;
  Sum<KAHAN> q1;
  float n4 = 0.0f;
  const int b0 = (2 * byl) * SW + 2 * bxl;      // tile index of pixel (2*bxl - 1, 2*byl - 1) incl. apron
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int idx = b0 + (j >> 2) * SW + (j & 3);
    q1.add(sQ[idx]);
    n4 += sI[idx].w;
  }
  const float n = mx(n4, 1.0f);
  return Math<false>::divs(q1.get(), n, Math<false>::rcpn(n));
}
template <bool KAHAN>
__device__ __forceinline__ float var4x4(const float4 *sI, const float *sQ, int byl, int bxl, float mu4) {
  constexpr int SW = TW + 2;
  Sum<KAHAN> q4;
  float n4 = 0.0f;
  const int b0 = (2 * byl) * SW + 2 * bxl;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int idx = b0 + (j >> 2) * SW + (j & 3);
    const float w = sI[idx].w, d = sQ[idx] - mu4;
    q4.add(w * d * d);
    n4 += w;
  }
  const float n = mx(n4, 1.0f);
  return Math<false>::divs(q4.get(), n, Math<false>::rcpn(n));
}

template <bool AUX, bool KAHAN, bool BUF>
__global__ __launch_bounds__(NT) void k_temporal(TemporalArgs a) {
  constexpr bool FAST = false;
  using F = Math<FAST>;
  constexpr int SW = TW + 2, SH = TH + 2, SN = SW * SH;
  __shared__ float4 sI[SN];        // demodulated illumination, valid flag
  __shared__ float sL[SN];         // lum(illum)
  __shared__ float sQ[SN];         // E * qf16(lum(illum))
  __shared__ float sP[NP][NT];     // feature inputs P0..P4 (+ decoded normal) per pixel
  const int tid = threadIdx.x;
  const int2 org = tile_origin<TW, TH>(a.W, a.H);
  const int x0 = org.x, y0 = org.y;
  const float E = a.E;
  const int lx = tid & 31, ly = tid >> 5;
  const int px = x0 + lx, py = y0 + ly;
  const bool in_img = inside(px, py, a.W, a.H);
  // the wave runs at a raised priority while it issues its loads (G-buffer, apron, reprojection gathers), so
  // that other waves' arithmetic on the SIMD does not hold them back; back to the base priority after them
  __builtin_amdgcn_s_setprio(1);

  // this pixel's G-buffer (also its own staged texel)
  // all four loads leave before any of them is waited for (the scheduler had put the first one's wait, and
  // its unpacking, between it and the other three)
  // (streaming: non-temporal loads, the pass reads each of these texels once)
  uint2 ndr = make_uint2(0u, 0xbc000000u), mvr = make_uint2(0u, 0u), rdr = mvr, mraw = mvr;   // nd = (0, 0, 0, -1)
  if (in_img) {
    ndr = ld_nt(texel<BUF>(a.nd.p, a.nd.pitch, px, py)); mvr = ld_nt(texel<BUF>(a.mv.p, a.mv.pitch, px, py));
    rdr = ld_nt(texel<BUF>(a.rad.p, a.rad.pitch, px, py)); mraw = ld_nt(texel<BUF>(a.mat.p, a.mat.pitch, px, py));
    __builtin_amdgcn_sched_barrier(0);
  }
  // the 84 apron texels of the staged tile (threads 0..83): their normalDepth, radiance and material are
  // loaded here, with the pixel's own and before any wait, so that no load waits between the reprojection
  // and the barrier
  int atx = 0, aty = 0;
  bool a_in = false;
  uint2 a_nd = make_uint2(0u, 0u), a_rd = a_nd, a_mt = a_nd;
  if (tid < 2 * SW + 2 * TH) {
    if (tid < SW) { atx = tid; aty = 0; }
    else if (tid < 2 * SW) { atx = tid - SW; aty = SH - 1; }
    else if (tid < 2 * SW + TH) { atx = 0; aty = tid - 2 * SW + 1; }
    else { atx = SW - 1; aty = tid - 2 * SW - TH + 1; }
    const int qx = x0 - 1 + atx, qy = y0 - 1 + aty;
    a_in = inside(qx, qy, a.W, a.H);
    if (a_in) {
      a_nd = ldr<BUF>(a.nd, qx, qy);
      a_rd = ldr<BUF>(a.rad, qx, qy);
      a_mt = ldr<BUF>(a.mat, qx, qy);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  const float4 nd = unpack4(ndr), mv = unpack4(mvr), rd = unpack4(rdr), mr = unpack4(mraw);
  const bool use_hist = in_img && !a.reset && nd.w >= 0.0f;

  // reprojection fetches (hush.js:144-170): the four bilinear taps of the decoded key, colour and moments
  // and the nearest half-res par0 texel, issued before the tile staging so that their latency overlaps the
  // staging's arithmetic; colour, moments and par0 stay packed until used
  float fx = 0.0f, fy = 0.0f;
  int bix = 0, biy = 0, cx = 0, cy = 0;
  float4 k4[4];
  uint2 c4r[4], m4r[4], pqr = make_uint2(0u, 0u);
  // every lane gathers (the taps are clamped into the surfaces, so a lane without history reads valid
  // texels it then ignores): with no branch around them, each later wait counts only the loads it needs
  {
    const float ppx = (float)px + mv.x, ppy = (float)py + mv.y;
    const float bx = floorf(ppx), by = floorf(ppy);
    fx = ppx - bx; fy = ppy - by;
    bix = f2i(bx); biy = f2i(by);
    cx = f2i(floorf(ppx + 0.5f)); cy = f2i(floorf(ppy + 0.5f));
    if constexpr (BUF) {
      const __amdgpu_buffer_rsrc_t rk = hist_rsrc(a.pkd.p, a.pkd.pitch, a.H, 16);
      const __amdgpu_buffer_rsrc_t rc = hist_rsrc(a.pcol.p, a.pcol.pitch, a.H, 8);
      const __amdgpu_buffer_rsrc_t rm = hist_rsrc(a.pmom.p, a.pmom.pitch, a.H, 8);
      const __amdgpu_buffer_rsrc_t rp = hist_rsrc(a.ppar0.p, a.ppar0.pitch, a.H2, 8);
      // colour and moments: a row's two horizontal taps in one 16-byte load at qx0 = clamp(bix, 0, W - 1) (the
      // launcher has checked that both surfaces hold at least one texel of padding per row, so texel qx0 + 1
      // lies inside the row).  Tap 0 is the low half.  Tap 1 is the high half when 0 <= bix <= W - 2, else the
      // low half: bix < 0 puts tap 0 outside, and texel 0 is tap 1's clamped texel; bix >= W - 1 puts tap 1
      // outside, where the clamped tap is texel W - 1.  So every tap holds exactly the clamped texel that a
      // separate load would read, and the padding's contents are never used.
      const uint32_t qx0 = clampi(bix, 0, a.W - 1);
      const bool lo1 = (bix < 0) | (bix > a.W - 2);
#pragma unroll
      for (int r = 0; r < 2; r++) {
        // qy < 16384 and pitch < 2^23 fit the 24-bit multiply; the product is below 2^28 texels
        const uint32_t qy = clampi(addw(biy, r), 0, a.H - 1);
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const uint32_t qx = clampi(addw(bix, c), 0, a.W - 1);
          const auto kv = __builtin_amdgcn_raw_buffer_load_b128(rk, (int)((__umul24(qy, a.pkd.pitch) + qx) * 16u), 0, 0);
          k4[2 * r + c] = make_float4(__uint_as_float(kv[0]), __uint_as_float(kv[1]), __uint_as_float(kv[2]), __uint_as_float(kv[3]));
        }
        const auto cv = __builtin_amdgcn_raw_buffer_load_b128(rc, (int)((__umul24(qy, a.pcol.pitch) + qx0) * 8u), 0, 0);
        const auto mv4 = __builtin_amdgcn_raw_buffer_load_b128(rm, (int)((__umul24(qy, a.pmom.pitch) + qx0) * 8u), 0, 0);
        c4r[2 * r] = make_uint2(cv[0], cv[1]);
        c4r[2 * r + 1] = lo1 ? make_uint2(cv[0], cv[1]) : make_uint2(cv[2], cv[3]);
        m4r[2 * r] = make_uint2(mv4[0], mv4[1]);
        m4r[2 * r + 1] = lo1 ? make_uint2(mv4[0], mv4[1]) : make_uint2(mv4[2], mv4[3]);
      }
      const uint32_t hx = clampi(cx, 0, a.W - 1) >> 1, hy = clampi(cy, 0, a.H - 1) >> 1;
      const auto pv = __builtin_amdgcn_raw_buffer_load_b64(rp, (int)((__umul24(hy, a.ppar0.pitch) + hx) * 8u), 0, 0);
      pqr = make_uint2(pv[0], pv[1]);
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const size_t qx = clampi(addw(bix, i & 1), 0, a.W - 1), qy = clampi(addw(biy, i >> 1), 0, a.H - 1);
        k4[i] = a.pkd.p[qy * a.pkd.pitch + qx];
        c4r[i] = a.pcol.p[qy * a.pcol.pitch + qx];
        m4r[i] = a.pmom.p[qy * a.pmom.pitch + qx];
      }
      const size_t hx = clampi(cx, 0, a.W - 1) >> 1, hy = clampi(cy, 0, a.H - 1) >> 1;
      pqr = a.ppar0.p[hy * a.ppar0.pitch + hx];
    }
  }

  // the staging's arithmetic runs while the gathers are in flight (its inputs arrived with the first loads)
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_setprio(0);
  // tile staging (hush.js:287-294): each thread its own pixel, threads 0..83 also an apron texel.
  // radiance / max(albedo, 1e-3): the quotient through the exact reciprocal (binary16 numerators, divisors in
  // [1e-3, 65504]: the domain where divs equals IEEE division, special operands included)
  {
    float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (in_img && nd.w >= 0.0f) {
      const float bx = mx(mr.x, 1e-3f), by = mx(mr.y, 1e-3f), bz = mx(mr.z, 1e-3f);
      e = make_float4(F::divs(rd.x, bx, F::rcpn(bx)), F::divs(rd.y, by, F::rcpn(by)), F::divs(rd.z, bz, F::rcpn(bz)), 1.0f);
    }
    const int c = (ly + 1) * SW + lx + 1;
    const float l = lum(e.x, e.y, e.z);
    sI[c] = e; sL[c] = l; sQ[c] = E * qf16(l);
  }
  if (tid < 2 * SW + 2 * TH) {
    float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (a_in && unpack2(a_nd.y).y >= 0.0f) {
      const float4 r = unpack4(a_rd), m = unpack4(a_mt);
      const float bx = mx(m.x, 1e-3f), by = mx(m.y, 1e-3f), bz = mx(m.z, 1e-3f);
      e = make_float4(F::divs(r.x, bx, F::rcpn(bx)), F::divs(r.y, by, F::rcpn(by)), F::divs(r.z, bz, F::rcpn(bz)), 1.0f);
    }
    const int c = aty * SW + atx;
    const float l = lum(e.x, e.y, e.z);
    sI[c] = e; sL[c] = l; sQ[c] = E * qf16(l);
  }

  // reprojection (hush.js:144-187) ahead of the staging barrier: it needs no LDS, and consuming the gathered taps
  // here keeps only its results live across the staging barrier and the 3x3 statistics
  const float3 n = make_float3(nd.x, nd.y, nd.z);
  uint32_t mbits = 0;
    // reprojection (hush.js:144-187)
    bool has = false;
    float hcx = 0.0f, hcy = 0.0f, hcz = 0.0f;
    float4 hm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float v = 0.0f, alpha = 1.0f, gamma = 1.0f;
    if (use_hist) {
      const int W = a.W, H = a.H;
      const float wts[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
      float4 c4[4], m4[4];
#pragma unroll
      for (int i = 0; i < 4; i++) { c4[i] = unpack4(c4r[i]); m4[i] = unpack4(m4r[i]); }
      const float2 pq = unpack2(pqr.x);
      // a rejected tap enters the sums with weight 0: the stored history is finite (binary16 stores clamp to
      // +-65504), so it adds a zero, which leaves every running sum unchanged bit for bit (each starts at +0
      // and never becomes -0); branch-free, with no divergent block per tap
      float sumW = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const bool ok = (wts[i] > 0.0f) & inside(addw(bix, i & 1), addw(biy, i >> 1), W, H) & check(k4[i], n, mv.w, mv.z);
        mbits |= (uint32_t)ok << i;
        const float w = ok ? wts[i] : 0.0f;
        sumW += w;
        hcx += mulh(w, c4[i].x, a.nz); hcy += mulh(w, c4[i].y, a.nz); hcz += mulh(w, c4[i].z, a.nz);
        hm.x += mulh(w, m4[i].x, a.nz); hm.y += mulh(w, m4[i].y, a.nz); hm.z += mulh(w, m4[i].z, a.nz);
        hm.w += mulh(w, m4[i].w, a.nz);
      }
      if (sumW >= 0.01f) {
        const float rs = F::rcpn(sumW);
        hcx = F::divsf(hcx, sumW, rs); hcy = F::divsf(hcy, sumW, rs); hcz = F::divsf(hcz, sumW, rs);
        hm.x = F::divsf(hm.x, sumW, rs); hm.y = F::divsf(hm.y, sumW, rs); hm.z = F::divsf(hm.z, sumW, rs); hm.w = F::divsf(hm.w, sumW, rs);
        v = sumW; has = true;
      } else {
        float n3 = 0.0f;
        hcx = hcy = hcz = 0.0f;
        hm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        // the nine keys in one batch of loads (clamped; a texel outside the image fails `inside`), then the
        // colour and moments of the consistent ones in a second batch
        float4 kq[9];
#pragma unroll
        for (int j = 0; j < 9; j++)
          kq[j] = ldf<BUF>(a.pkd, clampi(addw(cx, j % 3 - 1), 0, W - 1), clampi(addw(cy, j / 3 - 1), 0, H - 1));
        uint32_t okm = 0;
#pragma unroll
        for (int j = 0; j < 9; j++)
          okm |= (uint32_t)(inside(addw(cx, j % 3 - 1), addw(cy, j / 3 - 1), W, H) & check(kq[j], n, mv.w, mv.z)) << j;
        uint2 cq[9], mq[9];
#pragma unroll
        for (int j = 0; j < 9; j++) {
          cq[j] = mq[j] = make_uint2(0u, 0u);
          if (okm >> j & 1u) {
            const int qx = addw(cx, j % 3 - 1), qy = addw(cy, j / 3 - 1);
            cq[j] = ldr<BUF>(a.pcol, qx, qy); mq[j] = ldr<BUF>(a.pmom, qx, qy);
          }
        }
#pragma unroll
        for (int j = 0; j < 9; j++) {
          if (okm >> j & 1u) {
            n3 += 1.0f;
            const float4 cc = unpack4(cq[j]), mm = unpack4(mq[j]);
            hcx += cc.x; hcy += cc.y; hcz += cc.z;
            hm.x += mm.x; hm.y += mm.y; hm.z += mm.z; hm.w += mm.w;
          }
        }
        if (n3 > 0.0f) {
          const float rn = F::rcpn(n3);
          hcx = F::divsf(hcx, n3, rn); hcy = F::divsf(hcy, n3, rn); hcz = F::divsf(hcz, n3, rn);
          hm.x = F::divsf(hm.x, n3, rn); hm.y = F::divsf(hm.y, n3, rn); hm.z = F::divsf(hm.z, n3, rn); hm.w = F::divsf(hm.w, n3, rn);
          v = 0.1f; has = true;
          mbits |= 0x10;
        }
      }
      alpha = pq.x; gamma = pq.y;
    }
    has = has && hm.w > 0.0f;
    if (has) mbits |= 0x20;
  __syncthreads();

  // feature inputs go straight to LDS (sky / outside: zeros with depth -1, hush.js:112, 297)
  float *P = &sP[0][pslot(tid)];
  bool p_written = false;
  if (in_img) {
    if (nd.w < 0.0f) {                                   // sky (hush.js:115-119)
      st<BUF>(a.outA, px, py, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
      st_nt(texel<BUF>(a.outMom.p, a.outMom.pitch, px, py), make_uint2(0u, 0u));
      const uint32_t zid = pack_rtz(-1.0f, mv.z);
      st_nt(texel<BUF>(a.outKey.p, a.outKey.pitch, px, py), make_uint2(0u, zid));
      const float3 d0 = oct_dec(0.0f, 0.0f);
      st_nt(texel<BUF>(a.outKD.p, a.outKD.pitch, px, py), make_float4(d0.x, d0.y, d0.z, __uint_as_float(zid)));
      mbits = 0x80;
      if (AUX) a.aux[((size_t)py * a.W + px) * TEMPORAL_AUX] = -1.0f;
    } else {
      const int c0 = (ly + 1) * SW + lx + 1;
      const float4 ce = sI[c0];
      const float l = sL[c0];
      // 3x3 stats (hush.js:125-142)
      Sum<KAHAN> a1x, a1y, a1z;
      float cnt = 0.0f;
      // t.w is exactly 1 or 0 and an invalid texel is staged as (0, 0, 0, 0), so t.x * t.w is t.x bit for bit
#pragma unroll
      for (int j = 0; j < 9; j++) {
        float4 t = sI[c0 + (j / 3 - 1) * SW + (j % 3 - 1)];
        a1x.add(t.x); a1y.add(t.y); a1z.add(t.z);
        cnt += t.w;
      }
      const float s1x = a1x.get(), s1y = a1y.get(), s1z = a1z.get();
      cnt = mx(cnt, 1.0f);
      const float rcnt = F::rcpn(cnt);
      const float mux = F::divs(s1x, cnt, rcnt), muy = F::divs(s1y, cnt, rcnt), muz = F::divs(s1z, cnt, rcnt);
      const float slm = lum(mux, muy, muz);
      Sum<KAHAN> avx, avy, avz, avl;
      // centred sums, each term weighted by its validity t.w; when all nine texels of every active lane are
      // valid (t.w == 1), t.w * d * d is d * d bit for bit (1 * x is x, infinities and NaN included), and the
      // wave takes the form without the weights (a wave-uniform branch)
      auto centred = [&](auto weighted) {
#pragma unroll
        for (int j = 0; j < 9; j++) {
          int idx = c0 + (j / 3 - 1) * SW + (j % 3 - 1);
          float4 t = sI[idx];
          float dx = t.x - mux, dy = t.y - muy, dz = t.z - muz;
          float dl = sL[idx] - slm;
          if constexpr (decltype(weighted)::value) {
            avx.add(t.w * dx * dx); avy.add(t.w * dy * dy); avz.add(t.w * dz * dz);
            avl.add(t.w * dl * dl);
          } else {
            avx.add(dx * dx); avy.add(dy * dy); avz.add(dz * dz);
            avl.add(dl * dl);
          }
        }
      };
      if (__all(cnt == 9.0f)) centred(std::false_type{});
      else centred(std::true_type{});
      const float svx = avx.get(), svy = avy.get(), svz = avz.get();
      float svar = avl.get();
      const float sdx = F::sqrtn(mx(F::divs(svx, cnt, rcnt), 1e-12f)), sdy = F::sqrtn(mx(F::divs(svy, cnt, rcnt), 1e-12f)),
                  sdz = F::sqrtn(mx(F::divs(svz, cnt, rcnt), 1e-12f));
      svar = F::divs(svar, cnt, rcnt);
      // temporal blend (hush.js:189-211)
      float cr = ce.x, cg = ce.y, cb = ce.z;
      float nmx = l, nmy = l * l, nfast = l, len1 = 1.0f, lag = 0.0f;
      if (has) {
        float gx = gamma * sdx, gy = gamma * sdy, gz = gamma * sdz;
        float hlx = mn(mx(hcx, mux - gx), mux + gx), hly = mn(mx(hcy, muy - gy), muy + gy), hlz = mn(mx(hcz, muz - gz), muz + gz);
        len1 = mn(hm.w + 1.0f, 32.0f);
        float aa = mx(alpha, F::rcpn(len1));
        cr = hlx + (ce.x - hlx) * aa; cg = hly + (ce.y - hly) * aa; cb = hlz + (ce.z - hlz) * aa;
        nmx = hm.x + (l - hm.x) * aa;
        nmy = hm.y + (l * l - hm.y) * aa;
        nfast = hm.z + (l - hm.z) * mx(1.0f / 6.0f, F::rcpn(len1));
        float lh = lum(hcx, hcy, hcz);
        // the divisor is 0 or normal (near cancellation the sum is exact, a multiple of at least 2^-44) and the
        // numerator 0 or at least 2^-96 (luminances of demodulated binary16 texels and of their averages)
        const float ld = F::sqrtn(mx(F::divs(svar, cnt, rcnt), 1e-12f)) + 0.02f * lh + 1e-3f;
        lag = F::divs(fabsf(slm - lh), ld, F::rcpn(ld));
      } else {
        v = 0.0f;
      }
      float tvar = mx(nmy - nmx * nmx, 0.0f);
      float4 A = make_float4(cr, cg, cb, svar + (tvar - svar) * mn(len1 / 4.0f, 1.0f));
      if (AUX) {
        float *x = a.aux + ((size_t)py * a.W + px) * TEMPORAL_AUX;
        const int cxq = clampi(f2i(floorf((float)px + mv.x + 0.5f)), 0, a.W - 1) >> 1;
        const int cyq = clampi(f2i(floorf((float)py + mv.y + 0.5f)), 0, a.H - 1) >> 1;
        const float v[TEMPORAL_AUX] = {has ? 1.0f : 0.0f, ce.x, ce.y, ce.z, mux, muy, muz, sdx, sdy, sdz, hcx, hcy, hcz,
                                       hm.x, hm.y, hm.w, l, svar, (float)(cyq * a.W2 + cxq), 0.0f};
#pragma unroll
        for (int k = 0; k < TEMPORAL_AUX; k++) x[k] = v[k];
      }
      float2 oc = oct_enc(n);
      st<BUF>(a.outA, px, py, A);
      st_nt(texel<BUF>(a.outMom.p, a.outMom.pitch, px, py), pack4(make_float4(nmx, nmy, nfast, len1)));
      // the key as stored, and its exact decode for the next frame's check and this frame's fine level; mom,
      // key and KD are stored non-temporal (only the next frame reads them, after this frame's other passes)
      const uint32_t okp = pack_rtz(oc.x, oc.y), zid = pack_rtz(nd.w, mv.z);
      st_nt(texel<BUF>(a.outKey.p, a.outKey.pitch, px, py), make_uint2(okp, zid));
      const float2 ok = unpack2(okp);
      const float3 dec = oct_dec(ok.x, ok.y);
      st_nt(texel<BUF>(a.outKD.p, a.outKD.pitch, px, py), make_float4(dec.x, dec.y, dec.z, __uint_as_float(zid)));
      // feature inputs from the stored values (hush.js:212-222)
      float aqx = qf16(A.x), aqy = qf16(A.y), aqz = qf16(A.z), aqw = qf16(A.w);
      float m1 = E * qf16(nmx);
      float m2 = E * E * qf16(nmy);
      P[P0X * NT] = E * lum(aqx, aqy, aqz); P[P0Z * NT] = mx(m2 - m1 * m1, 0.0f); P[P0W * NT] = len1;
      P[P1X * NT] = qf16(v); P[P1Y * NT] = qf16(lag); P[P1Z * NT] = E * qf16(nfast); P[P1W * NT] = m1;
      float ocx = qf16(oc.x), ocy = qf16(oc.y);
      P[P2Z * NT] = qf16(nd.w); P[P2W * NT] = qf16(F::sqrtn(mv.x * mv.x + mv.y * mv.y));
      float3 ns = (ocx == ok.x && ocy == ok.y) ? dec : oct_dec_m<FAST>(ocx, ocy);
      P[NSX * NT] = ns.x; P[NSY * NT] = ns.y; P[NSZ * NT] = ns.z;
      P[P3A * NT] = __uint_as_float(mraw.x); P[P3B * NT] = __uint_as_float(mraw.y);
      P[P4X * NT] = aqx; P[P4Y * NT] = aqy; P[P4Z * NT] = aqz; P[P4W * NT] = aqw;
      p_written = true;
    }
    if (a.mask) *texel<BUF>(a.mask, a.mask_pitch, px, py) = (uint8_t)mbits;
  }
  if (!p_written) {
#pragma unroll
    for (int k = 0; k < NP; k++) P[k * NT] = k == P2Z ? -1.0f : 0.0f;
  }
  __syncthreads();

  // block phase (hush.js:226-282, 301-306): the 16x4 half-res texels of this tile, split over the four
  // waves by output, each over all 64 texels: F0.xy (luminance features and the 4x4 variance), F0.zw (the
  // remaining luminance, lag and fast-moment features), F2 (block count, motion, 4x4 mean, block variance,
  // albedo) and F1 + F3 + the decoded guide (normal, depth and material features).  Each forms the 2x2
  // sums it needs in the same order, so every feature is the one-wave result bit for bit; which wave takes
  // which part rotates with the workgroup index, so that the heavier parts spread over the SIMDs.
  const int role = ((tid >> 6) - (int)blockIdx.x) & 3;
  const int bxl = tid & 15, byl = (tid & 63) >> 4;
  const int X = (x0 >> 1) + bxl, Y = (y0 >> 1) + byl;
  if (X >= a.W2 || Y >= a.H2) return;
  // A sky or outside pixel's feature inputs are stored as zeros (its depth as -1), so it enters every sum as
  // +0, which leaves the running sums unchanged bit for bit (each starts at +0 and never becomes -0); only its
  // depth and staged luminance are masked.  Straight-line code, no divergent block per pixel.
  float nv = 0.0f;
  bool ok[4];
  int t4[4], tl4[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    t4[j] = pslot(byl * 64 + (j >> 1) * 32 + 2 * bxl + (j & 1));
    tl4[j] = (2 * byl + (j >> 1) + 1) * SW + 2 * bxl + (j & 1) + 1;        // the pixel's staged tile index
    ok[j] = sP[P2Z][t4[j]] >= 0.0f;
    nv += ok[j] ? 1.0f : 0.0f;
  }
  if (nv == 0.0f) {
    if (role == 0 || role == 3) {
      texel<BUF>((uint2 *)a.F0.p, 2 * a.F0.pitch, 2 * X + (role == 3), Y)[0] = make_uint2(0u, 0u);
    } else if (role == 1) {
      stu<BUF>(a.F2, X, Y, make_uint4(0, 0, 0, 0));
    } else {
      stu<BUF>(a.F1, X, Y, make_uint4(0, 0, 0, 0));
      const uint4 g3 = make_uint4(pk(0.0f, 0.0f), pk(-1.0f, 0.0f), 0u, 0u);
      stu<BUF>(a.F3, X, Y, g3);
      const float3 d0 = oct_dec(0.0f, 0.0f);
      stf<BUF>(a.outGH, X, Y, make_float4(d0.x, d0.y, d0.z, __uint_as_float(g3.y)));
    }
    return;
  }
  const float k = F::rcpn(nv);
  // F0 as two 8-byte halves of its 16-byte texel
  uint2 *const f0h = texel<BUF>((uint2 *)a.F0.p, 2 * a.F0.pitch, 2 * X + (role == 3), Y);
  if (role == 0) {
    Sum<KAHAN> qS0[3];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int t = t4[j];
      const float p0[3] = {sP[P0X][t], ok[j] ? sQ[tl4[j]] : 0.0f, sP[P0Z][t]};
#pragma unroll
      for (int c = 0; c < 3; c++) qS0[c].add(p0[c]);
    }
    float S0[3];
#pragma unroll
    for (int c = 0; c < 3; c++) S0[c] = qS0[c].get() * k;
    const float mu4 = mean4x4<KAHAN>(sI, sQ, byl, bxl);
    const float var4 = var4x4<KAHAN>(sI, sQ, byl, bxl, mu4);
    f0h[0] = make_uint2(pk(log1p_<FAST>(S0[0]), log1p_<FAST>(S0[1])), pk(log1p_<FAST>(S0[2]), log1p_<FAST>(var4)));
  } else if (role == 3) {
    Sum<KAHAN> qS03, qS1[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int t = t4[j];
      qS03.add(sP[P0W][t]);
#pragma unroll
      for (int c = 0; c < 4; c++) qS1[c].add(sP[P1X + c][t]);
    }
    float S1[4];
#pragma unroll
    for (int c = 0; c < 4; c++) S1[c] = qS1[c].get() * k;
    const float S03 = qS03.get() * k;
    const float fd = S1[3] + 1e-3f;        // 0 or normal (a sum with 1e-3 is exact near cancellation)
    f0h[0] = make_uint2(pk(S03 / 32.0f, S1[0]), pk(log1p_<FAST>(S1[1]), log1p_<FAST>(F::divs(fabsf(S1[2] - S1[3]), fd, F::rcpn(fd)))));
  } else if (role == 1) {
    Sum<KAHAN> qL, qAC[4], qMV;
    float la[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int t = t4[j];
      la[j] = sP[P0X][t];
      qL.add(la[j]);
#pragma unroll
      for (int c = 0; c < 4; c++) qAC[c].add(sP[P4X + c][t]);
      qMV.add(sP[P2W][t]);
    }
    float AC[4];
#pragma unroll
    for (int c = 0; c < 4; c++) AC[c] = qAC[c].get() * k;
    const float MV = qMV.get() * k;
    const float Lacc = qL.get() * k;
    Sum<KAHAN> qVA;
#pragma unroll
    for (int j = 0; j < 4; j++) qVA.add(ok[j] ? (la[j] - Lacc) * (la[j] - Lacc) : 0.0f);
    const float VA = qVA.get() * k;
    const float mu4 = mean4x4<KAHAN>(sI, sQ, byl, bxl);
    stu<BUF>(a.F2, X, Y, make_uint4(pk(1.0f - nv / 4.0f, log1p_<FAST>(MV)), pk(log1p_<FAST>(mu4), log1p_<FAST>(VA)),
                                    pk(AC[0], AC[1]), pk(AC[2], AC[3])));
  } else {
    Sum<KAHAN> qS3[4], qNx, qNy, qNz, qZ;
    float zs[4], nsx[4], nsy[4], nsz[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int t = t4[j];
      zs[j] = ok[j] ? sP[P2Z][t] : 0.0f;
      nsx[j] = sP[NSX][t]; nsy[j] = sP[NSY][t]; nsz[j] = sP[NSZ][t];
      const float2 m01 = unpack2(__float_as_uint(sP[P3A][t])), m23 = unpack2(__float_as_uint(sP[P3B][t]));
      const float p3[4] = {m01.x, m01.y, m23.x, m23.y};
#pragma unroll
      for (int c = 0; c < 4; c++) qS3[c].add(p3[c]);
      qNx.add(nsx[j]); qNy.add(nsy[j]); qNz.add(nsz[j]);
      qZ.add(zs[j]);
    }
    float S3[4];
#pragma unroll
    for (int c = 0; c < 4; c++) S3[c] = qS3[c].get() * k;
    const float Nx = qNx.get() * k, Ny = qNy.get() * k, Nz = qNz.get() * k, Z = qZ.get() * k;
    // sqrtn is sqrtf for |N|^2 >= 2^-96 or 0; a smaller |N|^2 has a root below 2^-48, and so has sqrtn's
    // result there, which the 1e-6 floor replaces either way
    const float ln = mx(F::sqrtn(Nx * Nx + Ny * Ny + Nz * Nz), 1e-6f);
    const float rln = F::rcpn(ln);
    const float3 Nn = make_float3(F::divsf(Nx, ln, rln), F::divsf(Ny, ln, rln), F::divsf(Nz, ln, rln));
    float zmax = -1e30f, zmin = 1e30f, mind = 1e30f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float d = dot3(make_float3(nsx[j], nsy[j], nsz[j]), Nn);
      zmax = ok[j] ? mx(zmax, zs[j]) : zmax; zmin = ok[j] ? mn(zmin, zs[j]) : zmin;
      mind = ok[j] ? mn(mind, d) : mind;
    }
    const float zd = mx(zmin, 1e-6f);     // in [1e-6, 65504]; the numerator is a difference of binary16 depths
    const float dz = mn(mx(F::divs(zmax - zmin, zd, F::rcpn(zd)), 0.0f), 1.0f);
    const float dn = mn(mx(1.0f - mind, 0.0f), 2.0f);
    const float2 oc = oct_enc_m<FAST>(Nn);
    stu<BUF>(a.F1, X, Y, make_uint4(pk(oc.x, oc.y), pk(dz, dn), pk(S3[0], S3[1]), pk(S3[2], S3[3])));
    const uint4 g3 = make_uint4(pk(oc.x, oc.y), pk(Z, nv), 0u, 0u);
    stu<BUF>(a.F3, X, Y, g3);
    const float2 go = unpack2(g3.x);
    const float3 gd = oct_dec(go.x, go.y);
    stf<BUF>(a.outGH, X, Y, make_float4(gd.x, gd.y, gd.z, __uint_as_float(g3.y)));
  }
}

// ============================================================================ edge-aware a-trous
// filt (atrousFn, hush.js:431-476) over an LDS tile (SW texels per row) holding the source (rgb, var) and
// the decoded guide (normal, depth) of clamp-to-edge texels; taps outside the image are skipped.

template <int R> struct Kern;
template <> struct Kern<1> { static constexpr double h[3] = {0.25, 0.5, 0.25}; };
template <> struct Kern<2> { static constexpr double h[5] = {1.0 / 16, 1.0 / 4, 3.0 / 8, 1.0 / 4, 1.0 / 16}; };

// sqrt(d2) for the squared tap distances of a (2R+1)^2 kernel
__host__ __device__ constexpr double tap_dist(int d2) {
  return d2 == 1 ? 1.0 : d2 == 2 ? 1.4142135623730951 : d2 == 4 ? 2.0 : d2 == 5 ? 2.23606797749979
       : d2 == 8 ? 2.8284271247461903 : 0.0;
}

// CHECK = false: the caller guarantees that every tap lies inside the image (an interior tile).  sL (when
// LUM): lum() of each staged texel's rgb, the value the taps would compute.
template <int R, int STEP, bool PRE, int SW, bool FAST, bool CHECK = true, bool LUM = false>
__device__ __forceinline__ float4 filt(const float4 *sA, const float4 *sG, int tx, int ty, int qx, int qy, int w, int h,
                                       float s, float cc, const float *sL = nullptr) {
  using F = Math<FAST>;
  const int c = ty * SW + tx;
  const float4 gp = sG[c], center = sA[c];
  float gv;
  if (PRE) {
    gv = 0.0625f * (sA[c - SW - 1].w + sA[c - SW + 1].w + sA[c + SW - 1].w + sA[c + SW + 1].w)
       + 0.125f * (sA[c - SW].w + sA[c - 1].w + sA[c + 1].w + sA[c + SW].w)
       + 0.25f * center.w;
  } else {
    gv = center.w;
  }
  const float lc = LUM ? sL[c] : lum(center.x, center.y, center.z);
  // s / sd: sd lies in [2e-6, 256] (a square root of at least 1e-12 of binary16 variances, plus 1e-6) and s is
  // 0.25 exp(clamp(o, +-5)) or the coarse level's 0.25, so the shared-reciprocal quotient is IEEE's
  const float sd = F::sqrtn(mx(gv, 1e-12f)) + 1e-6f;
  const float sl = FAST ? F::div(s, sd) : Math<false>::divs(s, sd, Math<false>::rcpn(sd));
  const float zl = sG[c - 1].w, zr = sG[c + 1].w, zu = sG[c - SW].w, zd = sG[c + SW].w;
  const float gx = mn(zr > 0.0f ? fabsf(zr - gp.w) : 1e9f, zl > 0.0f ? fabsf(zl - gp.w) : 1e9f);
  const float gy = mn(zd > 0.0f ? fabsf(zd - gp.w) : 1e9f, zu > 0.0f ? fabsf(zu - gp.w) : 1e9f);
  const float zg = mn(mx(gx, gy), gp.w);
  float iz[9];
#pragma unroll
  for (int d2 = 1; d2 <= 2 * R * R; d2++)
    if (tap_dist(d2) > 0.0) iz[d2] = F::rcpn(zg * (float)(tap_dist(d2) * STEP) + 1e-3f);
  const float cw = (float)(Kern<R>::h[R] * Kern<R>::h[R]) * cc;
  float sw = cw;
  float scx = center.x * cw, scy = center.y * cw, scz = center.z * cw;
  float sv = center.w * cw * cw;
#pragma unroll
  for (int dy = -R; dy <= R; dy++)
#pragma unroll
    for (int dx = -R; dx <= R; dx++) {
      if (dx == 0 && dy == 0) continue;
      const int o = dy * STEP * SW + dx * STEP;
      const float4 gq = sG[c + o];
      if ((!CHECK || inside(qx + dx * STEP, qy + dy * STEP, w, h)) && gq.w >= 0.0f) {
        const float4 cq = sA[c + o];
        const float lq = LUM ? sL[c + o] : lum(cq.x, cq.y, cq.z);
        const float ww = (float)(Kern<R>::h[dx + R] * Kern<R>::h[dy + R])
                       * pow128(mx(dot3(make_float3(gq.x, gq.y, gq.z), make_float3(gp.x, gp.y, gp.z)), 0.0f))
                       * F::expn(-fabsf(gq.w - gp.w) * iz[dx * dx + dy * dy] - sl * fabsf(lq - lc));
        sw += ww;
        scx += ww * cq.x; scy += ww * cq.y; scz += ww * cq.z;
        sv += ww * ww * cq.w;
      }
    }
  const float rsw = F::rcpn(sw);
  return make_float4(F::divsf(scx, sw, rsw), F::divsf(scy, sw, rsw), F::divsf(scz, sw, rsw), F::div(sv, sw * sw));
}

// decoded half-res guide from an F3 texel: octDec(oct normal), block depth (hush.js:337)
// LDS guide (normal, depth) from a decoded half-res guide texel (w = binary16 (Z, count))
__device__ __forceinline__ float4 guide_gh(float4 g) { return make_float4(g.x, g.y, g.z, pair_of(g.w).x); }
// decoded guide from an rgba16float (oct.x, oct.y, depth, .) texel (hush.js:497, 566-567)
template <bool FAST>
__device__ __forceinline__ float4 guide_tex(float4 g) {
  float3 n = oct_dec_m<FAST>(g.x, g.y);
  return make_float4(n.x, n.y, n.z, g.z);
}

// ============================================================================ pass 2: infer + mid1
// INFER (hush.js:312-368): MLP 20-h-h-12 on the block features -> par0/par1/par2, then a-trous 3x3 step 1
// with prefilter of the block means, s = qf16(s1), c = qf16(c1) -> Bh[0].

template <int HID>
__device__ __forceinline__ void mlp_fp32(const float *__restrict__ w, int hid, const float *f, float *o) {
  const int h = HID > 0 ? HID : hid;
  const float *W1 = w, *B1 = W1 + h * 20, *W2 = B1 + h, *B2 = W2 + h * h, *W3 = B2 + h, *B3 = W3 + 12 * h;
  float h1[HID > 0 ? HID : 256], h2[HID > 0 ? HID : 256];
#pragma unroll
  for (int j = 0; j < h; j++) {
    float acc = B1[j];
#pragma unroll
    for (int i = 0; i < 20; i++) acc = __builtin_fmaf(f[i], W1[j * 20 + i], acc);
    h1[j] = mx(acc, 0.0f);
  }
#pragma unroll
  for (int j = 0; j < h; j++) {
    float acc = B2[j];
#pragma unroll
    for (int i = 0; i < h; i++) acc = __builtin_fmaf(h1[i], W2[j * h + i], acc);
    h2[j] = mx(acc, 0.0f);
  }
#pragma unroll
  for (int j = 0; j < 12; j++) {
    float acc = B3[j];
#pragma unroll
    for (int i = 0; i < h; i++) acc = __builtin_fmaf(h2[i], W3[j * h + i], acc);
    o[j] = acc;
  }
}

// The FP32 MLP for hidden widths above 64, over the weights zero-padded to H (a multiple of 32) and stored
// input-major (W1t [20][H], b1 [H], W2t [H][H], b2 [H], W3t [H][12], b3 [12]): the units in chunks of 32
// accumulators held in registers, the hidden activations in per-lane memory, each read once per chunk.
// Every unit's sum runs over the same terms in the same order as mlp_fp32's (bias, then inputs ascending).
__device__ __forceinline__ void mlp_fp32_chunked(const float *__restrict__ w, int H, const float *f, float *o) {
  const float *W1 = w, *B1 = W1 + H * 20, *W2 = B1 + H, *B2 = W2 + H * H, *W3 = B2 + H, *B3 = W3 + 12 * H;
  float h1[256], h2[256];
  for (int j0 = 0; j0 < H; j0 += 32) {
    float acc[32];
#pragma unroll
    for (int k = 0; k < 32; k++) acc[k] = B1[j0 + k];
#pragma unroll
    for (int i = 0; i < 20; i++)
#pragma unroll
      for (int k = 0; k < 32; k++) acc[k] = __builtin_fmaf(f[i], W1[i * H + j0 + k], acc[k]);
#pragma unroll
    for (int k = 0; k < 32; k++) h1[j0 + k] = mx(acc[k], 0.0f);
  }
  for (int j0 = 0; j0 < H; j0 += 32) {
    float acc[32];
#pragma unroll
    for (int k = 0; k < 32; k++) acc[k] = B2[j0 + k];
    for (int i = 0; i < H; i++) {
      const float x = h1[i];
#pragma unroll
      for (int k = 0; k < 32; k++) acc[k] = __builtin_fmaf(x, W2[i * H + j0 + k], acc[k]);
    }
#pragma unroll
    for (int k = 0; k < 32; k++) h2[j0 + k] = mx(acc[k], 0.0f);
  }
  float acc[12];
#pragma unroll
  for (int j = 0; j < 12; j++) acc[j] = B3[j];
  for (int i = 0; i < H; i++) {
    const float x = h2[i];
#pragma unroll
    for (int j = 0; j < 12; j++) acc[j] = __builtin_fmaf(x, W3[i * 12 + j], acc[j]);
  }
#pragma unroll
  for (int j = 0; j < 12; j++) o[j] = acc[j];
}

// shader-f16 MLP of the reference's f16 option (hush.js:690, 699-702): binary16 inputs, weights,
// products and sums, each rounded to nearest even
__device__ __forceinline__ void mlp_f16(const float *__restrict__ w, int h, const float *f, float *o) {
  const float *W1 = w, *B1 = W1 + h * 20, *W2 = B1 + h, *B2 = W2 + h * h, *W3 = B2 + h, *B3 = W3 + 12 * h;
  _Float16 fh[20], h1[256], h2[256];
  for (int i = 0; i < 20; i++) fh[i] = (_Float16)f[i];
  for (int j = 0; j < h; j++) {
    _Float16 acc = (_Float16)B1[j];
    for (int i = 0; i < 20; i++) { _Float16 p = fh[i] * (_Float16)W1[j * 20 + i]; acc = acc + p; }
    h1[j] = acc > (_Float16)0 ? acc : (_Float16)0;
  }
  for (int j = 0; j < h; j++) {
    _Float16 acc = (_Float16)B2[j];
    for (int i = 0; i < h; i++) { _Float16 p = h1[i] * (_Float16)W2[j * h + i]; acc = acc + p; }
    h2[j] = acc > (_Float16)0 ? acc : (_Float16)0;
  }
  for (int j = 0; j < 12; j++) {
    _Float16 acc = (_Float16)B3[j];
    for (int i = 0; i < h; i++) { _Float16 p = h2[i] * (_Float16)W3[j * h + i]; acc = acc + p; }
    o[j] = (float)acc;
  }
}

// mlp_f16 at hidden width 16 on packed binary16 arithmetic: two units per instruction (v_pk_mul_f16,
// v_pk_add_f16), each half its own product and sum rounded to nearest even in the same order as mlp_f16's,
// so the same bits; the weights as binary16 pairs (MfmaPack::hp)
typedef _Float16 half2x __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void mlp_f16_16(const MfmaPack *__restrict__ P, const float *f, float *o) {
  const uint32_t *__restrict__ hp = P->hp;
  auto hw = [&](int k) { return __builtin_bit_cast(half2x, hp[k]); };
  _Float16 h1[16], h2[16];
#pragma unroll
  for (int p = 0; p < 8; p++) {
    half2x acc = hw(160 + p);
#pragma unroll
    for (int i = 0; i < 20; i++) {
      const _Float16 x = (_Float16)f[i];
      const half2x prod = half2x{x, x} * hw(p * 20 + i);
      acc = acc + prod;
    }
    h1[2 * p] = acc.x > (_Float16)0 ? acc.x : (_Float16)0;
    h1[2 * p + 1] = acc.y > (_Float16)0 ? acc.y : (_Float16)0;
  }
#pragma unroll
  for (int p = 0; p < 8; p++) {
    half2x acc = hw(296 + p);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const half2x prod = half2x{h1[i], h1[i]} * hw(168 + p * 16 + i);
      acc = acc + prod;
    }
    h2[2 * p] = acc.x > (_Float16)0 ? acc.x : (_Float16)0;
    h2[2 * p + 1] = acc.y > (_Float16)0 ? acc.y : (_Float16)0;
  }
#pragma unroll
  for (int p = 0; p < 6; p++) {
    half2x acc = hw(400 + p);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const half2x prod = half2x{h2[i], h2[i]} * hw(304 + p * 16 + i);
      acc = acc + prod;
    }
    o[2 * p] = (float)acc.x;
    o[2 * p + 1] = (float)acc.y;
  }
}

// activations (hush.js:357-366) and the four half-res outputs of one texel
template <bool FAST>
__device__ __forceinline__ void infer_outputs(const InferArgs &a, const float4 *sA, const float4 *sG, int tx, int ty,
                                              int X, int Y, const float *o) {
  using F = Math<FAST>;
  constexpr int SW = TW + 2;
  const float alpha = act_alpha(sigmoid<FAST>(o[0]));
  const float gamma = act_gamma(sigmoid<FAST>(o[1]));
  const float beta = sigmoid<FAST>(o[2]);
  const float kappa = sigmoid<FAST>(o[3]);
  float s[4], c[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    s[k] = 0.25f * F::expm(mn(mx(o[4 + k], -5.0f), 5.0f));
    c[k] = F::expm(mn(mx(o[8 + k], -4.0f), 4.0f));
  }
  st<true>(a.out0, X, Y, make_float4(alpha, gamma, beta, kappa));
  st<true>(a.out1, X, Y, make_float4(s[0], s[1], s[2], s[3]));
  st<true>(a.out2, X, Y, make_float4(c[0], c[1], c[2], c[3]));
  st<true>(a.outMid, X, Y, filt<1, 1, true, SW, FAST>(sA, sG, tx, ty, X, Y, a.W2, a.H2, qf16(s[1]), qf16(c[1])));
}

__device__ __forceinline__ void unpack_features(uint4 f0, uint4 f1, uint4 f2, float *f) {
  uint32_t u[10] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w, f2.x, f2.y};
#pragma unroll
  for (int i = 0; i < 10; i++) {
    float2 t = unpack2(u[i]);
    f[2 * i] = t.x; f[2 * i + 1] = t.y;
  }
}

// Each pass's tile body takes its workgroup index n within the pass and an LDS block of the size its
// *_lds() gives (in float4 units), so that one launch can run two independent passes side by side.
constexpr int infer_lds() { return 2 * (TW + 2) * (TH + 2); }

template <int HID, int MODE, bool FAST>   // MODE 0: fp32, 1: shader-f16 arithmetic
__device__ __forceinline__ void infer_tile(const InferArgs &a, int n, float4 *lds) {
  constexpr int AP = 1, SW = TW + 2, SH = TH + 2, SN = SW * SH;
  float4 *sA = lds, *sG = lds + SN;
  const int tid = threadIdx.x;
  const int2 org = tile_origin<TW, TH>(a.W2, a.H2, n);
  const int x0 = org.x, y0 = org.y;
  for (int i = tid; i < SN; i += NT) {
    int qx = clampi(x0 - AP + i % SW, 0, a.W2 - 1), qy = clampi(y0 - AP + i / SW, 0, a.H2 - 1);
    uint4 g2 = ldu32(a.F2, qx, qy);
    float2 a01 = unpack2(g2.z), a23 = unpack2(g2.w);
    sA[i] = make_float4(a01.x, a01.y, a23.x, a23.y);
    sG[i] = guide_gh(ldf<true>(a.GH, qx, qy));
  }
  __syncthreads();
  const int lx = tid & 31, ly = tid >> 5;
  const int X = x0 + lx, Y = y0 + ly;
  if (!inside(X, Y, a.W2, a.H2)) return;
  const int tx = lx + AP, ty = ly + AP;
  if (sG[ty * SW + tx].w < 0.0f) {
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    st<true>(a.out0, X, Y, z); st<true>(a.out1, X, Y, z); st<true>(a.out2, X, Y, z); st<true>(a.outMid, X, Y, z);
    return;
  }
  float f[20];
  unpack_features(ldu32(a.F0, X, Y), ldu32(a.F1, X, Y), ldu32(a.F2, X, Y), f);
  float o[12];
  if constexpr (MODE == 0 && HID < 0) mlp_fp32_chunked(a.wp, a.hpad, f, o);
  else if constexpr (MODE == 0) mlp_fp32<HID>(HID > 0 ? a.wp : a.w, a.hidden, f, o);
  else if constexpr (HID == 16) mlp_f16_16((const MfmaPack *)a.wq, f, o);
  else mlp_f16(a.w, a.hidden, f, o);
  infer_outputs<FAST>(a, sA, sG, tx, ty, X, Y, o);
}
template <int HID, int MODE, bool FAST>
__global__ __launch_bounds__(NT) void k_infer(InferArgs a) {
  __shared__ float4 lds[infer_lds()];
  infer_tile<HID, MODE, FAST>(a, (int)blockIdx.x, lds);
}

// ============================================================================ MFMA forms of the MLP
// The MLP 20-16-16-12 of 64 texels per wave, in four groups of 16 texels.  Each layer is D = W . X^T on a
// 16x16 MFMA tile: A = weight fragment (16 units x K), B = activations (K x 16 texels), D = 16 units x 16
// texels with fp32 accumulation; lane l holds D[4(l/16)+i][l%16].  Weights are the hidden-width-16 blob only.

typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t pack_h2(float a, float b) {       // binary16 pair, round to nearest even
  half2v h = {(_Float16)a, (_Float16)b};
  return __builtin_bit_cast(uint32_t, h);
}
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf2(float a, float b) {      // bfloat16 pair, round to nearest even
  const bf16x2v h = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, h);
}
// four floats -> four fp8 (E4M3) bytes, saturated to +-448, round to nearest even
__device__ __forceinline__ uint32_t pack_fp8x4(float a, float b, float c, float d) {
  const float M = 448.0f;
  int r = __builtin_amdgcn_cvt_pk_fp8_f32(fminf(fmaxf(a, -M), M), fminf(fmaxf(b, -M), M), 0, false);
  r = __builtin_amdgcn_cvt_pk_fp8_f32(fminf(fmaxf(c, -M), M), fminf(fmaxf(d, -M), M), r, true);
  return (uint32_t)r;
}
// one lane per weight-fragment slot: converts the fp32-widened blob into the fragment sets with the
// conversion instructions the MFMAs read, so the formats agree by construction.  sw: the fp8 weight scale of
// each layer; sx: the fp8 static activation scale of each layer's input (powers of two).
__global__ __launch_bounds__(64) void k_pack_mlp(const float *__restrict__ w, MfmaPack *p, float sw1, float sw2,
                                                 float sw3, float sx1, float sx2, float sx3) {
  const int l = threadIdx.x, m = l & 15, kb = l >> 4;
  const float *W1 = w, *B1 = W1 + 320, *W2 = B1 + 16, *B2 = W2 + 256, *W3 = B2 + 16, *B3 = W3 + 192;
  auto w1 = [&](int mm, int k) { return k < 20 ? W1[mm * 20 + k] : 0.0f; };
  auto w2 = [&](int mm, int k) { return k < 16 ? W2[mm * 16 + k] : 0.0f; };
  auto w3 = [&](int mm, int k) { return (mm < 12 && k < 16) ? W3[mm * 16 + k] : 0.0f; };
  float v1[8], v[2][4];
#pragma unroll
  for (int i = 0; i < 8; i++) v1[i] = w1(m, 8 * kb + i);
#pragma unroll
  for (int i = 0; i < 4; i++) { v[0][i] = w2(m, 4 * kb + i); v[1][i] = w3(m, 4 * kb + i); }
  p->h1[l] = make_uint4(pack_h2(v1[0], v1[1]), pack_h2(v1[2], v1[3]), pack_h2(v1[4], v1[5]), pack_h2(v1[6], v1[7]));
  p->b1w[l] = make_uint4(pack_bf2(v1[0], v1[1]), pack_bf2(v1[2], v1[3]), pack_bf2(v1[4], v1[5]), pack_bf2(v1[6], v1[7]));
#pragma unroll
  for (int b = 0; b < 2; b++) {
    p->h[b][l] = make_uint2(pack_h2(v[b][0], v[b][1]), pack_h2(v[b][2], v[b][3]));
    p->b[b][l] = make_uint2(pack_bf2(v[b][0], v[b][1]), pack_bf2(v[b][2], v[b][3]));
  }
  // fp8: layer 1 takes k = 8kb .. 8kb+7 as the f16 form does; layers 2 and 3 put hidden unit 4kb+j at
  // k = 8kb+j (j < 4) and zero weights at 8kb+4 .. 8kb+7, so each lane's B fragment is its own D fragment of
  // the layer before, padded with zeros
  for (int k = l; k < 406; k += 64) {
    float u, v;
    if (k < 160) { u = W1[(k / 20) * 40 + k % 20]; v = W1[(k / 20) * 40 + 20 + k % 20]; }
    else if (k < 168) { u = B1[2 * (k - 160)]; v = B1[2 * (k - 160) + 1]; }
    else if (k < 296) { const int q = k - 168; u = W2[(q / 16) * 32 + q % 16]; v = W2[(q / 16) * 32 + 16 + q % 16]; }
    else if (k < 304) { u = B2[2 * (k - 296)]; v = B2[2 * (k - 296) + 1]; }
    else if (k < 400) { const int q = k - 304; u = W3[(q / 16) * 32 + q % 16]; v = W3[(q / 16) * 32 + 16 + q % 16]; }
    else { u = B3[2 * (k - 400)]; v = B3[2 * (k - 400) + 1]; }
    p->hp[k] = pack_h2(u, v);                      // binary16 weights widened to fp32: exact
  }
  const float r1 = 1.0f / sw1, r2 = 1.0f / sw2, r3 = 1.0f / sw3;     // powers of two: exact
  float q[3][8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    q[0][i] = w1(m, 8 * kb + i) * r1;
    q[1][i] = i < 4 ? w2(m, 4 * kb + i) * r2 : 0.0f;
    q[2][i] = i < 4 ? w3(m, 4 * kb + i) * r3 : 0.0f;
  }
#pragma unroll
  for (int L = 0; L < 3; L++)
    p->q[L][l] = make_uint2(pack_fp8x4(q[L][0], q[L][1], q[L][2], q[L][3]), pack_fp8x4(q[L][4], q[L][5], q[L][6], q[L][7]));
  if (l < 16) {
    p->b1[l] = B1[l]; p->b2[l] = B2[l]; p->b3[l] = l < 12 ? B3[l] : 0.0f;
    p->qb1[l] = B1[l] / sx2; p->qb2[l] = B2[l] / sx3;                  // powers of two: exact
  }
  if (l == 0) {
    p->sw[0] = sw1; p->sw[1] = sw2; p->sw[2] = sw3;
    p->sx[0] = sx1; p->sx[1] = sx2; p->sx[2] = sx3;
    p->xclamp = pack_h2(fminf(448.0f * sx1, 65504.0f), fminf(448.0f * sx1, 65504.0f));   // exact for sx1 >= 2^-20
    p->qc[0] = sx1 * sw1 / sx2; p->qc[1] = sx2 * sw2 / sx3; p->qc[2] = sx3 * sw3;
  }
}

// bias + ReLU on a D fragment (units 4kb+i)
__device__ __forceinline__ f32x4 bias_relu(f32x4 d, const float *b, int kb) {
  f32x4 r;
#pragma unroll
  for (int i = 0; i < 4; i++) r[i] = mx(d[i] + b[4 * kb + i], 0.0f);
  return r;
}

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// eight binary16 features (four pairs) -> eight bfloat16, round to nearest even (each binary16 value is exact
// in fp32)
__device__ __forceinline__ uint4 bf16x8_of(uint4 x) {
  const float2 a = unpack2(x.x), b = unpack2(x.y), c = unpack2(x.z), d = unpack2(x.w);
  return make_uint4(pack_bf2(a.x, a.y), pack_bf2(b.x, b.y), pack_bf2(c.x, c.y), pack_bf2(d.x, d.y));
}

// The MFMA chains of one wave's 64 texels as four independent 16-texel tiles, interleaved layer by layer so
// the matrix cores overlap one tile's latency with the others.  xf[g]: lane (kb, tl)'s layer-1 B fragment of
// tile g, features 8kb .. 8kb+7 of its texel as binary16 pairs (zero beyond feature 19); d[g] receives units
// 4kb .. 4kb+3 of the output layer, bias included.
//   binary16 / bfloat16: layer 1 is one v_mfma_f32_16x16x32_f16 / _bf16 (K = 20 padded to 32); layers 2 and
//     3 are 16x16x16, whose B fragment (k = 4kb .. 4kb+3) is exactly the D fragment of the layer before, so
//     the hidden activations never leave the registers.
//   fp8 (E4M3): three v_mfma_f32_16x16x32_fp8_fp8 with a power-of-two scale per layer for the weights (sw)
//     and a static one per layer for its input (sx, set on the host: no reductions in the kernel); layers 2
//     and 3 use half of K (each lane's 4 hidden units and 4 zeros), so the hidden values stay in their lanes.
template <int MODE>
__device__ __forceinline__ void mlp_mfma_x4(const MfmaPack *__restrict__ P, const uint4 *xf, f32x4 *d, int lane) {
  const int kb = lane >> 4, tl = lane & 15;
  const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
  if constexpr (MODE == MLP_FP16_MFMA || MODE == MLP_BF16_MFMA) {
    uint2 hb[4];
    if constexpr (MODE == MLP_FP16_MFMA) {
      const f16x8 w1 = __builtin_bit_cast(f16x8, P->h1[lane]);
#pragma unroll
      for (int g = 0; g < 4; g++) d[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w1, __builtin_bit_cast(f16x8, xf[g]), z, 0, 0, 0);
    } else {
      const bf16x8 w1 = __builtin_bit_cast(bf16x8, P->b1w[lane]);
#pragma unroll
      for (int g = 0; g < 4; g++)
        d[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, __builtin_bit_cast(bf16x8, bf16x8_of(xf[g])), z, 0, 0, 0);
    }
#pragma unroll
    for (int L = 0; L < 2; L++) {
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const f32x4 h = bias_relu(d[g], L == 0 ? P->b1 : P->b2, kb);
        hb[g] = MODE == MLP_FP16_MFMA ? make_uint2(pack_h2(h[0], h[1]), pack_h2(h[2], h[3]))
                                      : make_uint2(pack_bf2(h[0], h[1]), pack_bf2(h[2], h[3]));
      }
      if constexpr (MODE == MLP_FP16_MFMA) {
        const f16x4 wl = __builtin_bit_cast(f16x4, P->h[L][lane]);
#pragma unroll
        for (int g = 0; g < 4; g++) d[g] = __builtin_amdgcn_mfma_f32_16x16x16f16(wl, __builtin_bit_cast(f16x4, hb[g]), z, 0, 0, 0);
      } else {
        const s16x4 wl = __builtin_bit_cast(s16x4, P->b[L][lane]);
#pragma unroll
        for (int g = 0; g < 4; g++) d[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wl, __builtin_bit_cast(s16x4, hb[g]), z, 0, 0, 0);
      }
    }
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int i = 0; i < 4; i++) d[g][i] = d[g][i] + P->b3[4 * kb + i];
  } else {
    // layer-1 input: the binary16 features clamped to +-448 sx[0] (packed binary16 min / max), then one scaled
    // conversion per pair (E4M3 of x / sx[0], round to nearest even; it does not saturate by itself)
    const half2v cl = __builtin_bit_cast(half2v, P->xclamp), ncl = -cl;
    const float s1 = P->sx[0];
    auto cvt4 = [&](uint32_t lo, uint32_t hi) {
      const half2v a = __builtin_elementwise_min(__builtin_elementwise_max(__builtin_bit_cast(half2v, lo), ncl), cl);
      const half2v b = __builtin_elementwise_min(__builtin_elementwise_max(__builtin_bit_cast(half2v, hi), ncl), cl);
      s16x2 r = {0, 0};
      r = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(r, a, s1, false);
      r = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(r, b, s1, true);
      return __builtin_bit_cast(uint32_t, r);
    };
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const uint2 xb = make_uint2(cvt4(xf[g].x, xf[g].y), cvt4(xf[g].z, xf[g].w));
      d[g] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8((long)__builtin_bit_cast(uint64_t, P->q[0][lane]),
                                                        (long)__builtin_bit_cast(uint64_t, xb), z, 0, 0, 0);
    }
    // layers 2 and 3: lane (kb, tl) feeds its own units 4kb .. 4kb+3 at k = 8kb .. 8kb+3 and zeros above
    // (the weights there are zero too), so no value crosses lanes
    // the next layer's E4M3 input directly: (d sx sw + b) / sx' = d qc + qb (sx' the next layer's input
    // scale, folded into qc and the prescaled bias qb), ReLU and saturation at 448 in one median
#pragma unroll
    for (int L = 1; L < 3; L++) {
      const float sc = P->qc[L - 1];
      const float *bias = L == 1 ? P->qb1 : P->qb2;
#pragma unroll
      for (int g = 0; g < 4; g++) {
        float h[4];
#pragma unroll
        for (int i = 0; i < 4; i++) h[i] = __builtin_amdgcn_fmed3f(__builtin_fmaf(d[g][i], sc, bias[4 * kb + i]), 0.0f, 448.0f);
        int r = __builtin_amdgcn_cvt_pk_fp8_f32(h[0], h[1], 0, false);
        r = __builtin_amdgcn_cvt_pk_fp8_f32(h[2], h[3], r, true);
        const uint2 xb = make_uint2((uint32_t)r, 0u);
        d[g] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8((long)__builtin_bit_cast(uint64_t, P->q[L][lane]),
                                                          (long)__builtin_bit_cast(uint64_t, xb), z, 0, 0, 0);
      }
    }
    const float sc = P->qc[2];
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int i = 0; i < 4; i++) d[g][i] = __builtin_fmaf(d[g][i], sc, P->b3[4 * kb + i]);
  }
}

// layer-1 B fragment of lane (kb, tl) from the feature surfaces: F0 (kb 0), F1 (kb 1), F2.xy (kb 2), zero
__device__ __forceinline__ uint4 feature_frag(const TexU &F0, const TexU &F1, const TexU &F2, int x, int y, int kb, bool live) {
  if (!live || kb == 3) return make_uint4(0u, 0u, 0u, 0u);
  if (kb == 2) {
    const uint2 v = *(const uint2 *)texel<true>(F2.p, F2.pitch, x, y);
    return make_uint4(v.x, v.y, 0u, 0u);
  }
  const TexU &F = kb == 0 ? F0 : F1;
  return *texel<true>(F.p, F.pitch, x, y);
}

// MFMA forms in the frame's inference pass.  Lane (kb, tl) of wave v loads its B fragments of the wave's four
// 16-texel tiles straight from F0 / F1 / F2 (tile g = texels 16(g&1) .. +15 of tile row 2v + (g>>1)), the
// four MFMA chains run interleaved, and the D fragments (+ bias) go through LDS at a 48-byte texel stride
// (16-byte stores and loads, conflict free) so each lane ends with the 12 outputs of its own texel.
// LDS: the staged tile, then per wave the 64 x 3 output float4s
template <int MODE>
constexpr int infer_mfma_lds() { return infer_lds() + 4 * 64 * 3; }
template <int MODE, bool FAST>
__device__ __forceinline__ void infer_mfma_tile(const InferArgs &a, int n, float4 *lds) {
  constexpr int AP = 1, SW = TW + 2, SH = TH + 2, SN = SW * SH, NI = (SN + NT - 1) / NT;
  float4 *sA = lds, *sG = lds + SN;
  const int tid = threadIdx.x;
  const int2 org = tile_origin<TW, TH>(a.W2, a.H2, n);
  const int x0 = org.x, y0 = org.y;
  const int lx = tid & 31, ly = tid >> 5;
  const int X = x0 + lx, Y = y0 + ly;
  const int wv = tid >> 6, lane = tid & 63, kb = lane >> 4, tl = lane & 15;
  const MfmaPack *__restrict__ P = (const MfmaPack *)a.wq;
  float4 (*sO)[64 * 3] = (float4 (*)[64 * 3])(lds + 2 * SN);
  // every load before the first wait: this lane's B fragments, then this thread's staged texels (clamped
  // addresses past the staged block: loaded and discarded)
  uint4 xf[4];
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int gx = x0 + 16 * (g & 1) + tl, gy = y0 + 2 * wv + (g >> 1);
    xf[g] = feature_frag(a.F0, a.F1, a.F2, gx, gy, kb, inside(gx, gy, a.W2, a.H2));
  }
  uint4 g2r[NI];
  float4 ghr[NI];
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = mn(tid + k * NT, SN - 1);
    const int qx = clampi(x0 - AP + i % SW, 0, a.W2 - 1), qy = clampi(y0 - AP + i / SW, 0, a.H2 - 1);
    g2r[k] = ldu32(a.F2, qx, qy);
    ghr[k] = ldf<true>(a.GH, qx, qy);
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = tid + k * NT;
    if (i < SN) {
      const float2 a01 = unpack2(g2r[k].z), a23 = unpack2(g2r[k].w);
      sA[i] = make_float4(a01.x, a01.y, a23.x, a23.y);
      sG[i] = guide_gh(ghr[k]);
    }
  }
  f32x4 d[4];
  mlp_mfma_x4<MODE>(P, xf, d, lane);
  if (kb < 3) {
#pragma unroll
    for (int g = 0; g < 4; g++)
      sO[wv][(16 * g + tl) * 3 + kb] = make_float4(d[g][0], d[g][1], d[g][2], d[g][3]);
  }
  __syncthreads();
  if (!inside(X, Y, a.W2, a.H2)) return;
  const int tx = lx + AP, ty = ly + AP;
  if (sG[ty * SW + tx].w < 0.0f) {
    const float4 zr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    st<true>(a.out0, X, Y, zr); st<true>(a.out1, X, Y, zr); st<true>(a.out2, X, Y, zr); st<true>(a.outMid, X, Y, zr);
    return;
  }
  float o[12];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const float4 v = sO[wv][lane * 3 + j];
    o[4 * j] = v.x; o[4 * j + 1] = v.y; o[4 * j + 2] = v.z; o[4 * j + 3] = v.w;
  }
  infer_outputs<FAST>(a, sA, sG, tx, ty, X, Y, o);
}
template <int MODE, bool FAST>
__global__ __launch_bounds__(NT) void k_infer_mfma(InferArgs a) {
  __shared__ float4 lds[infer_mfma_lds<MODE>()];
  infer_mfma_tile<MODE, FAST>(a, (int)blockIdx.x, lds);
}

// The MLP's 12 outputs (before the activations, bias included) of every half-res texel, read from F0..F2
// in place with the binary16 / bfloat16 MFMA forms: the training forward's MFMA option.  Texels in raster
// order, 64 per wave in four 16-texel tiles; lane (kb, tl) of tile g loads the B fragment of texel
// 64w + 16g + tl as k_infer_mfma does.  Sky texels are computed too (the caller masks them).
template <int MODE>
__global__ __launch_bounds__(NT) void k_mlp_tex(TexU F0, TexU F1, TexU F2, int W2, int H2, float *__restrict__ out,
                                                 const void *__restrict__ wq) {
  const MfmaPack *__restrict__ P = (const MfmaPack *)wq;
  const int tid = threadIdx.x, lane = tid & 63, kb = lane >> 4, tl = lane & 15;
  const long base = (long)blockIdx.x * NT + (tid & ~63), n = (long)W2 * H2;
  uint4 xf[4];
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const long r = base + 16 * g + tl;
    const bool live = r < n;
    xf[g] = feature_frag(F0, F1, F2, live ? (int)(r % W2) : 0, live ? (int)(r / W2) : 0, kb, live);
  }
  f32x4 d[4];
  mlp_mfma_x4<MODE>(P, xf, d, lane);
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const long r = base + 16 * g + tl;
    if (r < n && kb < 3)
      *(float4 *)(out + r * 12 + 4 * kb) = make_float4(d[g][0], d[g][1], d[g][2], d[g][3]);
  }
}

// The MLP alone over n feature rows (20 binary16 each, row-major, 40 B) -> n x 12 fp32 outputs.  Used to
// time the MLP stage in isolation and by training tools.  The MFMA forms read their B fragments and write
// their D fragments straight from / to global memory (no LDS): lane l of group g serves row
// 64w + 16g + l%16, k block l/16, and writes outputs 4(l/16) .. +3 of that row as one 16-byte store.
template <int MODE>
__global__ __launch_bounds__(NT) void k_mlp(const uint32_t *__restrict__ feat, float *__restrict__ out, int n,
                                             const float *__restrict__ w, const void *__restrict__ wq) {
  const int tid = threadIdx.x;
  if constexpr (MODE == MLP_FP32 || MODE == MLP_F16) {
    const long r = (long)blockIdx.x * NT + tid;
    if (r >= n) return;
    float f[20], o[12];
    const uint2 *row = (const uint2 *)(feat + r * 10);
#pragma unroll
    for (int j = 0; j < 5; j++) {
      const uint2 v = row[j];
      const float2 a = unpack2(v.x), b = unpack2(v.y);
      f[4 * j] = a.x; f[4 * j + 1] = a.y; f[4 * j + 2] = b.x; f[4 * j + 3] = b.y;
    }
    if (MODE == MLP_FP32) mlp_fp32<16>(w, 16, f, o);
    else mlp_f16_16((const MfmaPack *)wq, f, o);
    float4 *dst = (float4 *)(out + r * 12);
    dst[0] = make_float4(o[0], o[1], o[2], o[3]);
    dst[1] = make_float4(o[4], o[5], o[6], o[7]);
    dst[2] = make_float4(o[8], o[9], o[10], o[11]);
  } else {
    const MfmaPack *__restrict__ P = (const MfmaPack *)wq;
    const int lane = tid & 63, kb = lane >> 4, tl = lane & 15;
    const long base = ((long)blockIdx.x * NT + (tid & ~63));          // first row of this wave
    // the wave's 64 rows as four 16-row tiles: every load first (rows are 40 B, 8-byte aligned: words
    // 4kb .. 4kb+3 as two 8-byte loads), then the four MFMA chains interleaved
    uint4 xf[4];
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const long r = base + 16 * g + tl;
      const bool live = r < n && kb < 3;
      const uint2 *row = (const uint2 *)(feat + (live ? r : 0) * 10);
      const uint2 lo = live ? row[2 * kb] : make_uint2(0u, 0u);
      const uint2 hi = (live && kb < 2) ? row[2 * kb + 1] : make_uint2(0u, 0u);
      xf[g] = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
    f32x4 d[4];
    mlp_mfma_x4<MODE>(P, xf, d, lane);
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const long r = base + 16 * g + tl;
      if (r < n && kb < 3)
        *(float4 *)(out + r * 12 + 4 * kb) = make_float4(d[g][0], d[g][1], d[g][2], d[g][3]);
    }
  }
}

// ============================================================================ passes 3, 4, 6: a-trous
// atrousShader (hush.js:482-525): mid2 (R1 step 2) and mid3 (R1 step 4) at half res with the F3 guide
// and learned s/c from par1/par2; coarse2 (R2 step 2, prefilter) at quarter res with the CG guide.

// GUIDE_PACKED: guide w = binary16 (Z, count) (GH); else w = depth (CGD)
template <int R, int STEP>
constexpr int atrous_lds() {
  constexpr int AP = R * STEP > 1 ? R * STEP : 1;
  return 2 * (TW + 2 * AP) * (TH + 2 * AP);
}
template <int R, int STEP, bool PRE, bool GUIDE_PACKED, int COMP, bool FAST>
__device__ __forceinline__ void atrous_tile(const AtrousArgs &a, int n, float4 *lds) {
  constexpr int AP = R * STEP > 1 ? R * STEP : 1;
  constexpr int SW = TW + 2 * AP, SH = TH + 2 * AP, SN = SW * SH, NI = (SN + NT - 1) / NT;
  float4 *sA = lds, *sG = lds + SN;
  const int tid = threadIdx.x;
  const int2 org = tile_origin<TW, TH>(a.w, a.h, n);
  const int x0 = org.x, y0 = org.y;
  const int lx = tid & 31, ly = tid >> 5;
  const int X = x0 + lx, Y = y0 + ly;
  // Every load before the first wait: the pixel's own parameters, then the NI staged texels of this thread.
  // Addresses are clamped (a lane outside the level, or past the staged block, loads a texel it discards),
  // so no branch surrounds a load.
  uint2 p1r = make_uint2(0u, 0u), p2r = p1r;
  if (COMP >= 0) {
    const int cX = clampi(X, 0, a.w - 1), cY = clampi(Y, 0, a.h - 1);
    p1r = *texel<true>(a.par1.p, a.par1.pitch, cX, cY);
    p2r = *texel<true>(a.par2.p, a.par2.pitch, cX, cY);
  }
  uint2 ar[NI];
  float4 gr[NI];
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = mn(tid + k * NT, SN - 1);
    const int qx = clampi(x0 - AP + i % SW, 0, a.w - 1), qy = clampi(y0 - AP + i / SW, 0, a.h - 1);
    ar[k] = *texel<true>(a.inA.p, a.inA.pitch, qx, qy);
    gr[k] = *texel<true>(a.guide.p, a.guide.pitch, qx, qy);
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = tid + k * NT;
    if (i < SN) {
      sA[i] = unpack4(ar[k]);
      sG[i] = GUIDE_PACKED ? guide_gh(gr[k]) : gr[k];
    }
  }
  __syncthreads();
  if (!inside(X, Y, a.w, a.h)) return;
  const int tx = lx + AP, ty = ly + AP;
  if (sG[ty * SW + tx].w < 0.0f) { st<true>(a.out, X, Y, make_float4(0.0f, 0.0f, 0.0f, 0.0f)); return; }
  float s, c;
  if (COMP < 0) { s = a.s; c = 1.0f; }
  else {
    const float4 p1 = unpack4(p1r), p2 = unpack4(p2r);
    s = COMP == 0 ? p1.x : COMP == 1 ? p1.y : COMP == 2 ? p1.z : p1.w;
    c = COMP == 0 ? p2.x : COMP == 1 ? p2.y : COMP == 2 ? p2.z : p2.w;
  }
  // interior tiles (every tap of every pixel inside the level) skip the per-tap bounds tests
  const bool interior = x0 >= AP && y0 >= AP && x0 + TW + AP <= a.w && y0 + TH + AP <= a.h;
  if (interior) st<true>(a.out, X, Y, filt<R, STEP, PRE, SW, FAST, false>(sA, sG, tx, ty, X, Y, a.w, a.h, s, c));
  else st<true>(a.out, X, Y, filt<R, STEP, PRE, SW, FAST>(sA, sG, tx, ty, X, Y, a.w, a.h, s, c));
}
template <int R, int STEP, bool PRE, bool GUIDE_PACKED, int COMP, bool FAST>
__global__ __launch_bounds__(NT) void k_atrous(AtrousArgs a) {
  __shared__ float4 lds[atrous_lds<R, STEP>()];
  atrous_tile<R, STEP, PRE, GUIDE_PACKED, COMP, FAST>(a, (int)blockIdx.x, lds);
}

// ============================================================================ pass 5: coarse1
// COARSE1 (hush.js:372-422): count-weighted 2x2 half -> quarter down() of the block means and guides into
// the tile, the coarse guide CG written once per texel, then a-trous 5x5 step 1 with prefilter -> C[1].

template <bool FAST>
__device__ __forceinline__ void down(const TexU &F2, const TexF &GH, int QX, int QY, int W2, int H2, float4 &col, float4 &g) {
  float N = 0.0f, vr = 0.0f, z = 0.0f, clx = 0.0f, cly = 0.0f, clz = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f;
  // every tap's guide and level texel are loaded at once (clamped into the level); a tap outside the level or
  // with no valid pixel takes weight 0 and adds exact zeros (the staged values are finite and every sum
  // starts at +0), so the sums are those of the skipping form bit for bit
  float4 ghv[4];
  uint4 g2v[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int hx = mn(QX * 2 + (j & 1), W2 - 1), hy = mn(QY * 2 + (j >> 1), H2 - 1);
    ghv[j] = ldf<true>(GH, hx, hy);
    g2v[j] = ldu32(F2, hx, hy);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int hx = QX * 2 + (j & 1), hy = QY * 2 + (j >> 1);
    const float4 gh = ghv[j];
    const float2 oz = pair_of(gh.w);
    const uint4 g2 = g2v[j];
    const float2 a01 = unpack2(g2.z), a23 = unpack2(g2.w);
    const float w = ((hx < W2) & (hy < H2) & (oz.y > 0.0f)) ? oz.y : 0.0f;
    N += w;
    clx += w * a01.x; cly += w * a01.y; clz += w * a23.x;
    vr += w * a23.y;
    nx += w * gh.x; ny += w * gh.y; nz += w * gh.z;
    z += w * oz.x;
  }
  if (N == 0.0f) { col = make_float4(0.0f, 0.0f, 0.0f, 0.0f); g = make_float4(0.0f, 0.0f, -1.0f, 0.0f); return; }
  using F = Math<false>;   // stored level values: IEEE under both policies (they set the slope sl)
  float ln = mx(F::sqrtn(nx * nx + ny * ny + nz * nz), 1e-6f);   // below 2^-96 the 1e-6 floor decides, as in block()
  const float rln = F::rcpn(ln), rN = F::rcpn(N);
  float2 o = oct_enc_m<FAST>(make_float3(F::divsf(nx, ln, rln), F::divsf(ny, ln, rln), F::divsf(nz, ln, rln)));
  col = make_float4(qf16(F::divsf(clx, N, rN)), qf16(F::divsf(cly, N, rN)), qf16(F::divsf(clz, N, rN)),
                    qf16(F::divsf(F::divsf(vr, N, rN), N, rN)));
  g = make_float4(qf16(o.x), qf16(o.y), qf16(F::divsf(z, N, rN)), 0.0f);
}

constexpr int coarse1_lds() { return 2 * (TW + 4) * (TH + 4); }
template <bool FAST>
__device__ __forceinline__ void coarse1_tile(const Coarse1Args &a, int n, float4 *lds) {
  constexpr int AP = 2, SW = TW + 2 * AP, SH = TH + 2 * AP, SN = SW * SH;
  float4 *sA = lds, *sG = lds + SN;
  const int tid = threadIdx.x;
  const int2 org = tile_origin<TW, TH>(a.W4, a.H4, n);
  const int x0 = org.x, y0 = org.y;
  for (int i = tid; i < SN; i += NT) {
    int lqx = x0 - AP + i % SW, lqy = y0 - AP + i / SW;
    int qx = clampi(lqx, 0, a.W4 - 1), qy = clampi(lqy, 0, a.H4 - 1);
    float4 col, g;
    down<FAST>(a.F2, a.GH, qx, qy, a.W2, a.H2, col, g);
    sA[i] = col;
    sG[i] = guide_tex<FAST>(g);
    if (lqx == qx && lqy == qy && lqx >= x0 && lqy >= y0 && lqx < x0 + TW && lqy < y0 + TH) {
      const uint2 gp = pack4(g);
      *texel<true>(a.outG.p, a.outG.pitch, qx, qy) = gp;
      const float4 gs = unpack4(gp);
      const float3 d = oct_dec(gs.x, gs.y);
      stf<true>(a.outGD, qx, qy, make_float4(d.x, d.y, d.z, gs.z));
    }
  }
  __syncthreads();
  const int lx = tid & 31, ly = tid >> 5;
  const int X = x0 + lx, Y = y0 + ly;
  if (!inside(X, Y, a.W4, a.H4)) return;
  const int tx = lx + AP, ty = ly + AP;
  if (sG[ty * SW + tx].w < 0.0f) { st<true>(a.outC, X, Y, make_float4(0.0f, 0.0f, 0.0f, 0.0f)); return; }
  st<true>(a.outC, X, Y, filt<2, 1, true, SW, FAST>(sA, sG, tx, ty, X, Y, a.W4, a.H4, a.s, 1.0f));
}
template <bool FAST>
__global__ __launch_bounds__(NT) void k_coarse1(Coarse1Args a) {
  __shared__ float4 lds[coarse1_lds()];
  coarse1_tile<FAST>(a, (int)blockIdx.x, lds);
}

// ============================================================================ two passes in one launch
// The coarse passes depend only on their own inputs: coarse1 on the temporal pass (like infer + mid1) and
// coarse2 on coarse1 (mid3 on mid2).  Running each beside a pass of the main chain in one launch fills the
// small coarse grids into that launch instead of a second stream, whose cross-stream waits leave the GPU
// idle for several microseconds each.  Workgroups [0, nA) run the coarse pass's tiles, [nApad, ...) the
// other pass's; nApad = nA rounded up to a multiple of 8 keeps every tile on the XCD its own launch would
// give it.  The LDS block is the larger of the two passes' blocks.
constexpr int imax(int a, int b) { return a > b ? a : b; }

template <int HID, int MODE, bool FAST>
__global__ __launch_bounds__(NT) void k_coarse1_infer(Coarse1Args c, InferArgs a, int nA, int nApad) {
  __shared__ float4 lds[imax(coarse1_lds(), infer_lds())];
  const int b = (int)blockIdx.x;
  if (b < nApad) {
    if (b < nA) coarse1_tile<FAST>(c, b, lds);
    return;
  }
  infer_tile<HID, MODE, FAST>(a, b - nApad, lds);
}
template <int MODE, bool FAST>
__global__ __launch_bounds__(NT) void k_coarse1_infer_mfma(Coarse1Args c, InferArgs a, int nA, int nApad) {
  __shared__ float4 lds[imax(coarse1_lds(), infer_mfma_lds<MODE>())];
  const int b = (int)blockIdx.x;
  if (b < nApad) {
    if (b < nA) coarse1_tile<FAST>(c, b, lds);
    return;
  }
  infer_mfma_tile<MODE, FAST>(a, b - nApad, lds);
}
template <bool FAST>
__global__ __launch_bounds__(NT) void k_coarse2_mid3(AtrousArgs c, AtrousArgs m, int nA, int nApad) {
  __shared__ float4 lds[imax(atrous_lds<2, 2>(), atrous_lds<1, 4>())];
  const int b = (int)blockIdx.x;
  if (b < nApad) {
    if (b < nA) atrous_tile<2, 2, true, false, -1, FAST>(c, b, lds);
    return;
  }
  atrous_tile<1, 4, false, true, 3, FAST>(m, b - nApad, lds);
}

// ============================================================================ pass 7: fine + final
// FINE_FINAL (hush.js:530-622): a-trous 3x3 step 1 with prefilter of A0 (guide key[cur]) with s0/c0 ->
// histColor[cur]; joint-bilateral upsample of the mid (Bh[0]) and coarse (C[0]) levels; recombine with
// kappa and beta; remodulate -> X.

// joint-bilateral weight of one upsampling tap; the depth term's divisor 0.05 z + 1e-3 is the pixel's own,
// shared by its eight taps (jd, rjd = rcp(jd))
template <bool FAST>
__device__ __forceinline__ float jbw(float frx, float fry, int j, float3 nq, float zq, float3 n, float z, float jd,
                                     float rjd) {
  float wb = ((j & 1) == 1 ? frx : 1.0f - frx) * ((j >> 1) == 1 ? fry : 1.0f - fry);
  float dn = mx(dot3(nq, n), 0.0f);
  float dn2 = dn * dn;
  float dn4 = dn2 * dn2;
  return wb * Math<FAST>::expn(Math<FAST>::divsf(-fabsf(zq - z), jd, rjd)) * dn4 * dn4;
}

// The upsampling taps of a 32x8 tile fall in an 18x6 window of the half-res level (origin x0/2 - 1,
// y0/2 - 1: pixel 2k takes texels k-1, k and pixel 2k+1 texels k, k+1) and a 10x4 window of the
// quarter-res level (origin x0/4 - 1, y0/4 - 1), and its par texels are the 16x4 half-res texels
// (x0/2, y0/2) ..; all three are staged in LDS once per tile.
constexpr int HWW = TW / 2 + 2, HWH = TH / 2 + 2, HWN = HWW * HWH;      // 18 x 6
constexpr int QWW = TW / 4 + 2, QWH = TH / 4 + 2, QWN = QWW * QWH;      // 10 x 4
constexpr int PWN = (TW / 2) * (TH / 2);                                // 16 x 4

// O32: every surface the pass touches spans less than 2 GiB (checked by the launcher), so each access is a
// 32-bit byte offset from the surface base (the scalar-base address form) in place of 64-bit address
// arithmetic; larger surfaces take the 64-bit form.
template <bool FAST, bool O32>
__global__ __launch_bounds__(NT) void k_fine_final(FineArgs a) {
  using F = Math<FAST>;
  constexpr int AP = 1, SW = TW + 2, SH = TH + 2, SN = SW * SH;
  __shared__ float4 sA[SN], sG[SN];
  __shared__ float sL[SN];                // lum(A0 rgb)
  __shared__ float4 sHG[HWN], sHM[HWN];   // half-res: decoded guide (normal, depth); mid rgb + F2.z low
  __shared__ float2 sHA[HWN];             // half-res: F2.z high, F2.w low
  __shared__ float4 sQG[QWN], sQC[QWN];   // quarter-res: decoded guide; coarse
  __shared__ float4 sPar[PWN];            // par1.x, par2.x, par0.z, par0.w
  const int tid = threadIdx.x;
  const int2 org = tile_origin<TW, TH>(a.W, a.H);
  const int x0 = org.x, y0 = org.y;
  const int hx0 = (x0 >> 1) - 1, hy0 = (y0 >> 1) - 1, qx0 = (x0 >> 2) - 1, qy0 = (y0 >> 2) - 1;
  const int lx = tid & 31, ly = tid >> 5;
  const int px = x0 + lx, py = y0 + ly;
  // Every load of the tile is issued before the first wait: the staged A0 / key texels (two per thread for
  // threads 0..83), one window texel (threads 0..211), and the pixel's own material and emission, which the
  // pixel phase after the barrier reads.
  // The wave issues them at a raised priority (as in the temporal pass).
  __builtin_amdgcn_s_setprio(1);
  uint2 a0r[2];
  float4 kdr[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int i = tid + k * NT;
    a0r[k] = make_uint2(0u, 0u); kdr[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < SN) {
      const int qx = clampi(x0 - AP + i % SW, 0, a.W - 1), qy = clampi(y0 - AP + i / SW, 0, a.H - 1);
      a0r[k] = *texel<O32>(a.A0.p, a.A0.pitch, qx, qy);
      kdr[k] = *texel<O32>(a.kd.p, a.kd.pitch, qx, qy);
    }
  }
  // windows: texels outside the level are staged clamped and never read (the taps skip them)
  float4 w4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint4 wu = make_uint4(0u, 0u, 0u, 0u);
  uint2 wm = make_uint2(0u, 0u), wp1 = wm, wp2 = wm;
  if (tid < HWN) {
    const int qx = clampi(hx0 + tid % HWW, 0, a.W2 - 1), qy = clampi(hy0 + tid / HWW, 0, a.H2 - 1);
    w4 = *texel<O32>(a.GH.p, a.GH.pitch, qx, qy);
    wu = *texel<O32>(a.F2.p, a.F2.pitch, qx, qy);
    wm = *texel<O32>(a.mid.p, a.mid.pitch, qx, qy);
  } else if (tid < HWN + QWN) {
    const int i = tid - HWN;
    const int qx = clampi(qx0 + i % QWW, 0, a.W4 - 1), qy = clampi(qy0 + i / QWW, 0, a.H4 - 1);
    w4 = *texel<O32>(a.cgd.p, a.cgd.pitch, qx, qy);
    wm = *texel<O32>(a.coarse.p, a.coarse.pitch, qx, qy);
  } else if (tid < HWN + QWN + PWN) {
    const int i = tid - HWN - QWN;
    const int qx = clampi((x0 >> 1) + i % (TW / 2), 0, a.W2 - 1), qy = clampi((y0 >> 1) + i / (TW / 2), 0, a.H2 - 1);
    wm = *texel<O32>(a.par0.p, a.par0.pitch, qx, qy);
    wp1 = *texel<O32>(a.par1.p, a.par1.pitch, qx, qy);
    wp2 = *texel<O32>(a.par2.p, a.par2.pitch, qx, qy);
  }
  const int cpx = clampi(px, 0, a.W - 1), cpy = clampi(py, 0, a.H - 1);
  const uint2 albr = ld_nt(texel<O32>(a.mat.p, a.mat.pitch, cpx, cpy)), emir = ld_nt(texel<O32>(a.emi.p, a.emi.pitch, cpx, cpy));
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_setprio(0);
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int i = tid + k * NT;
    if (i < SN) {
      const float4 a0 = unpack4(a0r[k]);
      sA[i] = a0;
      sL[i] = lum(a0.x, a0.y, a0.z);
      sG[i] = make_float4(kdr[k].x, kdr[k].y, kdr[k].z, pair_of(kdr[k].w).x);
    }
  }
  if (tid < HWN) {
    const float4 m = unpack4(wm);
    const float2 a01 = unpack2(wu.z), a23 = unpack2(wu.w);
    sHG[tid] = make_float4(w4.x, w4.y, w4.z, pair_of(w4.w).x);
    sHM[tid] = make_float4(m.x, m.y, m.z, a01.x);
    sHA[tid] = make_float2(a01.y, a23.x);
  } else if (tid < HWN + QWN) {
    const int i = tid - HWN;
    sQG[i] = w4;
    sQC[i] = unpack4(wm);
  } else if (tid < HWN + QWN + PWN) {
    const int i = tid - HWN - QWN;
    const float4 p0 = unpack4(wm);
    sPar[i] = make_float4(unpack4(wp1).x, unpack4(wp2).x, p0.z, p0.w);
  }
  __syncthreads();
  if (!inside(px, py, a.W, a.H)) return;
  const int tx = lx + AP, ty = ly + AP;
  const float4 alb = unpack4(albr), emi = unpack4(emir);
  const float4 gp = sG[ty * SW + tx];
  if (gp.w < 0.0f) {
    st_nt(texel<O32>(a.outHist.p, a.outHist.pitch, px, py), make_uint2(0u, 0u));
    *texel<O32>(a.outX.p, a.outX.pitch, px, py) = pack4(make_float4(emi.x, emi.y, emi.z, 0.0f));
    return;
  }
  // interior tiles (every tap of the fine filter and of both upsampling windows inside the image: all but
  // the border tiles) run the pixel phase without the per-tap bounds tests
  const bool interior = x0 >= 1 && y0 >= 1 && x0 + TW < a.W && y0 + TH < a.H && hx0 >= 0 && hy0 >= 0 &&
                        hx0 + HWW <= a.W2 && hy0 + HWH <= a.H2 && qx0 >= 0 && qy0 >= 0 && qx0 + QWW <= a.W4 &&
                        qy0 + QWH <= a.H4;
  auto pixel = [&](auto in_tile) {
  constexpr bool CHK = !decltype(in_tile)::value;
  const float4 pp = sPar[(ly >> 1) * (TW / 2) + (lx >> 1)];       // par1.x, par2.x, par0.z, par0.w
  const float4 A1 = filt<1, 1, true, SW, FAST, CHK, true>(sA, sG, tx, ty, px, py, a.W, a.H, pp.x, pp.y, sL);
  st_nt(texel<O32>(a.outHist.p, a.outHist.pitch, px, py), pack4(A1));
  const float fx = qf16(A1.x), fy = qf16(A1.y), fz = qf16(A1.z);
  const float3 n = make_float3(gp.x, gp.y, gp.z);
  const float z = gp.w;
  const float jd = 0.05f * z + 1e-3f, rjd = F::rcpn(jd);
  // half-res mid level (hush.js:585-602).  Both upsampling loops are branch-free: a tap outside the level
  // (its window slot holds a clamped, finite copy) or on a sky texel (staged level values zero) gets weight
  // 0 and adds exact zeros, which leave the sums unchanged bit for bit (each starts at +0)
  const float hpx = ((float)px + 0.5f) / 2.0f - 0.5f, hpy = ((float)py + 0.5f) / 2.0f - 0.5f;
  const float hbx = floorf(hpx), hby = floorf(hpy);
  float uBx = 0.0f, uBy = 0.0f, uBz = 0.0f, uAx = 0.0f, uAy = 0.0f, uAz = 0.0f, swh = 0.0f;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    int qx = (int)hbx + (j & 1), qy = (int)hby + (j >> 1);
    const int wi = (qy - hy0) * HWW + (qx - hx0);
    const float4 gh = sHG[wi];
    const bool ok = (!CHK || ((qx >= 0) & (qy >= 0) & (qx < a.W2) & (qy < a.H2))) & (gh.w >= 0.0f);
    const float wj = jbw<FAST>(hpx - hbx, hpy - hby, j, make_float3(gh.x, gh.y, gh.z), gh.w, n, z, jd, rjd);
    const float w = ok ? wj : 0.0f;
    const float4 m = sHM[wi];
    const float2 aa = sHA[wi];
    swh += w;
    uBx += w * m.x; uBy += w * m.y; uBz += w * m.z;
    uAx += w * m.w; uAy += w * aa.x; uAz += w * aa.y;
  }
  float ix = fx, iy = fy, iz = fz;
  if (swh > 1e-4f) {
    const float rh = F::rcpn(swh);
    ix = F::divsf(uBx, swh, rh) + pp.w * (fx - F::divsf(uAx, swh, rh));
    iy = F::divsf(uBy, swh, rh) + pp.w * (fy - F::divsf(uAy, swh, rh));
    iz = F::divsf(uBz, swh, rh) + pp.w * (fz - F::divsf(uAz, swh, rh));
  }
  // quarter-res coarse level (hush.js:604-619)
  const float cpx = ((float)px + 0.5f) / 4.0f - 0.5f, cpy = ((float)py + 0.5f) / 4.0f - 0.5f;
  const float cbx = floorf(cpx), cby = floorf(cpy);
  float uCx = 0.0f, uCy = 0.0f, uCz = 0.0f, swq = 0.0f;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    int qx = (int)cbx + (j & 1), qy = (int)cby + (j >> 1);
    const int wi = (qy - qy0) * QWW + (qx - qx0);
    const float4 g = sQG[wi];
    const bool ok = (!CHK || ((qx >= 0) & (qy >= 0) & (qx < a.W4) & (qy < a.H4))) & (g.w >= 0.0f);
    const float wj = jbw<FAST>(cpx - cbx, cpy - cby, j, make_float3(g.x, g.y, g.z), g.w, n, z, jd, rjd);
    const float w = ok ? wj : 0.0f;
    swq += w;
    const float4 cc = sQC[wi];
    uCx += w * cc.x; uCy += w * cc.y; uCz += w * cc.z;
  }
  float upx = ix, upy = iy, upz = iz;
  if (swq > 1e-4f) {
    const float rq = F::rcpn(swq);
    upx = F::divsf(uCx, swq, rq); upy = F::divsf(uCy, swq, rq); upz = F::divsf(uCz, swq, rq);
  }
  ix = ix + (upx - ix) * pp.z;
  iy = iy + (upy - iy) * pp.z;
  iz = iz + (upz - iz) * pp.z;
  *texel<O32>(a.outX.p, a.outX.pitch, px, py) = pack4(make_float4(ix * alb.x + emi.x, iy * alb.y + emi.y, iz * alb.z + emi.z, 0.0f));
  };
  if (interior) pixel(std::true_type{});
  else pixel(std::false_type{});
}

// ============================================================================ pass 8: TAA
// TAA (hush.js:625-678): YCoCg variance clip (gamma 1.25) of a manually bilinear history, blend 0.1 in
// Karis tone-mapped space.

// Karis tone map and its inverse, IEEE under both policies: itm(tm(x)) lands on x exactly for most texels
// (the reset frame), where any negative error would drop a round-toward-zero store by a full binary16 ulp.
// The three quotients share one reciprocal, rcpn's (exact for a normal divisor).  tm's divisor 1 + lum(c) is
// 0 or normal: near lum = -1 the sum is exact, a multiple of 2^-24.  It is 0 for a colour with negative
// channels and luminance exactly -1; the quotient then follows IEEE (+-inf, or NaN for a zero channel) through
// v_div_fixup, which gives the special cases from the divisor and the numerator alone.  itm's divisor is at
// least 1e-4, or +inf when the colour is -inf after such a zero divisor (the fixup then gives IEEE's 0 or
// NaN).  Numerators are 0 or far above 2^-102 (binary16 texels and their tone-mapped values).
__device__ __forceinline__ float3 tm(float3 c) {
  using M = Math<false>;
  const float d = 1.0f + lum(c.x, c.y, c.z), r = M::rcpn(d);
  return make_float3(M::divs(c.x, d, r), M::divs(c.y, d, r), M::divs(c.z, d, r));
}
__device__ __forceinline__ float3 itm(float3 c) {
  using M = Math<false>;
  const float d = mx(1.0f - lum(c.x, c.y, c.z), 1e-4f), r = M::rcpn(d);
  return make_float3(M::divs(c.x, d, r), M::divs(c.y, d, r), M::divs(c.z, d, r));
}
__device__ __forceinline__ float3 to_ycocg(float3 c) {
  return make_float3(c.x * 0.25f + c.y * 0.5f + c.z * 0.25f, c.x * 0.5f - c.z * 0.5f, -c.x * 0.25f + c.y * 0.5f - c.z * 0.25f);
}
__device__ __forceinline__ float3 from_ycocg(float3 c) { return make_float3(c.x + c.y - c.z, c.x + c.z, c.x - c.y - c.z); }

template <bool FAST>
__global__ __launch_bounds__(NT) void k_taa(TaaArgs a) {
  using F = Math<FAST>;
  constexpr int AP = 1, SW = TW + 2, SH = TH + 2, SN = SW * SH;
  __shared__ float4 sX[SN];
  const int tid = threadIdx.x;
  const int2 org = tile_origin<TW, TH>(a.W, a.H);
  const int x0 = org.x, y0 = org.y;
  const int lx = tid & 31, ly = tid >> 5;
  const int px = x0 + lx, py = y0 + ly;
  // Loads first, waits later: the motion vector, the texel of X and an apron texel of the staged tile
  // (threads 0..83), then, as soon as the motion vector is in, the four bilinear history taps, so that their
  // latency overlaps the staging, its barrier and the 3x3 statistics.  Every address is clamped into the
  // image and every lane loads (lanes outside the image or without history discard the values): loads
  // under no lane-dependent branch let each wait count only the loads it needs.
  const bool own = px < a.W && py < a.H, use_hist = own && !a.reset;
  const int cpx = clampi(px, 0, a.W - 1), cpy = clampi(py, 0, a.H - 1);
  const uint2 mvr = a.mv.p[(size_t)cpy * a.mv.pitch + cpx];
  const uint2 xr = a.X.p[(size_t)cpy * a.X.pitch + cpx];
  int atx = 0, aty = 0;
  uint2 axr = make_uint2(0u, 0u);
  if (tid < 2 * SW + 2 * TH) {
    if (tid < SW) { atx = tid; aty = 0; }
    else if (tid < 2 * SW) { atx = tid - SW; aty = SH - 1; }
    else if (tid < 2 * SW + TH) { atx = 0; aty = tid - 2 * SW + 1; }
    else { atx = SW - 1; aty = tid - 2 * SW - TH + 1; }
    axr = a.X.p[(size_t)clampi(y0 - AP + aty, 0, a.H - 1) * a.X.pitch + clampi(x0 - AP + atx, 0, a.W - 1)];
  }
  __builtin_amdgcn_sched_barrier(0);
  const float szx = (float)a.W, szy = (float)a.H;
  const float2 mv = unpack2(mvr.x);
  // (px + 0.5 + mv) / size with the host's correctly rounded 1 / size: equal to the quotient for every finite
  // numerator; a non-finite one fails the range test below either way
  const float ux = Math<false>::divs((float)px + 0.5f + mv.x, szx, a.rW), uy = Math<false>::divs((float)py + 0.5f + mv.y, szy, a.rH);
  const bool in_uv = use_hist && ux >= 0.0f && uy >= 0.0f && ux <= 1.0f && uy <= 1.0f;
  const float ppx = ux * szx - 0.5f, ppy = uy * szy - 0.5f;
  const float bx = floorf(ppx), by = floorf(ppy);
  const float fx = ppx - bx, fy = ppy - by;
  uint2 hr4[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {       // f2i saturates (NaN -> 0) and the clamp keeps every tap in the image
    const int qx = clampi(addw(f2i(bx), j & 1), 0, a.W - 1), qy = clampi(addw(f2i(by), j >> 1), 0, a.H - 1);
    hr4[j] = a.prev.p[(size_t)qy * a.prev.pitch + qx];
  }
  __builtin_amdgcn_sched_barrier(0);
  // tile staging of to_ycocg(tm(X)), clamp to edge: each thread its own texel (whose tone-mapped colour it
  // keeps), threads 0..83 also an apron texel
  const float4 xc = unpack4(xr);
  const float3 cur = tm(make_float3(xc.x, xc.y, xc.z));
  {
    const float3 y = to_ycocg(cur);
    sX[(ly + AP) * SW + lx + AP] = make_float4(y.x, y.y, y.z, 0.0f);
  }
  if (tid < 2 * SW + 2 * TH) {
    const float4 x = unpack4(axr);
    const float3 y = to_ycocg(tm(make_float3(x.x, x.y, x.z)));
    sX[aty * SW + atx] = make_float4(y.x, y.y, y.z, 0.0f);
  }
  __syncthreads();
  if (!own) return;
  if (a.reset) {
    float3 r = itm(cur);
    st_nt(a.out.p + (size_t)py * a.out.pitch + px, pack4(make_float4(r.x, r.y, r.z, 1.0f)));
    return;
  }
  const int c0 = (ly + AP) * SW + lx + AP;
  float m1x = 0.0f, m1y = 0.0f, m1z = 0.0f;
#pragma unroll
  for (int j = 0; j < 9; j++) {
    float4 s = sX[c0 + (j / 3 - 1) * SW + (j % 3 - 1)];
    m1x += s.x; m1y += s.y; m1z += s.z;
  }
  const float r9 = 1.0f / 9.0f;                                // RN(1/9)
  m1x = F::divs(m1x, 9.0f, r9); m1y = F::divs(m1y, 9.0f, r9); m1z = F::divs(m1z, 9.0f, r9);
  float vx = 0.0f, vy = 0.0f, vz = 0.0f;
#pragma unroll
  for (int j = 0; j < 9; j++) {
    float4 s = sX[c0 + (j / 3 - 1) * SW + (j % 3 - 1)];
    float dx = s.x - m1x, dy = s.y - m1y, dz = s.z - m1z;
    vx += dx * dx; vy += dy * dy; vz += dz * dz;
  }
  const float sdx = F::sqrtn(mx(F::divs(vx, 9.0f, r9), 1e-12f)), sdy = F::sqrtn(mx(F::divs(vy, 9.0f, r9), 1e-12f)),
              sdz = F::sqrtn(mx(F::divs(vz, 9.0f, r9), 1e-12f));
  float3 res = cur;
  if (in_uv) {
    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int ox = j & 1, oy = j >> 1;
      float wb = (ox == 1 ? fx : 1.0f - fx) * (oy == 1 ? fy : 1.0f - fy);
      const float4 t = unpack4(hr4[j]);
      hx += mulh(wb, t.x, a.nz); hy += mulh(wb, t.y, a.nz); hz += mulh(wb, t.z, a.nz);
    }
    float3 hh = to_ycocg(tm(make_float3(hx, hy, hz)));
    float kx = 1.25f * sdx, ky = 1.25f * sdy, kz = 1.25f * sdz;
    hh.x = mn(mx(hh.x, m1x - kx), m1x + kx);
    hh.y = mn(mx(hh.y, m1y - ky), m1y + ky);
    hh.z = mn(mx(hh.z, m1z - kz), m1z + kz);
    float3 hr = from_ycocg(hh);
    res = make_float3(hr.x + (cur.x - hr.x) * 0.1f, hr.y + (cur.y - hr.y) * 0.1f, hr.z + (cur.z - hr.z) * 0.1f);
  }
  float3 r = itm(res);
  st_nt(a.out.p + (size_t)py * a.out.pitch + px, pack4(make_float4(r.x, r.y, r.z, 1.0f)));
}

// ============================================================================ launchers

// 1-D grids: tile_origin() maps the workgroup index to a tile
static dim3 grid_for(int w, int h) { return dim3(((w + TW - 1) / TW) * ((h + TH - 1) / TH)); }

// Per-pass timing: while set on the calling thread, the pass kernels launched from it record `start` as
// they begin and `stop` as they end (the kernel's own execution, hipExtLaunchKernelGGL's events), so that
// no marker packet, and no idle gap, sits between the timed passes.
static thread_local hipEvent_t t_ev_start = nullptr, t_ev_stop = nullptr;
void set_pass_events(hipEvent_t start, hipEvent_t stop) { t_ev_start = start; t_ev_stop = stop; }
template <typename K, typename... Args>
static void launch_pass_kernel(K kernel, dim3 grid, hipStream_t s, Args... args) {
  if (t_ev_start) hipExtLaunchKernelGGL(kernel, grid, dim3(NT), 0, s, t_ev_start, t_ev_stop, 0, args...);
  else hipLaunchKernelGGL(kernel, grid, dim3(NT), 0, s, args...);
}

#define HUSH_LAUNCH(KERNEL, GRID, ARGS)                                                  \
  do {                                                                                   \
    if (fast) launch_pass_kernel(KERNEL(true), GRID, s, ARGS);                          \
    else launch_pass_kernel(KERNEL(false), GRID, s, ARGS);                              \
  } while (0)

#define K_INFER16(F) k_infer<16, 0, F>
#define K_INFER8(F) k_infer<8, 0, F>
#define K_INFER32(F) k_infer<32, 0, F>
#define K_INFER64(F) k_infer<64, 0, F>
#define K_INFERC(F) k_infer<-1, 0, F>
#define K_INFER_F16(F) k_infer<0, 1, F>
#define K_INFER16_F16(F) k_infer<16, 1, F>
#define K_INFER_H(F) k_infer_mfma<MLP_FP16_MFMA, F>
#define K_INFER_B(F) k_infer_mfma<MLP_BF16_MFMA, F>
#define K_INFER_Q(F) k_infer_mfma<MLP_FP8_MFMA, F>
#define K_MID2(F) k_atrous<1, 2, false, true, 2, F>
#define K_MID3(F) k_atrous<1, 4, false, true, 3, F>
#define K_COARSE2(F) k_atrous<2, 2, true, false, -1, F>
#define K_COARSE1(F) k_coarse1<F>
#define K_FINE(F) k_fine_final<F, false>
#define K_FINE32(F) k_fine_final<F, true>
#define K_TAA(F) k_taa<F>

template <bool BUF>
static void launch_temporal_t(const TemporalArgs &a, hipStream_t s) {
  if (a.compensated) {
    if (a.aux) launch_pass_kernel(k_temporal<true, true, BUF>, grid_for(a.W, a.H), s, a);
    else launch_pass_kernel(k_temporal<false, true, BUF>, grid_for(a.W, a.H), s, a);
  } else {
    if (a.aux) launch_pass_kernel(k_temporal<true, false, BUF>, grid_for(a.W, a.H), s, a);
    else launch_pass_kernel(k_temporal<false, false, BUF>, grid_for(a.W, a.H), s, a);
  }
}
hipError_t launch_temporal(const TemporalArgs &a, bool, hipStream_t s) {
  // buffer descriptors take 32-bit offsets and record counts, passed as int, and the other accesses 32-bit
  // byte offsets from the surface base: every surface the pass touches must stay below 2 GiB (larger
  // surfaces take the 64-bit address form)
  const uint64_t lim = (uint64_t)1 << 31;
  auto small = [&](int pitch, int rows, int texel) { return (uint64_t)pitch * (uint64_t)rows * (uint64_t)texel < lim; };
  // the paired colour / moment taps also need a texel of padding per row of those two surfaces
  const bool buf = !a.wide && a.pcol.pitch > a.W && a.pmom.pitch > a.W &&
                   small(a.pkd.pitch, a.H, 16) && small(a.pcol.pitch, a.H, 8) && small(a.pmom.pitch, a.H, 8) &&
                   small(a.ppar0.pitch, a.H2, 8) && small(a.rad.pitch, a.H, 8) && small(a.nd.pitch, a.H, 8) &&
                   small(a.mv.pitch, a.H, 8) && small(a.mat.pitch, a.H, 8) && small(a.outA.pitch, a.H, 8) &&
                   small(a.outMom.pitch, a.H, 8) && small(a.outKey.pitch, a.H, 8) && small(a.outKD.pitch, a.H, 16) &&
                   small(a.F0.pitch, a.H2, 16) && small(a.F1.pitch, a.H2, 16) && small(a.F2.pitch, a.H2, 16) &&
                   small(a.F3.pitch, a.H2, 16) && small(a.outGH.pitch, a.H2, 16) && (!a.mask || small(a.mask_pitch, a.H, 1));
  if (buf) launch_temporal_t<true>(a, s);
  else launch_temporal_t<false>(a, s);
  return hipGetLastError();
}

hipError_t launch_infer(const InferArgs &a, int mode, bool fast, hipStream_t s) {
  dim3 g = grid_for(a.W2, a.H2);
  if (mode == MLP_FP32) {
    if (a.hpad == 16) HUSH_LAUNCH(K_INFER16, g, a);
    else if (a.hpad == 8) HUSH_LAUNCH(K_INFER8, g, a);
    else if (a.hpad == 32) HUSH_LAUNCH(K_INFER32, g, a);
    else if (a.hpad == 64) HUSH_LAUNCH(K_INFER64, g, a);
    else HUSH_LAUNCH(K_INFERC, g, a);
  } else if (mode == MLP_F16) {
    if (a.hidden <= 16 && a.wq) HUSH_LAUNCH(K_INFER16_F16, g, a);
    else HUSH_LAUNCH(K_INFER_F16, g, a);
  } else if (a.hidden != 16 || !a.wq) {
    return hipErrorInvalidValue;
  } else if (mode == MLP_FP16_MFMA) {
    HUSH_LAUNCH(K_INFER_H, g, a);
  } else if (mode == MLP_BF16_MFMA) {
    HUSH_LAUNCH(K_INFER_B, g, a);
  } else if (mode == MLP_FP8_MFMA) {
    HUSH_LAUNCH(K_INFER_Q, g, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pack_mlp(const float *w, MfmaPack *p, const float sw[3], const float sx[3], hipStream_t s) {
  hipLaunchKernelGGL(k_pack_mlp, dim3(1), dim3(64), 0, s, w, p, sw[0], sw[1], sw[2], sx[0], sx[1], sx[2]);
  return hipGetLastError();
}

hipError_t launch_mlp_tex(int mode, TexU F0, TexU F1, TexU F2, int W2, int H2, float *out, const void *wq, hipStream_t s) {
  const long n = (long)W2 * H2;
  const dim3 g((unsigned)((n + NT - 1) / NT));
  if (mode == MLP_FP16_MFMA) hipLaunchKernelGGL(k_mlp_tex<MLP_FP16_MFMA>, g, dim3(NT), 0, s, F0, F1, F2, W2, H2, out, wq);
  else if (mode == MLP_BF16_MFMA) hipLaunchKernelGGL(k_mlp_tex<MLP_BF16_MFMA>, g, dim3(NT), 0, s, F0, F1, F2, W2, H2, out, wq);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}
hipError_t launch_mlp(int mode, const uint32_t *feat, float *out, int n, const float *w, const void *wq, hipStream_t s) {
  dim3 g((n + NT - 1) / NT);
  switch (mode) {
    case MLP_FP32: hipLaunchKernelGGL(k_mlp<MLP_FP32>, g, dim3(NT), 0, s, feat, out, n, w, wq); break;
    case MLP_F16: hipLaunchKernelGGL(k_mlp<MLP_F16>, g, dim3(NT), 0, s, feat, out, n, w, wq); break;
    case MLP_FP16_MFMA: hipLaunchKernelGGL(k_mlp<MLP_FP16_MFMA>, g, dim3(NT), 0, s, feat, out, n, w, wq); break;
    case MLP_BF16_MFMA: hipLaunchKernelGGL(k_mlp<MLP_BF16_MFMA>, g, dim3(NT), 0, s, feat, out, n, w, wq); break;
    case MLP_FP8_MFMA: hipLaunchKernelGGL(k_mlp<MLP_FP8_MFMA>, g, dim3(NT), 0, s, feat, out, n, w, wq); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mid2(const AtrousArgs &a, bool fast, hipStream_t s) {
  HUSH_LAUNCH(K_MID2, grid_for(a.w, a.h), a);
  return hipGetLastError();
}
hipError_t launch_mid3(const AtrousArgs &a, bool fast, hipStream_t s) {
  HUSH_LAUNCH(K_MID3, grid_for(a.w, a.h), a);
  return hipGetLastError();
}
hipError_t launch_coarse1(const Coarse1Args &a, bool fast, hipStream_t s) {
  HUSH_LAUNCH(K_COARSE1, grid_for(a.W4, a.H4), a);
  return hipGetLastError();
}
hipError_t launch_coarse2(const AtrousArgs &a, bool fast, hipStream_t s) {
  HUSH_LAUNCH(K_COARSE2, grid_for(a.w, a.h), a);
  return hipGetLastError();
}
#define HUSH_LAUNCH_N(KERNEL, GRID, ...)                                                 \
  do {                                                                                   \
    if (fast) launch_pass_kernel(KERNEL(true), GRID, s, __VA_ARGS__);                   \
    else launch_pass_kernel(KERNEL(false), GRID, s, __VA_ARGS__);                       \
  } while (0)
#define K_C1_INFER16(F) k_coarse1_infer<16, 0, F>
#define K_C1_INFER8(F) k_coarse1_infer<8, 0, F>
#define K_C1_INFER32(F) k_coarse1_infer<32, 0, F>
#define K_C1_INFER64(F) k_coarse1_infer<64, 0, F>
#define K_C1_INFERC(F) k_coarse1_infer<-1, 0, F>
#define K_C1_INFER_F16(F) k_coarse1_infer<0, 1, F>
#define K_C1_INFER16_F16(F) k_coarse1_infer<16, 1, F>
#define K_C1_INFER_H(F) k_coarse1_infer_mfma<MLP_FP16_MFMA, F>
#define K_C1_INFER_B(F) k_coarse1_infer_mfma<MLP_BF16_MFMA, F>
#define K_C1_INFER_Q(F) k_coarse1_infer_mfma<MLP_FP8_MFMA, F>
#define K_C2_MID3(F) k_coarse2_mid3<F>

hipError_t launch_coarse1_infer(const Coarse1Args &c, const InferArgs &a, int mode, bool fast, hipStream_t s) {
  const int nA = (int)grid_for(c.W4, c.H4).x, nApad = (nA + 7) & ~7;
  const dim3 g((unsigned)nApad + grid_for(a.W2, a.H2).x);
  if (mode == MLP_FP32) {
    if (a.hpad == 16) HUSH_LAUNCH_N(K_C1_INFER16, g, c, a, nA, nApad);
    else if (a.hpad == 8) HUSH_LAUNCH_N(K_C1_INFER8, g, c, a, nA, nApad);
    else if (a.hpad == 32) HUSH_LAUNCH_N(K_C1_INFER32, g, c, a, nA, nApad);
    else if (a.hpad == 64) HUSH_LAUNCH_N(K_C1_INFER64, g, c, a, nA, nApad);
    else HUSH_LAUNCH_N(K_C1_INFERC, g, c, a, nA, nApad);
  } else if (mode == MLP_F16) {
    if (a.hidden <= 16 && a.wq) HUSH_LAUNCH_N(K_C1_INFER16_F16, g, c, a, nA, nApad);
    else HUSH_LAUNCH_N(K_C1_INFER_F16, g, c, a, nA, nApad);
  } else if (a.hidden != 16 || !a.wq) {
    return hipErrorInvalidValue;
  } else if (mode == MLP_FP16_MFMA) {
    HUSH_LAUNCH_N(K_C1_INFER_H, g, c, a, nA, nApad);
  } else if (mode == MLP_BF16_MFMA) {
    HUSH_LAUNCH_N(K_C1_INFER_B, g, c, a, nA, nApad);
  } else if (mode == MLP_FP8_MFMA) {
    HUSH_LAUNCH_N(K_C1_INFER_Q, g, c, a, nA, nApad);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_coarse2_mid3(const AtrousArgs &c, const AtrousArgs &m, bool fast, hipStream_t s) {
  const int nA = (int)grid_for(c.w, c.h).x, nApad = (nA + 7) & ~7;
  const dim3 g((unsigned)nApad + grid_for(m.w, m.h).x);
  HUSH_LAUNCH_N(K_C2_MID3, g, c, m, nA, nApad);
  return hipGetLastError();
}
hipError_t launch_fine_final(const FineArgs &a, bool fast, hipStream_t s) {
  const uint64_t lim = (uint64_t)1 << 31;
  auto small = [&](int pitch, int rows, int texel) { return (uint64_t)pitch * (uint64_t)rows * (uint64_t)texel < lim; };
  const bool o32 = !a.wide && small(a.A0.pitch, a.H, 8) && small(a.kd.pitch, a.H, 16) && small(a.par0.pitch, a.H2, 8) &&
                   small(a.par1.pitch, a.H2, 8) && small(a.par2.pitch, a.H2, 8) && small(a.F2.pitch, a.H2, 16) &&
                   small(a.GH.pitch, a.H2, 16) && small(a.cgd.pitch, a.H4, 16) && small(a.mid.pitch, a.H2, 8) &&
                   small(a.coarse.pitch, a.H4, 8) && small(a.mat.pitch, a.H, 8) && small(a.emi.pitch, a.H, 8) &&
                   small(a.outHist.pitch, a.H, 8) && small(a.outX.pitch, a.H, 8);
  if (o32) HUSH_LAUNCH(K_FINE32, grid_for(a.W, a.H), a);
  else HUSH_LAUNCH(K_FINE, grid_for(a.W, a.H), a);
  return hipGetLastError();
}
hipError_t launch_taa(const TaaArgs &a, bool fast, hipStream_t s) {
  HUSH_LAUNCH(K_TAA, grid_for(a.W, a.H), a);
  return hipGetLastError();
}

// ============================================================================ reduced-range math self-test
// one binary32 input per thread: inputs base .. base + NT * gridDim.x - 1 (as bit patterns)
__global__ __launch_bounds__(NT) void k_math_selftest(int which, uint32_t base, unsigned long long *bad,
                                                      uint32_t *first) {
  const uint32_t u = base + blockIdx.x * NT + threadIdx.x;
  const float x = __uint_as_float(u);
  float r, ref;
  if (which == 0) {
    if (!(x >= 0x1p-96f || u == 0u || x != x)) return;
    r = Math<false>::sqrtn(x); ref = sqrtf(x);
  } else if (which == 1) {
    if (!(x <= 0.0f || x != x)) return;
    r = Math<false>::expn(x); ref = expf(x);
  } else if (which == 2) {
    if (!(x >= __int_as_float(0xc2ce8ed0) && x <= __int_as_float(0x42b17218))) return;
    r = Math<false>::expm(x); ref = expf(x);
  } else if (which == 3) {
    const uint32_t e = (u >> 23) & 0xffu;
    if (!((e >= 1u && e <= 252u) || x != x)) return;          // normal, |x| < 2^126, or NaN
    r = Math<false>::rcpn(x); ref = 1.0f / x;
  } else if (which == 4) {
    if ((u & 0x7fffffffu) >= 1u && (u & 0x7fffffffu) < 0x00800000u) return;   // every input but a subnormal
    r = Math<false>::logn(x); ref = logf(x);
  } else if (which == 7) {
    // expw over every finite x <= 0: expf's bits from ln(2^-149) up; below, a value that any factor of at
    // most 1/2 takes to expf's 0
    if (!(x <= 0.0f) || x == -__builtin_inff()) return;
    const float r = Math<false>::expw(x), rf = expf(x);
    const bool ok = x >= __int_as_float(0xc2ce8ed0) ? __float_as_uint(r) == __float_as_uint(rf)
                                                   : (rf == 0.0f && r >= 0.0f && r * 0.5f == 0.0f);
    if (ok) return;
    atomicAdd(bad, 1ull);
    atomicMin(first, u);
    return;
  } else if (which == 6) {
    // the stored activations (binary16, truncated) of sigmoid against 1 / (1 + expf(-x)), every x
    const float s = sigmoid<false>(x), sr = 1.0f / (1.0f + expf(-x));
    const uint32_t a = pack_rtz(act_alpha(s), act_gamma(s)), ar = pack_rtz(act_alpha(sr), act_gamma(sr));
    const uint32_t b = pack_rtz(s, 0.0f), br = pack_rtz(sr, 0.0f);
    if (a == ar && b == br) return;
    atomicAdd(bad, 1ull);
    atomicMin(first, u);
    return;
  } else {
    // divs(x, d, RN(1/d)) against x / d for every numerator of its domain over divisors that cover its
    // domain: +-0, +-inf, NaN, the extremes 2^-126 and 2^126 - ulp, and significands with hard patterns at
    // several exponents
    const uint32_t ax = u & 0x7fffffffu;
    if (ax != 0u && ax < 0x0c800000u) return;                 // 0 < |x| < 2^-102
    const uint32_t dv[] = {0x00000000u, 0x80000000u, 0x7f800000u, 0xff800000u, 0x7fc00000u, 0x00800000u,
                           0x7e7fffffu, 0x3f800000u, 0xbf800000u, 0x3fffffffu, 0x3f800001u, 0x40400000u,
                           0x41100000u, 0x3a83126fu, 0x3c23d70au, 0x477fe000u, 0xc0e00001u, 0x2b7fffffu,
                           0x53000001u, 0x1d7ffffeu, 0x60abcdefu, 0x9f123457u, 0x3ec00000u, 0xbf3504f3u};
    unsigned nbad = 0;
    for (int k = 0; k < (int)(sizeof(dv) / sizeof(dv[0])); k++) {
      const float d = __uint_as_float(dv[k]), y = 1.0f / d;
      const float q = Math<false>::divs(x, d, y), qr = x / d;
      // finite non-zero operands need a normal quotient below 2^127 (underflow and overflow are outside)
      const bool fin = x == x && d == d && fabsf(x) < __builtin_inff() && fabsf(d) < __builtin_inff();
      if (fin && x != 0.0f && d != 0.0f && !(fabsf(qr) >= 0x1p-126f && fabsf(qr) < 0x1p127f)) continue;
      if (__float_as_uint(q) != __float_as_uint(qr) && !(q != q && qr != qr)) nbad++;
    }
    if (nbad == 0) return;
    atomicAdd(bad, (unsigned long long)nbad);
    atomicMin(first, u);
    return;
  }
  if (__float_as_uint(r) == __float_as_uint(ref) || (r != r && ref != ref)) return;
  atomicAdd(bad, 1ull);
  atomicMin(first, u);
}

hipError_t launch_math_selftest(int which, uint32_t base, uint32_t blocks, unsigned long long *bad, uint32_t *first,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_math_selftest, dim3(blocks), dim3(NT), 0, s, which, base, bad, first);
  return hipGetLastError();
}

}  // namespace hush
