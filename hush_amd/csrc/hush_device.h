// Device-side helpers shared by the Hush HIP kernels (gfx950).
//
// Surfaces live in HBM as row-major arrays of texels: rgba16float = 8 B (uint2, four binary16),
// rgba32uint = 16 B (uint4, eight packed binary16).  Every rgba16float store rounds toward zero
// (v_cvt_pkrtz_f16_f32 after a clamp to +-65504): the storage rounding hush.js:98 describes
// ("f16 rounding toward zero, like a rgba16float storage write").
//
// Arithmetic follows the reference's operation order (src/denoise/hush.js) with FMA contraction off
// (the whole library is compiled with -ffp-contract=off); where a multiply-add is fused on purpose it is
// written as __builtin_fmaf.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hush {

typedef _Float16 half2v __attribute__((ext_vector_type(2)));

struct Tex {            // rgba16float surface
  uint2 *p;
  int pitch;            // texels per row
};
struct TexU {           // rgba32uint surface
  uint4 *p;
  int pitch;
};
// Decoded guide surface, 16 B per texel: xyz = octDec of the stored binary16 octahedral normal (fp32,
// computed once by the producing pass with IEEE arithmetic), w = a binary16 pair (depth and id / block
// count) or an fp32 depth.  Consumers read it instead of decoding the same normal again.
struct TexF {
  float4 *p;
  int pitch;
};

__device__ __forceinline__ float2 unpack2(uint32_t u) {
  half2v h = __builtin_bit_cast(half2v, u);
  return make_float2((float)h.x, (float)h.y);
}
__device__ __forceinline__ float4 unpack4(uint2 v) {
  float2 a = unpack2(v.x), b = unpack2(v.y);
  return make_float4(a.x, a.y, b.x, b.y);
}
__device__ __forceinline__ float clamp_f16(float x) { return fminf(fmaxf(x, -65504.0f), 65504.0f); }
// two floats -> packed binary16, round toward zero
__device__ __forceinline__ uint32_t pack_rtz(float a, float b) {
  auto h = __builtin_amdgcn_cvt_pkrtz(clamp_f16(a), clamp_f16(b));
  return __builtin_bit_cast(uint32_t, h);
}
__device__ __forceinline__ uint2 pack4(float4 v) { return make_uint2(pack_rtz(v.x, v.y), pack_rtz(v.z, v.w)); }

// qf16 (hush.js:68): clamp, then truncate the f32 mantissa to 10 bits
__device__ __forceinline__ float qf16(float x) {
  return __uint_as_float(__float_as_uint(clamp_f16(x)) & 0xffffe000u);
}
// pk (hush.js:101): pack2x16float(qf16(a), qf16(b))
__device__ __forceinline__ uint32_t pk(float a, float b) { return pack_rtz(qf16(a), qf16(b)); }

// Host-side bound on a caller's surface (pitch in texels): the kernels take the pitch as an int and index
// some surfaces with 32-bit products.
inline bool hush_pitch_ok(size_t pitch_texels, int height) {
  return pitch_texels < ((size_t)1 << 23) && pitch_texels * (size_t)height < ((size_t)1 << 31);
}
__device__ __forceinline__ float4 ld(const Tex &t, int x, int y) { return unpack4(t.p[(size_t)y * t.pitch + x]); }
__device__ __forceinline__ void st(const Tex &t, int x, int y, float4 v) { t.p[(size_t)y * t.pitch + x] = pack4(v); }
__device__ __forceinline__ uint4 ldu(const TexU &t, int x, int y) { return t.p[(size_t)y * t.pitch + x]; }
__device__ __forceinline__ void stu(const TexU &t, int x, int y, uint4 v) { t.p[(size_t)y * t.pitch + x] = v; }

__device__ __forceinline__ float4 ldf(const TexF &t, int x, int y) { return t.p[(size_t)y * t.pitch + x]; }
__device__ __forceinline__ void stf(const TexF &t, int x, int y, float4 v) { t.p[(size_t)y * t.pitch + x] = v; }

// Texel (x, y) of a surface.  O32: the surface spans less than 2 GiB (checked by the launcher), so the byte
// offset from its base is a 32-bit value: one 24-bit multiply and a shift-add per access, and the
// scalar-base + 32-bit-offset address form, instead of 64-bit address arithmetic (y < 2^14 and pitch < 2^23
// fit the 24-bit multiply).
template <bool O32, typename T>
__device__ __forceinline__ T *texel(T *p, int pitch, int x, int y) {
  if constexpr (O32)
    return (T *)((char *)p + (__umul24((uint32_t)y, (uint32_t)pitch) + (uint32_t)x) * (uint32_t)sizeof(T));
  else
    return p + (size_t)y * pitch + x;
}
template <bool O32> __device__ __forceinline__ uint2 ldr(const Tex &t, int x, int y) { return *texel<O32>(t.p, t.pitch, x, y); }
template <bool O32> __device__ __forceinline__ void st(const Tex &t, int x, int y, float4 v) { *texel<O32>(t.p, t.pitch, x, y) = pack4(v); }
template <bool O32> __device__ __forceinline__ void stu(const TexU &t, int x, int y, uint4 v) { *texel<O32>(t.p, t.pitch, x, y) = v; }
template <bool O32> __device__ __forceinline__ void stf(const TexF &t, int x, int y, float4 v) { *texel<O32>(t.p, t.pitch, x, y) = v; }
template <bool O32> __device__ __forceinline__ float4 ldf(const TexF &t, int x, int y) { return *texel<O32>(t.p, t.pitch, x, y); }
// ld / ldu on 32-bit byte offsets (surfaces below 2 GiB)
__device__ __forceinline__ float4 ld32(const Tex &t, int x, int y) { return unpack4(*texel<true>(t.p, t.pitch, x, y)); }
__device__ __forceinline__ uint4 ldu32(const TexU &t, int x, int y) { return *texel<true>(t.p, t.pitch, x, y); }
// Non-temporal (streaming) accesses: the nt cache policy for a texel that this pass reads, or writes, and no
// later access in the frame needs from the caches (the G-buffer's own texels, outputs read only by the next
// frame), so that it competes less with the reused texels (history taps, staged neighbourhoods) in the L2.
typedef unsigned int u32x2v __attribute__((ext_vector_type(2)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 ld_nt(const uint2 *p) {
  const u32x2v v = __builtin_nontemporal_load((const u32x2v *)p);
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(uint2 *p, uint2 v) { __builtin_nontemporal_store(u32x2v{v.x, v.y}, (u32x2v *)p); }
__device__ __forceinline__ void st_nt(float4 *p, float4 v) {
  __builtin_nontemporal_store(f32x4v{v.x, v.y, v.z, v.w}, (f32x4v *)p);
}

// low / high binary16 of a packed pair kept in a float4 lane
__device__ __forceinline__ float2 pair_of(float w) { return unpack2(__float_as_uint(w)); }

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// f32 -> i32 toward zero, saturating (WGSL's conversion; the positive side stops at 2^31 - 128, the
// largest float below 2^31), NaN -> 0.  A C++ cast of an out-of-range value (a motion vector that
// overflowed binary16 to +-inf) is undefined, and the compiler then rewrites the clamps around it on the
// assumption that nothing overflows.
__device__ __forceinline__ int f2i(float x) {
  const int r = (int)fminf(fmaxf(x, -2147483648.0f), 2147483520.0f);
  return x == x ? r : 0;
}
// i32 addition with WGSL's two's-complement wrap-around (no signed-overflow assumption)
__device__ __forceinline__ int addw(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
__device__ __forceinline__ bool inside(int x, int y, int w, int h) { return (x >= 0) & (y >= 0) & (x < w) & (y < h); }

// WGSL min/max on non-NaN operands
__device__ __forceinline__ float mn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mx(float a, float b) { return a > b ? a : b; }

// lum (hush.js:31)
__device__ __forceinline__ float lum(float r, float g, float b) { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Tile origin of workgroup n in a 1-D grid of ceil(w/TW) * ceil(h/TH) workgroups.  The dispatcher hands
// workgroup n to XCD n % 8, and each of the 8 XCDs has its own L2; remapping n gives every XCD one
// contiguous band of tiles in raster order, so neighbouring tiles (which share apron texels) meet in the
// same L2 instead of fetching the apron again from HBM.  A launch that runs two passes passes n relative
// to its pass's first workgroup, which starts at a multiple of 8 (n keeps its XCD).
template <int TW, int TH>
__device__ __forceinline__ int2 tile_origin(int w, int h, int n) {
  const int gx = (w + TW - 1) / TW, gy = (h + TH - 1) / TH;
  const int T = gx * gy;
  const int q = T >> 3, r = T & 7, x = n & 7, i = n >> 3;
  const int L = x * q + (x < r ? x : r) + i;
  return make_int2((L % gx) * TW, (L / gx) * TH);
}
template <int TW, int TH>
__device__ __forceinline__ int2 tile_origin(int w, int h) { return tile_origin<TW, TH>(w, h, (int)blockIdx.x); }

// Math policies.  IEEE: correctly rounded division and square root, ocml expf / logf -- the operation-exact
// mirror of the oracle.  FAST: hardware reciprocal, square root, exp2 and log2 (about 1-2 ulp in fp32),
// used only where the result is a continuous value bound for a binary16 store; every decision the
// reference takes on a threshold (validity tests, sky tests, the history key) is computed with IEEE in
// both policies, so masks and ids stay bit-exact.
template <bool FAST> struct Math;
template <> struct Math<false> {
  static __device__ __forceinline__ float div(float a, float b) { return a / b; }
  // a / b for a denominator shared by several quotients, y = RN(1/b): q0 = RN(a y) is within an ulp of a / b,
  // the residual a - b q0 is exact in an FMA, and one correction RN(q0 + r y) is the correctly rounded
  // quotient (Markstein's theorem for a correctly rounded reciprocal).  v_div_fixup_f32, the last step of the
  // compiler's own IEEE division, then gives the result IEEE's sign (a zero quotient's sign can be lost when
  // the correction adds two zeros) and IEEE's special cases: a or b zero, infinite or NaN (there the
  // correction itself is NaN).  Bit-identical to a / b, in 4 VALU instructions instead of 11, for
  // a = +-0, +-inf, NaN or 2^-102 <= |a| < inf and b = +-0, +-inf, NaN or 2^-126 <= |b| < 2^126, with
  // 2^-126 <= |a / b| < 2^127 when both are finite and non-zero.  The residual's exactness sets the lower bound on a: b q0 is a
  // multiple of ulp(b) ulp(q0) >= 2^-47 |a|, so below |a| = 2^-102 the residual can fall under the
  // 2^-149 grid and the correction can miss by an ulp.  At the kernels' call sites the numerators are zero
  // or far above 2^-102 (sums and products of binary16 texels, which are multiples of 2^-24, and of their
  // quotients by bounded factors), or variance sums whose quotient is clamped at 1e-12 or goes into a
  // binary16 store, where an ulp of a value below 2^-102 does not survive.
  static __device__ __forceinline__ float divs(float a, float b, float y) {
    const float q0 = a * y;
    return __builtin_amdgcn_div_fixupf(__builtin_fmaf(__builtin_fmaf(-b, q0, a), y, q0), b, a);
  }
  // divs for finite operands (b non-zero): the copysign from q0 in place of the fixup gives the same bits
  // there, for a cheaper instruction.  The call sites whose operands come from clamped binary16 surfaces
  // (history, guides, filter sums) use it; those fed by caller inputs (radiance, normals, motion) use divs.
  static __device__ __forceinline__ float divsf(float a, float b, float y) {
    const float q0 = a * y;
    return __builtin_copysignf(__builtin_fmaf(__builtin_fmaf(-b, q0, a), y, q0), q0);
  }
  static __device__ __forceinline__ float rcp(float b) { return 1.0f / b; }
  static __device__ __forceinline__ float sqrt(float x) { return sqrtf(x); }
  static __device__ __forceinline__ float exp(float x) { return expf(x); }
  static __device__ __forceinline__ float log(float x) { return logf(x); }
  // Correctly rounded 1 / b for normal b with |b| < 2^126 (and NaN for NaN): the hardware reciprocal, then
  // one Newton step with an FMA residual, 3 VALU instructions instead of 11.  Zero, infinite and
  // subnormal b are outside the domain (the step returns NaN where 1 / b is infinite or zero).
  static __device__ __forceinline__ float rcpn(float b) {
    const float y = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
  }
  // The next three are sqrtf / expf with the range handling that their call sites' arguments never need
  // left out; on the stated domains they are bit-identical to sqrtf / expf (checked over every binary32
  // input of the domain by hush_debug_math_selftest, as is rcpn against 1 / b).
  // Square root of x >= 2^-96, +0, +inf or NaN: the hardware root y, then whichever of y and its two
  // neighbours has the residual x - y'^2 that brackets the exact root (sqrtf's own correction), without
  // the rescaling of arguments below 2^-96 and the +-0 / +inf pass-through (9 VALU instructions, not 17).
  static __device__ __forceinline__ float sqrtn(float x) {
    const float y = __builtin_amdgcn_sqrtf(x);
    const int iy = __float_as_int(y);
    const float ym = __int_as_float(iy - 1), yp = __int_as_float(iy + 1);
    const float vm = __builtin_fmaf(-ym, y, x), vp = __builtin_fmaf(-yp, y, x);
    const float r = vm <= 0.0f ? ym : y;
    return vp > 0.0f ? yp : r;
  }
  // e^x = 2^(x log2 e): x log2 e split into its nearest integer e and a remainder carried to ~48 bits
  // (log2 e as a two-float constant), 2^remainder in hardware, scaled by 2^e; results below 2^-149 flush
  // to zero.  expn: x <= 0 or NaN (no overflow test); expm: -103.28 <= x <= 88.72 (no range test).
  static __device__ __forceinline__ float exp_core(float x) {
    const float L2E = __int_as_float(0x3fb8aa3b), L2E_LO = __int_as_float(0x32a5705f);
    const float ph = x * L2E;
    float pl = __builtin_fmaf(x, L2E, -ph);
    const float e = __builtin_rintf(ph);
    pl = __builtin_fmaf(x, L2E_LO, pl);
    const float t = (ph - e) + pl;
    // e >= -200 keeps the conversion defined for every x (NaN included); below -149 the result is 0 anyway
    return __builtin_ldexpf(__builtin_amdgcn_exp2f(t), (int)fmaxf(e, -200.0f));
  }
  static __device__ __forceinline__ float expn(float x) {
    const float r = exp_core(x);
    return x < __int_as_float(0xc2ce8ed0) ? 0.0f : r;        // -103.27893 = ln(2^-149)
  }
  static __device__ __forceinline__ float expm(float x) { return exp_core(x); }
  // e^x for finite x <= 0 as a factor of a product whose other factor is at most 1/2 (SVGF's edge-stopping
  // weights: kernel weight x normal term <= 3/32): x clamped at -104 in place of expn's underflow test.  For
  // x >= ln(2^-149) = -103.28 it is expf(x) bit for bit; below, it is 0 or 2^-149 where expf is 0, and either
  // times a factor of at most 1/2 rounds to 0 (2^-150 ties to even), so the product is expf's exactly.
  static __device__ __forceinline__ float expw(float x) {
    const float L2E = __int_as_float(0x3fb8aa3b), L2E_LO = __int_as_float(0x32a5705f);
    const float xc = x > -104.0f ? x : -104.0f;
    const float ph = xc * L2E;
    float pl = __builtin_fmaf(xc, L2E, -ph);
    const float e = __builtin_rintf(ph);              // in [-151, 0]: the conversion below is exact
    pl = __builtin_fmaf(xc, L2E_LO, pl);
    const float t = (ph - e) + pl;
    return __builtin_ldexpf(__builtin_amdgcn_exp2f(t), (int)e);
  }
  // ln x for every x except a subnormal: the hardware log2, times ln 2 carried as a two-float constant,
  // infinities and NaN passed through (logf without its rescaling of subnormal arguments).  log1p_'s
  // 1 + x is never subnormal: near -1 the sum is exact, so it is 0 or at least 2^-24 in magnitude.
  static __device__ __forceinline__ float logn(float x) {
    const float L2 = __builtin_amdgcn_logf(x);
    const float LN2 = __int_as_float(0x3f317217), LN2_LO = __int_as_float(0x3377d1cf);
    const float h = L2 * LN2;
    const float lo = __builtin_fmaf(L2, LN2_LO, __builtin_fmaf(L2, LN2, -h));
    return fabsf(L2) < __builtin_inff() ? h + lo : L2;
  }
};
template <> struct Math<true> {
  static __device__ __forceinline__ float rcp(float b) { return __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float divs(float a, float, float y) { return a * y; }
  static __device__ __forceinline__ float divsf(float a, float, float y) { return a * y; }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float sqrtn(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float rcpn(float b) { return __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float expn(float x) { return exp(x); }
  static __device__ __forceinline__ float expm(float x) { return exp(x); }
  static __device__ __forceinline__ float expw(float x) { return exp(x); }
  // e^x = 2^t * 2^lo with t = fl(x log2 e) and lo its rounding error (x log2 e carried to ~48 bits)
  static __device__ __forceinline__ float exp(float x) {
    const float L2E = 1.44269502e+00f, L2E_LO = 1.92596299e-08f;
    const float t = x * L2E;
    const float lo = __builtin_fmaf(x, L2E, -t) + x * L2E_LO;
    const float e = __builtin_amdgcn_exp2f(t);
    return e == __builtin_inff() ? e : __builtin_fmaf(e, lo * 0.693147182f, e);   // no inf * lo on overflow
  }
  static __device__ __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.693147182f; }
  static __device__ __forceinline__ float logn(float x) { return log(x); }
};

// w * x for a binary16 value x widened to fp32, evaluated as RN(w x + (-0)): the exact product rounded once,
// the same bits as w * x (adding -0 changes neither a non-zero product nor the sign of a zero one, and
// infinities and NaN pass through).  nz must be -0.0f supplied at run time (a kernel argument): with a
// literal -0 the compiler folds the FMA back into a multiply, while an opaque addend keeps it, and then the
// widening and the multiply are one v_fma_mix_f32 that reads the binary16 operand as stored.
__device__ __forceinline__ float mulh(float w, float x, float nz) { return __builtin_fmaf(w, x, nz); }

template <bool FAST> __device__ __forceinline__ float log1p_(float x) { return Math<FAST>::logn(1.0f + x); }  // hush.js:47

__device__ __forceinline__ float pow128(float x) {                            // hush.js:70-73
  float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16, x64 = x32 * x32;
  return x64 * x64;
}

// 1 / (1 + e^-x) (hush.js:357-366).  IEEE: the reduced exponential and reciprocal on -x clamped to [-30, 87]
// (the constant first in mn / mx keeps a NaN), so 1 + e^t is normal and below 2^126.  Below -30, 1 + e^-x
// rounds to 1 either way; above 87 the exact value is below 2^-125 and this one returns 1 / (1 + e^87),
// about 1.6e-38, instead.  That difference never reaches a stored bit: 1/32 + (31/32) s and 0.5 + 7.5 s
// round to 1/32 and 0.5, and s itself truncates to binary16 zero (hush_debug_math_selftest checks the
// three stored activations against 1 / (1 + expf(-x)) for every binary32 x).
template <bool FAST> __device__ __forceinline__ float sigmoid(float x) {
  if constexpr (FAST) {
    return Math<true>::rcp(1.0f + Math<true>::exp(-x));
  } else {
    const float t = mx(-30.0f, mn(87.0f, -x));
    return Math<false>::rcpn(1.0f + Math<false>::expm(t));
  }
}
// the activations of the first three MLP outputs (hush.js:357-366): alpha, gamma, and beta / kappa
__device__ __forceinline__ float act_alpha(float s) { return 1.0f / 32.0f + (31.0f / 32.0f) * s; }
__device__ __forceinline__ float act_gamma(float s) { return 0.5f + 7.5f * s; }

// octEnc (hush.js:34-39)
// The shared reciprocal RN(1 / d) comes from rcpn (exact for normal d below 2^126, which d >= 1e-20 is
// unless the normal is not finite: then rcpn's NaN reaches only the quotients' special-case path, where
// v_div_fixup_f32 returns IEEE's result from d and the numerator alone).
__device__ __forceinline__ float2 oct_enc(float3 n) {
  float d = mx(fabsf(n.x) + fabsf(n.y) + fabsf(n.z), 1e-20f);
  const float r = Math<false>::rcpn(d);
  float vx = Math<false>::divs(n.x, d, r), vy = Math<false>::divs(n.y, d, r), vz = Math<false>::divs(n.z, d, r);
  float sx = vx >= 0.0f ? 1.0f : -1.0f, sy = vy >= 0.0f ? 1.0f : -1.0f;
  if (vz < 0.0f) return make_float2((1.0f - fabsf(vy)) * sx, (1.0f - fabsf(vx)) * sy);
  return make_float2(vx, vy);
}
// octDec (hush.js:40-46)
__device__ __forceinline__ float3 oct_dec(float ex, float ey) {
  float vx = ex, vy = ey, vz = 1.0f - fabsf(ex) - fabsf(ey);
  float t = mx(-vz, 0.0f);
  vx += vx >= 0.0f ? -t : t;
  vy += vy >= 0.0f ? -t : t;
  // the squared norm is at least 1/3 for every finite (ex, ey)
  float l = mx(Math<false>::sqrtn(vx * vx + vy * vy + vz * vz), 1e-20f);
  const float r = Math<false>::rcpn(l);            // l in [0.57, 2.3e5] for stored binary16 inputs
  return make_float3(Math<false>::divsf(vx, l, r), Math<false>::divsf(vy, l, r), Math<false>::divsf(vz, l, r));
}

// octDec with the policy's arithmetic (continuous uses only; validity tests use oct_dec)
template <bool FAST> __device__ __forceinline__ float3 oct_dec_m(float ex, float ey) {
  if (!FAST) return oct_dec(ex, ey);
  float vx = ex, vy = ey, vz = 1.0f - fabsf(ex) - fabsf(ey);
  float t = mx(-vz, 0.0f);
  vx += vx >= 0.0f ? -t : t;
  vy += vy >= 0.0f ? -t : t;
  float r = __builtin_amdgcn_rcpf(mx(__builtin_amdgcn_sqrtf(vx * vx + vy * vy + vz * vz), 1e-20f));
  return make_float3(vx * r, vy * r, vz * r);
}
// octEnc with the policy's arithmetic (continuous uses only; the history key uses oct_enc)
template <bool FAST> __device__ __forceinline__ float2 oct_enc_m(float3 n) {
  if (!FAST) return oct_enc(n);
  float r = __builtin_amdgcn_rcpf(mx(fabsf(n.x) + fabsf(n.y) + fabsf(n.z), 1e-20f));
  float vx = n.x * r, vy = n.y * r, vz = n.z * r;
  float sx = vx >= 0.0f ? 1.0f : -1.0f, sy = vy >= 0.0f ? 1.0f : -1.0f;
  if (vz < 0.0f) return make_float2((1.0f - fabsf(vy)) * sx, (1.0f - fabsf(vx)) * sy);
  return make_float2(vx, vy);
}

}  // namespace hush
