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

__device__ __forceinline__ float4 ld(const Tex &t, int x, int y) { return unpack4(t.p[(size_t)y * t.pitch + x]); }
__device__ __forceinline__ void st(const Tex &t, int x, int y, float4 v) { t.p[(size_t)y * t.pitch + x] = pack4(v); }
__device__ __forceinline__ uint4 ldu(const TexU &t, int x, int y) { return t.p[(size_t)y * t.pitch + x]; }
__device__ __forceinline__ void stu(const TexU &t, int x, int y, uint4 v) { t.p[(size_t)y * t.pitch + x] = v; }

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ bool inside(int x, int y, int w, int h) { return x >= 0 && y >= 0 && x < w && y < h; }

// WGSL min/max on non-NaN operands
__device__ __forceinline__ float mn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mx(float a, float b) { return a > b ? a : b; }

// lum (hush.js:31)
__device__ __forceinline__ float lum(float r, float g, float b) { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// octEnc (hush.js:34-39)
__device__ __forceinline__ float2 oct_enc(float3 n) {
  float d = mx(fabsf(n.x) + fabsf(n.y) + fabsf(n.z), 1e-20f);
  float vx = n.x / d, vy = n.y / d, vz = n.z / d;
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
  float l = mx(sqrtf(vx * vx + vy * vy + vz * vz), 1e-20f);
  return make_float3(vx / l, vy / l, vz / l);
}

__device__ __forceinline__ float log1p_(float x) { return logf(1.0f + x); }   // hush.js:47

__device__ __forceinline__ float pow128(float x) {                            // hush.js:70-73
  float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16, x64 = x32 * x32;
  return x64 * x64;
}

__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + expf(-x)); }

}  // namespace hush
