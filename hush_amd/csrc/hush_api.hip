// Host side of the Hush C ABI (include/hush.h): context, surface arena, pass sequencing.
// Mirrors the Hush class of src/denoise/hush.js:725-875 with HIP device buffers in place of GPUTextures.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "hush.h"
#include "hush_ctx.h"
#include "hush_internal.h"

using namespace hush;

namespace {

thread_local char g_err[512];



const char *kPassLabel[8] = {"temporal", "infer+mid", "infer+mid", "infer+mid", "coarse", "coarse", "fine+final", "taa"};
const char *kInputName[5] = {"radiance", "normalDepth", "motion", "material", "emission"};

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// halfToFloat (hush.js:706-711)
float half_to_float(uint16_t h) {
  int s = h >> 15, e = (h >> 10) & 31, m = h & 1023;
  float v;
  if (e == 0) v = std::ldexp((float)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = std::ldexp((float)(1024 + m), e - 25);
  return s ? -v : v;
}

}  // namespace



int hush_fail(hush_t *h, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(h ? h->err : g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hush_hip_fail(hush_t *h, hipError_t e, const char *what) {
  return hush_fail(h, HUSH_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}
#define fail hush_fail
#define hip_fail hush_hip_fail

// the MFMA forms are built for the 16-wide hidden layers of the 20-16-16-12 network (the built-in size)
static bool precision_built(int p, int hidden) {
  return p == HUSH_MLP_FP32 || p == HUSH_MLP_F16 || ((p == HUSH_MLP_FP16_MFMA || p == HUSH_MLP_BF16_MFMA || p == HUSH_MLP_FP8_MFMA) && hidden == 16);
}

// fp8 input scales are powers of two in [2^-20, 2^20] (so 448 s is an exact binary16 value below 2^8 s)
static const int FP8_SCALE_EXP = 20;
// smallest power of two s with m / s <= 448 (the largest fp8 E4M3 value), within the scale range; 1 for m == 0
static float pow2_over_448(double m) {
  if (!(m > 0.0)) return 1.0f;
  int e;
  if (std::frexp(m / 448.0, &e) == 0.5) e -= 1;       // m / 448 = f 2^e, f in [0.5, 1): 2^(e-1) suffices at f = 1/2
  return std::ldexp(1.0f, std::min(std::max(e, -FP8_SCALE_EXP), FP8_SCALE_EXP));
}
static float fp8_scale(const float *w, int n) {
  double m = 0.0;
  for (int i = 0; i < n; i++) m = std::fmax(m, std::fabs((double)w[i]));
  return pow2_over_448(m);
}
// Input scales of the fp8 form from the weights alone: features taken as bounded by 16 in magnitude (the
// logarithmic features of binary16 G-buffer values at unit exposure stay below log(1 + 65504) = 11.1, the
// normal, depth and count features below 2; material inputs and extreme exposures can exceed it, and such
// inputs saturate at 448 sx), each hidden unit by |b| + sum |w| x (the bound of its inputs)
static void fp8_default_input_scales(const float *w, float *sx) {
  const float *W1 = w, *B1 = w + 320, *W2 = w + 336, *B2 = w + 592;
  const double fb = 16.0;
  double m1 = 0.0, m2 = 0.0;
  for (int i = 0; i < 16; i++) {
    double a = std::fabs((double)B1[i]);
    for (int j = 0; j < 20; j++) a += std::fabs((double)W1[i * 20 + j]) * fb;
    m1 = std::fmax(m1, a);
  }
  for (int i = 0; i < 16; i++) {
    double a = std::fabs((double)B2[i]);
    for (int j = 0; j < 16; j++) a += std::fabs((double)W2[i * 16 + j]) * m1;
    m2 = std::fmax(m2, a);
  }
  sx[0] = pow2_over_448(fb); sx[1] = pow2_over_448(m1); sx[2] = pow2_over_448(m2);
}
static bool fp8_scale_ok(float x) {
  if (!(x > 0.0f) || !std::isfinite(x)) return false;
  int e;
  return std::frexp(x, &e) == 0.5f && e - 1 >= -FP8_SCALE_EXP && e - 1 <= FP8_SCALE_EXP;
}

extern "C" int32_t hush_hidden_of(size_t n) {
  for (long h = 1; h <= 256; h++)
    if ((size_t)(h * h + 34 * h + 12) == n) return (int32_t)h;
  return 0;
}

static void free_surfaces(hush_t *h) {
  if (h->arena) (void)hipFree(h->arena);
  h->arena = nullptr;
  std::memset(h->tex, 0, sizeof(h->tex));
  std::memset(h->F, 0, sizeof(h->F));
  std::memset(h->kd, 0, sizeof(h->kd));
  h->gh = TexF{};
  h->cgd = TexF{};
  h->mask = nullptr;
}

// allocates the 22 surfaces of hush.js:780-793 (plus the debug mask) in one HBM arena
static int alloc_surfaces(hush_t *h, int W, int H) {
  const int W2 = (W + 1) / 2, H2 = (H + 1) / 2, W4 = (W2 + 1) / 2, H4 = (H2 + 1) / 2;
  // pad: texels the row must hold beyond the width.  The colour and moment histories keep at least one: the
  // temporal pass reads a row's two horizontal reprojection taps as one 16-byte load, which at the right
  // edge covers texel W (its value is never used) and must stay inside the surface.
  struct Spec { int id, w, h, bpt, pad = 0; };
  std::vector<Spec> specs;
  for (int id : {S_A0, S_X, S_KEY0, S_KEY1, S_TAA0, S_TAA1}) specs.push_back({id, W, H, 8});
  for (int id : {S_MOM0, S_MOM1, S_HIST0, S_HIST1}) specs.push_back({id, W, H, 8, 1});
  for (int id : {S_BH0, S_BH1, S_PAR0_0, S_PAR0_1, S_PAR1, S_PAR2}) specs.push_back({id, W2, H2, 8});
  for (int id : {S_F0, S_F1, S_F2, S_F3}) specs.push_back({id, W2, H2, 16});
  for (int id : {S_C0, S_C1, S_CG}) specs.push_back({id, W4, H4, 8});
  specs.push_back({S_MASK, W, H, 1});
  for (int id : {S_KD0, S_KD1}) specs.push_back({id, W, H, 16});
  specs.push_back({S_GH, W2, H2, 16});
  specs.push_back({S_CGD, W4, H4, 16});
  size_t total = 0;
  std::vector<size_t> offs;
  for (auto &s : specs) {
    size_t pitch = (size_t)round_up(s.w + s.pad, 64) * s.bpt;
    offs.push_back(total);
    total += (pitch * s.h + 4095) / 4096 * 4096;
  }
  void *arena = nullptr;
  hipError_t e = hipMalloc(&arena, total);
  if (e != hipSuccess) return hip_fail(h, e, "hipMalloc of the denoiser surfaces");
  e = hipMemset(arena, 0, total);
  if (e != hipSuccess) { (void)hipFree(arena); return hip_fail(h, e, "hipMemset of the denoiser surfaces"); }
  h->arena = arena;
  for (size_t i = 0; i < specs.size(); i++) {
    auto &s = specs[i];
    char *p = (char *)arena + offs[i];
    int pitch = round_up(s.w + s.pad, 64);
    if (s.id >= S_F0 && s.id <= S_F3) h->F[s.id - S_F0] = TexU{(uint4 *)p, pitch};
    else if (s.id == S_MASK) { h->mask = (uint8_t *)p; h->mask_pitch = pitch; }
    else if (s.id == S_KD0 || s.id == S_KD1) h->kd[s.id - S_KD0] = TexF{(float4 *)p, pitch};
    else if (s.id == S_GH) h->gh = TexF{(float4 *)p, pitch};
    else if (s.id == S_CGD) h->cgd = TexF{(float4 *)p, pitch};
    else h->tex[s.id] = Tex{(uint2 *)p, pitch};
  }
  h->W = W; h->H = H; h->W2 = W2; h->H2 = H2; h->W4 = W4; h->H4 = H4;
  return HUSH_OK;
}

extern "C" int32_t hush_resize(hush_t *h, int32_t width, int32_t height) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_resize: null context");
  if (!(width > 0 && height > 0 && width <= HUSH_MAX_SIDE && height <= HUSH_MAX_SIDE))
    return fail(h, HUSH_ERR_SIZE, "Hush: bad size %dx%d", width, height);
  if (width == h->W && height == h->H) return HUSH_OK;
  hipError_t e = hipSetDevice(h->device);
  if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
  // the new surfaces first: if that fails the context keeps its size, surfaces and history
  void *old = h->arena;
  int r = alloc_surfaces(h, width, height);
  if (r != HUSH_OK) return r;
  if (old) {
    (void)hipDeviceSynchronize();      // work still in flight on the old surfaces
    (void)hipFree(old);
  }
  h->frame = 0;
  h->reset = true;
  return HUSH_OK;
}

extern "C" int32_t hush_create(int32_t device, int32_t width, int32_t height, float exposure, const uint16_t *weights,
                               size_t count, int32_t precision, hush_t **out) {
  if (!out) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_create: null output pointer");
  *out = nullptr;
  if (!weights) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_create: null weights");
  int hid = hush_hidden_of(count);
  if (!hid) return fail(nullptr, HUSH_ERR_WEIGHTS, "hush weights: %zu f16 values fit no MLP size", count);
  if (!(width > 0 && height > 0 && width <= HUSH_MAX_SIDE && height <= HUSH_MAX_SIDE))
    return fail(nullptr, HUSH_ERR_SIZE, "Hush: bad size %dx%d", width, height);
  if (precision < HUSH_MLP_FP32 || precision > HUSH_MLP_FP8_MFMA) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_create: unknown precision %d", precision);
  if (!precision_built(precision, hid))
    return fail(nullptr, HUSH_ERR_UNSUPPORTED, "hush_create: precision %d needs hidden width 16 (this blob: %d)", precision, hid);
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipSetDevice");
  hush_t *h = new (std::nothrow) hush_ctx();
  if (!h) return fail(nullptr, HUSH_ERR_DEVICE, "hush_create: out of host memory");
  h->device = device;
  h->exposure = exposure;
  h->hidden = hid;
  h->precision = precision;
  std::vector<float> wf(count);
  for (size_t i = 0; i < count; i++) wf[i] = half_to_float(weights[i]);
  e = hipMalloc((void **)&h->d_w, count * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(h->d_w, wf.data(), count * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hip_fail(nullptr, e, "weights upload");
    hush_destroy(h);
    return HUSH_ERR_DEVICE;
  }
  // The FP32 forms run in registers at 8, 16, 32 or 64 units: a blob of another width up to 64 is zero-padded
  // to the next of them.  A padded unit's sums are +0 (zero weights and bias), so its ReLU output is 0, and a
  // padded input adds +0 products after a real unit's own terms; that can only turn a -0 sum into +0, which
  // the ReLU maps to the same 0 and the activations (sigmoid, exp) map to the same value, so every output
  // is the unpadded MLP's bit for bit.
  // Wider blobs run the chunked form (units in chunks of 32) over the weights zero-padded to a multiple of 32
  // and stored input-major.
  h->hpad = hid <= 8 ? 8 : hid <= 16 ? 16 : hid <= 32 ? 32 : hid <= 64 ? 64 : (hid + 31) / 32 * 32;
  if (h->hpad == hid && hid <= 64) {
    h->d_wpad = h->d_w;
  } else if (h->hpad > 64) {
    const int H = h->hpad, n = hid;
    std::vector<float> wp((size_t)H * 20 + H + (size_t)H * H + H + 12 * H + 12, 0.0f);
    const float *W1 = wf.data(), *B1 = W1 + n * 20, *W2 = B1 + n, *B2 = W2 + n * n, *W3 = B2 + n, *B3 = W3 + 12 * n;
    float *P1 = wp.data(), *Q1 = P1 + H * 20, *P2 = Q1 + H, *Q2 = P2 + H * H, *P3 = Q2 + H, *Q3 = P3 + 12 * H;
    for (int j = 0; j < n; j++) {
      for (int i = 0; i < 20; i++) P1[i * H + j] = W1[j * 20 + i];
      Q1[j] = B1[j];
      for (int i = 0; i < n; i++) P2[i * H + j] = W2[j * n + i];
      Q2[j] = B2[j];
    }
    for (int j = 0; j < 12; j++) {
      for (int i = 0; i < n; i++) P3[i * 12 + j] = W3[j * n + i];
      Q3[j] = B3[j];
    }
    e = hipMalloc((void **)&h->d_wpad, wp.size() * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(h->d_wpad, wp.data(), wp.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hip_fail(nullptr, e, "padded weights upload");
      hush_destroy(h);
      return HUSH_ERR_DEVICE;
    }
  } else {
    const int H = h->hpad, n = hid;
    std::vector<float> wp((size_t)H * 20 + H + (size_t)H * H + H + 12 * H + 12, 0.0f);
    const float *W1 = wf.data(), *B1 = W1 + n * 20, *W2 = B1 + n, *B2 = W2 + n * n, *W3 = B2 + n, *B3 = W3 + 12 * n;
    float *P1 = wp.data(), *Q1 = P1 + H * 20, *P2 = Q1 + H, *Q2 = P2 + H * H, *P3 = Q2 + H, *Q3 = P3 + 12 * H;
    for (int j = 0; j < n; j++) {
      for (int i = 0; i < 20; i++) P1[j * 20 + i] = W1[j * 20 + i];
      Q1[j] = B1[j];
      for (int i = 0; i < n; i++) P2[j * H + i] = W2[j * n + i];
      Q2[j] = B2[j];
    }
    for (int j = 0; j < 12; j++) {
      for (int i = 0; i < n; i++) P3[j * H + i] = W3[j * n + i];
      Q3[j] = B3[j];
    }
    e = hipMalloc((void **)&h->d_wpad, wp.size() * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(h->d_wpad, wp.data(), wp.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hip_fail(nullptr, e, "padded weights upload");
      hush_destroy(h);
      return HUSH_ERR_DEVICE;
    }
  }
  // fragment-order weights of the MFMA forms (width 16), and the binary16 pairs of the f16 option's packed
  // form, which also runs narrower blobs over their weights zero-padded to 16 (the MFMA forms refuse them)
  if (hid <= 16) {
    float *w16 = h->d_w;
    if (hid == 16) {
      h->fp8_sw[0] = fp8_scale(wf.data(), 320); h->fp8_sw[1] = fp8_scale(wf.data() + 336, 256);
      h->fp8_sw[2] = fp8_scale(wf.data() + 608, 192);
      fp8_default_input_scales(wf.data(), h->fp8_sx);
    } else if (h->hpad == 16) {
      w16 = h->d_wpad;
    } else {                                      // padded to 8 for the FP32 forms: a width-16 copy to pack
      std::vector<float> wp(812, 0.0f);
      const int n = hid;
      const float *W1 = wf.data(), *B1 = W1 + n * 20, *W2 = B1 + n, *B2 = W2 + n * n, *W3 = B2 + n, *B3 = W3 + 12 * n;
      float *P1 = wp.data(), *Q1 = P1 + 320, *P2 = Q1 + 16, *Q2 = P2 + 256, *P3 = Q2 + 16, *Q3 = P3 + 192;
      for (int j = 0; j < n; j++) {
        for (int i = 0; i < 20; i++) P1[j * 20 + i] = W1[j * 20 + i];
        Q1[j] = B1[j];
        for (int i = 0; i < n; i++) P2[j * 16 + i] = W2[j * n + i];
        Q2[j] = B2[j];
      }
      for (int j = 0; j < 12; j++) {
        for (int i = 0; i < n; i++) P3[j * 16 + i] = W3[j * n + i];
        Q3[j] = B3[j];
      }
      w16 = nullptr;
      e = hipMalloc((void **)&w16, wp.size() * sizeof(float));
      if (e == hipSuccess) e = hipMemcpy(w16, wp.data(), wp.size() * sizeof(float), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        if (w16) (void)hipFree(w16);
        hip_fail(nullptr, e, "padded weights upload");
        hush_destroy(h);
        return HUSH_ERR_DEVICE;
      }
    }
    e = hipMalloc(&h->d_wq, sizeof(MfmaPack));
    if (e == hipSuccess) e = launch_pack_mlp(w16, (MfmaPack *)h->d_wq, h->fp8_sw, h->fp8_sx, nullptr);
    if (w16 != h->d_w && w16 != h->d_wpad) {
      if (e == hipSuccess) e = hipDeviceSynchronize();
      (void)hipFree(w16);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      hip_fail(nullptr, e, "MFMA weight packing");
      hush_destroy(h);
      return HUSH_ERR_DEVICE;
    }
  }
  for (int i = 0; i < 16 && e == hipSuccess; i++) e = hipEventCreate(&h->ev[i]);
  if (e != hipSuccess) {
    hip_fail(nullptr, e, "timing events");
    hush_destroy(h);
    return HUSH_ERR_DEVICE;
  }
  int r = hush_resize(h, width, height);
  if (r != HUSH_OK) {
    std::snprintf(g_err, sizeof(g_err), "%s", h->err);
    hush_destroy(h);
    return r;
  }
  *out = h;
  return HUSH_OK;
}

extern "C" void hush_destroy(hush_t *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  (void)hipDeviceSynchronize();
  free_surfaces(h);
  if (h->d_wpad && h->d_wpad != h->d_w) (void)hipFree(h->d_wpad);
  if (h->d_w) (void)hipFree(h->d_w);
  if (h->d_wq) (void)hipFree(h->d_wq);
  if (h->d_wscale) (void)hipFree(h->d_wscale);
  for (int i = 0; i < 16; i++)
    if (h->ev[i]) (void)hipEventDestroy(h->ev[i]);
  delete h;
}

extern "C" int32_t hush_reset(hush_t *h) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_reset: null context");
  h->reset = true;
  return HUSH_OK;
}

extern "C" int32_t hush_set_exposure(hush_t *h, float e) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_set_exposure: null context");
  h->exposure = e;
  return HUSH_OK;
}
extern "C" float hush_get_exposure(const hush_t *h) { return h ? h->exposure : NAN; }

extern "C" int32_t hush_get_size(const hush_t *h, int32_t *w, int32_t *hh) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_get_size: null context");
  if (w) *w = h->W;
  if (hh) *hh = h->H;
  return HUSH_OK;
}
extern "C" int32_t hush_hidden(const hush_t *h) { return h ? h->hidden : 0; }
extern "C" int64_t hush_frame(const hush_t *h) { return h ? h->frame : -1; }

extern "C" int32_t hush_set_precision(hush_t *h, int32_t p) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_set_precision: null context");
  if (p < HUSH_MLP_FP32 || p > HUSH_MLP_FP8_MFMA) return fail(h, HUSH_ERR_ARGUMENT, "hush_set_precision: unknown precision %d", p);
  if (!precision_built(p, h->hidden))
    return fail(h, HUSH_ERR_UNSUPPORTED, "hush_set_precision: precision %d needs hidden width 16 (this blob: %d)", p, h->hidden);
  h->precision = p;
  return HUSH_OK;
}
extern "C" int32_t hush_get_precision(const hush_t *h) { return h ? h->precision : -1; }

extern "C" int32_t hush_set_ieee_math(hush_t *h, int32_t on) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_set_ieee_math: null context");
  h->ieee = on != 0;
  return HUSH_OK;
}

extern "C" int32_t hush_set_fp8_scales(hush_t *h, const float *scales) {
  if (!h || !scales) return fail(h, HUSH_ERR_ARGUMENT, "hush_set_fp8_scales: null argument");
  if (!h->d_wq || h->hidden != 16) return fail(h, HUSH_ERR_UNSUPPORTED, "hush_set_fp8_scales: the fp8 form needs hidden width 16");
  for (int i = 0; i < 3; i++)
    if (!fp8_scale_ok(scales[i]))
      return fail(h, HUSH_ERR_ARGUMENT, "hush_set_fp8_scales: scale %d (%g) is not a power of two in [2^-20, 2^20]", i,
                  (double)scales[i]);
  hipError_t e = hipSetDevice(h->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();               // no frame in flight reads the old pack
  if (e != hipSuccess) return hip_fail(h, e, "hush_set_fp8_scales");
  for (int i = 0; i < 3; i++) h->fp8_sx[i] = scales[i];
  e = launch_pack_mlp(h->d_w, (MfmaPack *)h->d_wq, h->fp8_sw, h->fp8_sx, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(h, e, "hush_set_fp8_scales: weight packing");
  return HUSH_OK;
}

extern "C" int32_t hush_get_fp8_scales(const hush_t *h, float *scales) {
  if (!h || !scales) return HUSH_ERR_ARGUMENT;
  for (int i = 0; i < 3; i++) scales[i] = h->fp8_sx[i];
  return HUSH_OK;
}
extern "C" int32_t hush_get_ieee_math(const hush_t *h) { return h ? (h->ieee ? 1 : 0) : -1; }

extern "C" int32_t hush_set_compensated(hush_t *h, int32_t on) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_set_compensated: null context");
  h->compensated = on != 0;
  return HUSH_OK;
}
extern "C" int32_t hush_get_compensated(const hush_t *h) { return h ? (h->compensated ? 1 : 0) : -1; }

extern "C" const char *hush_last_error(const hush_t *h) { return h ? h->err : g_err; }

static Tex in_tex(const hush_surface &s) { return Tex{(uint2 *)s.ptr, (int)(s.pitch / 8)}; }

static hush_surface surf_of(const hush_t *h, int id) {
  hush_surface s{};
  if (id >= S_F0 && id <= S_F3) {
    s.ptr = h->F[id - S_F0].p; s.pitch = (size_t)h->F[id - S_F0].pitch * 16;
    s.width = h->W2; s.height = h->H2; s.format = HUSH_RGBA32U;
  } else if (id >= S_KD0 && id <= S_CGD) {
    const TexF &t = id == S_GH ? h->gh : id == S_CGD ? h->cgd : h->kd[id - S_KD0];
    s.ptr = t.p; s.pitch = (size_t)t.pitch * 16; s.format = HUSH_RGBA32F;
    s.width = id == S_GH ? h->W2 : id == S_CGD ? h->W4 : h->W;
    s.height = id == S_GH ? h->H2 : id == S_CGD ? h->H4 : h->H;
  } else if (id == S_MASK) {
    s.ptr = h->mask; s.pitch = (size_t)h->mask_pitch; s.width = h->W; s.height = h->H; s.format = HUSH_R8U;
  } else {
    s.ptr = h->tex[id].p; s.pitch = (size_t)h->tex[id].pitch * 8; s.format = HUSH_RGBA16F;
    bool half = id == S_BH0 || id == S_BH1 || id == S_PAR0_0 || id == S_PAR0_1 || id == S_PAR1 || id == S_PAR2;
    bool quarter = id == S_C0 || id == S_C1 || id == S_CG;
    s.width = half ? h->W2 : quarter ? h->W4 : h->W;
    s.height = half ? h->H2 : quarter ? h->H4 : h->H;
  }
  return s;
}

// kernel arguments of each pass at the context's current state
static TemporalArgs temporal_args(const hush_t *h, const hush_inputs *in) {
  const int cur = (int)(h->frame & 1), prev = cur ^ 1;
  const Tex *t = h->tex;
  TemporalArgs a;
  a.rad = in_tex(in->radiance); a.nd = in_tex(in->normal_depth); a.mv = in_tex(in->motion); a.mat = in_tex(in->material);
  a.pkd = h->kd[prev]; a.pcol = t[S_HIST0 + prev]; a.pmom = t[S_MOM0 + prev]; a.ppar0 = t[S_PAR0_0 + prev];
  a.outA = t[S_A0]; a.outMom = t[S_MOM0 + cur]; a.outKey = t[S_KEY0 + cur]; a.outKD = h->kd[cur];
  a.F0 = h->F[0]; a.F1 = h->F[1]; a.F2 = h->F[2]; a.F3 = h->F[3]; a.outGH = h->gh;
  a.mask = h->mask_on ? h->mask : nullptr; a.mask_pitch = h->mask_pitch; a.aux = h->train_aux; a.compensated = h->compensated ? 1 : 0;
  a.wide = h->wide_gathers ? 1 : 0;
  a.W = h->W; a.H = h->H; a.W2 = h->W2; a.H2 = h->H2; a.reset = h->reset ? 1 : 0; a.E = h->exposure;
  a.nz = -0.0f;
  return a;
}
static InferArgs infer_args(const hush_t *h) {
  const int cur = (int)(h->frame & 1);
  const Tex *t = h->tex;
  InferArgs a;
  a.F0 = h->F[0]; a.F1 = h->F[1]; a.F2 = h->F[2]; a.GH = h->gh;
  a.out0 = t[S_PAR0_0 + cur]; a.out1 = t[S_PAR1]; a.out2 = t[S_PAR2]; a.outMid = t[S_BH0];
  a.W2 = h->W2; a.H2 = h->H2; a.hidden = h->hidden; a.w = h->d_w; a.wq = h->d_wq; a.wscale = h->d_wscale;
  a.hpad = h->hpad; a.wp = h->d_wpad;
  return a;
}
static AtrousArgs mid_args(const hush_t *h, int pass) {      // pass 2 (mid2) or 3 (mid3)
  const Tex *t = h->tex;
  AtrousArgs a{};
  a.inA = t[pass == 2 ? S_BH0 : S_BH1]; a.guide = h->gh; a.out = t[pass == 2 ? S_BH1 : S_BH0];
  a.par1 = t[S_PAR1]; a.par2 = t[S_PAR2]; a.w = h->W2; a.h = h->H2; a.s = 0.0f;
  return a;
}
static Coarse1Args coarse1_args(const hush_t *h) {
  const Tex *t = h->tex;
  Coarse1Args a;
  a.F2 = h->F[2]; a.GH = h->gh; a.outC = t[S_C1]; a.outG = t[S_CG]; a.outGD = h->cgd;
  a.W4 = h->W4; a.H4 = h->H4; a.W2 = h->W2; a.H2 = h->H2; a.s = 0.25f;   // COARSE_S (hush.js:26)
  return a;
}
static AtrousArgs coarse2_args(const hush_t *h) {
  const Tex *t = h->tex;
  AtrousArgs a{};
  a.inA = t[S_C1]; a.guide = h->cgd; a.out = t[S_C0]; a.w = h->W4; a.h = h->H4; a.s = 0.25f;
  return a;
}
static FineArgs fine_args(const hush_t *h, const hush_inputs *in) {
  const int cur = (int)(h->frame & 1);
  const Tex *t = h->tex;
  FineArgs a;
  a.A0 = t[S_A0]; a.kd = h->kd[cur]; a.par0 = t[S_PAR0_0 + cur]; a.par1 = t[S_PAR1]; a.par2 = t[S_PAR2];
  a.F2 = h->F[2]; a.GH = h->gh; a.mid = t[S_BH0]; a.coarse = t[S_C0]; a.cgd = h->cgd;
  a.mat = in_tex(in->material); a.emi = in_tex(in->emission);
  a.outHist = t[S_HIST0 + cur]; a.outX = t[S_X];
  a.W = h->W; a.H = h->H; a.W2 = h->W2; a.H2 = h->H2; a.W4 = h->W4; a.H4 = h->H4;
  a.wide = h->wide_gathers ? 1 : 0;
  return a;
}
static TaaArgs taa_args(const hush_t *h, const hush_inputs *in) {
  const int cur = (int)(h->frame & 1), prev = cur ^ 1;
  const Tex *t = h->tex;
  TaaArgs a;
  a.X = t[S_X]; a.mv = in_tex(in->motion); a.prev = t[S_TAA0 + prev]; a.out = t[S_TAA0 + cur];
  a.W = h->W; a.H = h->H; a.reset = h->reset ? 1 : 0;
  a.rW = 1.0f / (float)h->W; a.rH = 1.0f / (float)h->H;
  a.nz = -0.0f;
  return a;
}

// one pass of the frame at the context's current state (hush.js:858-870)
hipError_t hush_run_pass(hush_t *h, int pass, const hush_inputs *in, hipStream_t s) {
  const bool fast = !h->ieee;
  switch (pass) {
    case 0: return launch_temporal(temporal_args(h, in), fast, s);
    case 1: return launch_infer(infer_args(h), h->precision, fast, s);
    case 2: return launch_mid2(mid_args(h, 2), fast, s);
    case 3: return launch_mid3(mid_args(h, 3), fast, s);
    case 4: return launch_coarse1(coarse1_args(h), fast, s);
    case 5: return launch_coarse2(coarse2_args(h), fast, s);
    case 6: return launch_fine_final(fine_args(h, in), fast, s);
    case 7: return launch_taa(taa_args(h, in), fast, s);
    default:
      return hipErrorInvalidValue;
  }
}

int hush_check_inputs(hush_t *h, const hush_inputs *in) {
  const hush_surface *s[5] = {&in->radiance, &in->normal_depth, &in->motion, &in->material, &in->emission};
  for (int k = 0; k < 5; k++) {
    if (!s[k]->ptr) return fail(h, HUSH_ERR_INPUT, "Hush.denoise: missing input '%s'", kInputName[k]);
    if (s[k]->width != h->W || s[k]->height != h->H)
      return fail(h, HUSH_ERR_INPUT, "Hush.denoise: '%s' is %dx%d, the denoiser %dx%d (call resize)", kInputName[k],
                  s[k]->width, s[k]->height, h->W, h->H);
    if (s[k]->format != HUSH_RGBA16F)
      return fail(h, HUSH_ERR_INPUT, "Hush.denoise: '%s' must be rgba16float", kInputName[k]);
    if (s[k]->pitch % 8 != 0 || s[k]->pitch < (size_t)h->W * 8 || ((uintptr_t)s[k]->ptr & 7) != 0)
      return fail(h, HUSH_ERR_INPUT, "Hush.denoise: '%s' has pitch %zu bytes; rows need 8-byte alignment and >= %d bytes",
                  kInputName[k], s[k]->pitch, h->W * 8);
    if (!hush_pitch_ok(s[k]->pitch / 8, h->H))
      return fail(h, HUSH_ERR_INPUT, "Hush.denoise: '%s' has pitch %zu bytes; pitch x height must stay below 2^31 texels",
                  kInputName[k], s[k]->pitch);
  }
  return HUSH_OK;
}

extern "C" int32_t hush_encode(hush_t *h, const hush_inputs *in, void *stream, hush_surface *out) {
  if (!h || !in) return fail(h, HUSH_ERR_ARGUMENT, "hush_encode: null argument");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(h->device);
  if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
  if (h->timing) {
    // timed frames run the eight passes as eight launches, each kernel recording its own start and stop
    for (int p = 0; p < 8; p++) {
      set_pass_events(h->ev[2 * p], h->ev[2 * p + 1]);
      e = hush_run_pass(h, p, in, s);
      set_pass_events(nullptr, nullptr);
      if (e != hipSuccess) return hip_fail(h, e, kPassLabel[p]);
    }
    h->ev_ready = true;
  } else {
    // six launches on the caller's stream: temporal; coarse1 beside infer + mid1; mid2; coarse2 beside mid3;
    // fine + final; TAA (the coarse level needs only the temporal pass's outputs)
    const bool fast = !h->ieee;
    e = hush_run_pass(h, 0, in, s);
    if (e == hipSuccess) e = launch_coarse1_infer(coarse1_args(h), infer_args(h), h->precision, fast, s);
    if (e == hipSuccess) e = hush_run_pass(h, 2, in, s);
    if (e == hipSuccess) e = launch_coarse2_mid3(coarse2_args(h), mid_args(h, 3), fast, s);
    for (int p = 6; p <= 7 && e == hipSuccess; p++) e = hush_run_pass(h, p, in, s);
    if (e != hipSuccess) return hip_fail(h, e, "hush_encode");
  }
  const int cur = (int)(h->frame & 1);
  h->reset = false;
  h->frame++;
  if (out) *out = surf_of(h, S_TAA0 + cur);
  return HUSH_OK;
}

extern "C" int32_t hush_denoise(hush_t *h, const hush_inputs *in, void *stream, hush_surface *out) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_denoise: null context");
  if (!in) return fail(h, HUSH_ERR_INPUT, "Hush.denoise: missing input '%s'", kInputName[0]);
  int r = hush_check_inputs(h, in);
  if (r != HUSH_OK) return r;
  return hush_encode(h, in, stream, out);
}

extern "C" int32_t hush_output(const hush_t *h, hush_surface *out) {
  if (!h || !out) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_output: null argument");
  *out = surf_of(h, S_TAA0 + (int)((h->frame + 1) & 1));
  return HUSH_OK;
}

extern "C" int32_t hush_mlp_forward(hush_t *h, int32_t precision, const void *features, float *out, int64_t n, void *stream) {
  if (!h || (n > 0 && (!features || !out))) return fail(h, HUSH_ERR_ARGUMENT, "hush_mlp_forward: null argument");
  if (precision < HUSH_MLP_FP32 || precision > HUSH_MLP_FP8_MFMA) return fail(h, HUSH_ERR_ARGUMENT, "hush_mlp_forward: unknown precision %d", precision);
  if (h->hidden != 16 || !precision_built(precision, h->hidden))
    return fail(h, HUSH_ERR_UNSUPPORTED, "hush_mlp_forward: built for hidden width 16 (this blob: %d)", h->hidden);
  if (n < 0 || n > (int64_t)1 << 30) return fail(h, HUSH_ERR_ARGUMENT, "hush_mlp_forward: bad row count %lld", (long long)n);
  if (((uintptr_t)features & 7) != 0 || ((uintptr_t)out & 15) != 0)
    return fail(h, HUSH_ERR_ARGUMENT, "hush_mlp_forward: features must be 8-byte and outputs 16-byte aligned");
  if (n == 0) return HUSH_OK;
  hipError_t e = hipSetDevice(h->device);
  if (e == hipSuccess) e = launch_mlp(precision, (const uint32_t *)features, out, (int)n, h->d_w, h->d_wq, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(h, e, "hush_mlp_forward");
  return HUSH_OK;
}

extern "C" int32_t hush_set_timing(hush_t *h, int32_t enabled) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_set_timing: null context");
  h->timing = enabled != 0;
  h->ev_ready = false;
  return HUSH_OK;
}

extern "C" int32_t hush_pass_times(hush_t *h, float *ms, const char **labels) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_pass_times: null context");
  if (labels)
    for (int i = 0; i < 8; i++) labels[i] = kPassLabel[i];
  if (!ms) return HUSH_OK;
  if (!h->ev_ready) return fail(h, HUSH_ERR_ARGUMENT, "hush_pass_times: no timed frame (enable timing first)");
  hipError_t e = hipEventSynchronize(h->ev[15]);
  if (e != hipSuccess) return hip_fail(h, e, "hipEventSynchronize");
  for (int i = 0; i < 8; i++) {
    e = hipEventElapsedTime(&ms[i], h->ev[2 * i], h->ev[2 * i + 1]);
    if (e != hipSuccess) return hip_fail(h, e, "hipEventElapsedTime");
  }
  return HUSH_OK;
}

extern "C" int32_t hush_debug_surface(hush_t *h, int32_t id, hush_surface *out) {
  if (!h || !out) return fail(h, HUSH_ERR_ARGUMENT, "hush_debug_surface: null argument");
  if (id < 0 || id >= S_COUNT) return fail(h, HUSH_ERR_ARGUMENT, "hush_debug_surface: unknown surface %d", id);
  *out = surf_of(h, id);
  return HUSH_OK;
}

extern "C" int32_t hush_debug_set_state(hush_t *h, int64_t frame, int32_t reset) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_debug_set_state: null context");
  h->frame = frame;
  h->reset = reset != 0;
  return HUSH_OK;
}

extern "C" int32_t hush_debug_run_pass(hush_t *h, int32_t pass, const hush_inputs *in, void *stream) {
  if (!h || !in) return fail(h, HUSH_ERR_ARGUMENT, "hush_debug_run_pass: null argument");
  if (pass < 0 || pass > 7) return fail(h, HUSH_ERR_ARGUMENT, "hush_debug_run_pass: unknown pass %d", pass);
  int r = hush_check_inputs(h, in);
  if (r != HUSH_OK) return r;
  hipError_t e = hipSetDevice(h->device);
  if (e == hipSuccess) e = hush_run_pass(h, pass, in, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(h, e, kPassLabel[pass]);
  return HUSH_OK;
}

extern "C" int32_t hush_set_debug_mask(hush_t *h, int32_t enabled) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_set_debug_mask: null context");
  h->mask_on = enabled != 0;
  return HUSH_OK;
}

extern "C" int32_t hush_debug_wide_gathers(hush_t *h, int32_t enabled) {
  if (!h) return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_debug_wide_gathers: null context");
  h->wide_gathers = enabled != 0;
  return HUSH_OK;
}

extern "C" int32_t hush_debug_math_selftest(int32_t which, uint64_t *mismatches, uint32_t *first) {
  if (which < 0 || which > 7 || !mismatches || !first)
    return fail(nullptr, HUSH_ERR_ARGUMENT, "hush_debug_math_selftest: bad argument");
  struct Counters { unsigned long long bad; uint32_t first, pad; } host = {0ull, 0xffffffffu, 0u};
  Counters *dev = nullptr;
  hipError_t e = hipMalloc((void **)&dev, sizeof(Counters));
  if (e != hipSuccess) return hip_fail(nullptr, e, "hush_debug_math_selftest: hipMalloc");
  e = hipMemcpy(dev, &host, sizeof(Counters), hipMemcpyHostToDevice);
  // 2^32 inputs in 16 launches of 2^28 (2^20 workgroups of 256)
  for (uint32_t k = 0; k < 16 && e == hipSuccess; k++)
    e = launch_math_selftest(which, k << 28, 1u << 20, &dev->bad, &dev->first, nullptr);
  if (e == hipSuccess) e = hipMemcpy(&host, dev, sizeof(Counters), hipMemcpyDeviceToHost);
  (void)hipFree(dev);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hush_debug_math_selftest");
  *mismatches = host.bad;
  *first = host.first;
  return HUSH_OK;
}
