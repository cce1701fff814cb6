// Node.js N-API addon over the Hush C ABI (include/hush.h), the SVGF baseline's (include/svgf.h) and the
// display post's (include/post.h).
//
// The JavaScript `Hush` class (hush_amd/node/hush.js) keeps the reference's API
// (src/denoise/hush.js:725-875); this addon is the thin native layer under it: context calls map 1:1 onto
// hush_* entry points, and device surfaces travel as plain objects {ptr: BigInt, pitch, width, height}.
// It also offers the HIP buffer plumbing a Node caller needs to feed the denoiser (alloc / upload /
// download / free of rgba16float surfaces).
#define NAPI_VERSION 8
#include <node_api.h>

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "hush.h"
#include "post.h"
#include "svgf.h"

namespace {

#define NAPI_CALL(env, call)                                     \
  do {                                                           \
    if ((call) != napi_ok) {                                     \
      napi_throw_error((env), nullptr, "hush addon: " #call);   \
      return nullptr;                                            \
    }                                                            \
  } while (0)

napi_value throw_err(napi_env env, const char *msg) {
  napi_throw_error(env, nullptr, msg);
  return nullptr;
}

napi_value throw_status(napi_env env, hush_t *h, int code) {
  napi_value err, msg, c;
  const char *m = hush_last_error(h);
  napi_create_string_utf8(env, m ? m : "hush error", NAPI_AUTO_LENGTH, &msg);
  napi_create_error(env, nullptr, msg, &err);
  napi_create_int32(env, code, &c);
  napi_set_named_property(env, err, "code", c);
  napi_throw(env, err);
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  for (size_t i = argc; i < want; i++) napi_get_undefined(env, &argv[i]);
  return true;
}

// a JS handle owns one Holder; destroy() releases the context early, the finalizer releases the rest
struct Holder {
  hush_t *h;
};

hush_t *ctx_of(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) return nullptr;
  return static_cast<Holder *>(p)->h;
}

double num(napi_env env, napi_value v, double dflt = 0) {
  double d;
  return napi_get_value_double(env, v, &d) == napi_ok ? d : dflt;
}

napi_value make_surface(napi_env env, const hush_surface &s) {
  napi_value o, p, pitch, w, h, f;
  napi_create_object(env, &o);
  napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)s.ptr, &p);
  napi_create_double(env, (double)s.pitch, &pitch);
  napi_create_int32(env, s.width, &w);
  napi_create_int32(env, s.height, &h);
  napi_create_int32(env, s.format, &f);
  napi_set_named_property(env, o, "ptr", p);
  napi_set_named_property(env, o, "pitch", pitch);
  napi_set_named_property(env, o, "width", w);
  napi_set_named_property(env, o, "height", h);
  napi_set_named_property(env, o, "format", f);
  return o;
}

// {ptr: BigInt, pitch, width, height} -> hush_surface; a missing or null object leaves ptr = NULL
bool read_surface(napi_env env, napi_value o, hush_surface *s) {
  std::memset(s, 0, sizeof(*s));
  napi_valuetype t;
  napi_typeof(env, o, &t);
  if (t != napi_object) return true;
  napi_value v;
  bool lossless = false;
  uint64_t ptr = 0;
  if (napi_get_named_property(env, o, "ptr", &v) != napi_ok) return false;
  napi_typeof(env, v, &t);
  if (t == napi_bigint) napi_get_value_bigint_uint64(env, v, &ptr, &lossless);
  else if (t == napi_number) ptr = (uint64_t)num(env, v);
  s->ptr = (void *)(uintptr_t)ptr;
  napi_get_named_property(env, o, "pitch", &v);
  s->pitch = (size_t)num(env, v);
  napi_get_named_property(env, o, "width", &v);
  s->width = (int32_t)num(env, v);
  napi_get_named_property(env, o, "height", &v);
  s->height = (int32_t)num(env, v);
  s->format = HUSH_RGBA16F;
  if (napi_get_named_property(env, o, "format", &v) == napi_ok) {
    napi_typeof(env, v, &t);
    if (t == napi_number) s->format = (int32_t)num(env, v);
  }
  return true;
}

// hiddenOf(count)
napi_value HiddenOf(napi_env env, napi_callback_info info) {
  napi_value a[1], r;
  if (!get_args(env, info, 1, a)) return nullptr;
  NAPI_CALL(env, napi_create_int32(env, hush_hidden_of((size_t)num(env, a[0])), &r));
  return r;
}

void finalize_ctx(napi_env, void *data, void *) {
  Holder *hd = static_cast<Holder *>(data);
  if (!hd) return;
  if (hd->h) hush_destroy(hd->h);
  delete hd;
}

// create(device, width, height, exposure, weights: Uint16Array, precision) -> external
napi_value Create(napi_env env, napi_callback_info info) {
  napi_value a[6];
  if (!get_args(env, info, 6, a)) return nullptr;
  napi_typedarray_type tt;
  size_t len = 0;
  void *data = nullptr;
  if (napi_get_typedarray_info(env, a[4], &tt, &len, &data, nullptr, nullptr) != napi_ok || tt != napi_uint16_array)
    return throw_err(env, "hush addon: weights must be a Uint16Array of binary16 values");
  static const uint16_t dummy = 0;
  hush_t *h = nullptr;
  int r = hush_create((int)num(env, a[0]), (int)num(env, a[1], 0), (int)num(env, a[2], 0), (float)num(env, a[3], 1.0),
                      len ? (const uint16_t *)data : &dummy, len, (int)num(env, a[5], 0), &h);
  if (r != HUSH_OK) return throw_status(env, nullptr, r);
  napi_value ext;
  Holder *hd = new Holder{h};
  if (napi_create_external(env, hd, finalize_ctx, nullptr, &ext) != napi_ok) {
    hush_destroy(h);
    delete hd;
    return throw_err(env, "hush addon: napi_create_external failed");
  }
  return ext;
}

// destroy(h): releases the context now (the handle's finalizer then has nothing to do)
napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  void *p = nullptr;
  if (napi_get_value_external(env, a[0], &p) == napi_ok && p) {
    Holder *hd = static_cast<Holder *>(p);
    if (hd->h) hush_destroy(hd->h);
    hd->h = nullptr;
  }
  return nullptr;
}

napi_value Resize(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  int r = hush_resize(h, (int)num(env, a[1]), (int)num(env, a[2]));
  if (r != HUSH_OK) return throw_status(env, h, r);
  return nullptr;
}

napi_value Reset(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  int r = hush_reset(h);
  if (r != HUSH_OK) return throw_status(env, h, r);
  return nullptr;
}

napi_value SetExposure(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  int r = hush_set_exposure(h, (float)num(env, a[1], 1.0));
  if (r != HUSH_OK) return throw_status(env, h, r);
  return nullptr;
}

napi_value SetPrecision(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  int r = hush_set_precision(h, (int)num(env, a[1]));
  if (r != HUSH_OK) return throw_status(env, h, r);
  return nullptr;
}

napi_value SetCompensated(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  bool on = false;
  napi_coerce_to_bool(env, a[1], &a[1]);
  napi_get_value_bool(env, a[1], &on);
  int r = hush_set_compensated(h, on ? 1 : 0);
  if (r != HUSH_OK) return throw_status(env, h, r);
  return nullptr;
}

// setFp8Scales(h, [s0, s1, s2]): the fp8 form's static per-layer input scales (powers of two)
napi_value SetFp8Scales(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  bool arr = false;
  napi_is_array(env, a[1], &arr);
  uint32_t n = 0;
  if (arr) napi_get_array_length(env, a[1], &n);
  if (!arr || n != 3) return throw_err(env, "hush addon: fp8 scales must be an array of 3 numbers");
  float sc[3];
  for (uint32_t i = 0; i < 3; i++) {
    napi_value v;
    napi_get_element(env, a[1], i, &v);
    sc[i] = (float)num(env, v, 0.0);
  }
  int r = hush_set_fp8_scales(h, sc);
  if (r != HUSH_OK) return throw_status(env, h, r);
  return nullptr;
}

// getFp8Scales(h) -> [s0, s1, s2]
napi_value GetFp8Scales(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  float sc[3];
  int r = hush_get_fp8_scales(h, sc);
  if (r != HUSH_OK) return throw_status(env, h, r);
  napi_value o, v;
  napi_create_array_with_length(env, 3, &o);
  for (uint32_t i = 0; i < 3; i++) { napi_create_double(env, sc[i], &v); napi_set_element(env, o, i, v); }
  return o;
}

napi_value Info(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  int32_t w = 0, hh = 0;
  hush_get_size(h, &w, &hh);
  napi_value o, v;
  napi_create_object(env, &o);
  napi_create_int32(env, w, &v); napi_set_named_property(env, o, "width", v);
  napi_create_int32(env, hh, &v); napi_set_named_property(env, o, "height", v);
  napi_create_double(env, hush_get_exposure(h), &v); napi_set_named_property(env, o, "exposure", v);
  napi_create_int32(env, hush_hidden(h), &v); napi_set_named_property(env, o, "hidden", v);
  napi_create_double(env, (double)hush_frame(h), &v); napi_set_named_property(env, o, "frame", v);
  napi_create_int32(env, hush_get_precision(h), &v); napi_set_named_property(env, o, "precision", v);
  return o;
}

// denoise(h, inputs, stream: BigInt|0, checked: bool, timing: bool) -> output surface
napi_value Denoise(napi_env env, napi_callback_info info) {
  napi_value a[5];
  if (!get_args(env, info, 5, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  static const char *keys[5] = {"radiance", "normalDepth", "motion", "material", "emission"};
  hush_inputs in;
  hush_surface *dst[5] = {&in.radiance, &in.normal_depth, &in.motion, &in.material, &in.emission};
  for (int k = 0; k < 5; k++) {
    napi_value v;
    bool has = false;
    napi_has_named_property(env, a[1], keys[k], &has);
    if (has) napi_get_named_property(env, a[1], keys[k], &v);
    else napi_get_undefined(env, &v);
    if (!read_surface(env, v, dst[k])) return throw_err(env, "hush addon: bad surface object");
  }
  uint64_t stream = 0;
  bool lossless = false;
  napi_valuetype t;
  napi_typeof(env, a[2], &t);
  if (t == napi_bigint) napi_get_value_bigint_uint64(env, a[2], &stream, &lossless);
  bool checked = true, timing = false;
  napi_typeof(env, a[3], &t);
  if (t == napi_boolean) napi_get_value_bool(env, a[3], &checked);
  napi_typeof(env, a[4], &t);
  if (t == napi_boolean) napi_get_value_bool(env, a[4], &timing);
  hush_set_timing(h, timing ? 1 : 0);
  hush_surface out;
  int r = checked ? hush_denoise(h, &in, (void *)(uintptr_t)stream, &out) : hush_encode(h, &in, (void *)(uintptr_t)stream, &out);
  if (r != HUSH_OK) return throw_status(env, h, r);
  return make_surface(env, out);
}

napi_value Output(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  hush_surface out;
  int r = hush_output(h, &out);
  if (r != HUSH_OK) return throw_status(env, h, r);
  return make_surface(env, out);
}

// passTimes(h) -> [{label, ms}] of the last timed frame
napi_value PassTimes(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  hush_t *h = ctx_of(env, a[0]);
  float ms[8];
  const char *labels[8];
  int r = hush_pass_times(h, ms, labels);
  if (r != HUSH_OK) return throw_status(env, h, r);
  napi_value arr;
  napi_create_array_with_length(env, 8, &arr);
  for (uint32_t i = 0; i < 8; i++) {
    napi_value o, l, m;
    napi_create_object(env, &o);
    napi_create_string_utf8(env, labels[i], NAPI_AUTO_LENGTH, &l);
    napi_create_double(env, ms[i], &m);
    napi_set_named_property(env, o, "label", l);
    napi_set_named_property(env, o, "ms", m);
    napi_set_element(env, arr, i, o);
  }
  return arr;
}

// ---- device surface plumbing (rgba16float, 8 bytes per texel)

napi_value Alloc(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  int dev = (int)num(env, a[0]), w = (int)num(env, a[1]), hh = (int)num(env, a[2]);
  if (w <= 0 || hh <= 0) return throw_err(env, "hush addon: bad surface size");
  if (hipSetDevice(dev) != hipSuccess) return throw_err(env, "hush addon: hipSetDevice failed");
  hush_surface s{};
  s.pitch = ((size_t)w * 8 + 255) / 256 * 256;
  s.width = w; s.height = hh; s.format = HUSH_RGBA16F;
  if (hipMalloc(&s.ptr, s.pitch * hh) != hipSuccess) return throw_err(env, "hush addon: hipMalloc failed");
  return make_surface(env, s);
}

napi_value Free(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  hush_surface s;
  read_surface(env, a[0], &s);
  if (s.ptr) (void)hipFree(s.ptr);
  return nullptr;
}

// upload(surface, Uint16Array of width*height*4 binary16 values)
napi_value Upload(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  hush_surface s;
  read_surface(env, a[0], &s);
  napi_typedarray_type tt;
  size_t len = 0;
  void *data = nullptr;
  if (napi_get_typedarray_info(env, a[1], &tt, &len, &data, nullptr, nullptr) != napi_ok || tt != napi_uint16_array ||
      len != (size_t)s.width * s.height * 4)
    return throw_err(env, "hush addon: upload needs a Uint16Array of width*height*4 values");
  if (hipMemcpy2D(s.ptr, s.pitch, data, (size_t)s.width * 8, (size_t)s.width * 8, s.height, hipMemcpyHostToDevice) != hipSuccess)
    return throw_err(env, "hush addon: hipMemcpy2D upload failed");
  return nullptr;
}

// download(surface) -> Uint16Array (synchronises the device)
napi_value Download(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  hush_surface s;
  read_surface(env, a[0], &s);
  size_t n = (size_t)s.width * s.height * 4;
  napi_value ab, arr;
  void *data = nullptr;
  NAPI_CALL(env, napi_create_arraybuffer(env, n * 2, &data, &ab));
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy2D(data, (size_t)s.width * 8, s.ptr, s.pitch, (size_t)s.width * 8, s.height, hipMemcpyDeviceToHost) != hipSuccess)
    return throw_err(env, "hush addon: hipMemcpy2D download failed");
  NAPI_CALL(env, napi_create_typedarray(env, napi_uint16_array, n, ab, 0, &arr));
  return arr;
}

// ---- 8-bit colour targets of the display post (4 bytes per texel)

// alloc8(device, width, height) -> {ptr, pitch, width, height}
napi_value Alloc8(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  int dev = (int)num(env, a[0]), w = (int)num(env, a[1]), hh = (int)num(env, a[2]);
  if (w <= 0 || hh <= 0) return throw_err(env, "hush addon: bad surface size");
  if (hipSetDevice(dev) != hipSuccess) return throw_err(env, "hush addon: hipSetDevice failed");
  hush_surface s{};
  s.pitch = ((size_t)w * 4 + 255) / 256 * 256;
  s.width = w; s.height = hh; s.format = -1;
  if (hipMalloc(&s.ptr, s.pitch * hh) != hipSuccess) return throw_err(env, "hush addon: hipMalloc failed");
  return make_surface(env, s);
}

// download8(surface) -> Uint8Array of width*height*4 bytes (synchronises the device)
napi_value Download8(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  hush_surface s;
  read_surface(env, a[0], &s);
  size_t n = (size_t)s.width * s.height * 4;
  napi_value ab, arr;
  void *data = nullptr;
  NAPI_CALL(env, napi_create_arraybuffer(env, n, &data, &ab));
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy2D(data, (size_t)s.width * 4, s.ptr, s.pitch, (size_t)s.width * 4, s.height, hipMemcpyDeviceToHost) != hipSuccess)
    return throw_err(env, "hush addon: hipMemcpy2D download failed");
  NAPI_CALL(env, napi_create_typedarray(env, napi_uint8_array, n, ab, 0, &arr));
  return arr;
}

napi_value DeviceCount(napi_env env, napi_callback_info) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  napi_value r;
  napi_create_int32(env, n, &r);
  return r;
}

// ---- SVGF baseline (include/svgf.h; the reference's SVGF class, src/denoise/svgf.js:273-379)

struct SvHolder {
  svgf_t *s;
};

svgf_t *sv_of(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) return nullptr;
  return static_cast<SvHolder *>(p)->s;
}

napi_value sv_throw(napi_env env, svgf_t *s, int code) {
  napi_value err, msg, c;
  const char *m = svgf_last_error(s);
  napi_create_string_utf8(env, m ? m : "svgf error", NAPI_AUTO_LENGTH, &msg);
  napi_create_error(env, nullptr, msg, &err);
  napi_create_int32(env, code, &c);
  napi_set_named_property(env, err, "code", c);
  napi_throw(env, err);
  return nullptr;
}

void finalize_sv(napi_env, void *data, void *) {
  SvHolder *hd = static_cast<SvHolder *>(data);
  if (!hd) return;
  if (hd->s) svgf_destroy(hd->s);
  delete hd;
}

// svgfCreate(device, width, height, opts {sigmaL, sigmaN, sigmaZ, alphaColor, alphaMoments, taaAlpha, taa, iterations})
napi_value SvCreate(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return nullptr;
  svgf_options o;
  svgf_default_options(&o);
  napi_valuetype t;
  napi_typeof(env, a[3], &t);
  if (t == napi_object) {
    struct { const char *k; float *f; } fl[] = {{"sigmaL", &o.sigma_l}, {"sigmaN", &o.sigma_n}, {"sigmaZ", &o.sigma_z},
                                               {"alphaColor", &o.alpha_color}, {"alphaMoments", &o.alpha_moments},
                                               {"taaAlpha", &o.taa_alpha}};
    for (auto &e : fl) {
      bool has = false;
      napi_has_named_property(env, a[3], e.k, &has);
      if (has) { napi_value v; napi_get_named_property(env, a[3], e.k, &v); *e.f = (float)num(env, v, *e.f); }
    }
    bool has = false;
    napi_has_named_property(env, a[3], "iterations", &has);
    if (has) { napi_value v; napi_get_named_property(env, a[3], "iterations", &v); o.iterations = (int32_t)num(env, v, 5); }
    napi_has_named_property(env, a[3], "taa", &has);
    if (has) {
      napi_value v;
      bool b = true;
      napi_get_named_property(env, a[3], "taa", &v);
      napi_coerce_to_bool(env, v, &v);
      napi_get_value_bool(env, v, &b);
      o.taa = b ? 1 : 0;
    }
  }
  svgf_t *s = nullptr;
  int r = svgf_create((int)num(env, a[0]), (int)num(env, a[1], 0), (int)num(env, a[2], 0), &o, &s);
  if (r != HUSH_OK) return sv_throw(env, nullptr, r);
  napi_value ext;
  SvHolder *hd = new SvHolder{s};
  if (napi_create_external(env, hd, finalize_sv, nullptr, &ext) != napi_ok) {
    svgf_destroy(s);
    delete hd;
    return throw_err(env, "hush addon: napi_create_external failed");
  }
  return ext;
}

napi_value SvDestroy(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  void *p = nullptr;
  if (napi_get_value_external(env, a[0], &p) == napi_ok && p) {
    SvHolder *hd = static_cast<SvHolder *>(p);
    if (hd->s) svgf_destroy(hd->s);
    hd->s = nullptr;
  }
  return nullptr;
}

napi_value SvResize(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  svgf_t *s = sv_of(env, a[0]);
  int r = svgf_resize(s, (int)num(env, a[1]), (int)num(env, a[2]));
  if (r != HUSH_OK) return sv_throw(env, s, r);
  return nullptr;
}

napi_value SvReset(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  svgf_t *s = sv_of(env, a[0]);
  int r = svgf_reset(s);
  if (r != HUSH_OK) return sv_throw(env, s, r);
  return nullptr;
}

napi_value SvInfo(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  svgf_t *s = sv_of(env, a[0]);
  int32_t w = 0, hh = 0;
  svgf_get_size(s, &w, &hh);
  svgf_options o{};
  svgf_get_options(s, &o);
  napi_value r, v;
  napi_create_object(env, &r);
  napi_create_int32(env, w, &v); napi_set_named_property(env, r, "width", v);
  napi_create_int32(env, hh, &v); napi_set_named_property(env, r, "height", v);
  napi_create_double(env, (double)svgf_frame(s), &v); napi_set_named_property(env, r, "frame", v);
  napi_create_int32(env, o.iterations, &v); napi_set_named_property(env, r, "iterations", v);
  napi_create_double(env, o.sigma_l, &v); napi_set_named_property(env, r, "sigmaL", v);
  napi_create_double(env, o.sigma_n, &v); napi_set_named_property(env, r, "sigmaN", v);
  napi_create_double(env, o.sigma_z, &v); napi_set_named_property(env, r, "sigmaZ", v);
  napi_create_double(env, o.alpha_color, &v); napi_set_named_property(env, r, "alphaColor", v);
  napi_create_double(env, o.alpha_moments, &v); napi_set_named_property(env, r, "alphaMoments", v);
  napi_create_double(env, o.taa_alpha, &v); napi_set_named_property(env, r, "taaAlpha", v);
  napi_get_boolean(env, o.taa != 0, &v); napi_set_named_property(env, r, "taa", v);
  return r;
}

// svgfEncode(s, g, stream: BigInt|0, timing: bool) -> output surface
napi_value SvEncode(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return nullptr;
  svgf_t *s = sv_of(env, a[0]);
  static const char *keys[7] = {"radiance", "normalDepth", "motion", "material", "emission", "prevNormalDepth", "prevMotion"};
  svgf_inputs in;
  hush_surface *dst[7] = {&in.radiance, &in.normal_depth, &in.motion, &in.material, &in.emission, &in.prev_normal_depth,
                          &in.prev_motion};
  for (int k = 0; k < 7; k++) {
    napi_value v;
    bool has = false;
    napi_has_named_property(env, a[1], keys[k], &has);
    if (has) napi_get_named_property(env, a[1], keys[k], &v);
    else napi_get_undefined(env, &v);
    if (!read_surface(env, v, dst[k])) return throw_err(env, "hush addon: bad surface object");
  }
  uint64_t stream = 0;
  bool lossless = false, timing = false;
  napi_valuetype t;
  napi_typeof(env, a[2], &t);
  if (t == napi_bigint) napi_get_value_bigint_uint64(env, a[2], &stream, &lossless);
  napi_typeof(env, a[3], &t);
  if (t == napi_boolean) napi_get_value_bool(env, a[3], &timing);
  svgf_set_timing(s, timing ? 1 : 0);
  hush_surface out;
  int r = svgf_encode(s, &in, (void *)(uintptr_t)stream, &out);
  if (r != HUSH_OK) return sv_throw(env, s, r);
  return make_surface(env, out);
}

napi_value SvOutput(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  svgf_t *s = sv_of(env, a[0]);
  hush_surface out;
  int r = svgf_output(s, &out);
  if (r != HUSH_OK) return sv_throw(env, s, r);
  return make_surface(env, out);
}

napi_value SvPassTimes(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  svgf_t *s = sv_of(env, a[0]);
  float ms[8];
  const char *labels[8];
  int32_t n = 0;
  int r = svgf_pass_times(s, ms, labels, &n);
  if (r != HUSH_OK) return sv_throw(env, s, r);
  napi_value arr;
  napi_create_array_with_length(env, (size_t)n, &arr);
  for (int32_t i = 0; i < n; i++) {
    napi_value o, l, m;
    napi_create_object(env, &o);
    napi_create_string_utf8(env, labels[i], NAPI_AUTO_LENGTH, &l);
    napi_create_double(env, ms[i], &m);
    napi_set_named_property(env, o, "label", l);
    napi_set_named_property(env, o, "ms", m);
    napi_set_element(env, arr, (uint32_t)i, o);
  }
  return arr;
}

// ---- display post (include/post.h; the reference's Post class, src/demo/post.js:131-210)

struct PoHolder {
  post_t *p;
};

post_t *po_of(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) return nullptr;
  return static_cast<PoHolder *>(p)->p;
}

napi_value po_throw(napi_env env, post_t *p, int code) {
  napi_value err, msg, c;
  const char *m = post_last_error(p);
  napi_create_string_utf8(env, m ? m : "post error", NAPI_AUTO_LENGTH, &msg);
  napi_create_error(env, nullptr, msg, &err);
  napi_create_int32(env, code, &c);
  napi_set_named_property(env, err, "code", c);
  napi_throw(env, err);
  return nullptr;
}

void finalize_po(napi_env, void *data, void *) {
  PoHolder *hd = static_cast<PoHolder *>(data);
  if (!hd) return;
  if (hd->p) post_destroy(hd->p);
  delete hd;
}

double named_num(napi_env env, napi_value o, const char *k, double dflt) {
  bool has = false;
  napi_has_named_property(env, o, k, &has);
  if (!has) return dflt;
  napi_value v;
  napi_get_named_property(env, o, k, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b = false;
    napi_get_value_bool(env, v, &b);
    return b ? 1.0 : 0.0;
  }
  return num(env, v, dflt);
}

// postCreate(device, opts {bloom, vignette})
napi_value PoCreate(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  post_options o;
  post_default_options(&o);
  napi_valuetype t;
  napi_typeof(env, a[1], &t);
  if (t == napi_object) {
    o.bloom = (float)named_num(env, a[1], "bloom", o.bloom);
    o.vignette = (float)named_num(env, a[1], "vignette", o.vignette);
  }
  post_t *p = nullptr;
  int r = post_create((int)num(env, a[0]), &o, &p);
  if (r != HUSH_OK) return po_throw(env, nullptr, r);
  napi_value ext;
  PoHolder *hd = new PoHolder{p};
  if (napi_create_external(env, hd, finalize_po, nullptr, &ext) != napi_ok) {
    post_destroy(p);
    delete hd;
    return throw_err(env, "hush addon: napi_create_external failed");
  }
  return ext;
}

napi_value PoDestroy(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  void *p = nullptr;
  if (napi_get_value_external(env, a[0], &p) == napi_ok && p) {
    PoHolder *hd = static_cast<PoHolder *>(p);
    if (hd->p) post_destroy(hd->p);
    hd->p = nullptr;
  }
  return nullptr;
}

napi_value PoResize(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  post_t *p = po_of(env, a[0]);
  int r = post_resize(p, (int)num(env, a[1]), (int)num(env, a[2]));
  if (r != HUSH_OK) return po_throw(env, p, r);
  return nullptr;
}

// postEncode(p, {left, right, mid, split, split2, same, exposure, line, target, width, height, format}, stream)
napi_value PoEncode(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  post_t *p = po_of(env, a[0]);
  post_args pa{};
  struct { const char *k; hush_surface *s; } imgs[3] = {{"left", &pa.left}, {"right", &pa.right}, {"mid", &pa.mid}};
  for (auto &e : imgs) {
    napi_value v;
    bool has = false;
    napi_has_named_property(env, a[1], e.k, &has);
    if (has) napi_get_named_property(env, a[1], e.k, &v);
    else napi_get_undefined(env, &v);
    if (!read_surface(env, v, e.s)) return throw_err(env, "hush addon: bad surface object");
  }
  pa.split = (float)named_num(env, a[1], "split", 0.5);
  pa.split2 = (float)named_num(env, a[1], "split2", pa.split);
  pa.same = named_num(env, a[1], "same", 0) != 0.0;
  pa.exposure = (float)named_num(env, a[1], "exposure", 1.0);
  pa.line = (float)named_num(env, a[1], "line", 1.5);
  pa.format = (int32_t)named_num(env, a[1], "format", POST_BGRA8_UNORM);
  hush_surface tgt;
  {
    napi_value v;
    bool has = false;
    napi_has_named_property(env, a[1], "target", &has);
    if (has) napi_get_named_property(env, a[1], "target", &v);
    else napi_get_undefined(env, &v);
    if (!read_surface(env, v, &tgt)) return throw_err(env, "hush addon: bad target object");
  }
  pa.target = tgt.ptr;
  pa.target_pitch = tgt.pitch;
  pa.width = (int32_t)named_num(env, a[1], "width", tgt.width);
  pa.height = (int32_t)named_num(env, a[1], "height", tgt.height);
  uint64_t stream = 0;
  bool lossless = false;
  napi_valuetype t;
  napi_typeof(env, a[2], &t);
  if (t == napi_bigint) napi_get_value_bigint_uint64(env, a[2], &stream, &lossless);
  int r = post_encode(p, &pa, (void *)(uintptr_t)stream);
  if (r != HUSH_OK) return po_throw(env, p, r);
  return nullptr;
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char *name; napi_callback fn; } fns[] = {
      {"hiddenOf", HiddenOf}, {"create", Create}, {"destroy", Destroy}, {"resize", Resize}, {"reset", Reset},
      {"setExposure", SetExposure}, {"setPrecision", SetPrecision}, {"setCompensated", SetCompensated},
      {"setFp8Scales", SetFp8Scales}, {"getFp8Scales", GetFp8Scales}, {"info", Info}, {"denoise", Denoise},
      {"output", Output}, {"passTimes", PassTimes}, {"alloc", Alloc}, {"free", Free}, {"upload", Upload},
      {"download", Download}, {"deviceCount", DeviceCount},
      {"svgfCreate", SvCreate}, {"svgfDestroy", SvDestroy}, {"svgfResize", SvResize}, {"svgfReset", SvReset},
      {"svgfInfo", SvInfo}, {"svgfEncode", SvEncode}, {"svgfOutput", SvOutput}, {"svgfPassTimes", SvPassTimes},
      {"alloc8", Alloc8}, {"download8", Download8}, {"postCreate", PoCreate}, {"postDestroy", PoDestroy},
      {"postResize", PoResize}, {"postEncode", PoEncode},
  };
  for (auto &f : fns) {
    napi_value v;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v);
    napi_set_named_property(env, exports, f.name, v);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
