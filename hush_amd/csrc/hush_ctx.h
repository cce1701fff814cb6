// The denoiser context behind the opaque hush_t of include/hush.h, shared by the C ABI (hush_api.hip)
// and the training path (hush_train.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hush.h"
#include "hush_device.h"

// surface ids of the arena (the debug ids of hush_debug_surface)
enum hush_surface_id {
  S_A0, S_X, S_BH0, S_BH1, S_MOM0, S_MOM1, S_KEY0, S_KEY1, S_HIST0, S_HIST1, S_TAA0, S_TAA1,
  S_PAR0_0, S_PAR0_1, S_PAR1, S_PAR2, S_F0, S_F1, S_F2, S_F3, S_C0, S_C1, S_CG, S_MASK,
  S_KD0, S_KD1, S_GH, S_CGD, S_COUNT
};

struct hush_ctx {
  int device = 0;
  int W = 0, H = 0, W2 = 0, H2 = 0, W4 = 0, H4 = 0;
  int64_t frame = 0;
  bool reset = true;
  float exposure = 1.0f;
  int hidden = 0;
  int precision = HUSH_MLP_FP32;
  float *d_w = nullptr;
  float *d_wpad = nullptr;       // the weights zero-padded to hpad units (d_w itself when hpad == hidden)
  int hpad = 0;
  void *d_wq = nullptr;
  float *d_wscale = nullptr;
  float fp8_sw[3] = {1.0f, 1.0f, 1.0f};   // fp8 form: per-layer weight scales (from the weights)
  float fp8_sx[3] = {1.0f, 1.0f, 1.0f};   // fp8 form: per-layer static input scales
  void *arena = nullptr;
  hush::Tex tex[S_COUNT] = {};
  hush::TexU F[4] = {};
  hush::TexF kd[2] = {}, gh = {}, cgd = {};   // decoded guides (key of each frame, half-res, quarter-res)
  uint8_t *mask = nullptr;
  int mask_pitch = 0;
  bool mask_on = false;
  bool wide_gathers = false;        // debug: temporal's history gathers on 64-bit addresses at any size
  float *train_aux = nullptr;      // set by the differentiable frame around its temporal pass
  bool ieee = true;                 // IEEE math policy (the oracle's operation-exact mirror); false = FAST
  bool compensated = false;         // compensated (Neumaier) moment sums in the temporal statistics
  bool timing = false;              // timed frames: the eight passes one after another, events between them
  bool ev_ready = false;
  hipEvent_t ev[16] = {};            // per-pass timing: start and stop of each pass's kernel
  char err[512] = {0};
};

// error helpers (message into h->err, or the thread's create error when h is NULL)
__attribute__((visibility("hidden"))) int hush_fail(hush_t *h, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
__attribute__((visibility("hidden"))) int hush_hip_fail(hush_t *h, hipError_t e, const char *what);
// one pass of the frame at the context's current state (0 temporal .. 7 taa)
__attribute__((visibility("hidden"))) hipError_t hush_run_pass(hush_t *h, int pass, const hush_inputs *in, hipStream_t s);
// the five G-buffer inputs against the context's size, format, alignment and pitch bounds (message on h)
__attribute__((visibility("hidden"))) int hush_check_inputs(hush_t *h, const hush_inputs *in);
