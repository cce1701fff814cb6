// Kernel argument blocks and launchers of the Hush passes (passed by value: a launch captures its
// parameters, so a caller may change exposure or inputs before the previous frame has run).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hush_device.h"

namespace hush {

enum MlpMode { MLP_FP32 = 0, MLP_F16 = 1, MLP_FP16_MFMA = 2, MLP_BF16_MFMA = 3, MLP_FP8_MFMA = 4 };

struct TemporalArgs {
  Tex rad, nd, mv, mat;              // this frame's G-buffer
  Tex pkey, pcol, pmom, ppar0;       // history: key, histColor, mom, par0 of the previous frame
  Tex outA, outMom, outKey;
  TexU F0, F1, F2, F3;
  uint8_t *mask;                     // optional per-pixel reprojection bits (debug export)
  int mask_pitch;
  int W, H, W2, H2, reset;
  float E;                           // exposure
};

struct InferArgs {
  TexU F0, F1, F2, F3;
  Tex out0, out1, out2, outMid;      // par0[cur], par1, par2, Bh[0]
  int W2, H2, hidden;
  const float *w;                    // MLP weights widened to fp32 (hush.js:744)
  const void *wq;                    // packed low-precision weights for the MFMA forms
  const float *wscale;               // per-layer scales (FP8)
};

struct AtrousArgs {
  Tex inA, guideT;
  TexU guideF;
  Tex out, par1, par2;
  int w, h;
  float s;
};

struct Coarse1Args {
  TexU F2, F3;
  Tex outC, outG;
  int W4, H4, W2, H2;
  float s;
};

struct FineArgs {
  Tex A0, key, par0, par1, par2;
  TexU F2, F3;
  Tex mid, coarse, cguide, mat, emi;
  Tex outHist, outX;
  int W, H, W2, H2, W4, H4;
};

struct TaaArgs {
  Tex X, mv, prev, out;
  int W, H, reset;
};

hipError_t launch_temporal(const TemporalArgs &a, hipStream_t s);
hipError_t launch_infer(const InferArgs &a, int mode, hipStream_t s);
hipError_t launch_mid2(const AtrousArgs &a, hipStream_t s);
hipError_t launch_mid3(const AtrousArgs &a, hipStream_t s);
hipError_t launch_coarse1(const Coarse1Args &a, hipStream_t s);
hipError_t launch_coarse2(const AtrousArgs &a, hipStream_t s);
hipError_t launch_fine_final(const FineArgs &a, hipStream_t s);
hipError_t launch_taa(const TaaArgs &a, hipStream_t s);

}  // namespace hush
