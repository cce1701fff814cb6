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
  TexF pkd;                          // history: decoded key of the previous frame (normal, (depth, id))
  Tex pcol, pmom, ppar0;             // history: histColor, mom, par0 of the previous frame
  Tex outA, outMom, outKey;
  TexF outKD;                        // decoded key of this frame
  TexU F0, F1, F2, F3;
  TexF outGH;                        // decoded half-res guide (normal, packed (Z, count))
  uint8_t *mask;                     // optional per-pixel reprojection bits (debug export)
  int mask_pitch;
  int W, H, W2, H2, reset;
  float E;                           // exposure
};

struct InferArgs {
  TexU F0, F1, F2;
  TexF GH;
  Tex out0, out1, out2, outMid;      // par0[cur], par1, par2, Bh[0]
  int W2, H2, hidden;
  const float *w;                    // MLP weights widened to fp32 (hush.js:744)
  const void *wq;                    // packed low-precision weights for the MFMA forms
  const float *wscale;               // per-layer scales (FP8)
};

struct AtrousArgs {
  Tex inA;
  TexF guide;                        // GH (w = packed (Z, count)) or CGD (w = depth)
  Tex out, par1, par2;
  int w, h;
  float s;
};

struct Coarse1Args {
  TexU F2;
  TexF GH;
  Tex outC, outG;
  TexF outGD;                        // decoded coarse guide
  int W4, H4, W2, H2;
  float s;
};

struct FineArgs {
  Tex A0;
  TexF kd;                           // decoded key of this frame (the fine level's guide)
  Tex par0, par1, par2;
  TexU F2;
  TexF GH, cgd;
  Tex mid, coarse, mat, emi;
  Tex outHist, outX;
  int W, H, W2, H2, W4, H4;
};

struct TaaArgs {
  Tex X, mv, prev, out;
  int W, H, reset;
};

hipError_t launch_temporal(const TemporalArgs &a, bool fast, hipStream_t s);
hipError_t launch_infer(const InferArgs &a, int mode, bool fast, hipStream_t s);
hipError_t launch_mid2(const AtrousArgs &a, bool fast, hipStream_t s);
hipError_t launch_mid3(const AtrousArgs &a, bool fast, hipStream_t s);
hipError_t launch_coarse1(const Coarse1Args &a, bool fast, hipStream_t s);
hipError_t launch_coarse2(const AtrousArgs &a, bool fast, hipStream_t s);
hipError_t launch_fine_final(const FineArgs &a, bool fast, hipStream_t s);
hipError_t launch_taa(const TaaArgs &a, bool fast, hipStream_t s);

}  // namespace hush
