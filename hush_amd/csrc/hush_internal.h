// Kernel argument blocks and launchers of the Hush passes (passed by value: a launch captures its
// parameters, so a caller may change exposure or inputs before the previous frame has run).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hush_device.h"

namespace hush {

enum MlpMode { MLP_FP32 = 0, MLP_F16 = 1, MLP_FP16_MFMA = 2, MLP_BF16_MFMA = 3, MLP_FP8_MFMA = 4 };

// weight fragments of the MFMA MLP forms in lane order (hidden width 16), packed on the device
struct MfmaPack {
  uint4 h1[64];            // fp16 layer 1 for 16x16x32: unit l%16, k = 8(l/16) .. +7 (8 x 16-bit per lane)
  uint4 b1w[64];           // bf16 layer 1, same order
  uint2 h[2][64];          // fp16 layers 2, 3 for 16x16x16: k = 4(l/16) .. +3 (4 x 16-bit per lane)
  uint2 b[2][64];          // bf16 layers 2, 3
  uint2 q[3][64];          // fp8 E4M3: L1 k = 8(l/16) .. +7; L2, L3 units 4(l/16) .. +3 then 4 zeros
  float b1[16], b2[16], b3[16];
  float sw[3];             // fp8 weight scale per layer (power of two)
  float sx[3];             // fp8 static scale of each layer's input (power of two)
  uint32_t xclamp;         // binary16 pair (448 sx[0], 448 sx[0]), capped at 65504: the layer-1 input clamp
  float qc[3];             // sx[L] sw[L] / sx[L+1] (L < 2), sx[2] sw[2]: E4M3 sums -> next E4M3 input / output
  float qb1[16], qb2[16];  // b1 / sx[1], b2 / sx[2]
  // binary16 pairs of the shader-f16 form (units 2p, 2p + 1 of one input, for packed half arithmetic):
  // W1 [8][20], b1 [8], W2 [8][16], b2 [8], W3 [6][16], b3 [6]
  uint32_t hp[406];
};

constexpr int TEMPORAL_AUX = 20;

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
  float *aux;                        // optional temporal-blend inputs for the differentiable frame (W x H x TEMPORAL_AUX)
  int compensated;                   // Neumaier-compensated moment sums
  int wide;                          // history gathers on 64-bit addresses even below 4 GiB (debug)
  int W, H, W2, H2, reset;
  float E;                           // exposure
  float nz;                          // -0.0f, the addend of mulh (exact products with binary16 history texels)
};

struct InferArgs {
  TexU F0, F1, F2;
  TexF GH;
  Tex out0, out1, out2, outMid;      // par0[cur], par1, par2, Bh[0]
  int W2, H2, hidden;
  const float *w;                    // MLP weights widened to fp32 (hush.js:744)
  // the FP32 forms' register width (8, 16, 32 or 64; above 64 the chunked form's width, a multiple of 32) and
  // the weights zero-padded to it (input-major for the chunked form)
  int hpad;
  const float *wp;
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
  int wide;                          // 64-bit addresses even below 2 GiB per surface (debug)
};

struct TaaArgs {
  Tex X, mv, prev, out;
  int W, H, reset;
  float rW, rH;                      // RN(1 / W), RN(1 / H)
  float nz;                          // -0.0f, the addend of mulh (exact products with binary16 history texels)
};

// while set (non-null) on the calling thread, the next pass kernels record start / stop on these events
void set_pass_events(hipEvent_t start, hipEvent_t stop);
hipError_t launch_temporal(const TemporalArgs &a, bool fast, hipStream_t s);
hipError_t launch_mlp_tex(int mode, TexU F0, TexU F1, TexU F2, int W2, int H2, float *out, const void *wq, hipStream_t s);
hipError_t launch_math_selftest(int which, uint32_t base, uint32_t blocks, unsigned long long *bad, uint32_t *first,
                                hipStream_t s);
hipError_t launch_infer(const InferArgs &a, int mode, bool fast, hipStream_t s);
hipError_t launch_mid2(const AtrousArgs &a, bool fast, hipStream_t s);
hipError_t launch_mid3(const AtrousArgs &a, bool fast, hipStream_t s);
hipError_t launch_coarse1(const Coarse1Args &a, bool fast, hipStream_t s);
hipError_t launch_coarse2(const AtrousArgs &a, bool fast, hipStream_t s);
// coarse1 beside infer + mid1, coarse2 beside mid3: one launch each, same results as the separate passes
hipError_t launch_coarse1_infer(const Coarse1Args &c, const InferArgs &a, int mode, bool fast, hipStream_t s);
hipError_t launch_coarse2_mid3(const AtrousArgs &c, const AtrousArgs &m, bool fast, hipStream_t s);
hipError_t launch_fine_final(const FineArgs &a, bool fast, hipStream_t s);
hipError_t launch_taa(const TaaArgs &a, bool fast, hipStream_t s);
hipError_t launch_pack_mlp(const float *w, MfmaPack *p, const float sw[3], const float sx[3], hipStream_t s);
hipError_t launch_mlp(int mode, const uint32_t *feat, float *out, int n, const float *w, const void *wq, hipStream_t s);

}  // namespace hush
