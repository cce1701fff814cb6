/*
 * hush.h — C ABI of the MI355X (gfx950) Hush denoiser.
 *
 * Replaces the WebGPU `Hush` class of kanaworksai-hue/hush (src/denoise/hush.js:725-875, typed in
 * src/index.d.ts:36-60).  Device buffers stand in for GPUTextures: every surface is a row-major array of
 * rgba16float texels (four IEEE binary16 values, 8 bytes) in HBM with a byte pitch.  The five inputs have
 * the G-buffer layout of src/index.d.ts:2-18:
 *   radiance     rgb = noisy 1 spp radiance without the primary hit's emission
 *   normal_depth xyz = world normal, w = linear view depth (< 0 = background / sky)
 *   motion       xy = previous - current pixel position (pixels), z = object id (< 2048), w = previous depth
 *   material     rgb = albedo (1 for glass / lights / sky), a = roughness
 *   emission     rgb = emission of the primary hit
 * The output is rgba16float linear HDR (TAA-resolved), alpha 1.
 *
 * Errors: every call returns HUSH_OK (0) or a negative code; hush_last_error() gives the message, with
 * the reference's texts (hush.js:735, 743, 765, 816, 818).
 * Threading: one context per stream; calls on one context must be serialised by the caller.  Kernel
 * arguments are captured at launch, so unlike the reference (hush.js:811-812) a caller may change inputs
 * or exposure before the previous frame has executed.  The output of frame N is overwritten by frame N+2
 * (two alternating buffers, hush.js:796-798).
 * Streams: `stream` is a hipStream_t (NULL = the default stream) of the context's device.
 */
#ifndef HUSH_H
#define HUSH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hush_ctx hush_t;

enum hush_status {
  HUSH_OK = 0,
  HUSH_ERR_ARGUMENT = -1,     /* null pointer or invalid enum */
  HUSH_ERR_WEIGHTS = -2,      /* weight count fits no MLP width (hush.js:743) */
  HUSH_ERR_SIZE = -3,         /* non-positive size (hush.js:765), or a side above HUSH_MAX_SIDE */
  HUSH_ERR_INPUT = -4,        /* missing input or size mismatch (hush.js:816, 818) */
  HUSH_ERR_DEVICE = -5,       /* HIP runtime error (allocation, launch) */
  HUSH_ERR_UNSUPPORTED = -6   /* precision / width combination not built */
};

/* MLP arithmetic (the reference's `f16` option is HUSH_MLP_F16; the MFMA forms have no counterpart) */
enum hush_precision {
  HUSH_MLP_FP32 = 0,          /* plain FP32 MLP: the drop-in (hush.js:683-704 with f16 = false) */
  HUSH_MLP_F16 = 1,           /* binary16 arithmetic, like shader-f16 (hush.js:690, 699-702) */
  HUSH_MLP_FP16_MFMA = 2,     /* fp16 operands on MFMA (layer 1 v_mfma_f32_16x16x32_f16, layers 2-3
                                 16x16x16), fp32 accumulation */
  HUSH_MLP_BF16_MFMA = 3,     /* bf16 operands on MFMA (same shapes), fp32 accumulation */
  HUSH_MLP_FP8_MFMA = 4       /* fp8 E4M3 operands on MFMA (16x16x32), fp32 accumulation; static
                                 power-of-two scales per layer for the weights (from the weights) and for
                                 each layer's input (hush_set_fp8_scales) */
};
/* The MFMA forms need hidden width 16 (the built-in 20-16-16-12 network); other widths report
   HUSH_ERR_UNSUPPORTED for them. */

/* Largest width or height (WebGPU's maxTextureDimension2D is 8192 by default); the kernels' 32-bit texel
   and workgroup indices hold within this bound. */
#define HUSH_MAX_SIDE 16384

enum hush_format { HUSH_RGBA16F = 0, HUSH_RGBA32U = 1, HUSH_R8U = 2, HUSH_RGBA32F = 3 };

/* a device surface: `ptr` points at texel (0, 0), rows are `pitch` bytes apart */
typedef struct {
  void *ptr;
  size_t pitch;
  int32_t width, height;
  int32_t format;             /* enum hush_format; inputs and output are HUSH_RGBA16F */
} hush_surface;

/* the per-frame G-buffer (HushInputs, src/index.d.ts:11-17) */
typedef struct {
  hush_surface radiance, normal_depth, motion, material, emission;
} hush_inputs;

/* hiddenOf (hush.js:22-25): the MLP width h with h*h + 34h + 12 == count, h in [1, 256]; 0 if none */
int32_t hush_hidden_of(size_t count);

/* new Hush(device, {width, height, exposure, weights, f16}) (hush.js:731-761).  `weights` holds `count`
   binary16 values in the layout of hush-weights.bin (W1 b1 W2 b2 W3 b3, row-major (out, in),
   hush.js:20-21).  Allocates the history and level surfaces (hush.js:780-793) on `device`.
   On failure *out is NULL and hush_last_error(NULL) holds the message. */
int32_t hush_create(int32_t device, int32_t width, int32_t height, float exposure, const uint16_t *weights,
                    size_t count, int32_t precision, hush_t **out);

/* destroy() (hush.js:803-807) */
void hush_destroy(hush_t *h);

/* resize(width, height) (hush.js:764-794): no-op if unchanged, else reallocates and resets the history */
int32_t hush_resize(hush_t *h, int32_t width, int32_t height);

/* reset() (hush.js:801): the next frame starts from scratch */
int32_t hush_reset(hush_t *h);

/* the `exposure` field (index.d.ts:42-43): read at the next frame */
int32_t hush_set_exposure(hush_t *h, float exposure);
float hush_get_exposure(const hush_t *h);

/* width / height getters (index.d.ts:38-39), the MLP width and the frame counter */
int32_t hush_get_size(const hush_t *h, int32_t *width, int32_t *height);
int32_t hush_hidden(const hush_t *h);
int64_t hush_frame(const hush_t *h);

/* MLP precision of later frames (enum hush_precision) */
int32_t hush_set_precision(hush_t *h, int32_t precision);
int32_t hush_get_precision(const hush_t *h);

/* Static activation scales of the fp8 form (no counterpart in the reference): scales[L] divides the input
   of layer L (features, hidden 1, hidden 2) before its E4M3 conversion, which saturates at +-448.  Each must
   be a power of two in [2^-20, 2^20] (HUSH_ERR_ARGUMENT otherwise).  hush_create sets bounds derived from the
   weights for features of magnitude at most 16; calibrated values (the largest input of each layer over
   representative frames, as hush_amd/quant.py computes them) keep more of E4M3's range in use.  Takes
   effect for later frames and hush_mlp_forward calls; hush_get_fp8_scales reads the current ones. */
int32_t hush_set_fp8_scales(hush_t *h, const float *scales);
int32_t hush_get_fp8_scales(const hush_t *h, float *scales);

/* Math policy of later frames (no counterpart in the reference, whose arithmetic is the browser's):
   1 (default) = IEEE: correctly rounded division and square root, accurate exp / log throughout;
   0 = FAST: hardware reciprocal / square root / exp2 / log2 in the spatial filters, MLP activations and
   upsampling weights (about 1% of output values move by more than 1e-3 relative after a few frames).
   Both evaluate every threshold decision (reprojection validity, sky, history key) with IEEE arithmetic. */
int32_t hush_set_ieee_math(hush_t *h, int32_t on);
int32_t hush_get_ieee_math(const hush_t *h);

/* Compensated moment sums (no counterpart in the reference): 1 = the temporal pass's neighbourhood and
   block sums (3x3 mean and variances, 4x4 luminance moments, 2x2 block means) carry their rounding error
   (Neumaier summation); 0 (default) = plain FP32 sums in the reference's order, the drop-in. */
int32_t hush_set_compensated(hush_t *h, int32_t on);
int32_t hush_get_compensated(const hush_t *h);

/* denoise(encoder, inputs) (hush.js:813-822): validates the five inputs, enqueues the 8 passes on
   `stream` and returns the output surface in *out (may be NULL). */
int32_t hush_denoise(hush_t *h, const hush_inputs *in, void *stream, hush_surface *out);

/* encode(encoder, inputs) (hush.js:835-874): denoise() without the input checks */
int32_t hush_encode(hush_t *h, const hush_inputs *in, void *stream, hush_surface *out);

/* the `output` getter (hush.js:798): the last frame's output */
int32_t hush_output(const hush_t *h, hush_surface *out);

/* The MLP alone (no counterpart in the reference): n rows of 20 binary16 features (row-major, 40 bytes a
   row, 8-byte aligned device memory) -> n x 12 fp32 outputs (pre-activation; 16-byte aligned device memory),
   with the given precision.  HUSH_ERR_ARGUMENT for misaligned pointers.
   For timing the MLP stage in isolation and for training tools. Hidden width 16 only. */
int32_t hush_mlp_forward(hush_t *h, int32_t precision, const void *features, float *out, int64_t n, void *stream);

/* last error message of `h`, or of the calling thread's last failed hush_create when h is NULL */
const char *hush_last_error(const hush_t *h);

/* Per-pass timing, the counterpart of the optional `timer` argument (hush.js:833-856, gpu.js:25-77).
   When enabled, every frame records hipEvents around its 8 passes.  hush_pass_times waits for the last
   frame and writes 8 durations in milliseconds in dispatch order; labels[i] (optional) receive the
   reference's labels: temporal, infer+mid, infer+mid, infer+mid, coarse, coarse, fine+final, taa. */
int32_t hush_set_timing(hush_t *h, int32_t enabled);
int32_t hush_pass_times(hush_t *h, float *ms, const char **labels);

/* ---- parity and test hooks (no counterpart in the reference) ---- */

/* internal surfaces by id (hush.js:783-793): 0 A0, 1 X, 2 Bh0, 3 Bh1, 4 mom0, 5 mom1, 6 key0, 7 key1,
   8 histColor0, 9 histColor1, 10 taa0, 11 taa1, 12 par0_0, 13 par0_1, 14 par1, 15 par2, 16..19 F0..F3,
   20 C0, 21 C1, 22 CG, 23 reprojection mask (HUSH_R8U; written only while enabled); decoded guides
   (HUSH_RGBA32F, no counterpart in the reference: xyz = octDec of the stored normal): 24, 25 key of
   frame parity 0 / 1 (w = binary16 depth, id), 26 half-res guide (w = binary16 Z, count), 27 coarse
   guide (w = depth) */
int32_t hush_debug_surface(hush_t *h, int32_t id, hush_surface *out);
/* set the frame counter and the reset flag (teacher-forced single passes) */
int32_t hush_debug_set_state(hush_t *h, int64_t frame, int32_t reset);
/* run one pass (0 temporal .. 7 taa, dispatch order) at the current state without advancing */
int32_t hush_debug_run_pass(hush_t *h, int32_t pass, const hush_inputs *in, void *stream);
/* enable the per-pixel reprojection mask: bit i (0..3) = bilinear tap i valid, bit 4 = 3x3 fallback
   used, bit 5 = history used (has), bit 7 = sky */
int32_t hush_set_debug_mask(hush_t *h, int32_t enabled);
/* force the temporal and fine+final passes onto 64-bit addresses (the form taken when a surface they touch
   reaches 2 GiB; smaller ones use 32-bit byte offsets, the temporal pass's history gathers through buffer
   descriptors), so that tests can check that both forms give the same bits */
int32_t hush_debug_wide_gathers(hush_t *h, int32_t enabled);
/* self-test of the reduced-range square root, exponential, reciprocal and logarithm that the kernels use
   where their arguments allow it, against the toolchain's own sqrtf / expf / IEEE division / logf evaluated
   on the current device, over every binary32 input of the domain: which = 0 square root over x >= 2^-96,
   +0, +inf and NaN; 1 exponential over x <= 0 and NaN; 2 exponential over -103.27893 <= x <= 88.72283;
   3 reciprocal over normal x with |x| < 2^126 and NaN; 4 logarithm over every x but a subnormal; 5 the
   shared-reciprocal division x / d for every x of its domain (0, +-inf, NaN, |x| >= 2^-102) over 24 divisors
   (+-0, +-inf, NaN and normal ones; finite pairs with a normal quotient below 2^127), counted per pair;
   6 the MLP's sigmoid over every x, compared as the three activations it feeds are stored (binary16,
   truncated: 1/32 + 31/32 s, 0.5 + 7.5 s and s) against 1 / (1 + expf(-x)); 7 SVGF's a-trous weight
   exponential over every finite x <= 0 against expf: the same bits for x >= -103.27893, and below, a value
   that halves to 0 where expf is 0.
   *mismatches = number of inputs whose results differ in any bit (any NaN equals any NaN), *first = the
   smallest such input's bits (0xffffffff when there is none).  Synchronous; allocates 16 bytes of device
   memory for the counters. */
int32_t hush_debug_math_selftest(int32_t which, uint64_t *mismatches, uint32_t *first);

#ifdef __cplusplus
}
#endif

#endif /* HUSH_H */
