/*
 * hush_train.h — differentiable frame of the Hush denoiser on MI355X (gfx950): forward and hand-written
 * backward kernels over the same device surfaces as hush.h, for training the MLP weights.
 *
 * No counterpart in the reference (which ships trained weights only, src/denoise/hush-weights.bin); the
 * trained weights use its layout (W1 b1 W2 b2 W3 b3, row-major (out, in), hush.js:20-21), so the
 * reference's Hush loads them unchanged once rounded to binary16.
 *
 * The differentiable function of one frame.  The temporal pass and the coarse level (which has no learned
 * parameters) run as in hush_encode, then, in the context's arithmetic (FP32, or FP64 for gradient
 * checks) and without binary16 rounding between the passes:
 *   MLP (hush.js:348-366)  ->  mid1, mid2, mid3 a-trous (hush.js:431-476, 366, 753-754)
 *   fine a-trous on the temporal output  ->  level recombination (hush.js:560-620)  ->  TAA (hush.js:625-677)
 * The temporal output A0 is re-blended (hush.js:189-211) from the temporal pass's reprojected history
 * and the previous frame's alpha / gamma, which the MLP recomputes from that frame's stored features with
 * the same weights.  The gradient flows to the 812 weights through the fine / mid edge-stopping strengths
 * s and centre weights c, the detail gain kappa and the coarse blend beta of this frame, and through the
 * temporal blend to alpha / gamma of the previous frame (one step of backpropagation through time); the
 * reprojected history itself, the previous TAA output and this frame's MLP features are inputs.  The
 * backward recomputes every a-trous weight instead of storing it.
 *
 * Buffers passed here are plain device arrays, row-major, 4 values per pixel (rgba), of the context's
 * element type (float, or double when created with fp64 = 1).  Weights and their gradient: 812 values.
 */
#ifndef HUSH_TRAIN_H
#define HUSH_TRAIN_H

#include <stddef.h>
#include <stdint.h>

#include "hush.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hush_train_ctx hush_train_t;

/* attaches to a denoiser context (its size, device, history); hidden width 16 only */
int32_t hush_train_create(hush_t *h, int32_t fp64, hush_train_t **out);
void hush_train_destroy(hush_train_t *t);
const char *hush_train_last_error(const hush_train_t *t);

/* runs the temporal pass (also recording its blend inputs) and the coarse level for this frame's inputs at
   the context's state, and copies the stored temporal output A0 into the element type */
int32_t hush_train_prepare(hush_train_t *t, const hush_inputs *in, void *stream);

/* forward of the prepared frame with `weights` (812 values) -> `out` (W x H x 4) */
int32_t hush_train_forward(hush_train_t *t, const void *weights, void *out, void *stream);

/* loss = mean over pixels of sum over rgb |tm(out) - tm(target)|
          + lambda * |(tm(out) - tm(prev_out)) - (tm(target) - tm(prev_target))|   (second term only when
   prev_out and prev_target are given), tm(c) = c / (1 + max(lum(c), 0)); the sum is reduced in FP64.
   Writes dloss/dout to grad_out and the loss to *loss_dev (one double in device memory). */
int32_t hush_train_loss(hush_train_t *t, const void *out, const void *target, const void *prev_out,
                        const void *prev_target, double lambda, void *grad_out, double *loss_dev, void *stream);

/* hush_train_loss without a training context and without any allocation.  No counterpart in the
   reference: for callers that cannot allocate device memory (code under HIP graph capture, tight serving or
   training loops, harnesses that forbid hipMalloc).  Same loss, gradient and FP64 reduction as
   hush_train_loss (the same kernels over the same workgroups, so the same bits) for width x height pixels
   of float (fp64 = 0) or double (fp64 = 1) rgba values, on the current HIP device (`stream` must belong to
   it).  The per-workgroup partial sums live in a static device array of 512 KiB per GPU, held until the
   process ends: at most 16,777,216 pixels (4096 x 4096), larger sizes return HUSH_ERR_SIZE before any
   launch.  Calls that overlap on one GPU share that array and race.  Errors: hush_train_last_error(NULL). */
int32_t hush_loss_noalloc(int32_t fp64, int32_t width, int32_t height, const void *out, const void *target,
                          const void *prev_out, const void *prev_target, double lambda, void *grad_out,
                          double *loss_dev, void *stream);

/* backward of the last forward: grad_out (W x H x 4) -> grad_w (812 values, overwritten); grad_a0
   (optional, W x H x 4) receives the gradient with respect to the temporal output A0 */
int32_t hush_train_backward(hush_train_t *t, const void *weights, const void *grad_out, void *grad_w, void *grad_a0,
                            void *stream);

/* ends the frame: the fine level and the MLP's (alpha, gamma, beta, kappa) become the denoiser's history
   (rounded to binary16 as hush_encode stores them), `out` becomes the previous TAA output, the frame
   counter advances and the reset flag clears */
int32_t hush_train_commit(hush_train_t *t, const void *out, void *stream);

/* MLP arithmetic of later forwards (enum hush_precision): HUSH_MLP_FP32 (default), HUSH_MLP_FP16_MFMA or
   HUSH_MLP_BF16_MFMA (FP32 contexts only): the MFMA forms of hush.h run the MLP of this frame (and of the
   previous frame, for the temporal blend) on weights packed from the FP32 weights at every forward; the
   rest of the forward and the whole backward stay FP32, the MLP's weight gradient recomputing its hidden
   layers in FP32 (mixed-precision training: low-precision forward, FP32 master weights and gradients) */
int32_t hush_train_set_forward_precision(hush_train_t *t, int32_t precision);

/* A0 in the element type (W x H x 4): after a forward, the re-blended temporal output */
int32_t hush_train_a0(hush_train_t *t, void **ptr);

/* the MLP's 12 outputs before the activations (W/2 x H/2 x 12, element type, row-major; zero for sky
   blocks): after a forward, those of this frame */
int32_t hush_train_mlp_out(hush_train_t *t, void **ptr);

/* fixed = 1: later forwards take A0 as it is in that buffer (writable, for input gradient checks) instead
   of re-blending it, and the backward stops at A0; 0 (default): re-blend */
int32_t hush_train_set_a0_fixed(hush_train_t *t, int32_t fixed);

#ifdef __cplusplus
}
#endif

#endif /* HUSH_TRAIN_H */
