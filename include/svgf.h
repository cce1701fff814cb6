/*
 * svgf.h — C ABI of the SVGF baseline denoiser on MI355X (gfx950).
 *
 * Replaces the WebGPU `SVGF` class of kanaworksai-hue/hush (src/denoise/svgf.js:273-379): spatiotemporal
 * variance-guided filtering (Schied et al. 2017) followed by a TAA resolve.  Surfaces are the rgba16float
 * device surfaces of hush.h (8-byte texels, byte pitch); the G-buffer layout is the one of hush.h plus the
 * previous frame's normal_depth and motion, which the caller ping-pongs (svgf.js:9).
 *
 * Status codes and messages follow hush.h (enum hush_status, svgf_last_error).
 * Streams: `stream` is a hipStream_t of the context's device (NULL = default stream); calls on one
 * context must be serialised by the caller.
 */
#ifndef SVGF_H
#define SVGF_H

#include <stddef.h>
#include <stdint.h>

#include "hush.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct svgf_ctx svgf_t;

/* constructor options (svgf.js:278-281) */
typedef struct {
  float sigma_l;          /* luminance edge-stopping, default 4 */
  float sigma_n;          /* normal edge-stopping exponent, default 128 */
  float sigma_z;          /* depth edge-stopping, default 1 */
  float alpha_color;      /* temporal blend of colour, default 0.2 */
  float alpha_moments;    /* temporal blend of the moments, default 0.2 */
  float taa_alpha;        /* TAA blend, default 0.1 */
  int32_t taa;            /* 1 = TAA resolve on (default), 0 = tone-mapped pass-through */
  int32_t iterations;     /* a-trous iterations, default 5, at most 5 (the reference has uniform buffers
                             for five, svgf.js:294) */
} svgf_options;

/* the per-frame inputs (svgf.js:341) */
typedef struct {
  hush_surface radiance, normal_depth, motion, material, emission, prev_normal_depth, prev_motion;
} svgf_inputs;

/* the defaults of svgf.js:278-281 */
void svgf_default_options(svgf_options *opts);

/* new SVGF(device, width, height, opts) (svgf.js:274-295); `opts` NULL = defaults */
int32_t svgf_create(int32_t device, int32_t width, int32_t height, const svgf_options *opts, svgf_t **out);

/* destroy() (svgf.js:321-325) */
void svgf_destroy(svgf_t *s);

/* resize(width, height) (svgf.js:298-319): no-op if unchanged, else reallocates, frame 0, history reset */
int32_t svgf_resize(svgf_t *s, int32_t width, int32_t height);

/* reset() (svgf.js:329) */
int32_t svgf_reset(svgf_t *s);

int32_t svgf_get_size(const svgf_t *s, int32_t *width, int32_t *height);
int64_t svgf_frame(const svgf_t *s);
int32_t svgf_get_options(const svgf_t *s, svgf_options *opts);

/* encode(encoder, g) (svgf.js:343-378): enqueues temporal, variance, `iterations` a-trous passes and the
   resolve on `stream`; *out (may be NULL) receives this frame's output.  Unlike the reference, which binds
   whatever it is given, the seven inputs are checked (present, the context's size, rgba16float). */
int32_t svgf_encode(svgf_t *s, const svgf_inputs *in, void *stream, hush_surface *out);

/* the `output` getter (svgf.js:327), which reads taa[frame & 1] after encode() advanced the frame: the
   buffer of the frame before the last one (the last frame's output is what svgf_encode returns) */
int32_t svgf_output(const svgf_t *s, hush_surface *out);

const char *svgf_last_error(const svgf_t *s);

/* per-pass timing (the optional `timer`, svgf.js:342): passes in dispatch order with the reference's
   labels temporal, variance, atrous x iterations, resolve.  *count receives the number of passes. */
int32_t svgf_set_timing(svgf_t *s, int32_t enabled);
int32_t svgf_pass_times(svgf_t *s, float *ms, const char **labels, int32_t *count);

/* ---- parity and test hooks (no counterpart in the reference) ---- */
/* internal surfaces (svgf.js:314-318): 0, 1 histColor; 2, 3 histMoments; 4 tmpColor; 5, 6 atrous;
   7, 8 taa */
int32_t svgf_debug_surface(svgf_t *s, int32_t id, hush_surface *out);
int32_t svgf_debug_set_state(svgf_t *s, int64_t frame, int32_t reset);
/* one pass at the current state: 0 temporal, 1 variance, 2 .. 1 + iterations a-trous, 2 + iterations
   resolve */
int32_t svgf_debug_run_pass(svgf_t *s, int32_t pass, const svgf_inputs *in, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SVGF_H */
