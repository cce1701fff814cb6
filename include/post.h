/*
 * post.h — C ABI of the display post on MI355X (gfx950).
 *
 * Replaces the WebGPU `Post` class of kanaworksai-hue/hush (src/demo/post.js:131-210), the demo's display
 * pass after the denoisers: a split-screen composite of two or three linear HDR images times the exposure
 * (post.js:21-40), a six-level bloom pyramid (13-tap downsample with a Karis average on the first level,
 * post.js:42-66; additive 3x3 tent upsample, post.js:68-82), and the final pass (bloom mix, vignette, AgX
 * tone map with a "punchy" look, sRGB encode and the divider lines, post.js:84-127) into an 8-bit colour
 * target.  Surfaces are the rgba16float device surfaces of hush.h; the target is a caller-owned
 * device surface of 4-byte texels (RGBA8 or BGRA8 unorm, the canvas formats the reference renders to).
 *
 * Status codes follow hush.h (enum hush_status), with the message in post_last_error.
 * Streams: `stream` is a hipStream_t of the context's device (NULL = default stream); calls on one
 * context must be serialised by the caller.
 */
#ifndef POST_H
#define POST_H

#include <stddef.h>
#include <stdint.h>

#include "hush.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct post_ctx post_t;

/* the canvas format (post.js:131's `format`): byte order of the 8-bit target */
enum post_format { POST_RGBA8_UNORM = 0, POST_BGRA8_UNORM = 1 };

/* number of bloom levels (post.js:129) */
#define POST_LEVELS 6

/* constructor options (post.js:132): bloom mix 0.03, vignette strength 0.35 */
typedef struct {
  float bloom;
  float vignette;
} post_options;

/* the arguments of encode() (post.js:179).  `mid.ptr` NULL means mid = right (the reference's default);
   the caller passes split2 = split when it has no second divider (the reference's default). */
typedef struct {
  hush_surface left, right, mid;   /* linear HDR, rgba16float, all of one size */
  float split, split2;             /* divider positions in [0, 1] of the width */
  int32_t same;                    /* 1: each third shows the middle third of its own image */
  float exposure;
  float line;                      /* divider half-width in target pixels (default 1.5); <= 0: none */
  void *target;                    /* the colour target: width x height 4-byte texels */
  size_t target_pitch;             /* bytes per row */
  int32_t width, height;           /* target size (the canvas; may differ from the images') */
  int32_t format;                  /* enum post_format */
} post_args;

/* the defaults of post.js:132 */
void post_default_options(post_options *opts);

/* new Post(device, format, {bloom, vignette}) (post.js:131-147); `opts` NULL = defaults.  The HDR and
   bloom surfaces are allocated at the first encode (or post_resize) for the images' size. */
int32_t post_create(int32_t device, const post_options *opts, post_t **out);

/* destroy() (post.js:165-168) */
void post_destroy(post_t *p);

/* resize(w, h) (post.js:149-158): the HDR surface and the levels max(1, w >> i) x max(1, h >> i),
   i = 1..6; no-op when unchanged.  encode() calls it with the images' size. */
int32_t post_resize(post_t *p, int32_t width, int32_t height);

/* encode(encoder, args) (post.js:179-209): composite, six downsamples, five additive upsamples, final;
   13 kernels on `stream`.  Checks the images (present, one size, rgba16float) and the target. */
int32_t post_encode(post_t *p, const post_args *a, void *stream);

const char *post_last_error(const post_t *p);

/* ---- parity and test hooks (no counterpart in the reference) ---- */
/* internal surfaces: 0 the composited HDR image, 1..6 the bloom levels (valid after an encode) */
int32_t post_debug_surface(post_t *p, int32_t id, hush_surface *out);

#ifdef __cplusplus
}
#endif

#endif /* POST_H */
