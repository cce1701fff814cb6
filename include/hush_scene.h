/*
 * hush_scene.h — GPU render of the seeded procedural G-buffer sequences that drive the denoiser's
 * workloads (no counterpart in the reference, whose inputs come from its WebGPU path tracer; the
 * generator follows that tracer's G-buffer conventions, src/pt/pathtracer.wgsl.js:505-577, 657-661).
 *
 * The host computes a frame's set-up — cameras, primitive positions and motions, per-16x16-tile primitive
 * lists — with hs_frame_setup of hush_amd/libhush_scene.so; hush_scene_render draws every pixel of the crop
 * window on the GPU (FP64, the CPU renderer's operation order) into five rgba16float surfaces and an
 * optional float32 clean target.
 */
#ifndef HUSH_SCENE_H
#define HUSH_SCENE_H

#include <stddef.h>
#include <stdint.h>

#include "hush.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hush_scene_ctx hush_scene_t;

/* one frame's set-up in hs_frame_setup's layout (host memory) */
typedef struct {
  const double *cams;        /* 2 x 14: current, previous camera (pos, fwd, right, up, tan(fov/2) x and y) */
  const double *prims;       /* n_prims x 20: type, id, pos, motion, half extents, albedo, roughness, emission */
  const int32_t *off;        /* ceil(w/16) * ceil(h/16) + 1 tile offsets into lst */
  const int32_t *lst;        /* n_list primitive indices */
  int32_t n_prims, n_list;
  uint64_t seed;             /* the scene's seed (hs_seed) */
  int32_t frame;
  int32_t width, height;     /* full image */
  int32_t x0, y0, w, h;      /* crop window */
} hush_scene_frame;

int32_t hush_scene_create(int32_t device, hush_scene_t **out);
void hush_scene_destroy(hush_scene_t *g);
const char *hush_scene_last_error(const hush_scene_t *g);

/* draws the crop window into out->radiance .. out->emission (w x h rgba16float surfaces) and, if clean is
   not NULL, the noise-free radiance (w x h x 4 float32, `clean_pitch` bytes per row); stream-ordered on
   `stream` (one stream per context: the set-up is staged in the context's device memory) */
int32_t hush_scene_render(hush_scene_t *g, const hush_scene_frame *fr, const hush_inputs *out, float *clean,
                          size_t clean_pitch, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HUSH_SCENE_H */
