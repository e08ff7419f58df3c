/* vt_sampler.h -- the Z++ sampler (main.swift:122-305, 497-515) on the device: the 1D / 2D / 3D draws
 * a pixel's sampling event produces in Z mode, for every pixel of an image.  Exported by libvt.so. */
#ifndef VT_SAMPLER_H
#define VT_SAMPLER_H

#include <stdint.h>

#include "vt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One event of every pixel: the pixel's key is z_pixel_key(x, y, frame_index, sample_index,
 * temporal_model) (main.swift:497-506; temporal_model 0..3 as sampler_z_temporal); the event is
 * sampler_event(path_vertex, stream) or, with candidate_bits > 0, sampler_candidate(path_vertex,
 * stream, candidate, candidate_bits) (main.swift:288-293); then `draws` draws of `kind` (1: rand_f,
 * 2: rand_f2, 3: rand_f3; main.swift:244-275) follow.  out (device memory) receives width x height x
 * draws x kind floats, pixel-major (y * width + x), draw by draw. */
typedef struct vt_z_query {
    uint32_t width, height;
    uint32_t frame_index, sample_index, temporal_model;
    uint32_t path_vertex, stream, candidate, candidate_bits;
    uint32_t draws, kind;
} vt_z_query;
vt_status vt_z_draws(const vt_z_query *query, float *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* VT_SAMPLER_H */
