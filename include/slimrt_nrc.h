/* slimrt_nrc: C ABI of the neural radiance cache (no counterpart in the reference; see include/slimrt.h for the
 * ray tracer itself).
 *
 * A small MLP maps a hit point's encoded position, normal, view direction and material to the radiance that
 * leaves it (the sum of the remaining path segments of the reference integrator, src/slim/renderer/ray_tracer.h:
 * 42-105).  Inference and training run as fused BF16 MFMA kernels with FP32 accumulation on gfx950.
 *
 * Raw query record (14 floats): position xyz, normal xyz, view direction xyz, roughness, metalness, albedo rgb.
 * Network: 96 encoded inputs -> 4 hidden layers of 64 (ReLU) -> RGB; 19,728 float parameters.  All buffers
 * passed to infer/train are device pointers; the context's stream orders the work.
 */
#ifndef SLIMRT_NRC_H
#define SLIMRT_NRC_H

#include <stdint.h>

#include "slimrt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slimrt_nrc slimrt_nrc;

#define SLIMRT_NRC_PARAMS 19728u
#define SLIMRT_NRC_RAW 14u

/* Creates a cache with seeded He-uniform weights.  Positions are mapped to [-1, 1]^3 by
 * (p - center) * inv_half before the 12 frequency bands. */
int slimrt_nrc_create(slimrt_nrc **out, uint32_t seed, const float center[3], const float inv_half[3]);
int slimrt_nrc_destroy(slimrt_nrc *nrc);
int slimrt_nrc_set_stream(slimrt_nrc *nrc, void *hip_stream);
/* Float32 master parameters (host arrays of SLIMRT_NRC_PARAMS).  set resets the Adam state. */
int slimrt_nrc_get_params(slimrt_nrc *nrc, float *params);
int slimrt_nrc_set_params(slimrt_nrc *nrc, const float *params);
/* The whole training state, for checkpoints (no reference counterpart): the float32 master parameters, Adam's
 * first and second moments (host arrays of SLIMRT_NRC_PARAMS floats each) and its step count.  A cache whose state
 * was set from another's get_state takes the same next Adam steps, bit for bit.  Any pointer of get_state may be
 * NULL; set_state needs all three arrays. */
int slimrt_nrc_get_state(slimrt_nrc *nrc, float *params, float *m, float *v, uint32_t *step);
int slimrt_nrc_set_state(slimrt_nrc *nrc, const float *params, const float *m, const float *v, uint32_t step);
/* out[i] (3 floats) = cache(raw[i]) for i < n. */
int slimrt_nrc_infer(slimrt_nrc *nrc, const float *d_raw, uint32_t n, float *d_out);
/* Loss and gradient of the relative L2 loss over n samples (targets: 3 floats each), without an update;
 * grad (host, SLIMRT_NRC_PARAMS floats) and loss (host) may be NULL.  Synchronous. */
int slimrt_nrc_gradients(slimrt_nrc *nrc, const float *d_raw, const float *d_targets, uint32_t n, float *grad, float *loss);
/* One Adam step (beta1 0.9, beta2 0.99, eps 1e-8) with learning rate lr on n samples.  Asynchronous unless loss
 * is non-NULL, in which case the step's loss is returned. */
int slimrt_nrc_train_step(slimrt_nrc *nrc, const float *d_raw, const float *d_targets, uint32_t n, float lr, float *loss);
const char *slimrt_nrc_last_error(const slimrt_nrc *nrc);

/* Queue variants for work produced on the device: the count is read from device memory when the kernels run,
 * so a frame needs no host round trip.  infer_queue: out[i] = cache(raw[i]) for i < min(*d_count, cap).
 * train_queue: `steps` Adam steps of `batch` samples (a multiple of 64) drawn from the min(*d_count, cap)
 * samples in raw/targets - spread evenly over the buffer when it holds steps*batch or more, else cycled.
 * d_losses (device, `steps` floats) may be NULL; otherwise it receives each step's loss.  Asynchronous. */
int slimrt_nrc_infer_queue(slimrt_nrc *nrc, const float *d_raw, const uint32_t *d_count, uint32_t cap, float *d_out);
int slimrt_nrc_train_queue(slimrt_nrc *nrc, const float *d_raw, const float *d_targets, const uint32_t *d_count,
                           uint32_t cap, uint32_t steps, uint32_t batch, float lr, float *d_losses);

/* ---------------------------------------------------------------- rendering with the cache
 * slimrt_render_nrc renders a beauty frame (render_mode SLIMRT_MODE_BEAUTY, output canvas or radiance) in which
 * every pixel traces the reference integrator's first cache_depth bounces and, at the next hit, takes the
 * cache's answer for the rest of the path, when that hit's gate cell is confident; other pixels trace the full
 * path.  One in train_one_in waves (64 pixels, chosen per frame) traces full paths and records one training
 * sample per shaded hit: the hit record and the radiance the rest of its path returned.  The frame then trains
 * the cache (train_steps Adam steps of train_batch samples) and folds the samples' luminance into the gate
 * table: per hashed cell of side gate_cell_size, an exponential moving average (weight 1/16) of the mean and
 * mean square.  A cell is confident after gate_min_frames frames with samples when its standard deviation is
 * at most gate_threshold * (mean + 0.01); gate_threshold <= 0 turns the gate off (every pixel queries).
 * Pixels and depths are device buffers as for slimrt_render_device; the cache's work is put on the renderer's
 * stream (slimrt_set_stream), which this call also sets as the cache's stream.  Asynchronous. */
#define SLIMRT_NRC_RESET_GATE 1u

typedef struct slimrt_nrc_render_params {
    uint32_t cache_depth;      /* bounces traced before the cache answers; 0 = the cache shades the primary hit */
    uint32_t train_one_in;     /* 0 = no training waves */
    uint32_t train_steps;
    uint32_t train_batch;      /* multiple of 64 */
    float learning_rate;
    float gate_threshold;
    uint32_t gate_min_frames;
    float gate_cell_size;
    uint32_t gate_cells_log2;  /* table size; 10..24 */
    uint32_t frame;            /* selects the training waves */
    uint32_t flags;            /* SLIMRT_NRC_RESET_GATE clears the gate table first */
    uint32_t reserved;
} slimrt_nrc_render_params;

/* Frame statistics of the last slimrt_render_nrc call (synchronizes): [0] queries answered by the cache,
 * [1] training samples recorded, [2] pixels that reached the cache vertex but traced on (gate closed),
 * [3] training pixels, [4] and [5] the losses of the frame's first and last training steps (float bits; 0 without
 * training), [6], [7] zero. */
int slimrt_nrc_frame_stats(slimrt_context *ctx, uint32_t stats[8]);
/* The last frame's training samples (synchronizes): *count receives their number; the first min(count, cap) are
 * copied to the host arrays raw (14 floats each), targets (3 floats: the radiance the rest of the path
 * returned), pixel (canvas offset of the sample's path) and vertex (1 = primary hit). */
int slimrt_nrc_frame_samples(slimrt_context *ctx, float *raw, float *targets, uint32_t *pixel, uint32_t *vertex,
                             uint32_t cap, uint32_t *count);
int slimrt_render_nrc(slimrt_context *ctx, const slimrt_render_params *params, slimrt_nrc *nrc,
                      const slimrt_nrc_render_params *nrc_params, void *d_pixels, void *d_depths);

#ifdef __cplusplus
}
#endif

#endif
