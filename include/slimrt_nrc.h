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
/* out[i] (3 floats) = cache(raw[i]) for i < n. */
int slimrt_nrc_infer(slimrt_nrc *nrc, const float *d_raw, uint32_t n, float *d_out);
/* Loss and gradient of the relative L2 loss over n samples (targets: 3 floats each), without an update;
 * grad (host, SLIMRT_NRC_PARAMS floats) and loss (host) may be NULL.  Synchronous. */
int slimrt_nrc_gradients(slimrt_nrc *nrc, const float *d_raw, const float *d_targets, uint32_t n, float *grad, float *loss);
/* One Adam step (beta1 0.9, beta2 0.99, eps 1e-8) with learning rate lr on n samples.  Asynchronous unless loss
 * is non-NULL, in which case the step's loss is returned. */
int slimrt_nrc_train_step(slimrt_nrc *nrc, const float *d_raw, const float *d_targets, uint32_t n, float lr, float *loss);
const char *slimrt_nrc_last_error(const slimrt_nrc *nrc);

#ifdef __cplusplus
}
#endif

#endif
