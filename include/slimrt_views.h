/* slimrt_views: scene updates on the device and many camera views traced in one launch.
 *
 * The reference renders one viewport per call (RayTracingRenderer::render, src/slim/renderer/renderer.h:86-118,
 * via renderOnGPU, renderer_GPU.h:66-94) after rebuilding the scene BVH on the host every frame
 * (Scene::updateAABBs / updateBVH, src/slim/scene/scene.h:165-191, called from the apps' OnUpdate).  Here a
 * frame's geometry transforms go to the device once, a one-workgroup kernel recomputes the world AABBs and
 * refits or rebuilds the scene BVH there, and one launch traces every view's tiles (one workgroup per tile, or
 * a persistent grid fed by a device work counter).  The whole frame can be captured in a HIP graph and
 * replayed.
 */
#ifndef SLIMRT_VIEWS_H
#define SLIMRT_VIEWS_H

#include <stdint.h>

#include "slimrt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scene BVH update modes.  REFIT keeps the uploaded tree's topology and recomputes every node's bounds from the
 * new world AABBs (leaves first, then parents).  REBUILD runs the reference's SAH build on the device and gives
 * the same nodes, leaf order and height as slimrt_host_build_bvh (scene BVH of up to 128 geometries, leaf size
 * 1).  Hits are the same either way: a refitted tree is a valid BVH over the same boxes. */
#define SLIMRT_SCENE_REFIT 0u
#define SLIMRT_SCENE_REBUILD 1u

/* Replaces every geometry's record (count must equal the scene's geometry count; material and mesh ids are
 * checked) and updates the scene BVH on the device, on the context's stream (asynchronous). */
int slimrt_update_scene_device(slimrt_context *ctx, const slimrt_geometry *geometries, uint32_t count, uint32_t mode);

/* Copies the scene BVH as the device holds it (nodes in the asset layout of slimrt_bvh_node, leaf ids) to host
 * arrays of node_cap and leaf_cap entries; *node_count receives the node count.  Synchronizes the stream. */
int slimrt_download_scene_bvh(slimrt_context *ctx, slimrt_bvh_node *nodes, uint32_t node_cap, uint32_t *leaf_ids,
                              uint32_t leaf_cap, uint32_t *node_count);

typedef struct slimrt_views slimrt_views;

/* Launch shape of the multi-view kernel; a workgroup is one wave.  tile_shape: its tile of samples, 0 = 8x8,
 * 1 = 16x4, 2 = 32x2, 3 = 4x16.  persistent = 0: one workgroup per tile of every view in one launch (tile i on
 * workgroup i); persistent = 1: a grid of 4 * blocks_per_cu (1..4) workgroups per compute unit (blocks_per_cu
 * waves per SIMD) takes tiles_per_grab (1..8) tiles at a time from a device work counter until all views are
 * done.
 * node_width: the mesh BVH walk, 2 (sibling pairs), 4 (the SLIMRT_BVH4 collapse) or 8 (the SLIMRT_BVH8 collapse),
 * or 0 = as the views' flags ask (8 when every view has SLIMRT_BVH8, else 4 when every view has SLIMRT_BVH4 or
 * SLIMRT_BVH8). */
typedef struct slimrt_views_config {
    uint32_t tile_shape;
    uint32_t blocks_per_cu;
    uint32_t tiles_per_grab;
    uint32_t persistent;
    uint32_t node_width;
} slimrt_views_config;

/* A set of views (render params as for slimrt_render_device; beauty or debug modes, canvas or radiance output)
 * with their device output buffers, traced against the context's scene. */
int slimrt_views_create(slimrt_context *ctx, const slimrt_render_params *views, uint32_t count, void *const *d_pixels,
                        void *const *d_depths, slimrt_views **out);
int slimrt_views_destroy(slimrt_views *v);
int slimrt_views_configure(slimrt_views *v, const slimrt_views_config *cfg);
int slimrt_views_get_config(const slimrt_views *v, slimrt_views_config *cfg);
/* One frame: when geometries is not NULL, the transform update of slimrt_update_scene_device (scene_mode) first;
 * then every view.  use_graph != 0 records the frame into a HIP graph on first use (again after a change of
 * configuration or of the update kind) and replays it; the geometries are staged in pinned host memory that the
 * graph's copy reads.  Asynchronous; the frame's GPU time is available from slimrt_views_last_ms. */
int slimrt_views_frame(slimrt_views *v, const slimrt_geometry *geometries, uint32_t count, uint32_t scene_mode,
                       uint32_t use_graph);
int slimrt_views_last_ms(slimrt_views *v, float *ms);
/* Times every launch configuration (tile shape x grid or persistent form x node width 2, 4 or 8) on eager frames
 * (reps each, median) and keeps the fastest. */
int slimrt_views_autotune(slimrt_views *v, uint32_t reps, slimrt_views_config *best, float *best_ms);
const char *slimrt_views_last_error(const slimrt_views *v);

/* Times the frame forms of one view on eager frames into the given device buffers (reps each after three warm-up
 * frames, so that a form that orders its tiles by cost is timed in that order; median) and returns the fastest form's flags (params->flags with the form bits replaced) and its median time.
 * Forms: the per-sample megakernel over 8-wide nodes (SLIMRT_BVH8, when the meshes fit it), and node width 2 (the
 * uploaded BVHs) or 4 (SLIMRT_BVH4) x (the per-sample megakernel, or the SLIMRT_WAVEFRONT variants: run to the end
 * or one launch per bounce, octant-sorted or not, in 1 or 4 bands, first hits traced by the shading kernel or by a
 * separate traversal kernel).
 * The wavefront forms give the megakernel's samples bit for bit; node widths 4 and 8 differ only on exact ties. */
int slimrt_render_autotune(slimrt_context *ctx, const slimrt_render_params *params, void *d_pixels, void *d_depths,
                           uint32_t reps, uint32_t *best_flags, float *best_ms);

#ifdef __cplusplus
}
#endif

#endif
