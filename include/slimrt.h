/* slimrt: C ABI of the MI355X per-pixel ray tracer.
 *
 * This is the boundary the reference's GPU path crosses (src/slim/renderer/renderer_GPU.h): scene upload once,
 * per-frame uploads of what moves, one render call into the canvas' pixel and depth buffers.  Plain C: no
 * torch or HIP types in any signature; streams are passed as void*.
 *
 * Every function returns SLIMRT_OK (0) or a negative SLIMRT_E_* code; slimrt_last_error(ctx) then holds a
 * message.  Calls on one context must not overlap (the reference is single-threaded too,
 * src/slim/renderer/renderer.h:24-35); distinct contexts are independent.
 *
 * Struct layouts use 32-bit integers (the layout of the reference's asset files, where u32 is 4 bytes).
 */
#ifndef SLIMRT_H
#define SLIMRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLIMRT_OK 0
#define SLIMRT_E_ARG (-1)        /* invalid argument */
#define SLIMRT_E_DEVICE (-2)     /* no usable GPU, or a HIP runtime error */
#define SLIMRT_E_MEMORY (-3)     /* device allocation failed, or beyond the context's memory limit */
#define SLIMRT_E_STATE (-4)      /* call out of order (e.g. render before a scene upload) */
#define SLIMRT_E_OVERFLOW (-5)   /* a traversal stack would have overflowed; results of that frame are invalid */

/* slimrt_create flags */
#define SLIMRT_DEVICE_GPU 0u     /* render with the gfx950 kernels (fails with SLIMRT_E_DEVICE without a GPU) */
#define SLIMRT_DEVICE_CPU 1u     /* render on host threads with the same per-pixel code (the CPU side of use_GPU) */

/* render_mode values: RenderMode (src/slim/core/base.h:517-524) */
#define SLIMRT_MODE_NORMALS 0u
#define SLIMRT_MODE_NORMAL_MAP 1u
#define SLIMRT_MODE_BEAUTY 2u
#define SLIMRT_MODE_DEPTH 3u
#define SLIMRT_MODE_MIP_LEVEL 4u
#define SLIMRT_MODE_UVS 5u

/* output values */
#define SLIMRT_OUTPUT_CANVAS 0u    /* Canvas::setPixel with opacity -1: RGBA = clamp(c)^2, 1 (src/slim/draw/canvas.h:128-169) */
#define SLIMRT_OUTPUT_RADIANCE 1u  /* the colour renderPixel returns (tone-mapped, unclamped), alpha 1 */
#define SLIMRT_OUTPUT_HITS 2u      /* primary closest hit per sample: int32 geometry index (-1 miss), int32 triangle id
                                      (-1 unless a mesh), float distance, 0 - in the four 32-bit lanes of the pixel */

/* render flags */
#define SLIMRT_RAY_CONES 1u        /* keep the primary ray's cone scaling through hits, for mip selection by ray cones
                                      (surface_shader.h:67-77); without it the hit record's copy in
                                      SceneTracer::hitGeometries (scene_tracer.h:129) zeroes it, as in the reference */
#define SLIMRT_WAVEFRONT 2u        /* beauty frames: after the first hit, the paths that bounce go to a compacted ray
                                      queue and one launch runs each queued path to its end, instead of the
                                      per-sample loop of one launch; same results bit for bit */
#define SLIMRT_WAVEFRONT_PER_BOUNCE 4u /* with SLIMRT_WAVEFRONT: one launch per bounce, the queue compacted again
                                          after every bounce */
#define SLIMRT_WAVEFRONT_SORT 8u   /* with SLIMRT_WAVEFRONT: sort each queue by direction octant before its launch */
/* with SLIMRT_WAVEFRONT: the frame in 2^log2n horizontal bands; a band's bounces run on a second stream while
 * the next band's first hits are traced, so the bounce launches' latency tails overlap the primary work */
#define SLIMRT_WAVEFRONT_BANDS(log2n) (((uint32_t)(log2n) & 3u) << 4)
#define SLIMRT_BVH4 64u            /* trace meshes through 4-wide BVHs collapsed from the uploaded ones (same leaves
                                      and triangles; only exact ties between triangles can resolve differently) */
#define SLIMRT_WAVEFRONT_COUNT 128u /* with SLIMRT_WAVEFRONT: also count the paths of the frame (per iteration,
                                       spawned, retired) for slimrt_wavefront_counts; costs atomics, for checks */
#define SLIMRT_WAVEFRONT_SPLIT 256u /* with SLIMRT_WAVEFRONT: the first hits are found by a traversal-only kernel
                                       (fewer registers, more waves in flight) that hands hit records to the
                                       shading kernel through device memory */
#define SLIMRT_BVH8 1024u          /* per-sample frames (no SLIMRT_WAVEFRONT) and the CPU context: trace meshes through
                                      8-wide BVHs collapsed from the uploaded ones, with 8-bit child boxes on a grid
                                      per node (128-byte nodes) that contain the exact boxes; every leaf repeats the
                                      exact leaf-box test before its triangles, so results are those of the
                                      sibling-pair walk up to exact ties between triangles.  Meshes need at most 65536
                                      such nodes and leaves of at most 254 triangles */
#define SLIMRT_TILE_ORDER_FIXED 512u /* per-sample frames (no SLIMRT_WAVEFRONT): dispatch the sample tiles in image
                                        order.  Without it a frame whose longest tiles bound its length (the longest
                                        tile three frames before lasted at least 3/4 of that frame's summed tile
                                        durations over the resident waves) dispatches its costliest tiles first and
                                        renders the first 1 in 32 of them as two half waves (with SLIMRT_BVH8 the
                                        first 1 in 1024 as four quarter waves); other frames, and the first three at
                                        a tile count, keep the image order.  Results are the same either way.  The
                                        cost order allocates its buffers on a context's first frame at a larger tile
                                        count and rotates them from frame to frame (each frame's first workgroups
                                        count and place for the next ones): callers that capture frames in a HIP
                                        graph pass this flag */

typedef struct slimrt_context slimrt_context;

/* Geometry (src/slim/core/transform.h:59-65) with its Transform (transform.h:5-10); orientation is the
 * quaternion (axis xyz, amount) of OrientationUsingQuaternion. */
typedef struct {
    float orientation[4];
    float position[3];
    float scale[3];
    uint32_t type;           /* GeometryType (base.h:399-410) */
    uint32_t material_id;
    uint32_t id;             /* mesh index for GeometryType_Mesh */
    uint32_t flags;          /* GEOMETRY_IS_VISIBLE | _SHADOWING | _TRANSPARENT (base.h:148-150) */
} slimrt_geometry;

/* Material (src/slim/scene/material.h:93-117) */
typedef struct {
    float albedo[3];
    float roughness;
    float emission[3];
    float ior;
    float reflectivity[3];
    float metalness;
    float uv_repeat[2];
    float normal_magnitude;
    uint32_t flags;          /* MATERIAL_* (base.h:142-146) */
    uint32_t texture_count;
    uint32_t texture_ids[2];
    uint32_t brdf;           /* BRDFType (base.h:624-629) */
} slimrt_material;

/* Light (src/slim/scene/material.h:17-25) */
typedef struct {
    float color[3];
    float intensity;
    float position_or_direction[3];
    uint32_t flags;          /* Light_IsDirectional = 1, Light_IsShadowing = 2 */
} slimrt_light;

/* BVHNode (src/slim/scene/bvh.h:5-15) in the 32-byte layout of the reference's asset files:
 * AABB (min, max), first_index (child pair, or first leaf item), leaf_count (0 for inner nodes), depth, flags. */
typedef struct {
    float min[3];
    float max[3];
    uint32_t first_index;
    uint16_t leaf_count;
    uint8_t depth;
    uint8_t flags;
} slimrt_bvh_node;

/* Triangle (src/slim/scene/mesh.h:19-24), 128 bytes. */
typedef struct {
    float local_to_tangent[9];   /* rows X, Y, Z */
    float position[3], normal[3], n1[3], n2[3], n3[3];
    float uv1[2], uv2[2], uv3[2];
    float uv_coverage, padding;
} slimrt_triangle;

/* Mesh (src/slim/scene/mesh.h:26-45): object-space AABB, BVH and leaf-ordered triangles. */
typedef struct {
    float aabb_min[3], aabb_max[3];
    const slimrt_bvh_node *nodes;
    uint32_t node_count, height;
    const slimrt_triangle *triangles;
    uint32_t triangle_count, normals_count, uvs_count;
} slimrt_mesh;

/* TextureMip (src/slim/core/texture.h:13-16): (width+1)*(height+1) texel quads of 12 bytes each, or, when the
 * texture's flags select texel storage (SLIMRT_TEXELS_*), (width+2)*(height+2) texels of that format. */
typedef struct {
    uint32_t width, height;
    const uint8_t *texel_quads;
} slimrt_texture_mip;

/* Texel storage, bits 24-25 of a texture's flags (no reference counterpart: the reference stores texel quads).
 * The texel formats hold a mip as a grid of (width+2)*(height+2) texels, rows top to bottom, with a one-texel
 * border: texel (x, y) of that grid is the top-left corner of quad (x, y), so the border holds what the quads'
 * outer corners hold (wrapped or clamped texels, or for cube maps the neighbouring faces' texels).
 *   SLIMRT_TEXELS_RGBA8:   4 bytes per texel (R, G, B, A; A unused), sampled as u8 / 255 like the quads, so a
 *                          texture converted from consistent quads samples to the same bits (3x less memory).
 *   SLIMRT_TEXELS_RGBA16F: 8 bytes per texel, four IEEE half floats, sampled as they are (HDR values allowed). */
#define SLIMRT_TEXELS_RGBA8   (1u << 24)
#define SLIMRT_TEXELS_RGBA16F (2u << 24)
#define SLIMRT_TEXELS_MASK    (3u << 24)

/* Texture (src/slim/core/texture.h:44-46): mip 0 size, ImageFlags bits (base.h:1283-1295) | texel storage,
 * mips. */
typedef struct {
    uint32_t width, height, flags, mip_count;
    const slimrt_texture_mip *mips;
} slimrt_texture;

/* Everything initDataOnGPU uploads (src/slim/renderer/renderer_GPU.h:80-181). */
typedef struct {
    const slimrt_geometry *geometries;  uint32_t geometry_count;
    const slimrt_material *materials;   uint32_t material_count;
    const slimrt_light *lights;         uint32_t light_count;
    const slimrt_mesh *meshes;          uint32_t mesh_count;
    const slimrt_texture *textures;     uint32_t texture_count;
    const slimrt_bvh_node *scene_nodes; uint32_t scene_node_count, scene_height;
    const uint32_t *scene_leaf_ids;     /* bvh_leaf_geometry_indices, geometry_count entries */
} slimrt_scene_desc;

/* CameraRayProjection (src/slim/scene/camera.h:59-64) + RayTracerSettings (src/slim/renderer/ray_tracer.h:10-17)
 * + canvas dimensions (CanvasData, src/slim/draw/canvas.h:11-18). */
typedef struct {
    float inverted_camera_rotation[9];
    float start[3], right[3], down[3], camera_position[3];
    float C_start[2];
    float squared_distance_to_projection_plane;
    float sample_size;
    uint32_t width, height;     /* canvas pixels */
    uint32_t ssaa;              /* 1: 2x2 samples per pixel, stored as pixel quads (canvas.h:147) */
    uint32_t max_depth;
    uint32_t render_mode;
    int32_t skybox_color, skybox_radiance, skybox_irradiance;   /* texture ids, -1 for none */
    uint32_t output;            /* SLIMRT_OUTPUT_* */
    uint32_t flags;             /* SLIMRT_RAY_CONES | SLIMRT_WAVEFRONT (| _PER_BOUNCE | _SORT | _BANDS(n) | _COUNT |
                                   _SPLIT) | SLIMRT_BVH4 */
    float mip_level_colors[9][3];
} slimrt_render_params;

/* Pixel (src/slim/core/base.h:1091-1094) */
typedef struct { float r, g, b, opacity; } slimrt_pixel;

/* ---- context ------------------------------------------------------------------------------------------ */
/* Replaces the one-time GPU setup of RayTracingRenderer's constructor (renderer.h:36-61). */
int slimrt_create(slimrt_context **out, uint32_t flags);
int slimrt_destroy(slimrt_context *ctx);
const char *slimrt_last_error(const slimrt_context *ctx);
/* Launch subsequent work on this HIP stream (hipStream_t passed as void*); NULL selects the context's own. */
int slimrt_set_stream(slimrt_context *ctx, void *hip_stream);

/* ---- uploads ------------------------------------------------------------------------------------------ */
/* initDataOnGPU (src/slim/renderer/renderer_GPU.h:80-181).  Meshes and textures are concatenated into flat
 * device arrays; geometries may be re-uploaded per frame with the functions below. */
int slimrt_upload_scene(slimrt_context *ctx, const slimrt_scene_desc *scene);
/* uploadGeometries / uploadMaterials / uploadLights / uploadSceneBVH (renderer_GPU.h:70-78). */
int slimrt_upload_geometries(slimrt_context *ctx, const slimrt_geometry *geometries, uint32_t count);
int slimrt_upload_materials(slimrt_context *ctx, const slimrt_material *materials, uint32_t count);
int slimrt_upload_lights(slimrt_context *ctx, const slimrt_light *lights, uint32_t count);
int slimrt_upload_scene_bvh(slimrt_context *ctx, const slimrt_bvh_node *nodes, uint32_t node_count,
                            const uint32_t *leaf_ids, uint32_t height);

/* ---- rendering ---------------------------------------------------------------------------------------- */
/* renderOnGPU (renderer_GPU.h:48-68): render a frame and copy it into host canvas buffers (pixels: one
 * slimrt_pixel per sample; depths: one float per sample).  Synchronous, as the reference's download is. */
int slimrt_render(slimrt_context *ctx, const slimrt_render_params *params, slimrt_pixel *pixels, float *depths);
/* Same frame into device buffers, asynchronously on the context's stream (no reference counterpart: for
 * callers that keep the canvas in HBM).  Check the frame's status with slimrt_synchronize. */
int slimrt_render_device(slimrt_context *ctx, const slimrt_render_params *params, void *d_pixels, void *d_depths);
/* slimrt_render_device that allocates nothing (no counterpart in the reference): for callers that cannot
 * allocate device memory - code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Per-sample frames only (the megakernel at either node width; SLIMRT_WAVEFRONT forms need ray queues the library
 * allocates: SLIMRT_E_ARG).  With tiles costliest first (unless SLIMRT_TILE_ORDER_FIXED) the order kernel runs on
 * the context's stream right before the frame, from the tile durations the last such frame on this GPU recorded
 * (the image order when they belong to another tile count), so one stream holds the whole frame.  Those durations
 * and the order live in one static device array per GPU (1 MiB, every context on the GPU shares it; it keeps the
 * last frame's contents): frames of at most 131072 tiles of 32 x 2 samples (8,388,608 samples: 3840 x 2160, or
 * 1920 x 1080 with SSAA), larger ones fail with SLIMRT_E_ARG.  Calls that overlap on one GPU (other streams,
 * other contexts) race on that array and must be serialised by the caller.  The frame's samples equal
 * slimrt_render_device's bit for bit.  Asynchronous: check the status with slimrt_synchronize. */
int slimrt_render_device_noalloc(slimrt_context *ctx, const slimrt_render_params *params, void *d_pixels,
                                 void *d_depths);
/* Waits for the context's stream and reports the status of the frames since the last call
 * (SLIMRT_E_OVERFLOW if a traversal stack overflowed). */
int slimrt_synchronize(slimrt_context *ctx);
/* Status flags of the frames since the last call, then cleared (no reference counterpart; the reference has no
 * device-side checks, src/slim/core/base.h:20-26): SLIMRT_FLAG_STACK_OVERFLOW and SLIMRT_FLAG_QUEUE_OVERFLOW (also
 * reported as SLIMRT_E_OVERFLOW by the next slimrt_synchronize or synchronous call), SLIMRT_FLAG_NAN_SAMPLE: some
 * sample's radiance, canvas value or depth came out NaN (for example from NaN material or light inputs; the
 * reference passes such values through silently, and so do the samples: this flag only reports it).
 * Synchronizes the context's stream. */
#define SLIMRT_FLAG_STACK_OVERFLOW 1u
#define SLIMRT_FLAG_QUEUE_OVERFLOW 2u
#define SLIMRT_FLAG_NAN_SAMPLE 4u
int slimrt_frame_flags(slimrt_context *ctx, uint32_t *flags);
/* A limit on the device memory the context holds (no reference counterpart): every allocation it makes from
 * then on (scene arrays, canvases, queues, cache buffers) that would take its total past `bytes` fails with
 * SLIMRT_E_MEMORY, leaving what it already holds in place; 0 removes the limit.  slimrt_memory_usage reports the
 * bytes held and the limit (either pointer may be NULL). */
int slimrt_set_memory_limit(slimrt_context *ctx, uint64_t bytes);
int slimrt_memory_usage(slimrt_context *ctx, uint64_t *used, uint64_t *limit);
/* Renders one frame with work counters and returns them: rays (SceneTracer::trace calls), sibling-pair
 * visits, root-box tests, instance (leaf geometry) visits, triangle tests, mesh closest-hit attribute
 * fetches, texel-quad taps, samples.  Slower than slimrt_render; for measurement only. */
int slimrt_count_work(slimrt_context *ctx, const slimrt_render_params *params, uint64_t counters[8]);
/* Kernel time of the last slimrt_render / slimrt_render_device in milliseconds (HIP events on the stream). */
int slimrt_last_kernel_ms(slimrt_context *ctx, float *ms);

/* ---- host-side scene derivation (CPU; what the reference does on the host before uploading) ------------- */
/* Orientation<quat>::setRotation (src/slim/core/base.h:574-620): Euler x, y, z -> quaternion. */
int slimrt_host_orientation(const float euler[3], float quat_out[4]);
/* CameraRayProjection::reset (src/slim/scene/camera.h:69-83) for a Camera with OrientationUsing3x3Matrix
 * Euler angles, position and focal length; fills the projection fields and width/height/ssaa of params. */
int slimrt_host_projection(const float camera_euler[3], const float camera_position[3], float focal_length,
                           uint32_t width, uint32_t height, uint32_t ssaa, slimrt_render_params *params);
/* Scene::updateAABBs (src/slim/scene/scene.h:165-179): world AABBs (min xyz, max xyz per geometry). */
int slimrt_host_world_aabbs(const slimrt_geometry *geometries, uint32_t count, const slimrt_mesh *meshes,
                            uint32_t mesh_count, float *aabbs_out);
/* BVHBuilder::build (src/slim/scene/bvh_builder.h:195-263) over N boxes (min xyz, max xyz).  nodes_out holds
 * 2N - 1 nodes (at least 1); leaf_ids_out N entries. */
int slimrt_host_build_bvh(const float *aabbs, uint32_t count, uint32_t max_leaf_size, slimrt_bvh_node *nodes_out,
                          uint32_t *node_count_out, uint32_t *height_out, uint32_t *leaf_ids_out);
/* 4-wide BVH node as the SLIMRT_BVH4 walk reads it (128 bytes): the four children's boxes as structure of
 * arrays, then per child an inner node index (count 0) or a leaf's first triangle and its triangle count; an
 * unused slot has child 0xFFFFFFFF, count 0 and the empty box (+inf, -inf). */
typedef struct slimrt_bvh4_node {
    float min_x[4], min_y[4], min_z[4];
    float max_x[4], max_y[4], max_z[4];
    uint32_t child[4], count[4];
} slimrt_bvh4_node;

/* The collapse slimrt_upload_scene applies to every mesh BVH (no counterpart in the reference, whose meshes are
 * walked as sibling pairs, src/slim/scene/mesh_tracer.h:51-115): a 4-wide node's children are its BVH2 node's
 * two children with, while fewer than four, the inner child of largest surface area replaced by its own two.
 * Leaves and their triangle ranges are the BVH2's.  nodes_out receives at most node_cap nodes ((node_count + 1) / 2
 * suffices), root first; *height_out the 4-wide tree's height (1 for a root that is a leaf); *stack_out the mesh
 * stack entries its walk can hold at most (a visit pushes its k inner children and pops one).  The walk keeps
 * 16-bit node indices on its stack, so SLIMRT_BVH4 frames need every mesh at 65536 4-wide nodes or fewer. */
int slimrt_host_collapse_bvh4(const slimrt_bvh_node *nodes, uint32_t node_count, uint32_t triangle_count,
                              slimrt_bvh4_node *nodes_out, uint32_t node_cap, uint32_t *node_count_out,
                              uint32_t *height_out, uint32_t *stack_out);

/* 8-wide BVH node as the SLIMRT_BVH8 walk reads it (128 bytes; no reference counterpart).  Child box bounds are
 * grid points origin + q * 2^(e - 127) per axis, e the axis' byte of exps (bits 0-7 x, 8-15 y, 16-23 z) and q the
 * slot's byte of qlo / qhi (slot k: byte k & 3 of word k >> 2).  Per slot, count (same byte layout): 0 an inner
 * child (child = node index), 1-254 a leaf of that many triangles (child = its leaf record, in 16-byte units),
 * 255 an empty slot. */
typedef struct slimrt_bvh8_node {
    float origin[3];
    uint32_t exps;
    uint32_t child[8];
    uint32_t count[2];
    uint32_t qlo[3][2], qhi[3][2];
    uint32_t pad[6];
} slimrt_bvh8_node;

/* The 8-wide collapse slimrt_upload_scene applies to every mesh BVH with leaves of at most 254 triangles (no
 * counterpart in the reference): a node's children are its BVH2 node's two children with, while fewer than eight,
 * the inner child of largest surface area replaced by its own two; child boxes are quantised outward to 8 bits on
 * a per-node grid, each decoded bound at least two ulps beyond the exact one; slots are ordered for the walk's
 * octant order.  Leaf records (leaves_out, 16-byte units): the BVH2 leaf's exact box with its first triangle and
 * triangle count (min xyz, first, max xyz, count), then its triangles' 48-byte test records (the first two
 * columns of local_to_tangent's rows, position, normal), followed by a final record's worth of padding.  nodes
 * receive at most node_cap nodes, root first, and leaves_out at most leaf_cap units; *stack_out the mesh stack
 * entries the walk can hold at most.  SLIMRT_E_OVERFLOW when a capacity is too small (the counts are set; with
 * both capacities 0 and NULL outputs this is the size query). */
int slimrt_host_collapse_bvh8(const slimrt_bvh_node *nodes, uint32_t node_count, const slimrt_triangle *triangles,
                              uint32_t triangle_count, slimrt_bvh8_node *nodes_out, uint32_t node_cap,
                              uint32_t *node_count_out, uint32_t *leaves_out, uint32_t leaf_cap,
                              uint32_t *leaf_count_out, uint32_t *stack_out);

/* Presentation: Canvas::drawToWindow (src/slim/draw/canvas.h:119-126) for a canvas written with
 * SLIMRT_OUTPUT_CANVAS: the window content, one 0xRRGGBB word per pixel (getPixelContent / Pixel::asContent,
 * canvas.h:187-190, base.h:1166-1171: the mean of the 2x2 sample quad under SSAA, each channel 255 * sqrt(c)
 * truncated).  GPU context: device pointers, asynchronous on the context's stream; CPU context: host pointers. */
int slimrt_canvas_content(slimrt_context *ctx, const void *pixels, uint32_t width, uint32_t height, uint32_t ssaa,
                          void *content);

/* Picking: Selection::manipulate's ray cast (src/slim/scene/selection.h:36-85) on the uploaded scene.  The camera
 * ray through integer pixel (x, y) of the params' projection (CameraRayProjection::getRayDirectionAt, camera.h:67,
 * normalised) is traced for the closest hit (SceneTracer::trace), then tested against every point light's pick
 * sphere (radius intensity / 2 / 64; the last light hit wins, as in the reference).  geometry / light / triangle:
 * -1 when none (triangle: leaf-ordered, meshes only); distance from the traced ray's origin (INFINITY when no
 * geometry is hit); position: that hit in world space (0 when none).  Synchronous. */
typedef struct slimrt_pick_result {
    int32_t geometry, light, triangle;
    float distance;
    float position[3];
} slimrt_pick_result;
int slimrt_pick(slimrt_context *ctx, const slimrt_render_params *params, int32_t x, int32_t y, slimrt_pick_result *out);

/* The path counters of the last frame rendered with SLIMRT_WAVEFRONT | SLIMRT_WAVEFRONT_COUNT (no counterpart in
 * the reference): counts[i] for i < 32 = paths that ran iteration i of the integrator loop (0 = the camera ray;
 * src/slim/renderer/ray_tracer.h:42-105), counts[32] = bounce rays spawned into a queue, counts[33] = samples
 * finished.  Every iteration either spawns or finishes, so sum(counts[0..31]) = counts[32] + counts[33].
 * n >= 34.  Synchronizes the context's stream. */
int slimrt_wavefront_counts(slimrt_context *ctx, uint32_t *counts, uint32_t n);

/* The tile costs of the last per-sample frame rendered without SLIMRT_TILE_ORDER_FIXED (no counterpart in the
 * reference).  A tile is the 32 x 2 samples one wave renders, tiles numbered in image order (row major over
 * ceil(w / 32) x ceil(h / 2) tiles of the sample grid, w and h doubled with SSAA).  costs[i] = tile i's wall-clock
 * duration in ticks of the device's 100 MHz real-time counter (a tile rendered as two half waves: twice the
 * duration of the half that ended last); order[k] = the tile dispatched k-th (the costliest first, or the image order).  Either pointer may
 * be NULL; *tile_count receives the number of tiles (0 before the first such frame); SLIMRT_E_ARG when cap is
 * smaller.  resident (may be NULL): the frame kernel's waves the device holds at once, which the choice of order
 * weighs the durations against.  Synchronizes the context's stream. */
int slimrt_tile_costs(slimrt_context *ctx, uint32_t *costs, uint32_t *order, uint32_t cap, uint32_t *tile_count,
                      uint32_t *resident);

/* The reference's texture converter (src/bmp2texture.cpp:24-361): the texel quads of an RGB image as it writes them
 * into .texture files.  rgb: width * height * 3 bytes, rows top to bottom.  flags: ImageFlags bits
 * (src/slim/core/base.h:1264-1277): linear 1 << 1 (no gamma 2.2), mipmap 1 << 4 (2 x 2 box-filtered mips down to
 * 4 texels), wrap 1 << 6 (border quads from the opposite edge, else clamped), normal 1 << 7 (renormalised normal
 * map), cubemap 1 << 8 (the image is a 4-face strip, a top and a bottom face, width = 6 * height; three mips with
 * seams stitched across faces).  Every mip's (w + 1) * (h + 1) quads go to quads_out in order (12 bytes each,
 * the layout of slimrt_texture_mip); *quad_count_out receives their total, *mip_count_out the mip count (at most
 * 16), mip_dims_out (2 per mip) their widths and heights, *flags_out the flags the converter stores.  With
 * quads_out NULL only the counts are returned; SLIMRT_E_OVERFLOW when quad_cap is smaller than the total. */
int slimrt_host_bake_texture(const uint8_t *rgb, uint32_t width, uint32_t height, uint32_t flags, uint8_t *quads_out,
                             size_t quad_cap, size_t *quad_count_out, uint32_t *mip_count_out, uint32_t *mip_dims_out,
                             uint32_t *flags_out);

/* A mip's texel quads (width * height texels, (width+1)*(height+1) quads) as the bordered texel grid of the
 * SLIMRT_TEXELS_* storage (no reference counterpart): texels_out receives (width+2)*(height+2) texels, RGBA8 (A =
 * 255) or RGBA16F (each u8 component c as the half nearest to the float c / 255, A = 1).  *mismatches_out (may be
 * NULL) receives the number of quads whose corners differ from the grid's texels: 0 means the RGBA8 texels sample
 * exactly as the quads do. */
int slimrt_host_texels_from_quads(const uint8_t *quads, uint32_t width, uint32_t height, uint32_t format,
                                  void *texels_out, uint64_t *mismatches_out);
/* IEEE half-float bits of f, rounded to nearest even (no reference counterpart; for building RGBA16F texels). */
uint16_t slimrt_host_float_to_half(float f);

/* BVHBuilder::buildMesh (src/slim/scene/bvh_builder.h:265-341): mesh BVH and leaf-ordered triangle records.
 * normals/uvs may be NULL (counts 0).  nodes_out: 2T - 1 nodes; triangles_out: T records; aabb_out: 6 floats
 * (componentwise min/max of the positions). */
int slimrt_host_build_mesh(const float *positions, uint32_t vertex_count, const uint32_t *position_indices,
                           const float *normals, uint32_t normal_count, const uint32_t *normal_indices,
                           const float *uvs, uint32_t uv_count, const uint32_t *uv_indices,
                           uint32_t triangle_count, uint32_t max_leaf_size,
                           slimrt_bvh_node *nodes_out, uint32_t *node_count_out, uint32_t *height_out,
                           slimrt_triangle *triangles_out, float aabb_out[6]);

/* ------------------------------------------------------------------------------------------ wireframe overlays
 * Boxes, BVH node boxes and mesh edges drawn into a canvas after a render, as the reference apps draw them
 * (drawBox, src/slim/draw/box.h:7-34; drawBVH, draw/bvh.h:8-28; drawMesh, draw/mesh.h:7-22; the scene-BVH view
 * of examples/03_BlinnPhong.cpp:194-209).  Each edge is taken to view space (Camera::internPos,
 * src/slim/scene/camera.h:47), clipped to the near/far planes and the frustum's sides
 * (Frustum::cullAndClipEdge, viewport/frustum.h:67-141), projected (projectPoint, frustum.h:180-188) and drawn
 * as an antialiased line (_drawLine, draw/line.h:70-273) whose samples are blended into the canvas by
 * Canvas::setPixel (canvas.h:128-185): depth-tested against the canvas depths, alpha-blended over or under.
 * pixels/depths: the canvas of a render (slimrt_pixel per sample, SSAA quads as SLIMRT_OUTPUT_CANVAS writes
 * them; a depth per sample); GPU context: device pointers, CPU context: host pointers.  The result equals the
 * reference's sequential drawing sample for sample on both.  GPU calls synchronise the context's stream once
 * (to size the sample lists) and are otherwise asynchronous on it. */

/* The viewport an overlay is drawn from: the camera (Camera: orientation from Euler angles as
 * OrientationUsing3x3Matrix::setRotation composes them, position, focal length), the canvas (width, height,
 * ssaa) and the frustum's clipping distances (the reference's defaults are 0.001 and 1000). */
typedef struct slimrt_overlay_view {
    float camera_euler[3];
    float camera_position[3];
    float focal_length;
    uint32_t width, height, ssaa;
    float near_distance, far_distance;
} slimrt_overlay_view;

/* A Transform (src/slim/core/transform.h:5-10): orientation quaternion (axis xyz, amount), position, scale. */
typedef struct slimrt_overlay_transform {
    float orientation[4];
    float position[3];
    float scale[3];
} slimrt_overlay_transform;

#define SLIMRT_BOX_ALL_SIDES 63u   /* BOX__ALL_SIDES (base.h:187); else BoxSide bits: left 1, bottom 2, back 4,
                                      right 8, top 16, front 32 (base.h:413-421) */

/* drawBox(Box{AABB{min, max}}, transform, viewport, color, opacity, line_width, sides). */
typedef struct slimrt_overlay_box {
    slimrt_overlay_transform transform;
    float min[3], max[3];      /* the box's local corners; drawBVH's and drawSelection's boxes are -1..1 */
    float color[3];
    float opacity;
    uint32_t line_width;       /* as the reference's u8 */
    uint32_t sides;            /* SLIMRT_BOX_ALL_SIDES or a BoxSide mask (the edges touching those sides) */
} slimrt_overlay_box;

/* Draws the boxes in order. */
int slimrt_draw_boxes(slimrt_context *ctx, const slimrt_overlay_view *view, const slimrt_overlay_box *boxes,
                      uint32_t count, void *pixels, void *depths);

/* drawMesh(mesh, transform, false, viewport, color, opacity, line_width): every edge (EdgeVertexIndices,
 * src/slim/scene/mesh.h:8-10, pairs of vertex indices) in order.  positions: vertex_count * 3 floats (host). */
int slimrt_draw_mesh_wireframe(slimrt_context *ctx, const slimrt_overlay_view *view, const float *positions,
                               uint32_t vertex_count, const uint32_t *edges, uint32_t edge_count,
                               const slimrt_overlay_transform *transform, const float color[3], float opacity,
                               uint32_t line_width, void *pixels, void *depths);

/* drawBVH's loop (draw/bvh.h:8-28) as box records for slimrt_draw_boxes: every node with min_depth <= depth <=
 * max_depth, in node order, as a -1..1 box scaled to the node's half extent and moved to its centre through
 * `transform`; leaves BrightMagenta at opacity * 2, the root BrightCyan and other inner nodes BrightGreen at
 * opacity / 2.  boxes_out: node_count entries at most; *count_out receives the number written. */
int slimrt_host_bvh_boxes(const slimrt_bvh_node *nodes, uint32_t node_count, const slimrt_overlay_transform *transform,
                          uint32_t min_depth, uint32_t max_depth, float opacity, uint32_t line_width,
                          slimrt_overlay_box *boxes_out, uint32_t *count_out);

/* The camera rotation of OrientationUsing3x3Matrix::setRotation(x, y, z) (src/slim/core/base.h:568-620,
 * math/mat3.h): rows X, Y, Z (9 floats). */
int slimrt_host_camera_rotation(const float euler[3], float rotation_out[9]);

#ifdef __cplusplus
}
#endif

#endif
