// Internal to libslimrt: the context behind the C ABI handle and the helpers the kernels' host code shares.
#pragma once

#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <unordered_map>
#include <cstring>

#include "slimrt_nrc_trace.h"
#include "../../include/slimrt.h"

namespace slimrt {

constexpr uint32_t kBlock = 256;      // render workgroups: 4 waves of 64
constexpr uint32_t kTile = 16;        // 16x16 samples per workgroup, 8x8 per wave
constexpr uint32_t kWaveTileX = 32, kWaveTileY = 2;   // the frame kernel's workgroups: one wave, 32x2 samples
static_assert(kWaveTileX * kWaveTileY == 64, "a wave tile holds 64 samples");

template <typename T>
struct DevBuf {
    T *ptr = nullptr;
    size_t cap = 0;   // elements
};


}  // namespace slimrt

using namespace slimrt;

struct slimrt_context {
    bool cpu = false;
    int device = 0;
    hipStream_t stream = nullptr, own_stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    std::string err;
    // device memory this context holds (every allocation, by pointer) and its limit (0: none)
    uint64_t mem_used = 0, mem_limit = 0;
    std::unordered_map<void *, size_t> allocs;
    uint32_t host_status = 0;   // kStatusNaN seen since the last slimrt_frame_flags (device word folded in)

    // host mirror of the device arrays (the CPU path renders from these)
    std::vector<Geo> geos;
    std::vector<Mat> mats;
    std::vector<Light> lights;
    std::vector<Node> snodes, mnodes;
    std::vector<uint32_t> leaf_ids;
    std::vector<MeshRec> meshes;
    std::vector<TriHit> tri_hit;
    std::vector<TriAttr> tri_attr;
    std::vector<TexRec> textures;
    std::vector<MipRec> mips;
    std::vector<uint8_t> quads;
    uint32_t scene_height = 0, mesh_height = 0, scene_flags = 0;
    std::vector<Node4> mnodes4;             // 4-wide mesh BVHs collapsed from mnodes at upload
    uint32_t mesh_height4 = 0;
    uint32_t mesh_stack4 = 0;       // mesh stack entries the 4-wide walk needs (largest over the meshes)
    uint32_t mesh_nodes4_max = 0;   // 4-wide nodes of the largest mesh (16-bit stack entries need <= 65536)
    std::vector<Node8> mnodes8;             // 8-wide mesh BVHs collapsed from mnodes at upload
    std::vector<uint4> leaf8;               // their leaf records (LeafHead + TriHit records)
    uint32_t mesh_stack8 = 0;       // mesh stack entries the 8-wide walk needs (largest over the meshes)
    uint32_t mesh_nodes8_max = 0;   // 8-wide nodes of the largest mesh
    uint32_t mesh_leaf_max = 0;     // triangles of the largest leaf (the 8-wide nodes hold counts up to 254)
    bool loaded = false;

    DevBuf<Geo> d_geos;
    DevBuf<Mat> d_mats;
    DevBuf<Light> d_lights;
    DevBuf<Node> d_snodes, d_mnodes;
    DevBuf<Node4> d_mnodes4;
    DevBuf<Node8> d_mnodes8;
    DevBuf<uint4> d_leaf8;
    DevBuf<uint32_t> d_leaf;
    DevBuf<MeshRec> d_meshes;
    DevBuf<TriHit> d_tri_hit;
    DevBuf<TriAttr> d_tri_attr;
    DevBuf<TexRec> d_tex;
    DevBuf<MipRec> d_mips;
    DevBuf<uint8_t> d_quads;
    DevBuf<float4> d_pixels;
    DevBuf<float> d_depths;
    DevBuf<slimrt_geometry> d_geo_stage;   // transform updates on their way to the scene update kernel
    DevBuf<uint8_t> ov_lines, ov_samples;   // overlay drawing: inputs and lines; the lines' canvas samples
    void *wf_queue_a = nullptr, *wf_queue_b = nullptr, *wf_accum = nullptr;   // wavefront ray queues, colours
    void *wf_hits = nullptr;   // first-hit records of the split primary pass (36 bytes per sample slot)
    uint32_t wf_hits_cap = 0;
    uint32_t *wf_counters = nullptr, wf_cap = 0;
    uint32_t *wf_stats = nullptr;   // kWfStats path counters (SLIMRT_WAVEFRONT_COUNT)
    int32_t *d_pick = nullptr;      // slimrt_pick's result
    hipStream_t side_stream = nullptr;          // a wavefront frame's bounce launches, overlapping the next band
    hipEvent_t wf_fork = nullptr, wf_join = nullptr, wf_band[8] = {};
    uint32_t *d_status = nullptr;
    unsigned long long *d_counters = nullptr;
    // per-sample frames dispatched by cost: cost-ordered frame g renders in order tile_order[g % 2] and writes its
    // tiles' durations (wall-clock ticks) to tile_cost[g % 3]; its first workgroups count frame g - 1's durations
    // per segment into tile_hist[(g + 1) % 2] and, from the counts frame g - 1 made, place frame g - 2's tiles into
    // tile_order[(g + 1) % 2].  *_tiles: the tile count a buffer holds (0: none).  tile_rendered follows the last
    // such frame, launched on order_stream.
    uint32_t *tile_cost[3] = {}, *tile_order[2] = {}, *tile_hist[2] = {};
    uint32_t tile_cap = 0, cost_tiles[3] = {}, hist_tiles[2] = {}, order_tiles[2] = {};
    bool order_used[2] = {};
    uint64_t tile_frame = 0;
    hipEvent_t tile_rendered = nullptr;
    hipStream_t order_stream = nullptr;
    // this device's copy of the no-allocation frames' scratch array (its device address, looked up at creation)
    uint32_t *noalloc_scratch = nullptr;
    uint32_t resident = 0;           // frame-kernel waves the device holds at once, for resident_key's configuration
    uint64_t resident_key = 0;

    // rendering with the cache: queues, training samples, gate table, frame counters
    float *q_raw = nullptr, *q_prefix = nullptr, *q_tp = nullptr, *q_out = nullptr;
    uint32_t *q_pixel = nullptr, q_cap = 0;
    float *t_raw = nullptr, *t_target = nullptr, *t_next = nullptr;
    uint32_t *t_link = nullptr, t_cap = 0;
    // the queue and sample capacities the last cache frame used (the buffers above may be larger, left from a
    // larger frame): counts read back are clamped to these
    uint32_t nrc_frame_q_cap = 0, nrc_frame_t_cap = 0;
    uint32_t *nrc_counters = nullptr;
    float *nrc_losses = nullptr;
    uint32_t nrc_last_steps = 0;
    float *cells = nullptr;
    uint32_t cells_log2 = 0;
};

namespace slimrt {


inline int fail(slimrt_context *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(ctx, e_ == hipErrorOutOfMemory ? SLIMRT_E_MEMORY : SLIMRT_E_DEVICE,        \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                        \
    } while (0)

// Every device allocation a context owns goes through these two, which keep its running total: with a limit set
// (slimrt_set_memory_limit), an allocation that would take the total past it fails as out of memory.
template <typename T>
inline hipError_t ctx_malloc(slimrt_context *c, T **p, size_t bytes) {
    if (c->mem_limit && (bytes > c->mem_limit || c->mem_used > c->mem_limit - bytes)) return hipErrorOutOfMemory;
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
        c->allocs[(void *)*p] = bytes;
        c->mem_used += bytes;
    }
    return e;
}

inline hipError_t ctx_free(slimrt_context *c, void *p) {
    if (!p) return hipSuccess;
    const auto it = c->allocs.find(p);
    if (it != c->allocs.end()) {
        c->mem_used -= it->second;
        c->allocs.erase(it);
    }
    return hipFree(p);
}

template <typename T>
inline int dev_upload(slimrt_context *c, DevBuf<T> &b, const T *src, size_t n) {
    if (n == 0) n = 1;   // keep a valid pointer for empty arrays
    if (b.cap < n) {
        if (b.ptr) HIP_TRY(c, ctx_free(c, b.ptr));
        b.ptr = nullptr;
        b.cap = 0;
        HIP_TRY(c, ctx_malloc(c, &b.ptr, n * sizeof(T)));
        b.cap = n;
    }
    if (src) HIP_TRY(c, hipMemcpyAsync(b.ptr, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return SLIMRT_OK;
}

template <typename T>
inline void dev_free(slimrt_context *c, DevBuf<T> &b) {
    if (b.ptr) (void)ctx_free(c, b.ptr);
    b.ptr = nullptr;
    b.cap = 0;
}

SLIM_HD Geo to_geo(const slimrt_geometry &g) {
    Geo o;
    o.qx = g.orientation[0]; o.qy = g.orientation[1]; o.qz = g.orientation[2]; o.qw = g.orientation[3];
    o.px = g.position[0]; o.py = g.position[1]; o.pz = g.position[2]; o.type = g.type;
    o.sx = g.scale[0]; o.sy = g.scale[1]; o.sz = g.scale[2]; o.material_id = g.material_id;
    o.isx = 1.0f / g.scale[0]; o.isy = 1.0f / g.scale[1]; o.isz = 1.0f / g.scale[2];   // Ray::localize (ray.h:30)
    o.id_flags = (g.id << 8) | (g.flags & 0xFFu);
    return o;
}

inline Mat to_mat(const slimrt_material &m) {
    Mat o;
    o.ar = m.albedo[0]; o.ag = m.albedo[1]; o.ab = m.albedo[2]; o.roughness = m.roughness;
    o.er = m.emission[0]; o.eg = m.emission[1]; o.eb = m.emission[2]; o.ior = m.ior;
    o.fr = m.reflectivity[0]; o.fg = m.reflectivity[1]; o.fb = m.reflectivity[2]; o.metalness = m.metalness;
    o.uv_ru = m.uv_repeat[0]; o.uv_rv = m.uv_repeat[1]; o.normal_magnitude = m.normal_magnitude;
    o.packed = (m.flags & 0xFFu) | ((m.texture_count & 0xFu) << 8) | ((m.brdf & 0xFu) << 12) |
               ((m.texture_ids[0] & 0xFFu) << 16) | ((m.texture_ids[1] & 0xFFu) << 24);
    return o;
}

inline Light to_light(const slimrt_light &l) {
    Light o;
    o.r = l.color[0]; o.g = l.color[1]; o.b = l.color[2]; o.intensity = l.intensity;
    o.x = l.position_or_direction[0]; o.y = l.position_or_direction[1]; o.z = l.position_or_direction[2];
    o.flags = l.flags;
    return o;
}

inline Node to_node(const slimrt_bvh_node &n) {   // asset layout -> device layout (bounds interleaved with indices)
    Node o;
    o.minx = n.min[0]; o.miny = n.min[1]; o.minz = n.min[2]; o.first_index = n.first_index;
    o.maxx = n.max[0]; o.maxy = n.max[1]; o.maxz = n.max[2];
    o.leaf_count_depth = (uint32_t)n.leaf_count | ((uint32_t)n.depth << 16) | ((uint32_t)n.flags << 24);
    return o;
}

inline SceneView host_view(const slimrt_context *c) {
    SceneView v;
    v.geos = c->geos.data(); v.mats = c->mats.data(); v.lights = c->lights.data();
    v.snodes = c->snodes.data(); v.leaf_ids = c->leaf_ids.data(); v.meshes = c->meshes.data();
    v.mnodes = c->mnodes.data(); v.tri_hit = c->tri_hit.data(); v.tri_attr = c->tri_attr.data();
    v.mnodes4 = c->mnodes4.data();
    v.mnodes8 = c->mnodes8.data(); v.leaf8 = c->leaf8.data();
    v.textures = c->textures.data(); v.mips = c->mips.data(); v.quads = c->quads.data();
    v.G = (uint32_t)c->geos.size(); v.M = (uint32_t)c->mats.size(); v.L = (uint32_t)c->lights.size();
    v.NM = (uint32_t)c->meshes.size(); v.NT = (uint32_t)c->textures.size(); v.flags = c->scene_flags;
    return v;
}

// Traversal stack entries for mesh BVHs: the declared height + 1 for the sibling-pair walk; for the wide walks
// what their collapse computed (pending siblings per level of the collapsed tree, plus the children of the node
// being visited).
inline uint32_t mesh_stack_depth(const slimrt_context *c, uint32_t width) {
    return width == 8 ? std::max(c->mesh_stack8, 1u) : width == 4 ? std::max(c->mesh_stack4, 1u) : c->mesh_height + 1;
}

// A workgroup's traversal stacks in LDS: kBlock columns of scene entries (32-bit), then kBlock columns of mesh
// entries (32-bit pair indices, or 16-bit node indices for the wide walks).
inline size_t stack_lds_bytes(uint32_t scene_depth, uint32_t mesh_depth, uint32_t width, uint32_t threads = kBlock) {
    return ((size_t)scene_depth * 4 + (size_t)mesh_depth * (width >= 4 ? 2 : 4)) * threads;
}
// Thread t's stacks in a workgroup of THREADS threads: one column each, entry i at i * THREADS.
template <uint32_t NW, uint32_t THREADS = kBlock>
__device__ __forceinline__ void bind_lds_stacks(TracerT<NW> &tr, uint32_t *lds, uint32_t t, uint32_t scene_depth,
                                                uint32_t mesh_depth) {
    tr.scene_stack = StackT<uint32_t>{lds + t, THREADS};
    tr.mesh_stack = StackT<typename TracerT<NW>::MeshEntry>{
        reinterpret_cast<typename TracerT<NW>::MeshEntry *>(lds + scene_depth * THREADS) + t, THREADS};
    tr.scene_depth = scene_depth;
    tr.mesh_depth = mesh_depth;
    tr.status = 0;
}

// Checks that a wide walk can run on the uploaded meshes (16-bit stack entries; 8-bit leaf counts for width 8).
inline int check_width(slimrt_context *c, uint32_t width) {
    if (width == 4 && c->mesh_nodes4_max > 65536)
        return fail(c, SLIMRT_E_ARG, "a mesh has more than 65536 4-wide nodes: render it with node width 2");
    if (width == 8 && (c->mesh_nodes8_max > 65536 || c->mesh_leaf_max > 254))
        return fail(c, SLIMRT_E_ARG, "a mesh has more than 65536 8-wide nodes or a leaf of more than 254 triangles: "
                                     "render it with node width 2 or 4");
    return SLIMRT_OK;
}

inline SceneView device_view(const slimrt_context *c) {
    SceneView v = host_view(c);
    v.geos = c->d_geos.ptr; v.mats = c->d_mats.ptr; v.lights = c->d_lights.ptr;
    v.snodes = c->d_snodes.ptr; v.leaf_ids = c->d_leaf.ptr; v.meshes = c->d_meshes.ptr;
    v.mnodes = c->d_mnodes.ptr; v.tri_hit = c->d_tri_hit.ptr; v.tri_attr = c->d_tri_attr.ptr;
    v.mnodes4 = c->d_mnodes4.ptr;
    v.mnodes8 = c->d_mnodes8.ptr; v.leaf8 = c->d_leaf8.ptr;
    v.textures = c->d_tex.ptr; v.mips = c->d_mips.ptr; v.quads = c->d_quads.ptr;
    return v;
}

// Whether the integrator's light loops run at all: point or directional lights, or emissive quads.
// The light terms a frame of this scene needs (kLightsNone, kLightsQuads or kLightsAll).
inline uint32_t scene_lights(const SceneView &v) {
    return v.L > 0 ? kLightsAll : ((v.flags & 1u) != 0 ? kLightsQuads : kLightsNone);
}

inline void update_scene_flags(slimrt_context *c) {   // Scene::Scene (scene.h:239-242)
    c->scene_flags = 0;
    for (const Geo &g : c->geos)
        if (g.type == GT_Quad && g.material_id < c->mats.size() && (mat_flags(c->mats[g.material_id]) & MAT_EMISSIVE))
            c->scene_flags = 1;
}

inline int validate_params(slimrt_context *c, const slimrt_render_params *p) {
    if (!p) return fail(c, SLIMRT_E_ARG, "params is NULL");
    if (!c->loaded) return fail(c, SLIMRT_E_STATE, "no scene uploaded");
    if (p->width == 0 || p->height == 0 || p->width > 65535 || p->height > 65535)
        return fail(c, SLIMRT_E_ARG, "canvas dimensions out of range");
    if (p->render_mode > 5 || p->output > 2 || p->max_depth > 255) return fail(c, SLIMRT_E_ARG, "invalid settings");
    const int32_t nt = (int32_t)c->textures.size();
    if (p->skybox_color >= nt || p->skybox_radiance >= nt || p->skybox_irradiance >= nt)
        return fail(c, SLIMRT_E_ARG, "skybox texture id out of range");
    return SLIMRT_OK;
}

// Path counters of a counted wavefront frame: paths per integrator iteration, then spawned, finished.
constexpr uint32_t kWfIterations = 32, kWfSpawned = 32, kWfFinished = 33, kWfStats = 34;

// Beauty frame with the bounces as a wavefront (slimrt_wavefront.hip).
int launch_wavefront(slimrt_context *c, const RenderParams &rp, float4 *pixels, float *depths, bool per_bounce, bool sort,
                     uint32_t bands);

// Reads and clears the device status word of the frames since the last check: a stack or queue overflow fails
// with SLIMRT_E_OVERFLOW; NaN samples are kept for slimrt_frame_flags.
inline int check_status(slimrt_context *c) {
    uint32_t st = 0;
    HIP_TRY(c, hipMemcpyAsync(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (st) {
        HIP_TRY(c, hipMemsetAsync(c->d_status, 0, sizeof st, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->host_status |= st & kStatusNaN;
        if (st & kStatusStack) return fail(c, SLIMRT_E_OVERFLOW, "traversal stack overflow");
        if (st & kStatusQueue) return fail(c, SLIMRT_E_OVERFLOW, "wavefront queue overflow");
    }
    return SLIMRT_OK;
}


}  // namespace slimrt
