// slimrt_views: the device-side scene update and the multi-view launch (include/slimrt_views.h).
//
// Scene update: one workgroup; its threads convert the new geometry records and compute their world AABBs
// (in parallel, into LDS), then one thread refits the uploaded tree bottom-up or rebuilds it with the SAH
// procedure of slimrt_sah.h from LDS scratch.  The scene BVH has at most a few hundred nodes, so this is a
// latency-bound few-microsecond step that saves the host round trip of the reference's per-frame rebuild.
//
// Views: one launch over every (view, tile) item, either one workgroup per tile or a persistent grid whose
// workgroups take items from one device counter until the counter passes the total; a tile is traced exactly
// as the single-view render kernel traces it, so every view's output is bit-identical to its own
// slimrt_render_device call.  Views of uneven size balance across the chip without a launch per view, and a
// frame (geometry copy, scene update, counter reset, trace) is four stream operations that a HIP graph captures
// once.  The counter is reset by a kernel, not a memset, so the captured graph holds no memset node.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "slimrt_context.h"
#include "slimrt_sah.h"
#include "slimrt_views_trace.h"
#include "../../include/slimrt_views.h"

namespace {

constexpr uint32_t kMaxRebuild = 128;

struct Carve {   // 16-byte aligned sub-allocations of one LDS block, identical on host (sizing) and device
    uint8_t *p;
    template <typename T>
    __host__ __device__ T *take(size_t n) {
        T *r = (T *)p;
        p += (n * sizeof(T) + 15) & ~(size_t)15;
        return r;
    }
};

__host__ __device__ inline SahScratch carve_rebuild(Carve &c, uint32_t N, SahNode *&out) {
    SahScratch s;
    s.boxes = c.take<SBox>(N);
    s.ids = c.take<uint32_t>(N);
    s.leaf_ids = c.take<uint32_t>(N);
    for (int a = 0; a < 3; a++) {
        s.sorted[a] = c.take<uint32_t>(N);
        s.left[a] = c.take<SBox>(N);
        s.right[a] = c.take<SBox>(N);
        s.left_area[a] = c.take<float>(N);
        s.right_area[a] = c.take<float>(N);
    }
    s.sort_stack = c.take<int32_t>(2 * N + 2);
    s.tasks = c.take<uint32_t>(4 * (N + 2));
    out = c.take<SahNode>(2 * N);
    return s;
}

size_t update_lds_bytes(uint32_t G, uint32_t mode) {
    Carve c{nullptr};
    if (mode == SLIMRT_SCENE_REBUILD) {
        SahNode *out;
        (void)carve_rebuild(c, G, out);
    } else {
        (void)c.take<SBox>(G);
    }
    return (size_t)c.p;
}

__device__ __forceinline__ SBox node_box(const Node &n) {
    SBox b;
    b.mn[0] = n.minx; b.mn[1] = n.miny; b.mn[2] = n.minz;
    b.mx[0] = n.maxx; b.mx[1] = n.maxy; b.mx[2] = n.maxz;
    return b;
}

__global__ __launch_bounds__(256) void scene_update_kernel(const slimrt_geometry *__restrict__ src, uint32_t G,
                                                           Geo *__restrict__ geos, const MeshRec *__restrict__ meshes,
                                                           Node *__restrict__ nodes, uint32_t *__restrict__ leaf_ids,
                                                           uint32_t node_count, uint32_t mode) {
    extern __shared__ __align__(16) uint8_t smem[];
    Carve carve{smem};
    SahNode *out = nullptr;
    SahScratch s;
    SBox *boxes;
    if (mode == SLIMRT_SCENE_REBUILD) {
        s = carve_rebuild(carve, G, out);
        boxes = s.boxes;
    } else {
        boxes = carve.take<SBox>(G);
    }
    for (uint32_t i = threadIdx.x; i < G; i += blockDim.x) {
        const slimrt_geometry g = src[i];
        geos[i] = to_geo(g);
        float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
        if (g.type == GT_Mesh) {
            const MeshRec m = meshes[g.id];
            mn[0] = m.minx; mn[1] = m.miny; mn[2] = m.minz;
            mx[0] = m.maxx; mx[1] = m.maxy; mx[2] = m.maxz;
        }
        boxes[i] = geometry_world_box(g, mn, mx);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (mode == SLIMRT_SCENE_REFIT) {   // children follow their parent, so a reverse sweep sees children first
        for (int32_t n = (int32_t)node_count - 1; n >= 0; n--) {
            Node nd = nodes[n];
            const uint32_t lc = nd.leaf_count_depth & 0xFFFFu;
            SBox b;
            if (lc) {
                b = boxes[leaf_ids[nd.first_index]];
                for (uint32_t k = 1; k < lc; k++) b = sbox_merge(b, boxes[leaf_ids[nd.first_index + k]]);
            } else {
                b = sbox_merge(node_box(nodes[nd.first_index]), node_box(nodes[nd.first_index + 1]));
            }
            nd.minx = b.mn[0]; nd.miny = b.mn[1]; nd.minz = b.mn[2];
            nd.maxx = b.mx[0]; nd.maxy = b.mx[1]; nd.maxz = b.mx[2];
            nodes[n] = nd;
        }
        return;
    }
    for (uint32_t i = 0; i < G; i++) s.ids[i] = i;
    uint32_t count, height;
    sah_build(s, G, 1, out, count, height);
    for (uint32_t n = 0; n < count; n++) {
        const SahNode &o = out[n];
        Node nd;
        nd.minx = o.mn[0]; nd.miny = o.mn[1]; nd.minz = o.mn[2];
        nd.maxx = o.mx[0]; nd.maxy = o.mx[1]; nd.maxz = o.mx[2];
        nd.first_index = o.first_index;
        nd.leaf_count_depth = o.leaf_count | (o.depth << 16);
        nodes[n] = nd;
    }
    for (uint32_t i = 0; i < G; i++) leaf_ids[i] = s.leaf_ids[i];
}

__global__ void counter_reset_kernel(uint32_t *__restrict__ counter) {
    if (threadIdx.x == 0) *counter = 0u;
}

int check_geometries(slimrt_context *c, const slimrt_geometry *g, uint32_t count) {
    if (!c->loaded) return fail(c, SLIMRT_E_STATE, "no scene uploaded");
    if (!g || count != c->geos.size()) return fail(c, SLIMRT_E_ARG, "geometry count differs from the scene's");
    for (uint32_t i = 0; i < count; i++) {
        if (g[i].material_id >= c->mats.size()) return fail(c, SLIMRT_E_ARG, "geometry material id out of range");
        if (g[i].type == GT_Mesh && g[i].id >= c->meshes.size()) return fail(c, SLIMRT_E_ARG, "geometry mesh id out of range");
        if (g[i].type != c->geos[i].type) return fail(c, SLIMRT_E_ARG, "a transform update cannot change a geometry's type");
    }
    return SLIMRT_OK;
}

int check_update_mode(slimrt_context *c, uint32_t mode) {
    const uint32_t G = (uint32_t)c->geos.size();
    if (mode == SLIMRT_SCENE_REBUILD) {
        if (G > kMaxRebuild) return fail(c, SLIMRT_E_ARG, "device rebuild takes at most 128 geometries (refit any number)");
        if (c->snodes.size() != (G > 1 ? 2 * G - 1 : 1)) return fail(c, SLIMRT_E_ARG, "device rebuild needs a leaf-size-1 scene BVH");
    } else if (mode != SLIMRT_SCENE_REFIT) {
        return fail(c, SLIMRT_E_ARG, "unknown scene update mode");
    }
    if (update_lds_bytes(G, mode) > 64 * 1024) return fail(c, SLIMRT_E_ARG, "too many geometries for the update kernel");
    return SLIMRT_OK;
}

// Host mirrors of a transform update (the device arrays are written by the kernel).
void mirror_geometries(slimrt_context *c, const slimrt_geometry *g, uint32_t mode) {
    for (size_t i = 0; i < c->geos.size(); i++) c->geos[i] = to_geo(g[i]);
    if (mode == SLIMRT_SCENE_REBUILD) {   // the rebuilt height is only known on the device: size stacks for the worst
        const uint32_t G = (uint32_t)c->geos.size();
        c->scene_height = std::max(c->scene_height, G > 1 ? G - 1 : 1u);
    }
}

int enqueue_update(slimrt_context *c, const slimrt_geometry *d_src, uint32_t mode) {
    const uint32_t G = (uint32_t)c->geos.size();
    hipLaunchKernelGGL(scene_update_kernel, dim3(1), dim3(256), update_lds_bytes(G, mode), c->stream, d_src, G,
                       c->d_geos.ptr, c->d_meshes.ptr, c->d_snodes.ptr, c->d_leaf.ptr, (uint32_t)c->snodes.size(), mode);
    HIP_TRY(c, hipGetLastError());
    return SLIMRT_OK;
}

}  // namespace

struct slimrt_views {
    slimrt_context *c = nullptr;
    std::vector<slimrt_render_params> params;
    std::vector<float4 *> pixels;
    std::vector<float *> depths;
    std::vector<ViewRec> recs;
    ViewRec *d_recs = nullptr;
    uint32_t total_tiles = 0;
    slimrt_views_config cfg{0, 2, 1, 0, 0};   // 16x16 tiles, one workgroup per tile, node width from the flags
    uint32_t *d_counter = nullptr;
    slimrt_geometry *h_stage = nullptr, *d_stage = nullptr;
    uint32_t scene_cap = 0;            // scene stack entries the launches size LDS for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t graph_key = -1;
    uint32_t generation = 0;           // bumped by every configuration change
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false, in_flight = false;
    std::string err;
};

namespace {

int vfail(slimrt_views *v, int code, const std::string &m) {
    if (v) v->err = m;
    return code;
}

#define VIEWS_TRY(v, expr)                                                                             \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return vfail(v, SLIMRT_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void wave_tile(uint32_t shape, uint32_t &wx, uint32_t &wy) {
    static const uint32_t W[4][2] = {{8, 8}, {16, 4}, {32, 2}, {4, 16}};
    wx = W[shape][0];
    wy = W[shape][1];
}

// Tile tables for the current tile shape.
int build_tables(slimrt_views *v) {
    uint32_t wx, wy;
    wave_tile(v->cfg.tile_shape, wx, wy);
    v->recs.resize(v->params.size());
    uint32_t base = 0;
    for (size_t i = 0; i < v->params.size(); i++) {
        ViewRec &r = v->recs[i];
        std::memcpy(&r.rp, &v->params[i], sizeof r.rp);
        r.rp.ssaa = r.rp.ssaa ? 1u : 0u;
        const uint32_t ws = r.rp.width << r.rp.ssaa, hs = r.rp.height << r.rp.ssaa;
        r.pixels = v->pixels[i];
        r.depths = v->depths[i];
        r.tiles_x = (ws + wx - 1) / wx;
        r.tiles = r.tiles_x * ((hs + wy - 1) / wy);
        r.tile_base = base;
        r.pad = 0;
        base += r.tiles;
    }
    v->total_tiles = base;
    VIEWS_TRY(v, hipMemcpyAsync(v->d_recs, v->recs.data(), v->recs.size() * sizeof(ViewRec), hipMemcpyHostToDevice,
                                v->c->stream));
    VIEWS_TRY(v, hipStreamSynchronize(v->c->stream));
    return SLIMRT_OK;
}

int enqueue_views(slimrt_views *v) {
    slimrt_context *c = v->c;
    hipLaunchKernelGGL(counter_reset_kernel, dim3(1), dim3(64), 0, c->stream, v->d_counter);
    VIEWS_TRY(v, hipGetLastError());
    int cus = 0;
    VIEWS_TRY(v, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const uint32_t grab = v->cfg.tiles_per_grab;
    const bool persistent = v->cfg.persistent != 0;
    const uint32_t blocks = persistent
        ? std::max(1u, std::min((uint32_t)cus * 4 * v->cfg.blocks_per_cu, (v->total_tiles + grab - 1) / grab))
        : v->total_tiles;
    uint32_t width = v->cfg.node_width;
    if (width == 0) {   // as the views ask: a wide walk when every view has its flag (SLIMRT_BVH8 before SLIMRT_BVH4)
        bool all8 = true, all4 = true;
        for (const slimrt_render_params &p : v->params) {
            all8 = all8 && (p.flags & SLIMRT_BVH8) != 0;
            all4 = all4 && (p.flags & (SLIMRT_BVH4 | SLIMRT_BVH8)) != 0;
        }
        width = all8 ? 8u : all4 ? 4u : 2u;
    }
    const uint32_t scene_depth = v->scene_cap + 1, mesh_depth = mesh_stack_depth(c, width);
    const size_t lds = stack_lds_bytes(scene_depth, mesh_depth, width, 64);
    if (lds + 16 > 64 * 1024) return vfail(v, SLIMRT_E_ARG, "BVH too deep for the LDS traversal stacks");
    if (check_width(c, width)) return vfail(v, SLIMRT_E_ARG, c->err);
    const SceneView sv = device_view(c);
    const uint32_t n = (uint32_t)v->recs.size();
    ViewsLaunch l{sv, (const ViewRec *)v->d_recs, v->d_counter, c->d_status, c->stream, lds, blocks, n, v->total_tiles,
                  grab, scene_depth, mesh_depth, v->cfg.tile_shape, width, persistent};
    // scenes without lights or emissive quads take the instantiation without light terms (no VGPR spills); with
    // emissive quads only, the one without shadow-ray walks measured slower than the full one on the 4K workload's
    // views (9.7 against 8.7 ms), so every lit scene takes the full one
    VIEWS_TRY(v, scene_lights(sv) == kLightsNone ? launch_views_none(l) : launch_views_all(l));
    VIEWS_TRY(v, hipGetLastError());
    return SLIMRT_OK;
}

int enqueue_frame(slimrt_views *v, bool update, uint32_t mode) {
    slimrt_context *c = v->c;
    if (update) {
        VIEWS_TRY(v, hipMemcpyAsync(v->d_stage, v->h_stage, c->geos.size() * sizeof(slimrt_geometry),
                                    hipMemcpyHostToDevice, c->stream));
        if (enqueue_update(c, v->d_stage, mode)) return vfail(v, SLIMRT_E_DEVICE, c->err);
    }
    return enqueue_views(v);
}

void drop_graph(slimrt_views *v) {
    if (v->exec) (void)hipGraphExecDestroy(v->exec);
    if (v->graph) (void)hipGraphDestroy(v->graph);
    v->exec = nullptr;
    v->graph = nullptr;
    v->graph_key = -1;
}

int frame_ms(slimrt_views *v, float *ms) {
    VIEWS_TRY(v, hipEventSynchronize(v->ev1));
    VIEWS_TRY(v, hipEventElapsedTime(ms, v->ev0, v->ev1));
    return SLIMRT_OK;
}

}  // namespace

extern "C" {

int slimrt_update_scene_device(slimrt_context *c, const slimrt_geometry *g, uint32_t count, uint32_t mode) {
    if (!c) return SLIMRT_E_ARG;
    if (c->cpu) return fail(c, SLIMRT_E_STATE, "device scene update needs a GPU context");
    int rc;
    if ((rc = check_geometries(c, g, count)) || (rc = check_update_mode(c, mode))) return rc;
    if (c->d_geo_stage.cap < count) {
        dev_free(c, c->d_geo_stage);
        HIP_TRY(c, ctx_malloc(c, &c->d_geo_stage.ptr, count * sizeof(slimrt_geometry)));
        c->d_geo_stage.cap = count;
    }
    HIP_TRY(c, hipMemcpyAsync(c->d_geo_stage.ptr, g, count * sizeof(slimrt_geometry), hipMemcpyHostToDevice, c->stream));
    if ((rc = enqueue_update(c, c->d_geo_stage.ptr, mode))) return rc;
    mirror_geometries(c, g, mode);
    return SLIMRT_OK;
}

int slimrt_download_scene_bvh(slimrt_context *c, slimrt_bvh_node *nodes, uint32_t node_cap, uint32_t *leaf_ids,
                              uint32_t leaf_cap, uint32_t *node_count) {
    if (!c || !nodes || !leaf_ids || !node_count) return SLIMRT_E_ARG;
    if (c->cpu) return fail(c, SLIMRT_E_STATE, "download_scene_bvh needs a GPU context");
    if (!c->loaded) return fail(c, SLIMRT_E_STATE, "no scene uploaded");
    const uint32_t n = (uint32_t)c->snodes.size(), G = (uint32_t)c->geos.size();
    *node_count = n;
    if (node_cap < n || leaf_cap < G) return fail(c, SLIMRT_E_ARG, "output arrays too small");
    std::vector<Node> dn(n);
    HIP_TRY(c, hipMemcpyAsync(dn.data(), c->d_snodes.ptr, n * sizeof(Node), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(leaf_ids, c->d_leaf.ptr, G * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < n; i++) {
        const Node &d = dn[i];
        slimrt_bvh_node &o = nodes[i];
        o.min[0] = d.minx; o.min[1] = d.miny; o.min[2] = d.minz;
        o.max[0] = d.maxx; o.max[1] = d.maxy; o.max[2] = d.maxz;
        o.first_index = d.first_index;
        o.leaf_count = (uint16_t)(d.leaf_count_depth & 0xFFFFu);
        o.depth = (uint8_t)((d.leaf_count_depth >> 16) & 0xFFu);
        o.flags = (uint8_t)(d.leaf_count_depth >> 24);
    }
    return SLIMRT_OK;
}

int slimrt_views_create(slimrt_context *c, const slimrt_render_params *views, uint32_t count, void *const *d_pixels,
                        void *const *d_depths, slimrt_views **out) {
    if (!c || !out || !views || !d_pixels || !d_depths || count == 0) return SLIMRT_E_ARG;
    *out = nullptr;
    if (c->cpu) return fail(c, SLIMRT_E_STATE, "views need a GPU context");
    for (uint32_t i = 0; i < count; i++) {
        int rc = validate_params(c, &views[i]);
        if (rc) return rc;
        if (!d_pixels[i] || !d_depths[i]) return fail(c, SLIMRT_E_ARG, "view output buffers are NULL");
    }
    auto *v = new slimrt_views();
    v->c = c;
    v->params.assign(views, views + count);
    for (uint32_t i = 0; i < count; i++) {
        v->pixels.push_back((float4 *)d_pixels[i]);
        v->depths.push_back((float *)d_depths[i]);
    }
    v->scene_cap = c->scene_height;
    const uint32_t G = (uint32_t)c->geos.size();
    if (hipMalloc(&v->d_recs, count * sizeof(ViewRec)) != hipSuccess || hipMalloc(&v->d_counter, 64) != hipSuccess ||
        hipHostMalloc(&v->h_stage, G * sizeof(slimrt_geometry), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&v->d_stage, G * sizeof(slimrt_geometry)) != hipSuccess || hipEventCreate(&v->ev0) != hipSuccess ||
        hipEventCreate(&v->ev1) != hipSuccess) {
        slimrt_views_destroy(v);
        return fail(c, SLIMRT_E_MEMORY, "views: device allocation failed");
    }
    int rc = build_tables(v);
    if (rc) {
        std::string m = v->err;
        slimrt_views_destroy(v);
        return fail(c, rc, m);
    }
    *out = v;
    return SLIMRT_OK;
}

int slimrt_views_destroy(slimrt_views *v) {
    if (!v) return SLIMRT_E_ARG;
    if (v->c) (void)hipStreamSynchronize(v->c->stream);
    drop_graph(v);
    if (v->d_recs) (void)hipFree(v->d_recs);
    if (v->d_counter) (void)hipFree(v->d_counter);
    if (v->d_stage) (void)hipFree(v->d_stage);
    if (v->h_stage) (void)hipHostFree(v->h_stage);
    if (v->ev0) (void)hipEventDestroy(v->ev0);
    if (v->ev1) (void)hipEventDestroy(v->ev1);
    delete v;
    return SLIMRT_OK;
}

int slimrt_views_configure(slimrt_views *v, const slimrt_views_config *cfg) {
    if (!v || !cfg) return SLIMRT_E_ARG;
    if (cfg->tile_shape > 3 || cfg->blocks_per_cu < 1 || cfg->blocks_per_cu > 4 || cfg->tiles_per_grab < 1 ||
        cfg->tiles_per_grab > 8 || cfg->persistent > 1 || (cfg->node_width != 0 && cfg->node_width != 2 &&
                                                            cfg->node_width != 4 && cfg->node_width != 8))
        return vfail(v, SLIMRT_E_ARG, "views config out of range");
    VIEWS_TRY(v, hipStreamSynchronize(v->c->stream));
    v->cfg = *cfg;
    v->generation++;
    drop_graph(v);
    return build_tables(v);
}

int slimrt_views_get_config(const slimrt_views *v, slimrt_views_config *cfg) {
    if (!v || !cfg) return SLIMRT_E_ARG;
    *cfg = v->cfg;
    return SLIMRT_OK;
}

int slimrt_views_frame(slimrt_views *v, const slimrt_geometry *g, uint32_t count, uint32_t mode, uint32_t use_graph) {
    if (!v) return SLIMRT_E_ARG;
    slimrt_context *c = v->c;
    const bool update = g != nullptr;
    if (update) {
        int rc;
        if ((rc = check_geometries(c, g, count)) || (rc = check_update_mode(c, mode))) return vfail(v, rc, c->err);
        if (v->in_flight) VIEWS_TRY(v, hipEventSynchronize(v->ev1));   // the previous frame's copy has read the stage
        std::memcpy(v->h_stage, g, count * sizeof(slimrt_geometry));
        mirror_geometries(c, g, mode);
        if (c->scene_height > v->scene_cap) {   // a rebuild may deepen the tree: stacks sized for the worst case
            v->scene_cap = c->scene_height;
            drop_graph(v);
        }
    }
    if (use_graph) {
        const int64_t key = (int64_t)(update ? 1 : 0) | ((int64_t)mode << 1) | ((int64_t)v->generation << 2);
        if (!v->exec || v->graph_key != key) {
            drop_graph(v);
            VIEWS_TRY(v, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
            int rc = enqueue_frame(v, update, mode);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) return vfail(v, SLIMRT_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            v->graph = graph;
            VIEWS_TRY(v, hipGraphInstantiate(&v->exec, v->graph, nullptr, nullptr, 0));
            v->graph_key = key;
        }
        VIEWS_TRY(v, hipEventRecord(v->ev0, c->stream));
        VIEWS_TRY(v, hipGraphLaunch(v->exec, c->stream));
        VIEWS_TRY(v, hipEventRecord(v->ev1, c->stream));
    } else {
        VIEWS_TRY(v, hipEventRecord(v->ev0, c->stream));
        int rc = enqueue_frame(v, update, mode);
        if (rc) return rc;
        VIEWS_TRY(v, hipEventRecord(v->ev1, c->stream));
    }
    v->timed = true;
    v->in_flight = true;
    return SLIMRT_OK;
}

int slimrt_views_last_ms(slimrt_views *v, float *ms) {
    if (!v || !ms) return SLIMRT_E_ARG;
    if (!v->timed) return vfail(v, SLIMRT_E_STATE, "no frame traced");
    return frame_ms(v, ms);
}

int slimrt_views_autotune(slimrt_views *v, uint32_t reps, slimrt_views_config *best, float *best_ms) {
    if (!v || reps == 0) return SLIMRT_E_ARG;
    slimrt_views_config keep = v->cfg, win = v->cfg;
    float win_ms = INFINITY;
    std::vector<slimrt_views_config> cands;
    const bool fits8 = v->c->mesh_nodes8_max <= 65536 && v->c->mesh_leaf_max <= 254;
    for (uint32_t width : {2u, 4u, 8u})
        for (uint32_t shape = 0; shape < 4 && (width == 2 || (width == 4 && v->c->mesh_nodes4_max <= 65536) ||
                                               (width == 8 && fits8)); shape++) {
            cands.push_back({shape, 1, 1, 0, width});
            for (uint32_t bpc = 1; bpc <= 2; bpc++)
                for (uint32_t grab : {1u, 4u}) cands.push_back({shape, bpc, grab, 1, width});
        }
    for (const slimrt_views_config &cfg : cands) {
        int rc = slimrt_views_configure(v, &cfg);
        if (rc) return rc;
        std::vector<float> t;
        for (uint32_t r = 0; r <= reps; r++) {   // one warm-up frame
            if ((rc = slimrt_views_frame(v, nullptr, 0, 0, 0))) return rc;
            float ms;
            if ((rc = frame_ms(v, &ms))) return rc;
            if (r) t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        const float med = t[t.size() / 2];
        if (med < win_ms) {
            win_ms = med;
            win = cfg;
        }
    }
    int rc = slimrt_views_configure(v, &win);
    if (rc) {
        (void)slimrt_views_configure(v, &keep);
        return rc;
    }
    if (best) *best = win;
    if (best_ms) *best_ms = win_ms;
    return check_status(v->c) ? vfail(v, SLIMRT_E_OVERFLOW, v->c->err) : SLIMRT_OK;
}

const char *slimrt_views_last_error(const slimrt_views *v) { return v ? v->err.c_str() : "null views"; }

int slimrt_render_autotune(slimrt_context *c, const slimrt_render_params *params, void *d_pixels, void *d_depths,
                           uint32_t reps, uint32_t *best_flags, float *best_ms) {
    if (!c || !params || !best_flags || reps == 0) return SLIMRT_E_ARG;
    const uint32_t base = params->flags & ~(uint32_t)(SLIMRT_WAVEFRONT | SLIMRT_WAVEFRONT_PER_BOUNCE |
                                                      SLIMRT_WAVEFRONT_SORT | SLIMRT_WAVEFRONT_BANDS(3) |
                                                      SLIMRT_WAVEFRONT_SPLIT | SLIMRT_WAVEFRONT_COUNT | SLIMRT_BVH4 |
                                                      SLIMRT_BVH8);
    std::vector<uint32_t> forms;
    // node width 8: the per-sample frame only, on meshes the 8-wide walk can hold
    if (c->mesh_nodes8_max <= 65536 && c->mesh_leaf_max <= 254) forms.push_back(SLIMRT_BVH8);
    for (uint32_t width : {0u, (uint32_t)SLIMRT_BVH4}) {   // node width 2 or 4
        if (width && c->mesh_nodes4_max > 65536) continue;   // meshes too large for the 4-wide walk
        forms.push_back(width);
        for (uint32_t split : {0u, (uint32_t)SLIMRT_WAVEFRONT_SPLIT})
            for (uint32_t bands_log2 : {0u, 2u})
                for (uint32_t v : {0u, (uint32_t)SLIMRT_WAVEFRONT_SORT, (uint32_t)SLIMRT_WAVEFRONT_PER_BOUNCE,
                                   (uint32_t)(SLIMRT_WAVEFRONT_PER_BOUNCE | SLIMRT_WAVEFRONT_SORT)})
                    forms.push_back(SLIMRT_WAVEFRONT | v | SLIMRT_WAVEFRONT_BANDS(bands_log2) | split | width);
    }
    slimrt_render_params p = *params;
    float win_ms = INFINITY;
    uint32_t win = base;
    for (uint32_t f : forms) {
        p.flags = base | f;
        std::vector<float> t;
        // three warm-up frames: a per-sample form renders its tiles costliest first from its fourth frame on
        for (uint32_t r = 0; r < reps + 3; r++) {
            int rc = slimrt_render_device(c, &p, d_pixels, d_depths);
            if (rc) return rc;
            float ms;
            if ((rc = slimrt_last_kernel_ms(c, &ms))) return rc;
            if (r >= 3) t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        if (t[t.size() / 2] < win_ms) {
            win_ms = t[t.size() / 2];
            win = p.flags;
        }
    }
    *best_flags = win;
    if (best_ms) *best_ms = win_ms;
    return check_status(c);
}

}  // extern "C"
