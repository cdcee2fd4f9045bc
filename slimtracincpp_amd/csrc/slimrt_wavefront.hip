// Glass and mirror bounces as a wavefront pipeline (render params flag SLIMRT_WAVEFRONT).
//
// The per-sample megakernel runs every path to its end inside one thread, so the few paths that keep bouncing
// through glass and mirrors hold their whole wave while the other lanes idle.  Here the first iteration of the
// reference integrator (renderPixelBeauty's loop, src/slim/renderer/ray_tracer.h:42-105) runs for every sample
// in the tile-shaped primary kernel, which without the bounce loop needs fewer registers (3 waves per SIMD
// instead of 2); paths that bounce are appended to a ray queue - one atomic per wave, lanes placed by ballot
// and popcount - and the bounce kernel runs the compacted queue: by default each path to its end in one launch,
// or (SLIMRT_WAVEFRONT_PER_BOUNCE) one iteration per launch with the survivors compacted again.  Optionally
// (SLIMRT_WAVEFRONT_SORT) a queue is sorted into the 8 direction octants first (counting sort: per-bin counts,
// prefix, scatter) so that a wave's rays head the same way through the BVHs.  A path's colour accumulates in a
// per-sample buffer in the integrator's own order, so every sample is bit-identical to the megakernel's.
//
// The bounce launches carry few rays, each with a long chain of dependent fetches, so their time is the latency
// of their slowest waves rather than throughput.  With SLIMRT_WAVEFRONT_BANDS the frame is cut into horizontal
// bands with queues of their own: the primary pass runs band after band on the caller's stream while each
// band's bounces run on a second stream as soon as its primary pass is done, so those tails overlap the next
// band's primary work (event fork and join, both streams at the default priority).
//
// A path in flight is 48 bytes: the bounced ray's origin and direction (as ray_reset leaves them), throughput,
// the pixel's ray-cone scaling factor, the remaining depth and the canvas offset.
#include <algorithm>
#include <string>

#include "slimrt_context.h"

namespace {

struct PathState {
    float ox, oy, oz, dx, dy, dz;
    float tr, tg, tb, scaling;
    uint32_t depth_left, pixel;
};
static_assert(sizeof(PathState) == 48, "path state layout");

enum : uint32_t { WF_COUNT_A = 0, WF_COUNT_B = 1, WF_BINS = 2, WF_CURSORS = 10, WF_COUNTERS = 18 };

// One iteration of the integrator loop (the body of render_beauty, in its order) after its closest-hit trace:
// geo and hit are what scene_trace returned for ray (whose origin it offset).  Returns whether the path continues,
// with ray, throughput and depth_left advanced to the bounce.
template <class TR>
__device__ __forceinline__ bool shade_iteration(const SceneView &s, const RenderParams &rp, Ray &ray, Hit &hit, int geo,
                                                float scaling, uint32_t &depth_left, color &tp, color &col,
                                                float &depth, TR &tr) {
    Surface sf;
    color cur = cs(0), next_tp = cs(0);
    bool bounced = false;
    sf.geo = geo;
    if (sf.geo >= 0) {
        hit.scaling_factor = keeps_ray_cones(rp) ? scaling : 0.0f;
        const Geo g = s.geos[sf.geo];
        const Mat m = s.mats[g.material_id];
        if ((mat_flags(m) & MAT_EMISSIVE) && g.type == GT_Quad && !hit.from_behind) {
            depth_left = 0;
            cur = color{m.er, m.eg, m.eb};
        } else {
            prepare_for_shading(sf, s, ray, hit, tr);
            if (depth_left == rp.max_depth) depth = depth_at(rp.proj, hit.position);
            for (uint32_t i = 0; i < s.L; i++) shade_from_light(sf, s.lights[i], s, tr, cur);
            if (s.flags & 1u) shade_from_emissive_quads(sf, s, cur);
            if (rp.sky_irradiance >= 0 && rp.sky_radiance >= 0) {
                sf.L = sf.N;
                sf.NdotL = 1.0f;
                color D = tex_sample_cube(s, (uint32_t)rp.sky_irradiance, sf.N.x, sf.N.y, sf.N.z);
                color S = tex_sample_cube(s, (uint32_t)rp.sky_radiance, sf.R.x, sf.R.y, sf.R.z);
                radiance_fraction(sf);
                cur = cmul_add(D, sf.Fd, cmul_add(sf.Fs, S, cur));
            }
            if ((mat_flags(sf.mat) & (MAT_REFLECTIVE | MAT_REFRACTIVE)) && --depth_left) {
                sf.F = schlick(clamp01(dot(sf.N, sf.R)), color{sf.mat.fr, sf.mat.fg, sf.mat.fb});
                next_tp = sf.refracted ? one_minus(sf.F) : sf.F;
                ray_reset(ray, hit.position, sf.RF);
                bounced = true;
            } else depth_left = 0;
        }
    } else {
        depth_left = 0;
        if (rp.sky_color >= 0) cur = tex_sample_cube(s, (uint32_t)rp.sky_color, ray.direction.x, ray.direction.y, ray.direction.z);
    }
    for (uint32_t i = 0; i < s.L; i++) {
        const Light l = s.lights[i];
        float density;
        if (light_glow(l, ray, hit.distance, density)) cur = cscale_add(color{l.r, l.g, l.b}, glow_factor(density), cur);
    }
    col = cmul_add(cur, tp, col);
    tp = tp * next_tp;
    return bounced;
}

template <class TR>
__device__ __forceinline__ bool beauty_iteration(const SceneView &s, const RenderParams &rp, Ray &ray, float scaling,
                                                 uint32_t &depth_left, color &tp, color &col, float &depth, TR &tr) {
    Hit hit;
    hit.scaling_factor = scaling;
    const int geo = scene_trace(s, ray, hit, false, INFINITY, tr);
    return shade_iteration(s, rp, ray, hit, geo, scaling, depth_left, tp, col, depth, tr);
}

// First-hit records of the split primary pass, structure of arrays over the frame's sample slots (tile * kBlock +
// thread, so a wave's 64 records are contiguous): the hit's local normal and distance; its uv, uv coverage and
// NdotRd; the geometry index with from_behind in bit 31 (all ones for a miss).  What shading reads of a Hit.
constexpr size_t kHitSlotBytes = 36;
struct HitSlots {
    float4 *a, *b;
    uint32_t *g;
};
inline HitSlots hit_slots(void *base, uint32_t cap) {
    char *p = (char *)base;
    return HitSlots{(float4 *)p, (float4 *)(p + (size_t)cap * 16), (uint32_t *)(p + (size_t)cap * 32)};
}
__device__ __forceinline__ void store_hit(const HitSlots &h, size_t i, const Hit &hit, int geo) {
    h.a[i] = make_float4(hit.normal.x, hit.normal.y, hit.normal.z, hit.distance);
    h.b[i] = make_float4(hit.uv.u, hit.uv.v, hit.uv_coverage, hit.NdotRd);
    h.g[i] = geo < 0 ? ~0u : ((uint32_t)geo | (hit.from_behind ? 0x80000000u : 0u));
}
__device__ __forceinline__ int load_hit(const HitSlots &h, size_t i, Hit &hit) {
    const float4 a = h.a[i], b = h.b[i];
    const uint32_t g = h.g[i];
    hit.normal = v3{a.x, a.y, a.z};
    hit.distance = a.w;
    hit.uv = v2{b.x, b.y};
    hit.uv_coverage = b.z;
    hit.NdotRd = b.w;
    hit.from_behind = g != ~0u && (g >> 31) != 0;
    hit.id = 0;
    hit.position = vs(0.0f);   // prepare_for_shading derives it from the ray and the distance
    return g == ~0u ? -1 : (int)(g & 0x7FFFFFFFu);
}

__device__ __forceinline__ void finish_sample(const RenderParams &rp, color col, float4 *pixels, uint32_t o,
                                              uint32_t &status) {
    float px[4];
    output_value(rp, color{tone_map(col.r), tone_map(col.g), tone_map(col.b)}, px);
    pixels[o] = make_float4(px[0], px[1], px[2], px[3]);
    status |= nan_status(rp, px, 0.0f);
}

__device__ __forceinline__ void push_path(PathState *q, uint32_t *count, uint32_t cap, bool cont, const Ray &ray,
                                          color tp, float scaling, uint32_t depth_left, uint32_t pixel, color col,
                                          float4 *accum, uint32_t *status) {
    const uint32_t slot = wave_append(cont, count);
    if (!cont) return;
    if (slot >= cap) {   // cannot happen: one entry per sample at most
        atomicOr(status, kStatusQueue);
        return;
    }
    PathState st;
    st.ox = ray.origin.x; st.oy = ray.origin.y; st.oz = ray.origin.z;
    st.dx = ray.direction.x; st.dy = ray.direction.y; st.dz = ray.direction.z;
    st.tr = tp.r; st.tg = tp.g; st.tb = tp.b; st.scaling = scaling;
    st.depth_left = depth_left;
    st.pixel = pixel;
    q[slot] = st;
    accum[pixel] = make_float4(col.r, col.g, col.b, 0.0f);
}

__global__ void wf_reset_kernel(uint32_t *__restrict__ counters, uint32_t n, uint32_t *__restrict__ stats) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0u;
    if (stats)
        for (uint32_t i = threadIdx.x; i < kWfStats; i += blockDim.x) stats[i] = 0u;
}

// Path counters (SLIMRT_WAVEFRONT_COUNT, stats != nullptr): one atomic per lane and event, for checks only.
__device__ __forceinline__ void count_iteration(uint32_t *stats, const RenderParams &rp, uint32_t depth_left) {
    if (stats) atomicAdd(stats + min(rp.max_depth - depth_left, kWfIterations - 1u), 1u);
}
__device__ __forceinline__ void count_end(uint32_t *stats, bool spawned) {
    if (stats) atomicAdd(stats + (spawned ? kWfSpawned : kWfFinished), 1u);
}

// Iteration 1 for every sample, tiles as in the render kernel.
// Tiles [tile_base, tile_base + tiles) of the frame (one band, or all of it).
// With SPLIT the closest hits come from wf_trace_kernel's records instead of a trace here.
template <bool WIDE, bool SPLIT>
__global__ __launch_bounds__(kBlock) void wf_primary_kernel(SceneView s, RenderParams rp, float4 *__restrict__ pixels,
                                                            float *__restrict__ depths, float4 *__restrict__ accum,
                                                            PathState *__restrict__ queue, uint32_t *__restrict__ counters,
                                                            uint32_t cap, uint32_t scene_depth, uint32_t mesh_depth,
                                                            uint32_t tiles_x, uint32_t tile_base, uint32_t tiles,
                                                            uint32_t *__restrict__ status,
                                                            uint32_t *__restrict__ stats, HitSlots hits) {
    extern __shared__ uint32_t lds[];
    const uint32_t local = blockIdx.x;
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    const uint32_t x = (tile % tiles_x) * kTile + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t y = (tile / tiles_x) * kTile + (wave >> 1) * 8u + (lane >> 3);
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;
    TracerT<WIDE ? 4u : 2u> tr;
    bind_lds_stacks(tr, lds, t, scene_depth, mesh_depth);
    tr.cnt = nullptr;
    const bool inside = x < ws && y < hs;
    bool cont = false;
    Ray ray;
    color tp = cs(1.0f), col = cs(0);
    float scaling = 0.0f, depth = INFINITY;
    uint32_t depth_left = rp.max_depth, o = 0;
    if (inside) {
        v3 dir;
        primary_ray(rp.proj, x, y, dir, scaling);
        const Projection &p = rp.proj;
        ray_reset(ray, v3{p.camera_position[0], p.camera_position[1], p.camera_position[2]}, normalized(dir));
        if (depth_left) {
            count_iteration(stats, rp, depth_left);
            if constexpr (SPLIT) {
                Hit hit;
                const int geo = load_hit(hits, (size_t)tile * kBlock + t, hit);
                hit.scaling_factor = scaling;
                ray_reset(ray, scale_add(ray.direction, kTraceOffset, ray.origin), ray.direction);   // as scene_trace
                cont = shade_iteration(s, rp, ray, hit, geo, scaling, depth_left, tp, col, depth, tr);
            } else {
                cont = beauty_iteration(s, rp, ray, scaling, depth_left, tp, col, depth, tr);
            }
        }
        count_end(stats, cont);
        o = canvas_offset(rp, x, y);
        depths[o] = depth;
        if (depth != depth) tr.status |= kStatusNaN;
        if (!cont) finish_sample(rp, col, pixels, o, tr.status);
        if (tr.status) atomicOr(status, tr.status);
    }
    push_path(queue, counters + WF_COUNT_A, cap, inside && cont, ray, tp, scaling, depth_left, o, col, accum, status);
}

// The closest hit of every sample's camera ray, for the split primary pass: traversal only, so the kernel holds
// fewer registers than the shading kernel and keeps more waves in flight to cover the walk's dependent fetches.
template <bool WIDE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
void wf_trace_kernel(SceneView s, RenderParams rp, HitSlots hits, uint32_t scene_depth, uint32_t mesh_depth,
                     uint32_t tiles_x, uint32_t tile_base, uint32_t tiles, uint32_t *__restrict__ status) {
    extern __shared__ uint32_t lds[];
    const uint32_t local = blockIdx.x;
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    const uint32_t x = (tile % tiles_x) * kTile + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t y = (tile / tiles_x) * kTile + (wave >> 1) * 8u + (lane >> 3);
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;
    TracerT<WIDE ? 4u : 2u> tr;
    bind_lds_stacks(tr, lds, t, scene_depth, mesh_depth);
    tr.cnt = nullptr;
    if (x < ws && y < hs && rp.max_depth) {
        v3 dir;
        float scaling;
        primary_ray(rp.proj, x, y, dir, scaling);
        const Projection &p = rp.proj;
        Ray ray;
        ray_reset(ray, v3{p.camera_position[0], p.camera_position[1], p.camera_position[2]}, normalized(dir));
        Hit hit;
        hit.scaling_factor = scaling;
        const int geo = scene_trace(s, ray, hit, false, INFINITY, tr);
        store_hit(hits, (size_t)tile * kBlock + t, hit, geo);
        if (tr.status) atomicOr(status, tr.status);
    }
}

// Further iterations over the queue `in` (count at in_count): with TO_END each path runs to its end; otherwise one
// iteration, appending continuing paths to `out`.
template <bool TO_END, bool WIDE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(TO_END && !WIDE ? 3 : 1)))
void wf_bounce_kernel(SceneView s, RenderParams rp, float4 *__restrict__ pixels, float4 *__restrict__ accum,
                      const PathState *__restrict__ in, const uint32_t *__restrict__ in_count,
                      PathState *__restrict__ out, uint32_t *__restrict__ out_count, uint32_t cap,
                      uint32_t scene_depth, uint32_t mesh_depth, uint32_t *__restrict__ status,
                      uint32_t *__restrict__ stats) {
    extern __shared__ uint32_t lds[];
    const uint32_t t = threadIdx.x;
    TracerT<WIDE ? 4u : 2u> tr;
    bind_lds_stacks(tr, lds, t, scene_depth, mesh_depth);
    tr.cnt = nullptr;
    const uint32_t n = min(*in_count, cap);
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += stride) {   // uniform trip count per workgroup
        const uint32_t i = base + t;
        const bool live = i < n;
        bool cont = false;
        Ray ray;
        color tp = cs(0), col = cs(0);
        float scaling = 0.0f, depth = 0.0f;
        uint32_t depth_left = 0, pixel = 0;
        if (live) {
            const PathState st = in[i];
            ray_reset(ray, v3{st.ox, st.oy, st.oz}, v3{st.dx, st.dy, st.dz});
            tp = color{st.tr, st.tg, st.tb};
            scaling = st.scaling;
            depth_left = st.depth_left;
            pixel = st.pixel;
            const float4 a = accum[pixel];
            col = color{a.x, a.y, a.z};
#pragma unroll 1
            for (;;) {   // one iteration, or (TO_END) every remaining one: a single inlined copy either way
                count_iteration(stats, rp, depth_left);
                cont = beauty_iteration(s, rp, ray, scaling, depth_left, tp, col, depth, tr);
                if (!TO_END || !cont) break;
                count_end(stats, true);
            }
            count_end(stats, cont);
            if (!cont) finish_sample(rp, col, pixels, pixel, tr.status);
        }
        if (!TO_END) push_path(out, out_count, cap, live && cont, ray, tp, scaling, depth_left, pixel, col, accum, status);
    }
    if (tr.status) atomicOr(status, tr.status);
}

__device__ __forceinline__ uint32_t octant(const PathState &st) {
    return (st.dx < 0.0f ? 1u : 0u) | (st.dy < 0.0f ? 2u : 0u) | (st.dz < 0.0f ? 4u : 0u);
}

// Counting sort of the queue by direction octant: counts per bin (one atomic per wave and bin) ...
__global__ __launch_bounds__(256) void wf_bin_count_kernel(const PathState *__restrict__ q, uint32_t *__restrict__ counters,
                                                           uint32_t cap) {
    const uint32_t n = min(counters[WF_COUNT_A], cap);
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t key = i < n ? octant(q[i]) : 8u;
#pragma unroll
        for (uint32_t b = 0; b < 8; b++) (void)wave_append(key == b, counters + WF_BINS + b);
    }
}

// ... exclusive prefix into the cursors; the sorted queue's count is the unsorted one's ...
__global__ void wf_bin_scan_kernel(uint32_t *__restrict__ counters) {
    if (threadIdx.x != 0) return;
    uint32_t acc = 0;
    for (uint32_t b = 0; b < 8; b++) {
        counters[WF_CURSORS + b] = acc;
        acc += counters[WF_BINS + b];
        counters[WF_BINS + b] = 0u;
    }
    counters[WF_COUNT_B] = counters[WF_COUNT_A];
}

// ... and the scatter into queue B, which empties queue A for the next iteration's appends.
__global__ __launch_bounds__(256) void wf_bin_scatter_kernel(const PathState *__restrict__ a, PathState *__restrict__ b,
                                                             uint32_t *__restrict__ counters, uint32_t cap) {
    const uint32_t n = min(counters[WF_COUNT_B], cap);
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        const uint32_t i = base + threadIdx.x;
        PathState st;
        uint32_t key = 8u;
        if (i < n) {
            st = a[i];
            key = octant(st);
        }
#pragma unroll
        for (uint32_t bin = 0; bin < 8; bin++) {
            const uint32_t slot = wave_append(key == bin, counters + WF_CURSORS + bin);
            if (key == bin && slot < cap) b[slot] = st;
        }
    }
}

__global__ void wf_clear_count_kernel(uint32_t *__restrict__ counters, uint32_t which) {
    if (threadIdx.x == 0) counters[which] = 0u;
}

}  // namespace

namespace slimrt {

// The bounce launches of one queue (queue A holds the paths to run; sorting moves them to B, emptying A).
template <bool WIDE>
int enqueue_bounces(slimrt_context *c, hipStream_t st, const SceneView &v, const RenderParams &rp, float4 *pixels,
                    float4 *accum, PathState *qa, PathState *qb, uint32_t *cnt, uint32_t cap, size_t lds,
                    uint32_t scene_depth, uint32_t mesh_depth, uint32_t blocks, bool per_bounce, bool sort,
                    uint32_t *stats) {
    auto sort_a_into_b = [&]() {
        hipLaunchKernelGGL(wf_bin_count_kernel, dim3(blocks), dim3(256), 0, st, qa, cnt, cap);
        hipLaunchKernelGGL(wf_bin_scan_kernel, dim3(1), dim3(64), 0, st, cnt);
        hipLaunchKernelGGL(wf_bin_scatter_kernel, dim3(blocks), dim3(256), 0, st, qa, qb, cnt, cap);
        hipLaunchKernelGGL(wf_clear_count_kernel, dim3(1), dim3(64), 0, st, cnt, (uint32_t)WF_COUNT_A);
    };
    if (rp.max_depth < 2) return SLIMRT_OK;
    if (!per_bounce) {
        if (sort) sort_a_into_b();
        hipLaunchKernelGGL((wf_bounce_kernel<true, WIDE>), dim3(blocks), dim3(kBlock), lds, st, v, rp, pixels, accum,
                           (const PathState *)(sort ? qb : qa), (const uint32_t *)(cnt + (sort ? WF_COUNT_B : WF_COUNT_A)),
                           sort ? qa : qb, cnt + (sort ? WF_COUNT_A : WF_COUNT_B), cap, scene_depth, mesh_depth,
                           c->d_status, stats);
        HIP_TRY(c, hipGetLastError());
        return SLIMRT_OK;
    }
    uint32_t cur_count = WF_COUNT_A, next_count = WF_COUNT_B;   // unsorted: ping-pong between A and B
    for (uint32_t it = 1; it < rp.max_depth; it++) {
        if (sort) {   // A -> sorted B (A emptied), then B -> A
            sort_a_into_b();
            hipLaunchKernelGGL((wf_bounce_kernel<false, WIDE>), dim3(blocks), dim3(kBlock), lds, st, v, rp, pixels, accum,
                               (const PathState *)qb, (const uint32_t *)(cnt + WF_COUNT_B), qa, cnt + WF_COUNT_A,
                               cap, scene_depth, mesh_depth, c->d_status, stats);
        } else {
            PathState *in = cur_count == WF_COUNT_A ? qa : qb, *out = cur_count == WF_COUNT_A ? qb : qa;
            hipLaunchKernelGGL(wf_clear_count_kernel, dim3(1), dim3(64), 0, st, cnt, next_count);
            hipLaunchKernelGGL((wf_bounce_kernel<false, WIDE>), dim3(blocks), dim3(kBlock), lds, st, v, rp, pixels, accum,
                               (const PathState *)in, (const uint32_t *)(cnt + cur_count), out, cnt + next_count,
                               cap, scene_depth, mesh_depth, c->d_status, stats);
            std::swap(cur_count, next_count);
        }
        HIP_TRY(c, hipGetLastError());
    }
    return SLIMRT_OK;
}

int launch_wavefront(slimrt_context *c, const RenderParams &rp, float4 *pixels, float *depths, bool per_bounce, bool sort,
                     uint32_t bands) {
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;
    const uint32_t tiles_x = (ws + kTile - 1) / kTile, tiles_y = (hs + kTile - 1) / kTile;
    bands = std::max(1u, std::min(bands, std::min(tiles_y, 8u)));
    const uint32_t slots = tiles_x * tiles_y * kBlock;   // queue entries: one per sample slot of every tile
    if (c->wf_cap < slots) {
        for (void *p : {(void *)c->wf_queue_a, (void *)c->wf_queue_b, (void *)c->wf_accum})
            if (p) HIP_TRY(c, ctx_free(c, p));
        c->wf_queue_a = c->wf_queue_b = nullptr;
        c->wf_accum = nullptr;
        c->wf_cap = 0;
        HIP_TRY(c, ctx_malloc(c, &c->wf_queue_a, (size_t)slots * sizeof(PathState)));
        HIP_TRY(c, ctx_malloc(c, &c->wf_queue_b, (size_t)slots * sizeof(PathState)));
        HIP_TRY(c, ctx_malloc(c, &c->wf_accum, (size_t)slots * sizeof(float4)));   // canvas offsets < ws * hs <= slots
        c->wf_cap = slots;
    }
    const bool split = (rp.flags & SLIMRT_WAVEFRONT_SPLIT) != 0;
    if (split && c->wf_hits_cap < slots) {
        if (c->wf_hits) HIP_TRY(c, ctx_free(c, c->wf_hits));
        c->wf_hits = nullptr;
        c->wf_hits_cap = 0;
        HIP_TRY(c, ctx_malloc(c, &c->wf_hits, (size_t)slots * kHitSlotBytes));
        c->wf_hits_cap = slots;
    }
    if (!c->wf_counters) HIP_TRY(c, ctx_malloc(c, &c->wf_counters, 8 * WF_COUNTERS * sizeof(uint32_t)));
    const uint32_t scene_depth = c->scene_height + 1, mesh_depth = mesh_stack_depth(c, wide_mesh_bvh(rp) ? 4u : 2u);
    const size_t lds = stack_lds_bytes(scene_depth, mesh_depth, wide_mesh_bvh(rp) ? 4u : 2u);
    if (lds > 160 * 1024) return fail(c, SLIMRT_E_ARG, "BVH too deep for the LDS traversal stacks");
    if (wide_mesh_bvh(rp) && check_width(c, 4)) return SLIMRT_E_ARG;
    int cus = 0;
    HIP_TRY(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const SceneView v = device_view(c);
    PathState *qa = (PathState *)c->wf_queue_a, *qb = (PathState *)c->wf_queue_b;
    const HitSlots hits = split ? hit_slots(c->wf_hits, c->wf_hits_cap) : HitSlots{nullptr, nullptr, nullptr};
    float4 *accum = (float4 *)c->wf_accum;
    uint32_t *cnt = c->wf_counters;
    const uint32_t blocks = (uint32_t)cus * 2;
    uint32_t *stats = nullptr;
    if (rp.flags & SLIMRT_WAVEFRONT_COUNT) {
        if (!c->wf_stats) HIP_TRY(c, ctx_malloc(c, &c->wf_stats, kWfStats * sizeof(uint32_t)));
        stats = c->wf_stats;
    }
    hipLaunchKernelGGL(wf_reset_kernel, dim3(1), dim3(256), 0, c->stream, cnt, bands * WF_COUNTERS, stats);
    HIP_TRY(c, hipGetLastError());
    hipStream_t side = c->stream;
    if (bands > 1) {   // fork: the side stream starts after everything queued on the caller's stream so far
        if (!c->side_stream) HIP_TRY(c, hipStreamCreateWithFlags(&c->side_stream, hipStreamNonBlocking));
        if (!c->wf_fork) HIP_TRY(c, hipEventCreateWithFlags(&c->wf_fork, hipEventDisableTiming));
        if (!c->wf_join) HIP_TRY(c, hipEventCreateWithFlags(&c->wf_join, hipEventDisableTiming));
        for (uint32_t b = 0; b < bands; b++)
            if (!c->wf_band[b]) HIP_TRY(c, hipEventCreateWithFlags(&c->wf_band[b], hipEventDisableTiming));
        side = c->side_stream;
        HIP_TRY(c, hipEventRecord(c->wf_fork, c->stream));
        HIP_TRY(c, hipStreamWaitEvent(side, c->wf_fork, 0));
    }
    for (uint32_t b = 0; b < bands; b++) {
        const uint32_t row0 = tiles_y * b / bands, row1 = tiles_y * (b + 1) / bands;
        const uint32_t tile_base = row0 * tiles_x, tiles = (row1 - row0) * tiles_x;
        const uint32_t base = tile_base * kBlock, cap = tiles * kBlock;
        uint32_t *bc = cnt + b * WF_COUNTERS;
#define WF_PRIMARY(WIDE, SPLIT)                                                                                 \
    hipLaunchKernelGGL((wf_primary_kernel<WIDE, SPLIT>), dim3(tiles), dim3(kBlock), lds, c->stream, v, rp,           \
                       pixels, depths, accum, qa + base, bc, cap, scene_depth, mesh_depth, tiles_x, tile_base, tiles, \
                       c->d_status, stats, hits)
#define WF_TRACE(WIDE)                                                                                          \
    hipLaunchKernelGGL((wf_trace_kernel<WIDE>), dim3(tiles), dim3(kBlock), lds, c->stream, v, rp, hits,               \
                       scene_depth, mesh_depth, tiles_x, tile_base, tiles, c->d_status)
        if (split) {
            if (wide_mesh_bvh(rp)) WF_TRACE(true); else WF_TRACE(false);
            HIP_TRY(c, hipGetLastError());
            if (wide_mesh_bvh(rp)) WF_PRIMARY(true, true); else WF_PRIMARY(false, true);
        } else {
            if (wide_mesh_bvh(rp)) WF_PRIMARY(true, false); else WF_PRIMARY(false, false);
        }
#undef WF_PRIMARY
#undef WF_TRACE
        HIP_TRY(c, hipGetLastError());
        if (bands > 1) {
            HIP_TRY(c, hipEventRecord(c->wf_band[b], c->stream));
            HIP_TRY(c, hipStreamWaitEvent(side, c->wf_band[b], 0));
        }
        int rc = wide_mesh_bvh(rp)
            ? enqueue_bounces<true>(c, side, v, rp, pixels, accum, qa + base, qb + base, bc, cap, lds, scene_depth,
                                    mesh_depth, blocks, per_bounce, sort, stats)
            : enqueue_bounces<false>(c, side, v, rp, pixels, accum, qa + base, qb + base, bc, cap, lds, scene_depth,
                                     mesh_depth, blocks, per_bounce, sort, stats);
        if (rc) return rc;
    }
    if (bands > 1) {   // join: the caller's stream continues after the last bounce launch
        HIP_TRY(c, hipEventRecord(c->wf_join, side));
        HIP_TRY(c, hipStreamWaitEvent(c->stream, c->wf_join, 0));
    }
    return SLIMRT_OK;
}

}  // namespace slimrt
