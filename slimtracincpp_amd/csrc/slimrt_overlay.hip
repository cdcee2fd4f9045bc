// slimrt_overlay: wireframe overlays drawn into a canvas (include/slimrt.h, "wireframe overlays").
//
// The reference draws edge after edge and sample after sample, blending each into the canvas (Canvas::setPixel,
// src/slim/draw/canvas.h:128-185), so what a canvas sample ends up holding depends on the order of the line
// samples that land on it.  The device form keeps that order without drawing sequentially:
//   1. one thread per edge: view space, clipping, projection: a canvas line (box_lines_kernel, mesh_lines_kernel);
//   2. one thread per line walks it and counts the samples it draws (count_kernel);
//   3. an exclusive scan gives each line the drawing-order index of its first sample;
//   4. the walk again writes every sample (canvas offset; opacity, depth and line) at its drawing-order index
//      (emit_kernel);
//   5. a radix sort by canvas offset, stable, keeps drawing order among the samples of one canvas offset;
//   6. one thread per run of equal offsets blends its samples into the canvas in that order (blend_kernel).
// Every canvas sample therefore sees the reference's sequence of setPixel calls, bit for bit.  Lines are walked
// twice rather than stored: a line is a few hundred samples at most, and the walk is a handful of flops each.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <cstring>
#include <string>
#include <vector>

#include "slimrt_context.h"
#include "slimrt_overlay.h"
#include "../../include/slimrt.h"

using namespace slimrt;

namespace {

constexpr uint32_t kOvBlock = 256;

OverlayView make_view(const slimrt_overlay_view *p) {
    OverlayView v;
    float r[9];
    slimrt_host_camera_rotation(p->camera_euler, r);
    v.cam_x = v3{r[0], r[1], r[2]};
    v.cam_y = v3{r[3], r[4], r[5]};
    v.cam_z = v3{r[6], r[7], r[8]};
    v.cam_position = v3{p->camera_position[0], p->camera_position[1], p->camera_position[2]};
    v.focal_length = p->focal_length;
    v.near_distance = p->near_distance;
    v.far_distance = p->far_distance;
    // Dimensions::update (base.h:1312-1323) and Frustum::Projection::update (viewport/frustum.h:25-27)
    const uint16_t w = (uint16_t)p->width, h = (uint16_t)p->height;
    v.width = w;
    v.height = h;
    v.f_width = (float)w;
    v.f_height = (float)h;
    v.h_width = v.f_width * 0.5f;
    v.h_height = v.f_height * 0.5f;
    v.width_over_height = v.f_width / v.f_height;
    const float height_over_width = v.f_height / v.f_width;
    v.proj_x = p->focal_length * height_over_width;
    v.proj_y = p->focal_length;
    v.ssaa = p->ssaa ? 1u : 0u;
    return v;
}

SLIM_HD quat to_quat(const slimrt_overlay_transform &t) {
    return quat{v3{t.orientation[0], t.orientation[1], t.orientation[2]}, t.orientation[3]};
}
SLIM_HD v3 f3(const float *p) { return v3{p[0], p[1], p[2]}; }

// One edge of one box as a canvas line (drawBox: Transform::externPos of the corners, then drawEdge).
SLIM_HD bool box_line(const OverlayView &v, const slimrt_overlay_box &b, int edge, OverlayLine &l) {
    int a, c;
    box_edge_verts(edge, a, c);
    const v3 mn = f3(b.min), mx = f3(b.max);
    const quat q = to_quat(b.transform);
    const v3 pos = f3(b.transform.position), sc = f3(b.transform.scale);
    l.col = color{b.color[0], b.color[1], b.color[2]};
    l.opacity = b.opacity;
    l.line_width = b.line_width & 0xFFu;
    return edge_line(v, transform_extern(q, pos, sc, box_corner(mn, mx, a)),
                     transform_extern(q, pos, sc, box_corner(mn, mx, c)), l);
}

struct MeshDraw {
    slimrt_overlay_transform t;
    color col;
    float opacity;
    uint32_t line_width;
};

SLIM_HD bool mesh_line(const OverlayView &v, const float *positions, const uint32_t *edges, uint32_t i,
                       const MeshDraw &m, OverlayLine &l) {
    const quat q = to_quat(m.t);
    const v3 pos = f3(m.t.position), sc = f3(m.t.scale);
    l.col = m.col;
    l.opacity = m.opacity;
    l.line_width = m.line_width & 0xFFu;
    return edge_line(v, transform_extern(q, pos, sc, f3(positions + 3 * (size_t)edges[2 * (size_t)i])),
                     transform_extern(q, pos, sc, f3(positions + 3 * (size_t)edges[2 * (size_t)i + 1])), l);
}

// ------------------------------------------------------------------------------------------------ CPU context
struct BlendPut {   // Canvas::setPixel straight into the canvas
    const OverlayView *v;
    float4 *pixels;
    float *depths;
    color col;
    SLIM_HD void operator()(int32_t x, int32_t y, float opacity, float depth) {
        const int64_t o = overlay_offset(*v, x, y);
        if (o >= 0) blend_sample(pixels[o], depths[o], col, opacity, depth);
    }
};

void cpu_draw_line(const OverlayView &v, const OverlayLine &l, float4 *pixels, float *depths) {
    BlendPut put{&v, pixels, depths, l.col};
    walk_line(v, l, put);
}

// ------------------------------------------------------------------------------------------------ device kernels
// Line i of the boxes: src[i] = box index << 4 | edge (the edges drawBox draws, in order).
__global__ __launch_bounds__(kOvBlock) void box_lines_kernel(OverlayView v, const slimrt_overlay_box *__restrict__ boxes,
                                                             const uint32_t *__restrict__ src, uint32_t n,
                                                             OverlayLine *__restrict__ lines, uint32_t *__restrict__ valid) {
    const uint32_t i = blockIdx.x * kOvBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = src[i];
    OverlayLine l;
    valid[i] = box_line(v, boxes[s >> 4], (int)(s & 15u), l) ? 1u : 0u;
    lines[i] = l;
}

__global__ __launch_bounds__(kOvBlock) void mesh_lines_kernel(OverlayView v, const float *__restrict__ positions,
                                                              const uint32_t *__restrict__ edges, uint32_t n, MeshDraw m,
                                                              OverlayLine *__restrict__ lines, uint32_t *__restrict__ valid) {
    const uint32_t i = blockIdx.x * kOvBlock + threadIdx.x;
    if (i >= n) return;
    OverlayLine l;
    valid[i] = mesh_line(v, positions, edges, i, m, l) ? 1u : 0u;
    lines[i] = l;
}

struct CountPut {
    const OverlayView *v;
    uint32_t n;
    SLIM_HD void operator()(int32_t x, int32_t y, float, float) { n += overlay_offset(*v, x, y) >= 0 ? 1u : 0u; }
};

// counts[i]: the canvas samples line i draws (0 for a culled line).
__global__ __launch_bounds__(kOvBlock) void count_kernel(OverlayView v, const OverlayLine *__restrict__ lines,
                                                         const uint32_t *__restrict__ valid, uint32_t n,
                                                         uint32_t *__restrict__ counts) {
    const uint32_t i = blockIdx.x * kOvBlock + threadIdx.x;
    if (i >= n) return;
    CountPut put{&v, 0u};
    if (valid[i]) walk_line(v, lines[i], put);
    counts[i] = put.n;
}

struct EmitPut {
    const OverlayView *v;
    uint32_t next, line;
    uint32_t *keys, *order, *line_of;
    float *opacity, *depth;
    SLIM_HD void operator()(int32_t x, int32_t y, float o, float z) {
        const int64_t off = overlay_offset(*v, x, y);
        if (off < 0) return;
        const uint32_t k = next++;
        keys[k] = (uint32_t)off;
        order[k] = k;
        line_of[k] = line;
        opacity[k] = o;
        depth[k] = z;
    }
};

// Writes line i's samples at first[i] onwards, in the order the reference draws them.
__global__ __launch_bounds__(kOvBlock) void emit_kernel(OverlayView v, const OverlayLine *__restrict__ lines,
                                                        const uint32_t *__restrict__ valid, uint32_t n,
                                                        const uint32_t *__restrict__ first, uint32_t *__restrict__ keys,
                                                        uint32_t *__restrict__ order, uint32_t *__restrict__ line_of,
                                                        float *__restrict__ opacity, float *__restrict__ depth) {
    const uint32_t i = blockIdx.x * kOvBlock + threadIdx.x;
    if (i >= n || !valid[i]) return;
    EmitPut put{&v, first[i], i, keys, order, line_of, opacity, depth};
    walk_line(v, lines[i], put);
}

// After the stable sort by canvas offset: the first sample of each run of one offset blends the whole run, in
// drawing order, into that canvas sample.
__global__ __launch_bounds__(kOvBlock) void blend_kernel(const OverlayLine *__restrict__ lines,
                                                         const uint32_t *__restrict__ keys,
                                                         const uint32_t *__restrict__ order, uint32_t n,
                                                         const uint32_t *__restrict__ line_of,
                                                         const float *__restrict__ opacity,
                                                         const float *__restrict__ depth, float4 *__restrict__ pixels,
                                                         float *__restrict__ depths) {
    const uint32_t i = blockIdx.x * kOvBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = keys[i];
    if (i > 0 && keys[i - 1] == key) return;
    float4 px = pixels[key];
    float d = depths[key];
    for (uint32_t j = i; j < n && keys[j] == key; j++) {
        const uint32_t f = order[j];
        blend_sample(px, d, lines[line_of[f]].col, opacity[f], depth[f]);
    }
    pixels[key] = px;
    depths[key] = d;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Grows a context buffer to `bytes` (the stream is drained first: earlier overlay launches may still read it).
int reserve(slimrt_context *c, DevBuf<uint8_t> &b, size_t bytes) {
    if (b.cap >= bytes) return SLIMRT_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (b.ptr) HIP_TRY(c, ctx_free(c, b.ptr));
    b.ptr = nullptr;
    b.cap = 0;
    HIP_TRY(c, ctx_malloc(c, &b.ptr, bytes));
    b.cap = bytes;
    return SLIMRT_OK;
}

struct Carver {
    uint8_t *base;
    size_t used;
    template <class T>
    T *take(size_t n) {
        T *p = (T *)(base + used);
        used += align256(n * sizeof(T));
        return p;
    }
};

// Steps 2-6 for `n` lines already in `lines` / `valid` (in the ov_lines buffer, after `used` bytes of it).
int device_draw_lines(slimrt_context *c, const OverlayView &v, OverlayLine *lines, uint32_t *valid, uint32_t n,
                      size_t used, float4 *pixels, float *depths) {
    if (n == 0) return SLIMRT_OK;
    const uint32_t grid = (n + kOvBlock - 1) / kOvBlock;
    size_t scan_tmp = 0;
    HIP_TRY(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp, (uint32_t *)nullptr, (uint32_t *)nullptr, (int)n,
                                                c->stream));
    const size_t need = used + 2 * align256((size_t)n * 4) + align256(scan_tmp) + 256;
    if (need > c->ov_lines.cap) return fail(c, SLIMRT_E_STATE, "overlay: line buffer sized too small");
    Carver cv{c->ov_lines.ptr, used};
    uint32_t *counts = cv.take<uint32_t>(n);
    uint32_t *first = cv.take<uint32_t>(n);
    void *tmp = cv.take<uint8_t>(scan_tmp);
    hipLaunchKernelGGL(count_kernel, dim3(grid), dim3(kOvBlock), 0, c->stream, v, lines, valid, n, counts);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipcub::DeviceScan::ExclusiveSum(tmp, scan_tmp, counts, first, (int)n, c->stream));
    uint32_t last[2];
    HIP_TRY(c, hipMemcpyAsync(&last[0], first + n - 1, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(&last[1], counts + n - 1, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint64_t total64 = (uint64_t)last[0] + last[1];
    if (total64 > 0x7FFFFFFFull) return fail(c, SLIMRT_E_ARG, "overlay: more than 2^31 line samples");
    const uint32_t total = (uint32_t)total64;
    if (total == 0) return SLIMRT_OK;

    const uint64_t samples = (uint64_t)v.width * v.height * (v.ssaa ? 4u : 1u);
    int end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) < samples) end_bit++;
    size_t sort_tmp = 0;
    HIP_TRY(c, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_tmp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                  (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)total, 0, end_bit,
                                                  c->stream));
    const size_t per = align256((size_t)total * 4);
    int rc = reserve(c, c->ov_samples, 7 * per + align256(sort_tmp) + 256);
    if (rc) return rc;
    Carver sv{c->ov_samples.ptr, 0};
    uint32_t *keys = sv.take<uint32_t>(total), *keys_sorted = sv.take<uint32_t>(total);
    uint32_t *order = sv.take<uint32_t>(total), *order_sorted = sv.take<uint32_t>(total);
    uint32_t *line_of = sv.take<uint32_t>(total);
    float *opacity = sv.take<float>(total), *depth = sv.take<float>(total);
    void *stmp = sv.take<uint8_t>(sort_tmp);
    hipLaunchKernelGGL(emit_kernel, dim3(grid), dim3(kOvBlock), 0, c->stream, v, lines, valid, n, first, keys, order,
                       line_of, opacity, depth);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipcub::DeviceRadixSort::SortPairs(stmp, sort_tmp, keys, keys_sorted, order, order_sorted, (int)total, 0,
                                                  end_bit, c->stream));
    hipLaunchKernelGGL(blend_kernel, dim3((total + kOvBlock - 1) / kOvBlock), dim3(kOvBlock), 0, c->stream, lines,
                       keys_sorted, order_sorted, total, line_of, opacity, depth, pixels, depths);
    HIP_TRY(c, hipGetLastError());
    return SLIMRT_OK;
}

int check_view(slimrt_context *c, const slimrt_overlay_view *view, const void *pixels, const void *depths) {
    if (!view || !pixels || !depths) return fail(c, SLIMRT_E_ARG, "overlay: NULL view or canvas");
    if (!view->width || !view->height || view->width > 65535 || view->height > 65535 ||
        (uint64_t)view->width * view->height * (view->ssaa ? 4u : 1u) > 0xFFFFFFFFull)
        return fail(c, SLIMRT_E_ARG, "overlay: canvas dimensions out of range");
    return SLIMRT_OK;
}

}  // namespace

extern "C" {

int slimrt_draw_boxes(slimrt_context *c, const slimrt_overlay_view *view, const slimrt_overlay_box *boxes,
                      uint32_t count, void *pixels, void *depths) {
    if (!c) return SLIMRT_E_ARG;
    int rc = check_view(c, view, pixels, depths);
    if (rc) return rc;
    if (count && !boxes) return fail(c, SLIMRT_E_ARG, "overlay: NULL boxes");
    if (count > (1u << 27)) return fail(c, SLIMRT_E_ARG, "overlay: too many boxes");
    const OverlayView v = make_view(view);
    std::vector<uint32_t> src;   // the edges each box draws, in drawing order
    src.reserve((size_t)count * 12);
    for (uint32_t b = 0; b < count; b++) {
        int e[12];
        const int k = box_edges(boxes[b].sides, e);
        for (int i = 0; i < k; i++) src.push_back(b << 4 | (uint32_t)e[i]);
    }
    const uint32_t n = (uint32_t)src.size();
    if (c->cpu) {
        for (uint32_t i = 0; i < n; i++) {
            OverlayLine l;
            if (box_line(v, boxes[src[i] >> 4], (int)(src[i] & 15u), l))
                cpu_draw_line(v, l, (float4 *)pixels, (float *)depths);
        }
        return SLIMRT_OK;
    }
    if (n == 0) return SLIMRT_OK;
    size_t scan_tmp = 0;
    HIP_TRY(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp, (uint32_t *)nullptr, (uint32_t *)nullptr, (int)n,
                                                c->stream));
    const size_t inputs = align256((size_t)count * sizeof(slimrt_overlay_box)) + align256((size_t)n * 4);
    const size_t line_bytes = align256((size_t)n * sizeof(OverlayLine)) + align256((size_t)n * 4);
    if ((rc = reserve(c, c->ov_lines, inputs + line_bytes + 2 * align256((size_t)n * 4) + align256(scan_tmp) + 256)))
        return rc;
    Carver cv{c->ov_lines.ptr, 0};
    slimrt_overlay_box *d_boxes = cv.take<slimrt_overlay_box>(count);
    uint32_t *d_src = cv.take<uint32_t>(n);
    OverlayLine *lines = cv.take<OverlayLine>(n);
    uint32_t *valid = cv.take<uint32_t>(n);
    HIP_TRY(c, hipMemcpyAsync(d_boxes, boxes, (size_t)count * sizeof(slimrt_overlay_box), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d_src, src.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(box_lines_kernel, dim3((n + kOvBlock - 1) / kOvBlock), dim3(kOvBlock), 0, c->stream, v, d_boxes,
                       d_src, n, lines, valid);
    HIP_TRY(c, hipGetLastError());
    return device_draw_lines(c, v, lines, valid, n, cv.used, (float4 *)pixels, (float *)depths);
}

int slimrt_draw_mesh_wireframe(slimrt_context *c, const slimrt_overlay_view *view, const float *positions,
                               uint32_t vertex_count, const uint32_t *edges, uint32_t edge_count,
                               const slimrt_overlay_transform *transform, const float color_[3], float opacity,
                               uint32_t line_width, void *pixels, void *depths) {
    if (!c) return SLIMRT_E_ARG;
    int rc = check_view(c, view, pixels, depths);
    if (rc) return rc;
    if (!transform || !color_ || (edge_count && (!positions || !edges)))
        return fail(c, SLIMRT_E_ARG, "overlay: NULL mesh arrays, transform or colour");
    if (edge_count > (1u << 28)) return fail(c, SLIMRT_E_ARG, "overlay: too many edges");
    for (uint32_t i = 0; i < 2 * edge_count; i++)
        if (edges[i] >= vertex_count) return fail(c, SLIMRT_E_ARG, "overlay: edge vertex index out of range");
    const OverlayView v = make_view(view);
    MeshDraw m{*transform, color{color_[0], color_[1], color_[2]}, opacity, line_width};
    const uint32_t n = edge_count;
    if (c->cpu) {
        for (uint32_t i = 0; i < n; i++) {
            OverlayLine l;
            if (mesh_line(v, positions, edges, i, m, l)) cpu_draw_line(v, l, (float4 *)pixels, (float *)depths);
        }
        return SLIMRT_OK;
    }
    if (n == 0) return SLIMRT_OK;
    size_t scan_tmp = 0;
    HIP_TRY(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp, (uint32_t *)nullptr, (uint32_t *)nullptr, (int)n,
                                                c->stream));
    const size_t inputs = align256((size_t)vertex_count * 12) + align256((size_t)n * 8);
    const size_t line_bytes = align256((size_t)n * sizeof(OverlayLine)) + align256((size_t)n * 4);
    if ((rc = reserve(c, c->ov_lines, inputs + line_bytes + 2 * align256((size_t)n * 4) + align256(scan_tmp) + 256)))
        return rc;
    Carver cv{c->ov_lines.ptr, 0};
    float *d_pos = cv.take<float>((size_t)vertex_count * 3);
    uint32_t *d_edges = cv.take<uint32_t>((size_t)n * 2);
    OverlayLine *lines = cv.take<OverlayLine>(n);
    uint32_t *valid = cv.take<uint32_t>(n);
    HIP_TRY(c, hipMemcpyAsync(d_pos, positions, (size_t)vertex_count * 12, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d_edges, edges, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(mesh_lines_kernel, dim3((n + kOvBlock - 1) / kOvBlock), dim3(kOvBlock), 0, c->stream, v, d_pos,
                       d_edges, n, m, lines, valid);
    HIP_TRY(c, hipGetLastError());
    return device_draw_lines(c, v, lines, valid, n, cv.used, (float4 *)pixels, (float *)depths);
}

}  // extern "C"
