// Host-side scene derivation: orientations, camera projection, world AABBs, SAH BVH builds and triangle records.
//
// These run on the CPU exactly where the reference runs them (Scene::updateAABBs / updateBVH every frame,
// BVHBuilder::buildMesh once per mesh), and must reproduce the reference bit for bit: BVH topology and node
// bounds are part of the parity contract.  Each function states the reference code it follows; floating-point
// expressions keep the reference's operation order (compiled without FMA contraction).
#include <math.h>
#include <string.h>
#include <stdint.h>
#include <vector>
#include <algorithm>

#include "../../include/slimrt.h"

namespace {

struct f3 { float x, y, z; };
inline f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 mul(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 lo(f3 a, f3 b) { return {a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z}; }
inline f3 hi(f3 a, f3 b) { return {a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z}; }
inline f3 cross(f3 a, f3 b) { return {(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)}; }
inline float sqlen(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline f3 scale_add(f3 a, float f, f3 b) { return {a.x * f + b.x, a.y * f + b.y, a.z * f + b.z}; }

struct q4 { f3 a; float w; };
// quat * vec3 (src/slim/math/quat.h:37-43)
inline f3 rotate(q4 q, f3 v) {
    f3 out = cross(q.a, v);
    f3 qqv = cross(q.a, out);
    out = scale_add(out, q.w, qqv);
    return scale_add(out, 2.0f, v);
}
// quat::AxisAngle (quat.h:78-84) followed by normalized() (quat.h:18-24, 33-35)
inline q4 axis_angle(f3 axis, float radians) {
    radians *= 0.5f;
    q4 q{mul(axis, sinf(radians)), cosf(radians)};
    float len = sqrtf(q.a.x * q.a.x + q.a.y * q.a.y + q.a.z * q.a.z + q.w * q.w);
    float f = 1.0f / len;
    return {mul(q.a, f), q.w * f};
}
// quat::operator*= (quat.h:56-65): this = this * rhs
inline q4 compose(q4 l, q4 r) {
    float x = l.a.x, X = r.a.x, y = l.a.y, Y = r.a.y, z = l.a.z, Z = r.a.z, w = l.w, W = r.w;
    return {{X * w + W * x + Z * y - Y * z, Y * w - Z * x + W * y + X * z, Z * w + Y * x - X * y + W * z},
            W * w - X * x - Y * y - Z * z};
}

struct m33 { f3 X, Y, Z; };
inline m33 mat_mul(const m33 &l, const m33 &r) {   // mat3::operator*= (mat3.h:272-285)
    m33 o;
    o.X = {l.X.x * r.X.x + l.X.y * r.Y.x + l.X.z * r.Z.x, l.X.x * r.X.y + l.X.y * r.Y.y + l.X.z * r.Z.y,
           l.X.x * r.X.z + l.X.y * r.Y.z + l.X.z * r.Z.z};
    o.Y = {l.Y.x * r.X.x + l.Y.y * r.Y.x + l.Y.z * r.Z.x, l.Y.x * r.X.y + l.Y.y * r.Y.y + l.Y.z * r.Z.y,
           l.Y.x * r.X.z + l.Y.y * r.Y.z + l.Y.z * r.Z.z};
    o.Z = {l.Z.x * r.X.x + l.Z.y * r.Y.x + l.Z.z * r.Z.x, l.Z.x * r.X.y + l.Z.y * r.Y.y + l.Z.z * r.Z.y,
           l.Z.x * r.X.z + l.Z.y * r.Y.z + l.Z.z * r.Z.z};
    return o;
}
inline m33 inverted(const m33 &m) {   // mat3::inverted (mat3.h:146-178): cofactors times 1/det
    const f3 X = m.X, Y = m.Y, Z = m.Z;
    float det = +X.x * (Y.y * Z.z - Z.y * Y.z) - Y.x * (X.y * Z.z - Z.y * X.z) + Z.x * (X.y * Y.z - Y.y * X.z);
    float f = 1.0f / det;
    m33 c{{+(Y.y * Z.z - Z.y * Y.z), -(X.y * Z.z - Z.y * X.z), +(X.y * Y.z - Y.y * X.z)},
          {-(Y.x * Z.z - Z.x * Y.z), +(X.x * Z.z - Z.x * X.z), -(X.x * Y.z - Y.x * X.z)},
          {+(Y.x * Z.y - Z.x * Y.y), -(X.x * Z.y - Z.x * X.y), +(X.x * Y.y - Y.x * X.y)}};
    return {mul(c.X, f), mul(c.Y, f), mul(c.Z, f)};
}

struct Box { f3 mn, mx; };
inline Box merge(const Box &a, const Box &b) { return {lo(a.mn, b.mn), hi(a.mx, b.mx)}; }   // AABB::operator+ (vec3.h:471)
inline float area(const Box &b) {   // AABB::area (vec3.h:512-515)
    f3 e = sub(b.mx, b.mn);
    return e.x * e.y + e.y * e.z + e.z * e.x;
}
inline float box_max(const Box &b, int axis) { return axis == 0 ? b.mx.x : axis == 1 ? b.mx.y : b.mx.z; }

// The SAH builder of src/slim/scene/bvh_builder.h:10-263, reorganised around std::vector scratch.  Its output is
// defined by the reference's exact procedure (iterative Lomuto quicksort per axis, prefix/suffix area sweep,
// first strict minimum, right subtree built first), so that procedure is kept step for step.
class SahBuilder {
public:
    explicit SahBuilder(uint32_t n) : boxes_(n), ids_(n), leaf_ids_(n), sort_stack_(2 * n + 2) {
        for (auto &ax : axes_) {
            ax.sorted.resize(n);
            ax.left.resize(n);
            ax.right.resize(n);
            ax.left_area.resize(n);
            ax.right_area.resize(n);
        }
    }
    std::vector<Box> boxes_;          // one per primitive
    std::vector<uint32_t> ids_;       // primitive order being partitioned
    std::vector<uint32_t> leaf_ids_;  // output: leaf order -> primitive

    // BVHBuilder::build (bvh_builder.h:195-263)
    void build(uint32_t N, uint32_t max_leaf, slimrt_bvh_node *out, uint32_t &count, uint32_t &height) {
        height = 1;
        count = 1;
        memset(&out[0], 0, sizeof(out[0]));
        if (N <= max_leaf) {
            Box b{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
            for (uint32_t i = 0; i < N; i++) {
                leaf_ids_[i] = i;
                b = merge(b, boxes_[i]);
            }
            set_box(out[0], b);
            out[0].leaf_count = (uint16_t)N;
            return;
        }
        struct Task { uint32_t start, end, node, depth; };
        std::vector<Task> stack;
        stack.reserve(64);
        uint32_t middle = split(out, count, 0, 0, N);
        stack.push_back({0, middle, 1, 1});
        stack.push_back({middle, N, 2, 1});
        uint32_t leaves = 0;
        while (!stack.empty()) {
            Task t = stack.back();
            uint32_t n = t.end - t.start;
            if (n <= max_leaf) {
                out[t.node].depth = (uint8_t)t.depth;
                out[t.node].leaf_count = (uint16_t)n;
                out[t.node].first_index = leaves;
                for (uint32_t i = 0; i < n; i++) leaf_ids_[leaves + i] = ids_[t.start + i];
                leaves += n;
                stack.pop_back();
            } else {
                middle = split(out, count, t.node, t.start, t.end);
                uint32_t d = t.depth + 1;
                uint32_t first = out[t.node].first_index;
                out[first].depth = out[first + 1].depth = (uint8_t)d;
                stack.back() = {t.start, middle, first, d};
                stack.push_back({middle, t.end, first + 1, d});
                if (d > height) height = d;
            }
        }
        out[1].depth = out[2].depth = 1;
        Box root = merge(get_box(out[1]), get_box(out[2]));
        set_box(out[0], root);
    }

private:
    struct Axis {
        std::vector<uint32_t> sorted;
        std::vector<Box> left, right;
        std::vector<float> left_area, right_area;
        uint32_t left_count = 1;
        float cost = INFINITY;
    };
    Axis axes_[3];
    std::vector<int32_t> sort_stack_;

    static void set_box(slimrt_bvh_node &n, const Box &b) {
        n.min[0] = b.mn.x; n.min[1] = b.mn.y; n.min[2] = b.mn.z;
        n.max[0] = b.mx.x; n.max[1] = b.mx.y; n.max[2] = b.mx.z;
    }
    static Box get_box(const slimrt_bvh_node &n) { return {{n.min[0], n.min[1], n.min[2]}, {n.max[0], n.max[1], n.max[2]}}; }

    // What the quicksort below does to a range whose keys never decrease (none NaN), in one pass: with the pivot the
    // last element, every other element is either less (kept in place) or equal to it, so each step moves the
    // pivot to the front of its run of equal keys and that run's first element to the run's end, then continues
    // on the rest of the run and on the range before it.  The result is every maximal run of equal keys rotated
    // right by one.  Ranges arrive this way whenever a split hands its children down sorted along the chosen axis,
    // where the step-by-step procedure takes time quadratic in the range (a 262,144-triangle mesh: 37 s).
    // Returns false, leaving ids unchanged, when the range is not non-decreasing.
    bool sorted_range_result(uint32_t *ids, int32_t start, int32_t end, int axis) const {
        float prev = box_max(boxes_[ids[start]], axis);
        if (prev != prev) return false;
        for (int32_t i = start + 1; i <= end; i++) {
            const float k = box_max(boxes_[ids[i]], axis);
            if (k != k || k < prev) return false;
            prev = k;
        }
        for (int32_t run = start; run <= end;) {
            const float k = box_max(boxes_[ids[run]], axis);
            int32_t last = run;
            while (last < end && !(k < box_max(boxes_[ids[last + 1]], axis))) last++;
            if (last > run) std::rotate(ids + run, ids + last, ids + last + 1);
            run = last + 1;
        }
        return true;
    }

    // BVHPartition::partition (bvh_builder.h:15-113)
    void partition(Axis &ax, int axis, uint32_t N) {
        uint32_t *ids = ax.sorted.data();
        int32_t *st = sort_stack_.data();
        int32_t top = -1;
        st[++top] = 0;
        st[++top] = (int32_t)N - 1;
        while (top >= 0) {
            int32_t end = st[top--], start = st[top--], p = start;
            if (end - start >= 16 && sorted_range_result(ids, start, end, axis)) continue;
            const float pivot = box_max(boxes_[ids[end]], axis);
            for (int32_t i = start; i < end; i++)
                if (box_max(boxes_[ids[i]], axis) < pivot) std::swap(ids[i], ids[p++]);
            std::swap(ids[end], ids[p]);
            if (p - 1 > start) { st[++top] = start; st[++top] = p - 1; }
            if (p + 1 < end) { st[++top] = p + 1; st[++top] = end; }
        }
        Box L = boxes_[ids[0]], R = boxes_[ids[N - 1]];
        ax.left[0] = L;
        ax.right[N - 1] = R;
        for (uint32_t i = 0, j = N - 1; i < N; i++, j--) {
            if (i) {
                L = merge(L, boxes_[ids[i]]);
                R = merge(R, boxes_[ids[j]]);
                ax.left[i] = L;
                ax.right[j] = R;
            }
            ax.left_area[i] = area(L);
            ax.right_area[j] = area(R);
        }
        ax.left_count = 1;
        ax.cost = INFINITY;
        for (uint32_t cur = 0, nl = 1, nr = N - 1; cur < N - 1; cur++, nl++, nr--) {
            float c = ax.left_area[cur] * (float)nl + ax.right_area[cur + 1] * (float)nr;
            if (ax.cost > c) { ax.cost = c; ax.left_count = cur + 1; }
        }
    }

    // BVHBuilder::splitNode (bvh_builder.h:159-193)
    uint32_t split(slimrt_bvh_node *out, uint32_t &count, uint32_t node, uint32_t start, uint32_t end) {
        uint32_t N = end - start;
        out[node].first_index = count;
        slimrt_bvh_node &l = out[count++];
        slimrt_bvh_node &r = out[count++];
        memset(&l, 0, sizeof l);
        memset(&r, 0, sizeof r);
        float best = INFINITY;
        int chosen = 0;
        for (int a = 0; a < 3; a++) {
            std::copy(ids_.begin() + start, ids_.begin() + end, axes_[a].sorted.begin());
            partition(axes_[a], a, N);
            if (axes_[a].cost < best) { best = axes_[a].cost; chosen = a; }
        }
        Axis &ax = axes_[chosen];
        set_box(l, ax.left[ax.left_count - 1]);
        set_box(r, ax.right[ax.left_count]);
        std::copy(ax.sorted.begin(), ax.sorted.begin() + N, ids_.begin() + start);
        return start + ax.left_count;
    }
};

inline f3 load3(const float *p, uint32_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

}  // namespace

extern "C" {

int slimrt_host_orientation(const float euler[3], float q_out[4]) {
    if (!euler || !q_out) return SLIMRT_E_ARG;
    // Orientation<quat>::_update (src/slim/core/base.h:612-620)
    q4 q = euler[2] != 0.0f ? axis_angle({0, 0, 1}, euler[2]) : q4{{0, 0, 0}, 1.0f};
    if (euler[0] != 0.0f) q = compose(q, axis_angle({1, 0, 0}, euler[0]));
    if (euler[1] != 0.0f) q = compose(q, axis_angle({0, 1, 0}, euler[1]));
    q_out[0] = q.a.x; q_out[1] = q.a.y; q_out[2] = q.a.z; q_out[3] = q.w;
    return SLIMRT_OK;
}

// OrientationUsing3x3Matrix::setRotation (base.h:568-620, mat3.h:49-77); cos/sin of a float argument are evaluated
// in double there, as unqualified cos/sin resolve to the C library's double functions.
static m33 camera_matrix(const float e[3]) {
    m33 o{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (e[2] != 0.0f) {
        float c = (float)cos((double)e[2]), s = (float)sin((double)e[2]);
        o = {{c, -s, 0}, {s, c, 0}, {0, 0, 1}};
    }
    if (e[0] != 0.0f) {
        float c = (float)cos((double)e[0]), s = (float)sin((double)e[0]);
        o = mat_mul(o, {{1, 0, 0}, {0, c, -s}, {0, s, c}});
    }
    if (e[1] != 0.0f) {
        float c = (float)cos((double)e[1]), s = (float)sin((double)e[1]);
        o = mat_mul(o, {{c, 0, s}, {0, 1, 0}, {-s, 0, c}});
    }
    return o;
}

int slimrt_host_camera_rotation(const float euler[3], float rotation_out[9]) {
    if (!euler || !rotation_out) return SLIMRT_E_ARG;
    const m33 o = camera_matrix(euler);
    memcpy(rotation_out, &o, sizeof(float) * 9);
    return SLIMRT_OK;
}

// drawBVH (draw/bvh.h:8-28): per node in depth range, the -1..1 box under the transform scaled by the node's half
// extent (vec3 *= vec3) and moved to Transform::externPos of its centre.
int slimrt_host_bvh_boxes(const slimrt_bvh_node *nodes, uint32_t node_count, const slimrt_overlay_transform *t,
                          uint32_t min_depth, uint32_t max_depth, float opacity, uint32_t line_width,
                          slimrt_overlay_box *out, uint32_t *count_out) {
    if ((node_count && (!nodes || !out)) || !t || !count_out) return SLIMRT_E_ARG;
    const q4 q{{t->orientation[0], t->orientation[1], t->orientation[2]}, t->orientation[3]};
    const f3 scale{t->scale[0], t->scale[1], t->scale[2]}, pos{t->position[0], t->position[1], t->position[2]};
    const uint16_t lo = (uint16_t)min_depth, hi = (uint16_t)max_depth;   // drawBVH takes u16 depths
    uint32_t n = 0;
    for (uint32_t i = 0; i < node_count; i++) {
        const slimrt_bvh_node &nd = nodes[i];
        if (nd.depth < lo || nd.depth > hi) continue;
        const f3 mn{nd.min[0], nd.min[1], nd.min[2]}, mx{nd.max[0], nd.max[1], nd.max[2]};
        const f3 half = mul(sub(mx, mn), 0.5f);
        const f3 s = mul(scale, half);
        const f3 c = add(rotate(q, mul(scale, mul(add(mn, mx), 0.5f))), pos);
        slimrt_overlay_box &b = out[n++];
        memcpy(b.transform.orientation, t->orientation, sizeof b.transform.orientation);
        b.transform.position[0] = c.x; b.transform.position[1] = c.y; b.transform.position[2] = c.z;
        b.transform.scale[0] = s.x; b.transform.scale[1] = s.y; b.transform.scale[2] = s.z;
        for (int k = 0; k < 3; k++) { b.min[k] = -1.0f; b.max[k] = 1.0f; }
        // BrightMagenta leaves, BrightCyan root, BrightGreen inner nodes (base.h:822-831)
        const float leaf[3] = {0.75f, 0.0f, 0.75f}, root[3] = {0.0f, 0.75f, 0.75f}, inner[3] = {0.5f, 1.0f, 0.5f};
        memcpy(b.color, nd.leaf_count ? leaf : (i ? inner : root), sizeof b.color);
        b.opacity = nd.leaf_count ? opacity * 2.0f : opacity * 0.5f;
        b.line_width = line_width & 0xFFu;
        b.sides = SLIMRT_BOX_ALL_SIDES;
    }
    *count_out = n;
    return SLIMRT_OK;
}

int slimrt_host_projection(const float e[3], const float position[3], float focal_length, uint32_t width,
                           uint32_t height, uint32_t ssaa, slimrt_render_params *p) {
    if (!e || !position || !p || width == 0 || height == 0 || width > 65535 || height > 65535) return SLIMRT_E_ARG;
    const m33 o = camera_matrix(e);
    // Dimensions::update (base.h:1312-1323) and CameraRayProjection::reset (camera.h:69-83)
    const float fw = (float)(uint16_t)width, fh = (float)(uint16_t)height;
    const float hw = fw * 0.5f, hh = fh * 0.5f;
    p->width = width;
    p->height = height;
    p->ssaa = ssaa ? 1u : 0u;
    p->sample_size = ssaa ? 0.5f : 1.0f;
    float d = hh * focal_length;
    p->C_start[0] = (p->sample_size * 0.5f) - hw;
    p->C_start[1] = hh - (p->sample_size * 0.5f);
    m33 inv = inverted(o);
    memcpy(p->inverted_camera_rotation, &inv, sizeof(float) * 9);
    memcpy(p->camera_position, position, sizeof(float) * 3);
    f3 down = mul(f3{-o.Y.x, -o.Y.y, -o.Y.z}, p->sample_size);
    f3 right = mul(o.X, p->sample_size);
    f3 start = add(add(mul(o.X, p->C_start[0]), mul(o.Y, p->C_start[1])), mul(o.Z, d));
    memcpy(p->down, &down, 12);
    memcpy(p->right, &right, 12);
    memcpy(p->start, &start, 12);
    p->squared_distance_to_projection_plane = d * d;
    return SLIMRT_OK;
}

int slimrt_host_world_aabbs(const slimrt_geometry *g, uint32_t count, const slimrt_mesh *meshes, uint32_t mesh_count,
                            float *out) {
    if ((!g && count) || !out) return SLIMRT_E_ARG;
    for (uint32_t i = 0; i < count; i++) {   // Scene::updateAABB (scene.h:165-174) + Transform::externAABB (transform.h:33-48)
        Box b;
        if (g[i].type == 1) {
            if (g[i].id >= mesh_count) return SLIMRT_E_ARG;
            const slimrt_mesh &m = meshes[g[i].id];
            b = {{m.aabb_min[0], m.aabb_min[1], m.aabb_min[2]}, {m.aabb_max[0], m.aabb_max[1], m.aabb_max[2]}};
        } else {
            float e = g[i].type == 5 ? 0.577350259f : 1.0f;
            b = {{-e, -e, -e}, {e, e, e}};
            if (g[i].type == 6) b.mn.y = b.mx.y = 0.0f;
        }
        q4 q{{g[i].orientation[0], g[i].orientation[1], g[i].orientation[2]}, g[i].orientation[3]};
        f3 s{g[i].scale[0], g[i].scale[1], g[i].scale[2]}, t{g[i].position[0], g[i].position[1], g[i].position[2]};
        f3 mn{INFINITY, INFINITY, INFINITY}, mx{-INFINITY, -INFINITY, -INFINITY};
        for (int cx = 0; cx < 2; cx++)
            for (int cy = 0; cy < 2; cy++)
                for (int cz = 0; cz < 2; cz++) {
                    f3 p{cx ? b.mx.x : b.mn.x, cy ? b.mx.y : b.mn.y, cz ? b.mx.z : b.mn.z};
                    p = add(rotate(q, mul(s, p)), t);
                    mn = lo(mn, p);
                    mx = hi(mx, p);
                }
        float v[6] = {mn.x, mn.y, mn.z, mx.x, mx.y, mx.z};
        memcpy(out + 6 * i, v, sizeof v);
    }
    return SLIMRT_OK;
}

int slimrt_host_build_bvh(const float *aabbs, uint32_t count, uint32_t max_leaf, slimrt_bvh_node *nodes,
                          uint32_t *node_count, uint32_t *height, uint32_t *leaf_ids) {
    if (!aabbs || !nodes || !node_count || !height || !leaf_ids || count == 0 || max_leaf == 0) return SLIMRT_E_ARG;
    SahBuilder b(count);
    for (uint32_t i = 0; i < count; i++) {   // Scene::updateBVH (scene.h:181-191)
        b.boxes_[i] = {{aabbs[6 * i], aabbs[6 * i + 1], aabbs[6 * i + 2]}, {aabbs[6 * i + 3], aabbs[6 * i + 4], aabbs[6 * i + 5]}};
        b.ids_[i] = i;
    }
    b.build(count, max_leaf, nodes, *node_count, *height);
    memcpy(leaf_ids, b.leaf_ids_.data(), sizeof(uint32_t) * count);
    return SLIMRT_OK;
}

int slimrt_host_build_mesh(const float *pos, uint32_t V, const uint32_t *pidx, const float *nrm, uint32_t NN,
                           const uint32_t *nidx, const float *uv, uint32_t NU, const uint32_t *uidx, uint32_t T,
                           uint32_t max_leaf, slimrt_bvh_node *nodes, uint32_t *node_count, uint32_t *height,
                           slimrt_triangle *tris, float aabb_out[6]) {
    if (!pos || !pidx || !nodes || !node_count || !height || !tris || !aabb_out || T == 0 || max_leaf == 0)
        return SLIMRT_E_ARG;
    if ((NN && (!nrm || !nidx)) || (NU && (!uv || !uidx))) return SLIMRT_E_ARG;
    for (uint32_t i = 0; i < 3 * T; i++)
        if (pidx[i] >= V || (NN && nidx[i] >= NN) || (NU && uidx[i] >= NU)) return SLIMRT_E_ARG;
    f3 mn{INFINITY, INFINITY, INFINITY}, mx{-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v = 0; v < V; v++) { mn = lo(mn, load3(pos, v)); mx = hi(mx, load3(pos, v)); }
    float ab[6] = {mn.x, mn.y, mn.z, mx.x, mx.y, mx.z};
    memcpy(aabb_out, ab, sizeof ab);
    const float eps = 0.0001f;
    SahBuilder b(T);
    for (uint32_t i = 0; i < T; i++) {   // bvh_builder.h:269-303: triangle boxes, padded on flat axes
        f3 a = load3(pos, pidx[3 * i]), c = load3(pos, pidx[3 * i + 1]), d = load3(pos, pidx[3 * i + 2]);
        f3 lo3 = lo(lo(a, c), d), hi3 = hi(hi(a, c), d);
        float *l = &lo3.x, *h = &hi3.x;
        for (int k = 0; k < 3; k++) {
            float diff = h[k] - l[k];
            if (diff < 0) diff = -diff;
            if (diff < eps) { l[k] -= eps; h[k] += eps; }
        }
        b.boxes_[i] = {lo3, hi3};
        b.ids_[i] = i;
    }
    b.build(T, max_leaf, nodes, *node_count, *height);
    for (uint32_t i = 0; i < T; i++) {   // bvh_builder.h:306-340: triangle records in leaf order
        const uint32_t t = b.leaf_ids_[i];
        f3 v1 = load3(pos, pidx[3 * t]), v2 = load3(pos, pidx[3 * t + 1]), v3 = load3(pos, pidx[3 * t + 2]);
        m33 m{sub(v3, v1), sub(v2, v1), {0, 0, 0}};
        m.Z = cross(m.X, m.Y);
        float area_par = sqrtf(sqlen(m.Z));
        float f = 1.0f / area_par;
        m.Z = mul(m.Z, f);
        slimrt_triangle &tr = tris[i];
        memset(&tr, 0, sizeof tr);
        memcpy(tr.normal, &m.Z, 12);
        memcpy(tr.position, &v1, 12);
        m33 inv = inverted(m);
        memcpy(tr.local_to_tangent, &inv, 36);
        tr.uv_coverage = 1.0f;
        if (NN) {
            memcpy(tr.n1, nrm + 3 * nidx[3 * t], 12);
            memcpy(tr.n2, nrm + 3 * nidx[3 * t + 1], 12);
            memcpy(tr.n3, nrm + 3 * nidx[3 * t + 2], 12);
        }
        if (NU) {
            memcpy(tr.uv1, uv + 2 * uidx[3 * t], 8);
            memcpy(tr.uv2, uv + 2 * uidx[3 * t + 1], 8);
            memcpy(tr.uv3, uv + 2 * uidx[3 * t + 2], 8);
            float area_uv = (tr.uv2[0] - tr.uv1[0]) * (tr.uv3[1] - tr.uv1[1]) - (tr.uv3[0] - tr.uv1[0]) * (tr.uv2[1] - tr.uv1[1]);
            tr.uv_coverage = fabsf(area_uv / area_par);
        }
    }
    return SLIMRT_OK;
}

}  // extern "C"
