// Scene-level geometry work shared by the host and the device: a geometry's world AABB and the SAH build of
// the scene BVH over those boxes, written over caller-provided scratch so that one thread of a kernel can run
// it from LDS.  Both follow the reference step for step (their results are compared bit-for-bit with the host
// builder, which is itself compared with the reference):
//   world AABB   Scene::updateAABB (src/slim/scene/scene.h:165-179) + Transform::externAABB (transform.h:33-48)
//   SAH build    BVHBuilder::build / splitNode (src/slim/scene/bvh_builder.h:159-263) with
//                BVHPartition::partition (bvh_builder.h:15-113): iterative Lomuto quicksort of the ids per axis
//                by aabb.max, prefix/suffix area sweep, first strict minimum of A_L n_L + A_R n_R, first axis on
//                ties, children appended as a pair, right subtree processed first.
#pragma once

#include "slimrt_core.h"
#include "../../include/slimrt.h"

namespace slimrt {

struct SBox { float mn[3], mx[3]; };

SLIM_HD SBox sbox_merge(const SBox &a, const SBox &b) {   // AABB::operator+ (vec3.h:471)
    SBox r;
    for (int k = 0; k < 3; k++) {
        r.mn[k] = a.mn[k] < b.mn[k] ? a.mn[k] : b.mn[k];
        r.mx[k] = a.mx[k] > b.mx[k] ? a.mx[k] : b.mx[k];
    }
    return r;
}

SLIM_HD float sbox_area(const SBox &b) {   // AABB::area (vec3.h:512-515)
    const float ex = b.mx[0] - b.mn[0], ey = b.mx[1] - b.mn[1], ez = b.mx[2] - b.mn[2];
    return ex * ey + ey * ez + ez * ex;
}

// quat * vec3 (src/slim/math/quat.h:37-43): out = a x v; out = out * w + a x out; out * 2 + v
SLIM_HD void quat_rotate(const float q[4], const float v[3], float out[3]) {
    const float o0 = (q[1] * v[2]) - (q[2] * v[1]), o1 = (q[2] * v[0]) - (q[0] * v[2]), o2 = (q[0] * v[1]) - (q[1] * v[0]);
    const float c0 = (q[1] * o2) - (q[2] * o1), c1 = (q[2] * o0) - (q[0] * o2), c2 = (q[0] * o1) - (q[1] * o0);
    const float s0 = o0 * q[3] + c0, s1 = o1 * q[3] + c1, s2 = o2 * q[3] + c2;
    out[0] = s0 * 2.0f + v[0];
    out[1] = s1 * 2.0f + v[1];
    out[2] = s2 * 2.0f + v[2];
}

// World AABB of geometry g: its local box (mesh AABB, unit box, tetrahedron box, or the y-flat quad) scaled,
// rotated and translated corner by corner.  mesh_min/mesh_max: the mesh's local AABB when g is a mesh.
SLIM_HD SBox geometry_world_box(const slimrt_geometry &g, const float *mesh_min, const float *mesh_max) {
    SBox b;
    if (g.type == 1) {
        for (int k = 0; k < 3; k++) { b.mn[k] = mesh_min[k]; b.mx[k] = mesh_max[k]; }
    } else {
        const float e = g.type == 5 ? 0.577350259f : 1.0f;
        for (int k = 0; k < 3; k++) { b.mn[k] = -e; b.mx[k] = e; }
        if (g.type == 6) b.mn[1] = b.mx[1] = 0.0f;
    }
    SBox w;
    for (int k = 0; k < 3; k++) { w.mn[k] = INFINITY; w.mx[k] = -INFINITY; }
    for (int cx = 0; cx < 2; cx++)
        for (int cy = 0; cy < 2; cy++)
            for (int cz = 0; cz < 2; cz++) {
                const float p[3] = {g.scale[0] * (cx ? b.mx[0] : b.mn[0]), g.scale[1] * (cy ? b.mx[1] : b.mn[1]),
                                    g.scale[2] * (cz ? b.mx[2] : b.mn[2])};
                float r[3];
                quat_rotate(g.orientation, p, r);
                for (int k = 0; k < 3; k++) {
                    const float v = r[k] + g.position[k];
                    w.mn[k] = w.mn[k] < v ? w.mn[k] : v;
                    w.mx[k] = w.mx[k] > v ? w.mx[k] : v;
                }
            }
    return w;
}

// Scratch for sah_build over N boxes (all arrays of N, sort_stack of 2N + 2, tasks of N + 2).
struct SahScratch {
    SBox *boxes;
    uint32_t *ids, *leaf_ids;
    uint32_t *sorted[3];
    SBox *left[3], *right[3];
    float *left_area[3], *right_area[3];
    int32_t *sort_stack;
    uint32_t *tasks;   // 4 words each: start, end, node, depth
};

// Output node in the reference's asset layout.
struct SahNode { float mn[3], mx[3]; uint32_t first_index, leaf_count, depth; };

// BVHPartition::partition for one axis: returns the axis cost and the left count of its best split.
SLIM_HD float sah_partition(SahScratch &s, int axis, uint32_t N, uint32_t &left_count) {
    uint32_t *ids = s.sorted[axis];
    int32_t *st = s.sort_stack;
    int32_t top = -1;
    st[++top] = 0;
    st[++top] = (int32_t)N - 1;
    while (top >= 0) {
        const int32_t end = st[top--], start = st[top--];
        int32_t p = start;
        const float pivot = s.boxes[ids[end]].mx[axis];
        for (int32_t i = start; i < end; i++)
            if (s.boxes[ids[i]].mx[axis] < pivot) {
                const uint32_t t = ids[i]; ids[i] = ids[p]; ids[p] = t;
                p++;
            }
        { const uint32_t t = ids[end]; ids[end] = ids[p]; ids[p] = t; }
        if (p - 1 > start) { st[++top] = start; st[++top] = p - 1; }
        if (p + 1 < end) { st[++top] = p + 1; st[++top] = end; }
    }
    SBox L = s.boxes[ids[0]], R = s.boxes[ids[N - 1]];
    s.left[axis][0] = L;
    s.right[axis][N - 1] = R;
    for (uint32_t i = 0, j = N - 1; i < N; i++, j--) {
        if (i) {
            L = sbox_merge(L, s.boxes[ids[i]]);
            R = sbox_merge(R, s.boxes[ids[j]]);
            s.left[axis][i] = L;
            s.right[axis][j] = R;
        }
        s.left_area[axis][i] = sbox_area(L);
        s.right_area[axis][j] = sbox_area(R);
    }
    left_count = 1;
    float cost = INFINITY;
    for (uint32_t cur = 0, nl = 1, nr = N - 1; cur < N - 1; cur++, nl++, nr--) {
        const float c = s.left_area[axis][cur] * (float)nl + s.right_area[axis][cur + 1] * (float)nr;
        if (cost > c) { cost = c; left_count = cur + 1; }
    }
    return cost;
}

SLIM_HD void sah_set(SahNode &n, const SBox &b) {
    for (int k = 0; k < 3; k++) { n.mn[k] = b.mn[k]; n.mx[k] = b.mx[k]; }
}

SLIM_HD void sah_clear(SahNode &n) {
    for (int k = 0; k < 3; k++) { n.mn[k] = 0.0f; n.mx[k] = 0.0f; }
    n.first_index = n.leaf_count = n.depth = 0;
}

// BVHBuilder::splitNode: appends the two children of `node` and returns the split position.
SLIM_HD uint32_t sah_split(SahScratch &s, SahNode *out, uint32_t &count, uint32_t node, uint32_t start, uint32_t end) {
    const uint32_t N = end - start;
    out[node].first_index = count;
    SahNode &l = out[count++];
    SahNode &r = out[count++];
    sah_clear(l);
    sah_clear(r);
    float best = INFINITY;
    int chosen = 0;
    uint32_t lc[3];
    for (int a = 0; a < 3; a++) {
        for (uint32_t i = 0; i < N; i++) s.sorted[a][i] = s.ids[start + i];
        const float c = sah_partition(s, a, N, lc[a]);
        if (c < best) { best = c; chosen = a; }
    }
    sah_set(l, s.left[chosen][lc[chosen] - 1]);
    sah_set(r, s.right[chosen][lc[chosen]]);
    for (uint32_t i = 0; i < N; i++) s.ids[start + i] = s.sorted[chosen][i];
    return start + lc[chosen];
}

// BVHBuilder::build over s.boxes[0..N) with s.ids = 0..N-1; writes node_count nodes, the height and leaf_ids.
SLIM_HD void sah_build(SahScratch &s, uint32_t N, uint32_t max_leaf, SahNode *out, uint32_t &count, uint32_t &height) {
    height = 1;
    count = 1;
    sah_clear(out[0]);
    if (N <= max_leaf) {
        SBox b;
        for (int k = 0; k < 3; k++) { b.mn[k] = INFINITY; b.mx[k] = -INFINITY; }
        for (uint32_t i = 0; i < N; i++) {
            s.leaf_ids[i] = i;
            b = sbox_merge(b, s.boxes[i]);
        }
        sah_set(out[0], b);
        out[0].leaf_count = N;
        return;
    }
    uint32_t *tk = s.tasks;
    int32_t top = -1;
    uint32_t middle = sah_split(s, out, count, 0, 0, N);
    ++top; tk[4 * top] = 0; tk[4 * top + 1] = middle; tk[4 * top + 2] = 1; tk[4 * top + 3] = 1;
    ++top; tk[4 * top] = middle; tk[4 * top + 1] = N; tk[4 * top + 2] = 2; tk[4 * top + 3] = 1;
    uint32_t leaves = 0;
    while (top >= 0) {
        const uint32_t start = tk[4 * top], end = tk[4 * top + 1], node = tk[4 * top + 2], depth = tk[4 * top + 3];
        const uint32_t n = end - start;
        if (n <= max_leaf) {
            out[node].depth = depth;
            out[node].leaf_count = n;
            out[node].first_index = leaves;
            for (uint32_t i = 0; i < n; i++) s.leaf_ids[leaves + i] = s.ids[start + i];
            leaves += n;
            top--;
        } else {
            middle = sah_split(s, out, count, node, start, end);
            const uint32_t d = depth + 1, first = out[node].first_index;
            out[first].depth = out[first + 1].depth = d;
            tk[4 * top] = start; tk[4 * top + 1] = middle; tk[4 * top + 2] = first; tk[4 * top + 3] = d;
            ++top; tk[4 * top] = middle; tk[4 * top + 1] = end; tk[4 * top + 2] = first + 1; tk[4 * top + 3] = d;
            if (d > height) height = d;
        }
    }
    out[1].depth = out[2].depth = 1;
    SBox l, r;
    for (int k = 0; k < 3; k++) {
        l.mn[k] = out[1].mn[k]; l.mx[k] = out[1].mx[k];
        r.mn[k] = out[2].mn[k]; r.mx[k] = out[2].mx[k];
    }
    sah_set(out[0], sbox_merge(l, r));
}

}  // namespace slimrt
