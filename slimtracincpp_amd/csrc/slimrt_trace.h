// Two-level BVH traversal, shading and the per-pixel integrator (__host__ __device__).
//
// Traversal visits sibling pairs as the reference does (SceneTracer::trace, src/slim/scene/scene_tracer.h:26-98;
// MeshTracer::trace, src/slim/scene/mesh_tracer.h:51-134): both children of a pair are slab-tested, leaves are
// intersected in place, the nearer of two hit inner children is entered and the other pushed.  The traversal
// stacks live wherever the caller puts them: on the GPU in LDS, one column per thread (element i of a thread's
// stack at base[i * stride]), on the CPU in a local array (stride 1).
#pragma once

#include <type_traits>

#include "slimrt_core.h"

namespace slimrt {

// Bits of a traversal's status (TracerT::status) and of a context's device status word.
constexpr uint32_t kStatusStack = 1u;   // a traversal stack push would have exceeded its capacity
constexpr uint32_t kStatusQueue = 2u;   // a wavefront queue append would have exceeded its capacity
constexpr uint32_t kStatusNaN = 4u;     // a sample's radiance or depth came out NaN

template <class T>
struct StackT {
    T *base;
    uint32_t stride;
    SLIM_HD T &operator[](uint32_t i) const { return base[i * stride]; }
};
using Stack = StackT<uint32_t>;

struct Counters {   // work counters, accumulated only when the caller asks for them
    uint64_t rays, pairs, roots, instances, triangles, mesh_hits, taps, pixels;
};

// The traversal state of one sample.  NW, the meshes' node width (2: the sibling-pair walk, 4 or 8: the collapsed
// wide BVHs), is chosen at compile time, so a kernel carries only the mesh walk it uses.
template <uint32_t NW>
struct TracerT {
    static_assert(NW == 2 || NW == 4 || NW == 8, "node width 2, 4 or 8");
    static constexpr uint32_t kWidth = NW;
    static constexpr bool kWide = NW >= 4;
    // mesh stack entries: sibling-pair indices, or wide node indices relative to the mesh (at most 65536 nodes)
    using MeshEntry = typename std::conditional<(NW >= 4), uint16_t, uint32_t>::type;
    Stack scene_stack;
    StackT<MeshEntry> mesh_stack;
    uint32_t scene_depth, mesh_depth;   // capacities
    uint32_t status;                    // kStatus* bits: a push beyond a capacity, a NaN sample
    Ray aux_ray;                        // last localized ray (SceneTracer::aux_ray)
    Hit aux_hit;                        // SceneTracer::aux_hit
    Counters *cnt;                      // null: no counting
};
using Tracer = TracerT<2>;

#define SLIM_COUNT(tr, field) do { if ((tr).cnt) (tr).cnt->field++; } while (0)

SLIM_HD Node load_node(const Node *nodes, uint32_t i) { return nodes[i]; }

// The triangle test of MeshTracer::hitTriangles (mesh_tracer.h:23-49) over nb records tb, triangle ids id0, id0 + 1,
// ...; closest and the scratch hit th carry over between calls.  Returns true when an any-hit walk may stop.
template <class TR>
SLIM_HD bool test_triangles(const TriHit *tb, uint32_t nb, uint32_t id0, float &closest, const Ray &r, Hit &hit,
                            Hit &th, bool any_hit, bool &found, TR &tr) {
#pragma unroll
    for (uint32_t k = 0; k < 2; k++) {
        if (k >= nb) break;
        const TriHit &t = tb[k];
        SLIM_COUNT(tr, triangles);
        v3 tp{t.px, t.py, t.pz};
        if (hits_plane(r, tp, v3{t.nx, t.ny, t.nz}, th)) {
            v3 d = th.position - tp;
            float ux = t.xx * d.x + t.yx * d.y + t.zx * d.z;
            float uy = t.xy * d.x + t.yy * d.y + t.zy * d.z;
            if (ux < 0 || uy < 0 || (ux + uy) > 1 || th.distance >= closest) continue;
            closest = th.distance;
            hit = th;
            hit.uv = v2{ux, uy};
            hit.id = id0 + k;
            found = true;
            if (any_hit) return true;
        }
    }
    return false;
}

// MeshTracer::hitTriangles (mesh_tracer.h:23-49); triangle indices are relative to the mesh.  The records are
// fetched two at a time, so a leaf costs half as many dependent round trips to memory; they are tested in order
// as before.
template <class TR>
SLIM_HD bool hit_triangles(const TriHit *tris, uint32_t first, uint32_t n, float closest, const Ray &r, Hit &hit,
                           bool any_hit, TR &tr) {
    bool found = false;
    closest = fmin2(closest, hit.distance);
    Hit th = hit;
    for (uint32_t i0 = 0; i0 < n; i0 += 2) {
        TriHit tb[2];
#pragma unroll
        for (uint32_t k = 0; k < 2; k++)
            if (i0 + k < n) tb[k] = tris[i0 + k];
        if (test_triangles(tb, n - i0, first + i0, closest, r, hit, th, any_hit, found, tr)) break;
    }
    return found;
}

// The closest hit's UV coverage, interpolated UV and normal (mesh_tracer.h:36, 117-131).
template <class TR>
SLIM_HD void mesh_hit_attributes(const SceneView &s, const MeshRec &m, Hit &hit, TR &tr) {
    const TriAttr t = s.tri_attr[m.tri_base + hit.id];
    hit.uv_coverage = t.uv_coverage;
    if (!(m.normals_count | m.uvs_count)) return;
    float a = hit.uv.u, b = hit.uv.v, c = 1 - a - b;
    SLIM_COUNT(tr, mesh_hits);
    if (m.uvs_count) {
        hit.uv.u = t.uv3u * a + (t.uv2u * b + t.uv1u * c);
        hit.uv.v = t.uv3v * a + (t.uv2v * b + t.uv1v * c);
    }
    if (m.normals_count) {
        hit.normal.x = t.n3x * a + (t.n2x * b + t.n1x * c);
        hit.normal.y = t.n3y * a + (t.n2y * b + t.n1y * c);
        hit.normal.z = t.n3z * a + (t.n2z * b + t.n1z * c);
    }
}

template <class TR>
SLIM_HD bool mesh_trace4(const SceneView &s, const MeshRec &m, const Ray &r, Hit &hit, bool any_hit, TR &tr);
template <class TR>
SLIM_HD bool mesh_trace8(const SceneView &s, const MeshRec &m, const Ray &r, Hit &hit, bool any_hit, TR &tr);

// MeshTracer::trace (mesh_tracer.h:51-134)
template <class TR>
SLIM_HD bool mesh_trace(const SceneView &s, uint32_t mesh_id, const Ray &r, Hit &hit, bool any_hit, TR &tr) {
    const MeshRec m = s.meshes[mesh_id];
    if constexpr (TR::kWidth == 8) return mesh_trace8(s, m, r, hit, any_hit, tr);
    if constexpr (TR::kWidth == 4) return mesh_trace4(s, m, r, hit, any_hit, tr);
    const Node *nodes = s.mnodes + m.node_base;
    float ln, rn, lf, rf;
    bool found = false;
    SLIM_COUNT(tr, roots);
    const Node root = load_node(nodes, 0);
    if (!(hits_node(r, root, ln, lf) && ln < hit.distance)) return false;
    if (root.leaf_count_depth & 0xFFFFu) return hit_triangles(s.tri_hit + m.tri_base, 0, m.tri_count, lf, r, hit, any_hit, tr);
    uint32_t pair = root.first_index, top = 0;
    for (;;) {
        SLIM_COUNT(tr, pairs);
        const Node L = load_node(nodes, pair), R = load_node(nodes, pair + 1);
        bool hl = hits_node(r, L, ln, lf) && ln < hit.distance;
        bool hr = hits_node(r, R, rn, rf) && rn < hit.distance;
        bool go_l = false, go_r = false;
        if (hl) {
            uint32_t lc = L.leaf_count_depth & 0xFFFFu;
            if (lc) {
                if (hit_triangles(s.tri_hit + m.tri_base + L.first_index, L.first_index, lc, lf, r, hit, any_hit, tr)) {
                    found = true;
                    if (any_hit) break;
                }
            } else go_l = true;
        }
        if (hr) {
            uint32_t rc = R.leaf_count_depth & 0xFFFFu;
            if (rc) {
                if (hit_triangles(s.tri_hit + m.tri_base + R.first_index, R.first_index, rc, rf, r, hit, any_hit, tr)) {
                    found = true;
                    if (any_hit) break;
                }
            } else go_r = true;
        }
        if (go_l) {
            if (go_r) {
                uint32_t near_pair = L.first_index, far_pair = R.first_index;
                if (!any_hit && ln > rn) { near_pair = R.first_index; far_pair = L.first_index; }
                if (top < tr.mesh_depth) tr.mesh_stack[top++] = far_pair; else tr.status |= kStatusStack;
                pair = near_pair;
            } else pair = L.first_index;
        } else if (go_r) {
            pair = R.first_index;
        } else {
            if (top == 0) break;
            pair = tr.mesh_stack[--top];
        }
    }
    if (found && !any_hit) mesh_hit_attributes(s, m, hit, tr);
    return found;
}

// The 4-wide traversal of the same mesh: each visit slab-tests the (up to) four children, intersects the leaf
// children nearest first, enters the nearest inner child and pushes the others far to near.  It reaches the
// same closest hit as the sibling-pair walk (leaves keep their triangle ranges and each leaf's triangles are
// clipped to that leaf's box as before); only exact ties between triangles can resolve differently.
template <class TR>
SLIM_HD bool mesh_trace4(const SceneView &s, const MeshRec &m, const Ray &r, Hit &hit, bool any_hit, TR &tr) {
    const Node4 *nodes = s.mnodes4 + m.node4_base;
    bool found = false;
    uint32_t cur = 0, top = 0;
    SLIM_COUNT(tr, roots);
    for (;;) {
        SLIM_COUNT(tr, pairs);
        const Node4 n = nodes[cur];
        float nd[4] = {0.0f, 0.0f, 0.0f, 0.0f}, fd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        bool h[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
            h[i] = n.child[i] != ~0u &&
                   hits_box_bounds(r, n.minx[i], n.miny[i], n.minz[i], n.maxx[i], n.maxy[i], n.maxz[i], nd[i], fd[i]) &&
                   nd[i] < hit.distance;
        // leaves, nearest first (a found hit shortens hit.distance for the later ones)
        float key[4];
        uint32_t idx[4] = {0, 1, 2, 3};
#pragma unroll
        for (int i = 0; i < 4; i++) key[i] = h[i] ? nd[i] : INFINITY;
#define SLIM_CSWAP(a, b)                                                             \
    if (key[b] < key[a]) {                                                           \
        const float tk = key[a]; key[a] = key[b]; key[b] = tk;                       \
        const uint32_t ti = idx[a]; idx[a] = idx[b]; idx[b] = ti;                    \
    }
        SLIM_CSWAP(0, 1) SLIM_CSWAP(2, 3) SLIM_CSWAP(0, 2) SLIM_CSWAP(1, 3) SLIM_CSWAP(1, 2)
#undef SLIM_CSWAP
        bool stop = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {   // leaf children, nearest first
            const uint32_t i = idx[k];
            const uint32_t count = i == 0 ? n.count[0] : i == 1 ? n.count[1] : i == 2 ? n.count[2] : n.count[3];
            if (!(key[k] < hit.distance) || count == 0 || stop) continue;
            const uint32_t child = i == 0 ? n.child[0] : i == 1 ? n.child[1] : i == 2 ? n.child[2] : n.child[3];
            const float far_d = i == 0 ? fd[0] : i == 1 ? fd[1] : i == 2 ? fd[2] : fd[3];
            if (hit_triangles(s.tri_hit + m.tri_base + child, child, count, far_d, r, hit, any_hit, tr)) {
                found = true;
                if (any_hit) stop = true;
            }
        }
        if (stop) break;
#pragma unroll
        for (int k = 3; k >= 0; k--) {   // inner children, far to near: the nearest ends on top
            const uint32_t i = idx[k];
            const uint32_t count = i == 0 ? n.count[0] : i == 1 ? n.count[1] : i == 2 ? n.count[2] : n.count[3];
            if (!(key[k] < hit.distance) || count != 0) continue;
            const uint32_t child = i == 0 ? n.child[0] : i == 1 ? n.child[1] : i == 2 ? n.child[2] : n.child[3];
            if (top < tr.mesh_depth) tr.mesh_stack[top++] = (typename TR::MeshEntry)child; else tr.status |= kStatusStack;
        }
        if (top == 0) break;
        cur = tr.mesh_stack[--top];
    }
    if (found && !any_hit) mesh_hit_attributes(s, m, hit, tr);
    return found;
}

// Byte k of a pair of words (k = 0..7; the words hold slots 0-3 and 4-7).
SLIM_HD uint32_t slot_byte(const uint32_t w[2], uint32_t k) { return (w[k >> 2] >> (8 * (k & 3u))) & 0xFFu; }
// Word k of eight, for a k known only at run time (a select chain instead of indexed registers).
SLIM_HD uint32_t pick8(const uint32_t v[8], uint32_t k) {
    const uint32_t a = (k & 1u) ? v[1] : v[0], b = (k & 1u) ? v[3] : v[2], c = (k & 1u) ? v[5] : v[4],
                   d = (k & 1u) ? v[7] : v[6];
    const uint32_t ab = (k & 2u) ? b : a, cd = (k & 2u) ? d : c;
    return (k & 4u) ? cd : ab;
}
SLIM_HD float exp_step(uint32_t biased) {   // 2^(biased - 127) for biased in 1..254
    const uint32_t bits = biased << 23;
    float f;
    __builtin_memcpy(&f, &bits, 4);
    return f;
}

// The 8-wide walk of a mesh (Node8): every visit decodes and slab-tests the (up to) eight quantised child boxes,
// takes the hit leaves in the ray's octant order (each repeats the reference's exact leaf-box test from its leaf
// record, then tests the leaf's triangles, whose records follow it), and pushes the hit inner children so that the
// first in octant order is popped next.  The octant order stands in for sorting by entry distance (slots were
// filled at the collapse so that, for each ray octant, the children nearer the ray's entry side come first); the
// closest hit does not depend on the order, only exact ties between triangles can resolve differently.
template <class TR>
SLIM_HD bool mesh_trace8(const SceneView &s, const MeshRec &m, const Ray &r, Hit &hit, bool any_hit, TR &tr) {
    const Node8 *nodes = s.mnodes8 + m.node8_base;
    const uint4 *leafs = s.leaf8 + m.leaf8_base;
    const uint32_t mo = 7u ^ (r.sl | (r.sb << 1) | (r.sk << 2));   // slot s is visited at position s ^ mo
    bool found = false;
    uint32_t cur = 0, top = 0;
    SLIM_COUNT(tr, roots);
    for (;;) {
        SLIM_COUNT(tr, pairs);
        const Node8 n = nodes[cur];
        const float sx = exp_step(n.exps & 0xFFu), sy = exp_step((n.exps >> 8) & 0xFFu),
                    sz = exp_step((n.exps >> 16) & 0xFFu);
        // the near and far planes of every slot for this ray's octant (the slab test's swap, once per node)
        uint32_t nx[2], ny[2], nz[2], fx[2], fy[2], fz[2];
#pragma unroll
        for (uint32_t w = 0; w < 2; w++) {
            nx[w] = r.sl ? n.qhi[0][w] : n.qlo[0][w]; fx[w] = r.sl ? n.qlo[0][w] : n.qhi[0][w];
            ny[w] = r.sb ? n.qhi[1][w] : n.qlo[1][w]; fy[w] = r.sb ? n.qlo[1][w] : n.qhi[1][w];
            nz[w] = r.sk ? n.qhi[2][w] : n.qlo[2][w]; fz[w] = r.sk ? n.qlo[2][w] : n.qhi[2][w];
        }
        uint32_t hits = 0, leaves = 0;   // bit per position
#pragma unroll
        for (uint32_t k = 0; k < 8; k++) {
            // Every slot without a branch (an empty slot, count 255, is masked out).  Decoded bounds lie two ulps
            // beyond the exact ones (checked at the collapse), so the fused slab distances bracket the reference's
            // rounded ones on the exact box: a ray the exact box accepts is accepted.
            const float tnx = fmaf(fmaf((float)slot_byte(nx, k), sx, n.ox), r.rcp.x, r.scaled_origin.x);
            const float tny = fmaf(fmaf((float)slot_byte(ny, k), sy, n.oy), r.rcp.y, r.scaled_origin.y);
            const float tnz = fmaf(fmaf((float)slot_byte(nz, k), sz, n.oz), r.rcp.z, r.scaled_origin.z);
            const float tfx = fmaf(fmaf((float)slot_byte(fx, k), sx, n.ox), r.rcp.x, r.scaled_origin.x);
            const float tfy = fmaf(fmaf((float)slot_byte(fy, k), sy, n.oy), r.rcp.y, r.scaled_origin.y);
            const float tfz = fmaf(fmaf((float)slot_byte(fz, k), sz, n.oz), r.rcp.z, r.scaled_origin.z);
            // IEEE max / min (three-operand instructions): a NaN operand is dropped wherever it stands, where the
            // reference's ternaries keep or drop it by position, so this test accepts every box they accept
            const float nd = fmaxf(0.0f, fmaxf(tnx, fmaxf(tny, tnz))), fd = fminf(tfx, fminf(tfy, tfz));
            const uint32_t cnt = slot_byte(n.count, k);
            const bool h = nd <= fd && nd < hit.distance && cnt != 0xFFu;
            hits |= (uint32_t)h << (k ^ mo);
            leaves |= (uint32_t)(h && cnt != 0u) << (k ^ mo);
        }
        bool stop = false;
        for (uint32_t lm = leaves; lm && !stop; lm &= lm - 1u) {   // hit leaves, first position first
            const uint32_t k = (uint32_t)__builtin_ctz(lm) ^ mo;
            // the leaf record and its first two triangle records in one round trip (the leaf array ends with a
            // record's worth of padding, so the second read stays inside it for one-triangle leaves)
            const uint4 *rec = leafs + pick8(n.child, k);
            LeafHead h;
            TriHit tb[2];
            __builtin_memcpy(&h, rec, sizeof h);
            __builtin_memcpy(tb, rec + 2, sizeof tb);
            float nd, fd;
            if (!(hits_box_bounds(r, h.minx, h.miny, h.minz, h.maxx, h.maxy, h.maxz, nd, fd) && nd < hit.distance))
                continue;
            float closest = fmin2(fd, hit.distance);
            Hit th = hit;
            bool got = false;
            stop = test_triangles(tb, h.count, h.first, closest, r, hit, th, any_hit, got, tr);
            const TriHit *rest = reinterpret_cast<const TriHit *>(rec + 2);
            for (uint32_t i0 = 2; i0 < h.count && !stop; i0 += 2) {
#pragma unroll
                for (uint32_t j = 0; j < 2; j++)
                    if (i0 + j < h.count) tb[j] = rest[i0 + j];
                stop = test_triangles(tb, h.count - i0, h.first + i0, closest, r, hit, th, any_hit, got, tr);
            }
            found = found || got;
        }
        if (stop) break;
        for (uint32_t im = hits & ~leaves; im;) {   // inner children, last position first: the first ends on top
            const uint32_t p = 31u - (uint32_t)__builtin_clz(im);
            im &= ~(1u << p);
            if (top < tr.mesh_depth) tr.mesh_stack[top++] = (typename TR::MeshEntry)pick8(n.child, p ^ mo);
            else tr.status |= kStatusStack;
        }
        if (top == 0) break;
        cur = tr.mesh_stack[--top];
    }
    if (found && !any_hit) mesh_hit_attributes(s, m, hit, tr);
    return found;
}

// SceneTracer::hitGeometryInLocalSpace (scene_tracer.h:138-165)
template <class TR>
SLIM_HD bool hit_geometry_local(const SceneView &s, const Geo &g, const Ray &r, Hit &hit, bool any_hit, TR &tr) {
    SLIM_COUNT(tr, instances);
    ray_localize(tr.aux_ray, r.origin, r.direction, g);
    float n, f;
    bool box_hit;
    if (g.type == GT_Mesh) {
        const MeshRec &m = s.meshes[geo_id(g)];
        box_hit = hits_box_bounds(tr.aux_ray, m.minx, m.miny, m.minz, m.maxx, m.maxy, m.maxz, n, f);
    } else {
        float mx = g.type == GT_Tet ? kTetMax : 1.0f;
        float ylo = -mx, yhi = mx;
        if (g.type == GT_Quad) { ylo = -kEps; yhi = kEps; }
        box_hit = hits_box_bounds(tr.aux_ray, -mx, ylo, -mx, mx, yhi, mx, n, f);
    }
    if (!box_hit) return false;
    const bool transparent = (geo_flags(g) & GEO_TRANSPARENT) != 0;
    switch (g.type) {
        case GT_Quad: return hits_quad(tr.aux_ray, hit, transparent);
        case GT_Box: return hits_box(tr.aux_ray, hit, transparent);
        case GT_Sphere: return hits_sphere(tr.aux_ray, hit, transparent, nullptr);
        case GT_Tet: return hits_tet(tr.aux_ray, hit, transparent);
        case GT_Mesh: return mesh_trace(s, geo_id(g), tr.aux_ray, hit, any_hit, tr);
        default: return false;
    }
}

// SceneTracer::hitGeometries (scene_tracer.h:111-136); returns a geometry index or -1.
template <class TR>
SLIM_HD int hit_geometries(const SceneView &s, uint32_t first_leaf, uint32_t n, float closest, const Ray &r, Hit &hit,
                           bool any_hit, TR &tr) {
    int hit_geo = -1;
    const uint32_t vis = any_hit ? GEO_SHADOWING : GEO_VISIBLE;
    tr.aux_hit.distance = fmin2(closest + kEps, hit.distance);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t gi = s.leaf_ids[first_leaf + i];
        const Geo g = s.geos[gi];
        if (!(geo_flags(g) & vis)) continue;
        if (hit_geometry_local(s, g, r, tr.aux_hit, any_hit, tr)) {
            if (any_hit) return (int)gi;
            if (tr.aux_hit.distance < hit.distance) {
                hit_geo = (int)gi;
                hit = tr.aux_hit;
                hit.NdotRd = -(dot(hit.normal, tr.aux_ray.direction));
            }
        }
    }
    return hit_geo;
}

// SceneTracer::trace (scene_tracer.h:26-98); returns a geometry index or -1.  Offsets the ray origin.
template <class TR>
SLIM_HD int scene_trace(const SceneView &s, Ray &r, Hit &hit, bool any_hit, float max_distance, TR &tr) {
    SLIM_COUNT(tr, rays);
    ray_reset(r, scale_add(r.direction, kTraceOffset, r.origin), r.direction);
    hit.distance = max_distance;
    float ln, rn, lf, rf;
    SLIM_COUNT(tr, roots);
    const Node root = load_node(s.snodes, 0);
    if (!(hits_node(r, root, ln, lf) && ln < hit.distance)) return -1;
    if (root.leaf_count_depth & 0xFFFFu) return hit_geometries(s, 0, s.G, lf, r, hit, any_hit, tr);
    int closest = -1;
    uint32_t pair = root.first_index, top = 0;
    for (;;) {
        SLIM_COUNT(tr, pairs);
        const Node L = load_node(s.snodes, pair), R = load_node(s.snodes, pair + 1);
        bool hl = hits_node(r, L, ln, lf) && ln < hit.distance;
        bool hr = hits_node(r, R, rn, rf) && rn < hit.distance;
        bool go_l = false, go_r = false;
        if (hl) {
            uint32_t lc = L.leaf_count_depth & 0xFFFFu;
            if (lc) {
                int g = hit_geometries(s, L.first_index, lc, lf, r, hit, any_hit, tr);
                if (g >= 0) { closest = g; if (any_hit) break; }
            } else go_l = true;
        }
        if (hr) {
            uint32_t rc = R.leaf_count_depth & 0xFFFFu;
            if (rc) {
                int g = hit_geometries(s, R.first_index, rc, rf, r, hit, any_hit, tr);
                if (g >= 0) { closest = g; if (any_hit) break; }
            } else go_r = true;
        }
        if (go_l) {
            if (go_r) {
                uint32_t near_pair = L.first_index, far_pair = R.first_index;
                if (!any_hit && ln > rn) { near_pair = R.first_index; far_pair = L.first_index; }
                if (top < tr.scene_depth) tr.scene_stack[top++] = far_pair; else tr.status |= kStatusStack;
                pair = near_pair;
            } else pair = L.first_index;
        } else if (go_r) {
            pair = R.first_index;
        } else {
            if (top == 0) break;
            pair = tr.scene_stack[--top];
        }
    }
    return closest;
}

// ------------------------------------------------------------------ shading (src/slim/scene/material.h, renderer/surface_shader.h)
SLIM_HD float ggx_D(float roughness, float NdotH) {   // material.h:27-36
    float a = roughness * roughness;
    float denom = NdotH * NdotH * (a - 1.0f) + 1.0f;
    return a / (kPi * denom * denom);
}
SLIM_HD float ggx_G(float roughness, float NdotL, float NdotV) {   // material.h:38-55
    float a = roughness * roughness, k = a * 0.5f, omk = 1.0f - k;
    float denom = NdotV * omk + k;
    float result = NdotV / fmaxf(denom, kEps);
    denom = NdotL * omk + k;
    result *= NdotL / fmaxf(denom, kEps);
    return result;
}
SLIM_HD color schlick(float HdotL, color F0) {   // material.h:57-59
    float p = powf(1.0f - HdotL, 5.0f);
    return F0 + one_minus(F0) * p;
}
SLIM_HD float tone_map(float x) {   // toneMappedBaked, base.h:777-786
    x = x - 0.004f;
    if (x < 0.0f) x = 0.0f;
    float x2s = x * x * 6.2f;
    return (x2s + x * 0.5f) / (x2s + x * 1.7f + 0.06f);
}

struct Surface {
    int geo;
    Mat mat;
    color F, albedo_from_map, Fs, Fd;
    v3 P, N, V, L, R, RF;
    float Ld, Ld2, NdotL, NdotV;
    bool refracted;
};

SLIM_HD color sample_slot(const SceneView &s, const Mat &m, int slot, v2 uv, float coverage) {   // surface_shader.h:9-11
    return (s.NT && mat_tex_count(m) > (uint32_t)slot) ? tex_sample(s, mat_tex(m, slot), uv.u, uv.v, coverage) : cs(0);
}

SLIM_HD v3 rotate_normal(v3 normal, color smp, float magnitude) {   // material.h:83-86
    v3 N = normalized(scale_add(v3{smp.r, smp.b, smp.g}, 2.0f, vs(-1.0f)));
    quat q = axis_angle(normalized(v3{N.z, 0, -N.x}), acosf(N.y) * magnitude);
    return qrot(q, normal);
}

SLIM_HD void radiance_fraction(Surface &sf) {   // surface_shader.h:109-145
    const Mat &m = sf.mat;
    sf.Fs = cs(0);
    sf.Fd = color{m.ar, m.ag, m.ab} * sf.albedo_from_map;
    const uint32_t brdf = mat_brdf(m);
    if (brdf == BRDF_CookTorrance) {
        sf.Fd = sf.Fd * ((1.0f - m.metalness) * kOneOverPi);
        if (sf.NdotV > 0.0f && m.roughness > 0.0f) {
            v3 H = normalized(sf.L + sf.V);
            float NdotH = clamp01(dot(sf.N, H));
            float HdotL = clamp01(dot(H, sf.L));
            sf.F = schlick(HdotL, color{m.fr, m.fg, m.fb});
            float D = ggx_D(m.roughness, NdotH), G = ggx_G(m.roughness, sf.NdotL, sf.NdotV);
            sf.Fs = sf.F * (D * G / (4.0f * sf.NdotL * sf.NdotV));
            sf.Fd = sf.Fd * one_minus(sf.F);
        }
    } else {
        sf.Fd = sf.Fd * (m.roughness * kOneOverPi);
        if (brdf != BRDF_Lambert) {
            float spec, exponent;
            if (brdf == BRDF_Phong) { exponent = 4.0f; spec = clamp01(dot(sf.R, sf.L)); }
            else { exponent = 16.0f; spec = clamp01(dot(sf.N, normalized(sf.L + sf.V))); }
            if (spec > 0.0f) sf.Fs = color{m.fr, m.fg, m.fb} * (powf(spec, exponent) * (1.0f - m.roughness));
        }
    }
}

SLIM_HD bool facing_light(Surface &sf, const Light &l) {   // surface_shader.h:147-160
    if (l.flags & LIGHT_DIRECTIONAL) {
        sf.Ld = INFINITY;
        sf.Ld2 = 1.0f;
        sf.L = -v3{l.x, l.y, l.z};
    } else {
        sf.L = v3{l.x, l.y, l.z} - sf.P;
        sf.Ld2 = sqlen(sf.L);
        sf.Ld = sqrtf(sf.Ld2);
        sf.L = divs(sf.L, sf.Ld);
    }
    sf.NdotL = clamp01(dot(sf.L, sf.N));
    return sf.NdotL > 0.0f;
}

template <class TR>
SLIM_HD void shade_from_light(Surface &sf, const Light &l, const SceneView &s, TR &tr, color &c) {   // surface_shader.h:43-59
    if (facing_light(sf, l)) {
        bool lit = true;
        if (l.flags & LIGHT_SHADOWING) {
            Ray shadow;
            shadow.origin = sf.P;
            shadow.direction = sf.L;
            Hit sh;
            lit = scene_trace(s, shadow, sh, true, sf.Ld, tr) < 0;
        }
        if (lit) {
            radiance_fraction(sf);
            c = cmul_add(sf.Fs + sf.Fd, color{l.r, l.g, l.b} * (sf.NdotL * l.intensity / sf.Ld2), c);
        }
    }
}

template <class TR>
SLIM_HD void prepare_for_shading(Surface &sf, const SceneView &s, const Ray &r, Hit &hit, TR &tr) {   // surface_shader.h:61-107
    const Geo g = s.geos[sf.geo];
    sf.mat = s.mats[g.material_id];
    const Mat &m = sf.mat;
    hit.uv.u *= m.uv_ru;
    hit.uv.v *= m.uv_rv;
    if (hit.from_behind) hit.normal = -hit.normal;
    float cone = hit.distance * hit.scaling_factor;
    cone *= cone;
    cone *= cone;
    hit.uv_coverage *= cone / (m.uv_ru * m.uv_rv * hit.NdotRd * fabsf(dot(v3{1.0f - hit.normal.x, 1.0f - hit.normal.y, 1.0f - hit.normal.z}, v3{g.sx, g.sy, g.sz})));
    hit.position = ray_at(r, hit.distance);
    hit.normal = normalized(qrot(geo_quat(g), v3{g.sx, g.sy, g.sz} * hit.normal));
    const uint32_t mf = mat_flags(m);
    if (mat_tex_count(m) && (mf & MAT_NORMAL_MAP)) {
        if (s.NT && mat_tex_count(m) > 1) SLIM_COUNT(tr, taps);
        hit.normal = rotate_normal(hit.normal, sample_slot(s, m, 1, hit.uv, hit.uv_coverage), m.normal_magnitude);
    }
    sf.P = hit.position;
    sf.N = hit.normal;
    sf.R = sf.RF = reflected(r.direction, sf.N);
    sf.V = -r.direction;
    sf.NdotV = clamp01(dot(sf.N, sf.V));
    const bool albedo_map = mat_tex_count(m) && (mf & MAT_ALBEDO_MAP);
    if (albedo_map && s.NT) SLIM_COUNT(tr, taps);
    sf.albedo_from_map = albedo_map ? sample_slot(s, m, 0, hit.uv, hit.uv_coverage) : cs(1.0f);
    sf.refracted = (mf & MAT_REFRACTIVE) != 0;
    if (sf.refracted) {
        float ior = hit.from_behind ? (m.ior / kIorAir) : (kIorAir / m.ior);
        float c = ior * ior * (1.0f - (sf.NdotV * sf.NdotV));
        sf.refracted = c < 1.0f;
        if (sf.refracted) sf.RF = normalized(scale_add(sf.N, ior * sf.NdotV - sqrtf(1 - c), r.direction * ior));
    }
}

// Irradiance vector of an emissive quad (surface_shader.h:162-186)
SLIM_HD v3 area_light_vector(const Surface &sf, const Geo &g, v3 *quad_v) {
    const float sx = g.sx, sz = g.sz;
    if (sx == 0 || sz == 0) return vs(0.0f);
    quat q = geo_quat(g);
    v3 U = qrot(q, v3{sx < 0 ? -sx : sx, 0.0f, 0.0f});
    v3 W = qrot(q, v3{0.0f, 0.0f, sz < 0 ? -sz : sz});
    v3 p{g.px, g.py, g.pz};
    quad_v[0] = (p - U) - W;
    quad_v[1] = (p + U) - W;
    quad_v[2] = (p + U) + W;
    quad_v[3] = (p - U) + W;
    v3 u1 = normalized(quad_v[0] - sf.P), u2 = normalized(quad_v[1] - sf.P);
    v3 u3 = normalized(quad_v[2] - sf.P), u4 = normalized(quad_v[3] - sf.P);
    return ((cross(u1, u2) * (acosf(dot(u1, u2)) * 0.5f) + cross(u2, u3) * (acosf(dot(u2, u3)) * 0.5f)) +
            cross(u3, u4) * (acosf(dot(u3, u4)) * 0.5f)) + cross(u4, u1) * (acosf(dot(u4, u1)) * 0.5f);
}

// SurfaceShader::shadeFromEmissiveQuads (surface_shader.h:188-264): heuristic occlusion over all geometries.
SLIM_HD void shade_from_emissive_quads(Surface &sf, const SceneView &s, color &c) {
    v3 quad_v[4];
    for (uint32_t i = 0; i < s.G; i++) {
        const Geo eq = s.geos[i];
        if ((int)i == sf.geo || eq.type != GT_Quad || !(mat_flags(s.mats[eq.material_id]) & MAT_EMISSIVE)) continue;
        sf.L = v3{eq.px, eq.py, eq.pz} - sf.P;
        sf.NdotL = dot(sf.N, sf.L);
        if (sf.NdotL <= 0.0f) continue;
        sf.Ld2 = sqlen(sf.L);
        sf.Ld = sqrtf(sf.Ld2);
        float Ld_rcp = 1.0f / sf.Ld;
        sf.L = sf.L * Ld_rcp;
        sf.NdotL *= Ld_rcp;
        v3 Ro = scale_add(sf.L, kTraceOffset, sf.P);
        float emission_intensity = dot(sf.N, area_light_vector(sf, eq, quad_v));
        if (emission_intensity > 0) {
            bool skip = true;
            for (int k = 0; k < 4; k++)
                if (dot(sf.N, quad_v[k] - sf.P) >= 0) { skip = false; break; }
            if (skip) continue;
            float shaded = 1.0f, sq_dist;
            for (uint32_t j = 0; j < s.G; j++) {
                const Geo sg = s.geos[j];
                if (j == i || (int)j == sf.geo) continue;
                Ray sr;
                ray_localize(sr, Ro, sf.L, sg);
                Hit sh;
                sh.distance = INFINITY;
                sr.direction = normalized(sr.direction);   // the reciprocal stays that of the unnormalised direction
                float d = 1.0f;
                const bool transparent = (geo_flags(sg) & GEO_TRANSPARENT) != 0;
                if (sg.type == GT_Sphere) {
                    if (hits_sphere(sr, sh, transparent, &sq_dist))
                        d -= (1.0f - sqrtf(sq_dist)) / (sh.distance * emission_intensity * 3.0f);
                } else if (sg.type == GT_Quad) {
                    if (hits_quad(sr, sh, transparent))
                        d -= 3.0f * (1.0f - fmax2(fabsf(sh.position.x), fabsf(sh.position.z))) / (sh.distance * emission_intensity);
                }
                shaded = fmin2(shaded, d);
            }
            if (shaded > 0.0f) {
                radiance_fraction(sf);
                const Mat em = s.mats[eq.material_id];
                c = cmul_add(sf.Fd + sf.Fs, color{em.er, em.eg, em.eb} * (emission_intensity * shaded * 7.0f), c);
            }
        }
    }
}

SLIM_HD float depth_at(const Projection &p, v3 pos) {   // camera.h:66
    v3 d = pos - v3{p.camera_position[0], p.camera_position[1], p.camera_position[2]};
    return p.inv_rot[2] * d.x + p.inv_rot[5] * d.y + p.inv_rot[8] * d.z;
}

// Point-light glow: SceneTracer::hitLight + SphereTracer (scene_tracer.h:100-109, ray.h:314-346).
// Mutates max_distance exactly as the reference does.  On a hit returns true and the density integral.
SLIM_HD bool light_glow(const Light &l, const Ray &r, float &max_distance, float &density) {
    if (l.flags & LIGHT_DIRECTIONAL) return false;
    const float radius = l.intensity * kLightRadiusFactor;
    const float oor = kLightInvRadiusFactor / l.intensity;
    const float t_max = max_distance * oor;
    const v3 rc = (v3{l.x, l.y, l.z} - r.origin) * oor;
    const float b = dot(r.direction, rc);
    const float c = sqlen(rc) - 1;
    float h = b * b - c;
    if (h >= 0) {
        h = sqrtf(h);
        const float t_near = b - h, t_far = b + h;
        if (t_far > 0 && t_near < t_max) {
            max_distance = t_max * radius;
            const float tn = fmax2(t_near, 0), tf = fmin2(t_far, t_max);
            const float tn2 = tn * tn, tf2 = tf * tf;
            density = (c * tn - b * tn2 + tn * tn2 / 3.0f - (c * tf - b * tf2 + tf * tf2 / 3.0f)) * (3.0f / 4.0f);
            return true;
        }
    }
    return false;
}

// pow(density, 8.0f) * 4 evaluated in double as on the reference's host (ray_tracer.h:101)
SLIM_HD float glow_factor(float density) {
    double x = (double)density, x2 = x * x, x4 = x2 * x2;
    return (float)(x4 * x4 * 4.0);
}

SLIM_HD void primary_ray(const Projection &p, uint32_t x, uint32_t y, v3 &dir, float &scaling) {   // renderer_GPU.h:36-43
    float a = (float)(int32_t)x * p.sample_size + p.c_start[0];
    float b = (float)(-(int32_t)y) * p.sample_size + p.c_start[1];
    scaling = 1.0f / sqrtf(p.sq_dist + (a * a + b * b));
    dir = (v3{p.start[0], p.start[1], p.start[2]} + v3{p.down[0], p.down[1], p.down[2]} * (float)(int32_t)y) +
          v3{p.right[0], p.right[1], p.right[2]} * (float)(int32_t)x;
}

// The light terms a frame kernel instantiation carries: point and directional lights (with their shadow rays and
// glow) and emissive quads.  A scene without lights and without emissive quads runs kLightsNone, one with emissive
// quads only kLightsQuads: the loops left out would run zero times, and leaving them out spares a kernel their
// registers (the shadow rays' walks above all).
constexpr uint32_t kLightPoint = 1, kLightQuads = 2;
constexpr uint32_t kLightsNone = 0, kLightsQuads = kLightQuads, kLightsAll = kLightPoint | kLightQuads;

// renderPixelBeauty (src/slim/renderer/ray_tracer.h:19-108), with the light terms LIGHTS selects.
template <class TR, uint32_t LIGHTS = kLightsAll>
SLIM_HD void render_beauty(const SceneView &s, const RenderParams &rp, v3 direction, float scaling, TR &tr,
                           color &out, float &depth) {
    Surface sf;
    Ray ray;
    Hit hit;
    hit.scaling_factor = scaling;
    color col = cs(0), cur, next_tp = cs(0), tp = cs(1.0f);
    depth = INFINITY;
    const Projection &p = rp.proj;
    ray_reset(ray, v3{p.camera_position[0], p.camera_position[1], p.camera_position[2]}, normalized(direction));
    uint32_t depth_left = rp.max_depth;
    while (depth_left) {
        cur = cs(0);
        sf.geo = scene_trace(s, ray, hit, false, INFINITY, tr);
        if (sf.geo >= 0) {
            hit.scaling_factor = keeps_ray_cones(rp) ? scaling : 0.0f;
            const Geo g = s.geos[sf.geo];
            const Mat m = s.mats[g.material_id];
            if ((mat_flags(m) & MAT_EMISSIVE) && g.type == GT_Quad && !hit.from_behind) {
                depth_left = 0;
                cur = color{m.er, m.eg, m.eb};
            } else {
                prepare_for_shading(sf, s, ray, hit, tr);
                if (depth_left == rp.max_depth) depth = depth_at(p, hit.position);
                if constexpr ((LIGHTS & kLightPoint) != 0)
                    for (uint32_t i = 0; i < s.L; i++) shade_from_light(sf, s.lights[i], s, tr, cur);
                if constexpr ((LIGHTS & kLightQuads) != 0)
                    if (s.flags & 1u) shade_from_emissive_quads(sf, s, cur);
                if (rp.sky_irradiance >= 0 && rp.sky_radiance >= 0) {
                    sf.L = sf.N;
                    sf.NdotL = 1.0f;
                    if (tr.cnt) tr.cnt->taps += 2;
                    color D = tex_sample_cube(s, (uint32_t)rp.sky_irradiance, sf.N.x, sf.N.y, sf.N.z);
                    color S = tex_sample_cube(s, (uint32_t)rp.sky_radiance, sf.R.x, sf.R.y, sf.R.z);
                    radiance_fraction(sf);
                    cur = cmul_add(D, sf.Fd, cmul_add(sf.Fs, S, cur));
                }
                if ((mat_flags(sf.mat) & (MAT_REFLECTIVE | MAT_REFRACTIVE)) && --depth_left) {
                    sf.F = schlick(clamp01(dot(sf.N, sf.R)), color{sf.mat.fr, sf.mat.fg, sf.mat.fb});
                    next_tp = sf.refracted ? one_minus(sf.F) : sf.F;
                    ray_reset(ray, hit.position, sf.RF);
                } else depth_left = 0;
            }
        } else {
            depth_left = 0;
            if (rp.sky_color >= 0) SLIM_COUNT(tr, taps);
            if (rp.sky_color >= 0) cur = tex_sample_cube(s, (uint32_t)rp.sky_color, ray.direction.x, ray.direction.y, ray.direction.z);
        }
        if constexpr ((LIGHTS & kLightPoint) != 0) {
            for (uint32_t i = 0; i < s.L; i++) {
                const Light l = s.lights[i];
                float density;
                if (light_glow(l, ray, hit.distance, density)) cur = cscale_add(color{l.r, l.g, l.b}, glow_factor(density), cur);
            }
        }
        col = cmul_add(cur, tp, col);
        tp = tp * next_tp;
    }
    out = color{tone_map(col.r), tone_map(col.g), tone_map(col.b)};
}

// renderPixelDebugMode (src/slim/renderer/ray_tracer.h:110-142)
template <class TR>
SLIM_HD void render_debug(const SceneView &s, const RenderParams &rp, v3 direction, float scaling, TR &tr,
                          color &out, float &depth) {
    Surface sf;
    Ray ray;
    Hit hit;
    out = cs(0);
    depth = INFINITY;
    const Projection &p = rp.proj;
    ray_reset(ray, v3{p.camera_position[0], p.camera_position[1], p.camera_position[2]}, normalized(direction));
    sf.geo = scene_trace(s, ray, hit, false, INFINITY, tr);
    if (sf.geo >= 0) {
        hit.scaling_factor = keeps_ray_cones(rp) ? scaling : 0.0f;
        prepare_for_shading(sf, s, ray, hit, tr);
        depth = depth_at(p, hit.position);
        switch (rp.render_mode) {
            case RM_UVs: out = color{hit.uv.u, hit.uv.v, 1.0f}; break;
            case RM_Depth: out = cs(4.0f / hit.distance); break;
            case RM_Normals: out = color{hit.normal.x * 0.5f + 0.5f, hit.normal.y * 0.5f + 0.5f, hit.normal.z * 0.5f + 0.5f}; break;
            case RM_NormalMap: {
                if (s.NT && mat_tex_count(sf.mat) > 1) SLIM_COUNT(tr, taps);
                color n = sample_slot(s, sf.mat, 1, hit.uv, hit.uv_coverage);
                out = color{n.r * 0.5f + 0.5f, n.g * 0.5f + 0.5f, n.b * 0.5f + 0.5f};
            } break;
            case RM_MipLevel:
                if (s.NT) {
                    const TexRec t0 = s.textures[0];
                    uint32_t lvl = mip_level(hit.uv_coverage * (float)((uint64_t)t0.width * t0.height), t0.mip_count);
                    lvl = lvl > 8 ? 8 : lvl;
                    // a select over constant indices: a dynamic index would force a caller's private copy of
                    // the parameters (the multi-view kernel's) into scratch memory
#pragma unroll
                    for (uint32_t k = 0; k < 9; k++)
                        if (k == lvl) out = color{rp.mip_colors[k][0], rp.mip_colors[k][1], rp.mip_colors[k][2]};
                } else out = cs(0.5f);
                break;
            default: break;
        }
    }
}

// The value the canvas or the radiance buffer receives for a tone-mapped sample colour.
SLIM_HD void output_value(const RenderParams &rp, color c, float out_px[4]) {
    if (rp.output == 0) {   // Canvas::setPixel with opacity -1 (src/slim/draw/canvas.h:128-169)
        color k{clamp01(c.r), clamp01(c.g), clamp01(c.b)};
        out_px[0] = k.r * (k.r * 1.0f);
        out_px[1] = k.g * (k.g * 1.0f);
        out_px[2] = k.b * (k.b * 1.0f);
    } else {
        out_px[0] = c.r;
        out_px[1] = c.g;
        out_px[2] = c.b;
    }
    out_px[3] = 1.0f;
}

// One sample: primary ray, integrator, and the value the canvas or the radiance buffer receives.
template <class TR, uint32_t LIGHTS = kLightsAll>
SLIM_HD void render_sample(const SceneView &s, const RenderParams &rp, uint32_t x, uint32_t y, TR &tr,
                           float out_px[4], float &out_depth) {
    v3 dir;
    float scaling;
    primary_ray(rp.proj, x, y, dir, scaling);
    if (rp.output == 2) {   // primary closest hit: geometry index, leaf-ordered triangle id (meshes), distance
        Ray ray;
        Hit hit;
        const Projection &p = rp.proj;
        ray_reset(ray, v3{p.camera_position[0], p.camera_position[1], p.camera_position[2]}, normalized(dir));
        const int g = scene_trace(s, ray, hit, false, INFINITY, tr);
        const int tri = (g >= 0 && s.geos[g].type == GT_Mesh) ? (int)hit.id : -1;
        out_px[0] = __builtin_bit_cast(float, g);
        out_px[1] = __builtin_bit_cast(float, tri);
        out_px[2] = g >= 0 ? hit.distance : INFINITY;
        out_px[3] = 0.0f;
        out_depth = out_px[2];
        if (tr.cnt) tr.cnt->pixels++;
        return;
    }
    color c;
    if (rp.render_mode == RM_Beauty) render_beauty<TR, LIGHTS>(s, rp, dir, scaling, tr, c, out_depth);
    else render_debug(s, rp, dir, scaling, tr, c, out_depth);
    if (tr.cnt) tr.cnt->pixels++;
    output_value(rp, c, out_px);
}

// Selection::manipulate's pick (src/slim/scene/selection.h:36-85): the camera ray through integer pixel (x, y),
// direction CameraRayProjection::getRayDirectionAt (camera.h:67) normalised, traced for the closest hit; then
// every point light's pick sphere (radius intensity / 2 / 64, SphereTracer::hit, ray.h:317-334) against the hit
// distance, the last light hit winning.  out: geometry, light, triangle (meshes), distance, world position.
template <class TR>
SLIM_HD void pick(const SceneView &s, const RenderParams &rp, int32_t x, int32_t y, TR &tr, int32_t out_i[3],
                  float out_f[4]) {
    const Projection &p = rp.proj;
    const v3 dir = (v3{p.start[0], p.start[1], p.start[2]} + v3{p.down[0], p.down[1], p.down[2]} * (float)y) +
                   v3{p.right[0], p.right[1], p.right[2]} * (float)x;
    Ray ray;
    Hit hit;
    ray_reset(ray, v3{p.camera_position[0], p.camera_position[1], p.camera_position[2]}, normalized(dir));
    const int g = scene_trace(s, ray, hit, false, INFINITY, tr);
    int light = -1;
    float max_distance = hit.distance;
    for (uint32_t i = 0; i < s.L; i++) {
        const Light l = s.lights[i];
        if (l.flags & LIGHT_DIRECTIONAL) continue;
        const float half = l.intensity * 0.5f;
        const float radius = half * kLightRadiusFactor, oor = kLightInvRadiusFactor / half;
        const float t_max = max_distance * oor;
        const v3 rc = (v3{l.x, l.y, l.z} - ray.origin) * oor;
        const float b = dot(ray.direction, rc), c = sqlen(rc) - 1;
        float h = b * b - c;
        if (h >= 0) {
            h = sqrtf(h);
            if (b + h > 0 && b - h < t_max) {
                max_distance = t_max * radius;
                light = (int)i;
            }
        }
    }
    out_i[0] = g;
    out_i[1] = light;
    out_i[2] = (g >= 0 && s.geos[g].type == GT_Mesh) ? (int)hit.id : -1;
    out_f[0] = g >= 0 ? hit.distance : INFINITY;
    const v3 pos = g >= 0 ? ray_at(ray, hit.distance) : vs(0.0f);
    out_f[1] = pos.x;
    out_f[2] = pos.y;
    out_f[3] = pos.z;
}

// Window content of one pixel of a canvas (Canvas::getPixelContent, canvas.h:187-190, 254-265; Pixel::asContent,
// base.h:1166-1171): under SSAA the mean of its 2x2 sample quad, ((p0 + p1) + p2) + p3 times 0.25, or 0 when all
// four opacities are 0; else the pixel (0 when its opacity is 0).  0xRRGGBB, each channel 255 * sqrt(c) truncated,
// 255 above 1.
SLIM_HD uint32_t pixel_content(const float4 *p, bool ssaa) {
    float4 c = p[0];
    if (ssaa) {
        if (p[0].w == 0.0f && p[1].w == 0.0f && p[2].w == 0.0f && p[3].w == 0.0f) return 0u;
        c = make_float4((((p[0].x + p[1].x) + p[2].x) + p[3].x) * 0.25f, (((p[0].y + p[1].y) + p[2].y) + p[3].y) * 0.25f,
                        (((p[0].z + p[1].z) + p[2].z) + p[3].z) * 0.25f, 0.0f);
    } else if (c.w == 0.0f) {
        return 0u;
    }
    auto channel = [](float v) { return v > 1.0f ? 255u : (uint32_t)(uint8_t)(255.0f * sqrtf(v)); };
    return channel(c.x) << 16 | channel(c.y) << 8 | channel(c.z);
}

// kStatusNaN when a sample's value (radiance or canvas output; not the hit records of OUTPUT_HITS) or depth is NaN.
SLIM_HD uint32_t nan_status(const RenderParams &rp, const float px[4], float depth) {
    const bool nan = (rp.output != 2 && (px[0] != px[0] || px[1] != px[1] || px[2] != px[2])) || depth != depth;
    return nan ? kStatusNaN : 0u;
}

// Canvas offsets: samples are stored row-major, or as 2x2 quads per pixel under SSAA (canvas.h:147).
SLIM_HD uint32_t canvas_offset(const RenderParams &rp, uint32_t x, uint32_t y) {
    return rp.ssaa ? ((rp.width * (y >> 1) + (x >> 1)) * 4 + (2 * (y & 1)) + (x & 1)) : (rp.width * y + x);
}

}  // namespace slimrt
