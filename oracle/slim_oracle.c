/* CPU oracle for the SlimTracin per-pixel ray-tracing path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and only as the checker or as the CPU baseline.
 *
 * A plain-C restatement of the reference's algorithm, operation by operation, so that on the same
 * inputs it produces the same floating-point results as the reference compiled for baseline x86-64
 * (no FMA contraction: src/slim/core/base.h:84-99 expands fast_mul_add to a multiply and an add).
 * Every function cites the reference file:line it restates.  Pinning (what checks each part):
 *   - scene derivation (quaternions, AABBs, BVHs, triangle records, projection) and per-pixel
 *     results: compared bit-for-bit with the compiled reference (oracle/_ref) on seeded scenes and
 *     against committed golden vectors from it (tests/golden);
 *   - mesh BVH build: the reference's own monkey.mesh / cube.mesh assets, bit-for-bit; its one-pass result
 *     for sorted ranges (sorted_runs) against the step-by-step quicksort;
 *   - intersectors, slab test, BRDF terms, tone map, bilinear/mip/cube sampling: closed forms,
 *     double-precision roots and brute force in tests/test_oracle_math.py;
 *   - canvas window content and the wireframe overlays (boxes, BVH node boxes, mesh edges drawn into a
 *     rendered canvas): bit-for-bit with the compiled reference's own Canvas and draw functions on
 *     seeded scenes (tests/test_canvas_content.py, tests/test_overlay.py) and its golden overlays.
 * Parity unpinned: none of the above covers libm ULP differences (sinf/cosf/powf/acosf); the
 * oracle calls the host libm exactly where the reference does.
 *
 * Reference-compatibility note (src/slim/scene/scene_tracer.h:129, mesh_tracer.h:201): the closest-hit
 * copies `hit = aux_hit` / `hit = triangle_hit` overwrite the caller's `hit.scaling_factor`
 * (set at renderer.h:95 / renderer_GPU.h:38) with the tracer's never-written field, which is zero in
 * the zero-initialised tracer the apps use.  The oracle reproduces that data flow by default
 * (ray_cones = 0).  With ray_cones = 1 the caller's scaling factor is kept, which is the ray-cone
 * mip selection the reference describes (surface_shader.h:67-77); that mode is parity unpinned
 * against the reference (the reference never executes it) and is pinned by closed-form tests only.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "specfile.h"

/* ---------------------------------------------------------------- constants (src/slim/core/base.h:125-160) */
#define EPS 0.0001f
#define TRACE_OFFSET 0.0001f
#define PI_F 3.14159265f
#define ONE_OVER_PI (1.0f / PI_F)
#define SQRT3 1.73205080757f
#define UNIT_SPHERE_AREA_OVER_SIX ((4.0f * PI_F) / 6.0f)
#define TET_MAX (0.577350259f)
#define TET_MIN (0.288675159f)
#define COLOR_COMPONENT_TO_FLOAT 0.00392156862f
#define IOR_AIR 1.0003f
#define LIGHT_RADIUS_INTENSITY_FACTOR 64.0f
#define LIGHT_INTENSITY_RADIUS_FACTOR (1.0f / LIGHT_RADIUS_INTENSITY_FACTOR)

enum { GT_NONE = 0, GT_MESH, GT_GRID, GT_BOX, GT_CURVE, GT_TET, GT_QUAD, GT_SPHERE };
enum { BRDF_LAMBERT = 0, BRDF_BLINN, BRDF_PHONG, BRDF_CT };
enum { RM_NORMALS = 0, RM_NORMAL_MAP, RM_BEAUTY, RM_DEPTH, RM_MIP_LEVEL, RM_UVS };
enum { SIDE_NONE = 0, SIDE_LEFT = 1, SIDE_BOTTOM = 2, SIDE_BACK = 4, SIDE_RIGHT = 8, SIDE_TOP = 16, SIDE_FRONT = 32 };
#define MAT_REFLECTIVE 1u
#define MAT_REFRACTIVE 2u
#define MAT_EMISSIVE 4u
#define MAT_ALBEDO_MAP 16u
#define MAT_NORMAL_MAP 32u
#define GEO_VISIBLE 1u
#define GEO_SHADOWING 2u
#define GEO_TRANSPARENT 4u
#define LIGHT_DIRECTIONAL 1u
#define LIGHT_SHADOWING 2u
#define IMG_MIPMAP (1u << 4)

#define MAX2(a, b) ((a) > (b) ? (a) : (b))   /* base.h:117 */
#define MIN2(a, b) ((a) < (b) ? (a) : (b))   /* base.h:118 */

/* ---------------------------------------------------------------- vector algebra (src/slim/math/vec3.h) */
typedef struct { float x, y, z; } v3;
typedef struct { float u, v; } v2;
typedef struct { v3 axis; float w; } quat;
typedef struct { v3 X, Y, Z; } m3;          /* rows, as mat3 (src/slim/math/mat3.h:5-13) */
typedef struct { v3 min, max; } aabb;
typedef struct { float r, g, b; } color;

static inline v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vs(float s) { return V3(s, s, s); }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline v3 vdivs(v3 a, float s) { float f = 1.0f / s; return V3(a.x * f, a.y * f, a.z * f); } /* vec3.h:148 */
static inline v3 vsdiv(float s, v3 a) { return V3(s / a.x, s / a.y, s / a.z); }                       /* vec3.h:435 */
static inline v3 vssub(float s, v3 a) { return V3(s - a.x, s - a.y, s - a.z); }                       /* vec3.h:419 */
static inline float vdot(v3 a, v3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }               /* vec3.h:292 */
static inline v3 vcross(v3 a, v3 b) {                                                                   /* vec3.h:296 */
    return V3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
static inline float vsqlen(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }                         /* vec3.h:304 */
static inline float vlen(v3 a) { return sqrtf(vsqlen(a)); }
static inline v3 vnorm(v3 a) { return vdivs(a, vlen(a)); }                                              /* vec3.h:312 */
static inline v3 vscaleadd(v3 a, float f, v3 b) { return V3(a.x * f + b.x, a.y * f + b.y, a.z * f + b.z); } /* vec3.h:364 */
static inline v3 vmuladd(v3 a, v3 f, v3 b) { return V3(a.x * f.x + b.x, a.y * f.y + b.y, a.z * f.z + b.z); } /* vec3.h:380 */
static inline v3 vreflect(v3 d, v3 n) { return vscaleadd(n, -2 * vdot(d, n), d); }                      /* vec3.h:316 */
static inline v3 vmin(v3 a, v3 b) { return V3(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z); }
static inline v3 vmax(v3 a, v3 b) { return V3(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z); }
static inline float vminimum(v3 a) { return MIN2(a.x, MIN2(a.y, a.z)); }                               /* vec3.h:262 */
static inline float vmaximum(v3 a) { return MAX2(a.x, MAX2(a.y, a.z)); }                               /* vec3.h:277 */
static inline float vminimum_axis(v3 a, int *axis) {                                                     /* vec3.h:263 */
    float r = a.x; *axis = 1;
    if (a.y < r) { r = a.y; *axis = 2; }
    if (a.z < r) { r = a.z; *axis = 4; }
    return r;
}
static inline float vmaximum_axis(v3 a, int *axis) {                                                     /* vec3.h:278 */
    float r = a.x; *axis = 1;
    if (a.y > r) { r = a.y; *axis = 2; }
    if (a.z > r) { r = a.z; *axis = 4; }
    return r;
}
static inline float clamp01(float v) { float mn = v < 1.0f ? v : 1.0f; return mn > 0.0f ? mn : 0.0f; } /* base.h:231 */

static inline color C3(float r, float g, float b) { color c = {r, g, b}; return c; }
static inline color cs(float s) { return C3(s, s, s); }
static inline color cadd(color a, color b) { return C3(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline color cmul(color a, color b) { return C3(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline color cscale(color a, float s) { return C3(a.r * s, a.g * s, a.b * s); }
static inline color cmuladd(color a, color f, color b) { return C3(a.r * f.r + b.r, a.g * f.g + b.g, a.b * f.b + b.b); }
static inline color cscaleadd(color a, float f, color b) { return C3(a.r * f + b.r, a.g * f + b.g, a.b * f + b.b); }
static inline color c1minus(color a) { return C3(1.0f - a.r, 1.0f - a.g, 1.0f - a.b); }

/* quat (src/slim/math/quat.h) */
static inline quat Q(v3 axis, float w) { quat q = {axis, w}; return q; }
static inline quat qnorm(quat q) {                                                                        /* quat.h:18-35 */
    float len = sqrtf(q.axis.x * q.axis.x + q.axis.y * q.axis.y + q.axis.z * q.axis.z + q.w * q.w);
    float f = 1.0f / len;
    return Q(vscale(q.axis, f), q.w * f);
}
static inline quat qaxisangle(v3 axis, float radians) {                                                   /* quat.h:78-84 */
    radians *= 0.5f;
    return qnorm(Q(vscale(axis, sinf(radians)), cosf(radians)));
}
static inline quat qmul_assign(quat a, quat b) {                                                          /* quat.h:56-65 */
    float x = a.axis.x, X = b.axis.x, y = a.axis.y, Y = b.axis.y, z = a.axis.z, Z = b.axis.z, w = a.w, W = b.w;
    quat r;
    r.axis.x = X * w + W * x + Z * y - Y * z;
    r.axis.y = Y * w - Z * x + W * y + X * z;
    r.axis.z = Z * w + Y * x - X * y + W * z;
    r.w = W * w - X * x - Y * y - Z * z;
    return r;
}
static inline v3 qrot(quat q, v3 v) {                                                                     /* quat.h:37-43 */
    v3 out = vcross(q.axis, v);
    v3 qqv = vcross(q.axis, out);
    out = vscaleadd(out, q.w, qqv);
    return vscaleadd(out, 2.0f, v);
}
static inline quat qconj(quat q) { return Q(vneg(q.axis), q.w); }
/* Orientation<quat>::_update (src/slim/core/base.h:612-620) */
static quat quat_from_euler(float x, float y, float z) {
    quat q = z != 0.0f ? qaxisangle(V3(0, 0, 1), z) : Q(V3(0, 0, 0), 1.0f);
    if (x != 0.0f) q = qmul_assign(q, qaxisangle(V3(1, 0, 0), x));
    if (y != 0.0f) q = qmul_assign(q, qaxisangle(V3(0, 1, 0), y));
    return q;
}

/* mat3 (src/slim/math/mat3.h); cos/sin of a float argument resolve to the double functions there. */
static inline m3 M3(v3 X, v3 Y, v3 Z) { m3 m = {X, Y, Z}; return m; }
static inline m3 m3_identity(void) { return M3(V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)); }
static inline v3 m3_mulv(m3 m, v3 v) {                                                                    /* mat3.h:250-256 */
    return V3(m.X.x * v.x + m.Y.x * v.y + m.Z.x * v.z,
              m.X.y * v.x + m.Y.y * v.y + m.Z.y * v.z,
              m.X.z * v.x + m.Y.z * v.y + m.Z.z * v.z);
}
static inline m3 m3_mul(m3 l, m3 r) {                                                                     /* mat3.h:272-285 */
    m3 o;
    o.X.x = l.X.x * r.X.x + l.X.y * r.Y.x + l.X.z * r.Z.x;
    o.X.y = l.X.x * r.X.y + l.X.y * r.Y.y + l.X.z * r.Z.y;
    o.X.z = l.X.x * r.X.z + l.X.y * r.Y.z + l.X.z * r.Z.z;
    o.Y.x = l.Y.x * r.X.x + l.Y.y * r.Y.x + l.Y.z * r.Z.x;
    o.Y.y = l.Y.x * r.X.y + l.Y.y * r.Y.y + l.Y.z * r.Z.y;
    o.Y.z = l.Y.x * r.X.z + l.Y.y * r.Y.z + l.Y.z * r.Z.z;
    o.Z.x = l.Z.x * r.X.x + l.Z.y * r.Y.x + l.Z.z * r.Z.x;
    o.Z.y = l.Z.x * r.X.y + l.Z.y * r.Y.y + l.Z.z * r.Z.y;
    o.Z.z = l.Z.x * r.X.z + l.Z.y * r.Y.z + l.Z.z * r.Z.z;
    return o;
}
static inline m3 m3_scale(m3 m, float f) {
    return M3(vscale(m.X, f), vscale(m.Y, f), vscale(m.Z, f));
}
static inline m3 m3_inverted(m3 m) {                                                                      /* mat3.h:146-178 */
    v3 X = m.X, Y = m.Y, Z = m.Z;
    float det = +X.x * (Y.y * Z.z - Z.y * Y.z) - Y.x * (X.y * Z.z - Z.y * X.z) + Z.x * (X.y * Y.z - Y.y * X.z);
    m3 c = M3(V3(+(Y.y * Z.z - Z.y * Y.z), -(X.y * Z.z - Z.y * X.z), +(X.y * Y.z - Y.y * X.z)),
              V3(-(Y.x * Z.z - Z.x * Y.z), +(X.x * Z.z - Z.x * X.z), -(X.x * Y.z - Y.x * X.z)),
              V3(+(Y.x * Z.y - Z.x * Y.y), -(X.x * Z.y - Z.x * X.y), +(X.x * Y.y - Y.x * X.y)));
    return m3_scale(c, 1.0f / det);
}
static m3 m3_from_euler(float x, float y, float z) {                                                      /* base.h:612-620 */
    m3 m;
    if (z != 0.0f) {                                                                                      /* mat3.h:49-55 */
        float c = (float)cos((double)z), s = (float)sin((double)z);
        m = M3(V3(c, -s, 0), V3(s, c, 0), V3(0, 0, 1));
    } else m = m3_identity();
    if (x != 0.0f) {                                                                                      /* mat3.h:57-66 */
        float c = (float)cos((double)x), s = (float)sin((double)x);
        m = m3_mul(m, M3(V3(1, 0, 0), V3(0, c, -s), V3(0, s, c)));
    }
    if (y != 0.0f) {                                                                                      /* mat3.h:68-77 */
        float c = (float)cos((double)y), s = (float)sin((double)y);
        m = m3_mul(m, M3(V3(c, 0, s), V3(0, 1, 0), V3(-s, 0, c)));
    }
    return m;
}

static inline aabb aunion(aabb a, aabb b) { aabb r = {vmin(a.min, b.min), vmax(a.max, b.max)}; return r; } /* vec3.h:471 */
static inline float aarea(aabb a) {                                                                      /* vec3.h:512-515 */
    v3 e = vsub(a.max, a.min);
    return e.x * e.y + e.y * e.z + e.z * e.x;
}

/* ---------------------------------------------------------------- scene data */
typedef struct {
    quat q;
    v3 position, scale;
    uint32_t type, material_id, id, flags;
} Geo;

typedef struct {
    color albedo;
    float roughness, normal_magnitude;
    v2 uv_repeat;
    color emission;
    float ior;
    color reflectivity;
    float metalness;
    uint32_t flags, texture_count, texture_ids[2], brdf;
} Mat;

typedef struct { color color; v3 pos; float intensity; uint32_t flags; } Lite;

typedef struct { aabb box; uint32_t first_index, leaf_count, depth; } Node;

typedef struct {            /* Triangle (src/slim/scene/mesh.h:19-24), 128 bytes */
    m3 l2t;
    v3 position, normal, n1, n2, n3;
    v2 uv1, uv2, uv3;
    float uv_coverage, padding;
} Tri;

typedef struct {
    aabb box;
    Node *nodes;
    uint32_t node_count, height;
    Tri *tris;
    uint32_t tri_count, normals_count, uvs_count;
} MeshD;

/* fmt 0: texel quads (the reference's layout); 1: bordered RGBA8 texels; 2: bordered RGBA16F texels (the
 * product's texel storage, include/slimrt.h SLIMRT_TEXELS_*: texture flags bits 24-25). */
typedef struct { uint32_t w, h; const uint8_t *quads; uint32_t fmt; } Mip;
typedef struct { uint32_t width, height, flags, mip_count; Mip *mips; } Tex;

typedef struct {
    m3 inv_rot;
    v3 start, right, down, camera_position;
    v2 c_start;
    float sq_dist, sample_size;
} Projection;

typedef struct {
    SpecFile sf;
    int width, height, ssaa, max_depth, render_mode, sky_color, sky_radiance, sky_irradiance;
    int ray_cones;
    uint32_t G, M, L, NM, NT, scene_flags;
    Geo *geos;
    Mat *mats;
    Lite *lights;
    MeshD *meshes;
    Tex *texs;
    aabb *geo_aabbs;
    Node *snodes;
    uint32_t snode_count, sheight, *leaf_ids;
    Projection proj;
} Scene;

typedef struct { uint64_t rays, pairs, roots, insts, tris, mesh_hits, taps, pixels; } Counters;

/* ---------------------------------------------------------------- SAH BVH builder (src/slim/scene/bvh_builder.h) */
typedef struct {
    aabb *l_boxes, *r_boxes;
    float *l_areas, *r_areas;
    uint32_t left_node_count, *sorted;
    float surface_area;
} Partition;

typedef struct {
    Node *nodes;          /* per-primitive nodes: box + first_index = primitive id */
    Partition parts[3];
    uint32_t *node_ids, *leaf_ids;
    int32_t *sort_stack;
    struct It { uint32_t start, end, node_id, depth; } *iters;
} Builder;

static void builder_init(Builder *b, uint32_t n) {
    if (n == 0) n = 1;
    b->nodes = (Node *)calloc(n, sizeof(Node));
    b->node_ids = (uint32_t *)calloc(n, 4);
    b->leaf_ids = (uint32_t *)calloc(n, 4);
    b->sort_stack = (int32_t *)calloc(2 * n + 2, 4);
    b->iters = calloc(n + 1, sizeof(*b->iters));
    for (int a = 0; a < 3; a++) {
        b->parts[a].sorted = (uint32_t *)calloc(n, 4);
        b->parts[a].l_boxes = (aabb *)calloc(n, sizeof(aabb));
        b->parts[a].r_boxes = (aabb *)calloc(n, sizeof(aabb));
        b->parts[a].l_areas = (float *)calloc(n, 4);
        b->parts[a].r_areas = (float *)calloc(n, 4);
    }
}
static void builder_free(Builder *b) {
    free(b->nodes); free(b->node_ids); free(b->leaf_ids); free(b->sort_stack); free(b->iters);
    for (int a = 0; a < 3; a++) {
        free(b->parts[a].sorted); free(b->parts[a].l_boxes); free(b->parts[a].r_boxes);
        free(b->parts[a].l_areas); free(b->parts[a].r_areas);
    }
}

static inline float axis_max(const aabb *b, int axis) { return axis == 0 ? b->max.x : axis == 1 ? b->max.y : b->max.z; }

/* The quicksort below, applied to a range whose keys never decrease (none NaN), leaves every maximal run of
 * equal keys rotated right by one: each step finds every element but the run of the pivot's key less than the
 * pivot (kept in place), swaps the pivot with the first element of that run, and continues on the rest of the run
 * and on the range before it.  sorted_runs computes that result in one pass (returns 0, ids unchanged, when the
 * range is not non-decreasing); the step-by-step procedure takes quadratic time on such ranges, which every split
 * hands to its children along the chosen axis.  Pinned against the step-by-step procedure
 * (oracle_set_plain_quicksort) on meshes and on tie-heavy random boxes in tests/test_oracle_math.py. */
static int g_plain_quicksort = 0;
void oracle_set_plain_quicksort(int on) { g_plain_quicksort = on; }

static int sorted_runs(uint32_t *ids, const Node *nodes, int axis, int32_t start, int32_t end) {
    float prev = axis_max(&nodes[ids[start]].box, axis);
    if (prev != prev) return 0;
    for (int32_t i = start + 1; i <= end; i++) {
        float k = axis_max(&nodes[ids[i]].box, axis);
        if (k != k || k < prev) return 0;
        prev = k;
    }
    for (int32_t run = start; run <= end;) {
        float k = axis_max(&nodes[ids[run]].box, axis);
        int32_t last = run;
        while (last < end && !(k < axis_max(&nodes[ids[last + 1]].box, axis))) last++;
        uint32_t t = ids[last];
        for (int32_t i = last; i > run; i--) ids[i] = ids[i - 1];
        ids[run] = t;
        run = last + 1;
    }
    return 1;
}

/* BVHPartition::partition (bvh_builder.h:15-113) */
static void partition(Partition *p, int axis, const Node *nodes, int32_t *stack, uint32_t N) {
    uint32_t *ids = p->sorted;
    int32_t start = 0, end = (int32_t)N - 1, top = -1;
    stack[++top] = start;
    stack[++top] = end;
    while (top >= 0) {              /* iterative Lomuto quicksort by aabb.max[axis], pivot = last */
        end = stack[top--];
        start = stack[top--];
        if (!g_plain_quicksort && end - start >= 16 && sorted_runs(ids, nodes, axis, start, end)) continue;
        int32_t pi = start;
        float pivot = axis_max(&nodes[ids[end]].box, axis);
        for (int32_t i = start; i < end; i++) {
            if (axis_max(&nodes[ids[i]].box, axis) < pivot) {
                uint32_t t = ids[i]; ids[i] = ids[pi]; ids[pi] = t;
                pi++;
            }
        }
        uint32_t t = ids[end]; ids[end] = ids[pi]; ids[pi] = t;
        if (pi - 1 > start) { stack[++top] = start; stack[++top] = pi - 1; }
        if (pi + 1 < end) { stack[++top] = pi + 1; stack[++top] = end; }
    }
    uint32_t li = 0, ri = N - 1;
    aabb L = nodes[ids[li]].box, R = nodes[ids[ri]].box;
    p->l_boxes[li] = L;
    p->r_boxes[ri] = R;
    for (li = 0; li < N; li++, ri--) {
        if (li) {
            L = aunion(L, nodes[ids[li]].box);
            R = aunion(R, nodes[ids[ri]].box);
            p->l_boxes[li] = L;
            p->r_boxes[ri] = R;
        }
        p->l_areas[li] = aarea(L);
        p->r_areas[ri] = aarea(R);
    }
    uint32_t last = N - 1, next = 1, lcount = 1, rcount = N - 1;
    p->left_node_count = 1;
    p->surface_area = INFINITY;
    for (uint32_t cur = 0; cur < last; cur++, next++, lcount++, rcount--) {
        float cost = p->l_areas[cur] * (float)lcount + p->r_areas[next] * (float)rcount;
        if (p->surface_area > cost) { p->surface_area = cost; p->left_node_count = next; }
    }
}

/* BVHBuilder::splitNode (bvh_builder.h:159-193) */
static uint32_t split_node(Builder *b, Node *out_nodes, uint32_t *node_count, uint32_t node_index, uint32_t start, uint32_t end) {
    uint32_t N = end - start, *ids = b->node_ids + start;
    out_nodes[node_index].first_index = *node_count;
    Node *left = &out_nodes[(*node_count)++];
    Node *right = &out_nodes[(*node_count)++];
    memset(left, 0, sizeof *left);
    memset(right, 0, sizeof *right);
    float best = INFINITY;
    int chosen = 0;
    for (int a = 0; a < 3; a++) {
        Partition *p = &b->parts[a];
        for (uint32_t i = 0; i < N; i++) p->sorted[i] = ids[i];
        partition(p, a, b->nodes, b->sort_stack, N);
        if (p->surface_area < best) { best = p->surface_area; chosen = a; }
    }
    Partition *p = &b->parts[chosen];
    left->box = p->l_boxes[p->left_node_count - 1];
    right->box = p->r_boxes[p->left_node_count];
    for (uint32_t i = 0; i < N; i++) ids[i] = p->sorted[i];
    return start + p->left_node_count;
}

/* BVHBuilder::build (bvh_builder.h:195-263); b->nodes[0..N) and b->node_ids must be filled. */
static void bvh_build(Builder *b, Node *out, uint32_t *node_count, uint32_t *height, uint32_t N, uint32_t max_leaf) {
    *height = 1;
    *node_count = 1;
    Node *root = &out[0];
    memset(root, 0, sizeof *root);
    if (N <= max_leaf) {
        root->leaf_count = N;
        root->box.min = vs(INFINITY);
        root->box.max = vs(-INFINITY);
        for (uint32_t i = 0; i < N; i++) {
            b->leaf_ids[i] = b->nodes[i].first_index;
            root->box = aunion(root->box, b->nodes[i].box);
        }
        return;
    }
    uint32_t middle = split_node(b, out, node_count, 0, 0, N);
    struct It *stack = b->iters;
    struct It left = {0, middle, 1, 1}, right = {middle, N, 2, 1};
    stack[0] = left;
    stack[1] = right;
    int32_t sp = 1;
    uint32_t leaf_count = 0;
    while (sp >= 0) {
        left = stack[sp];
        Node *node = &out[left.node_id];
        uint32_t n = left.end - left.start;
        if (n <= max_leaf) {
            node->depth = left.depth;
            node->leaf_count = n;
            node->first_index = leaf_count;
            for (uint32_t i = 0; i < n; i++) b->leaf_ids[leaf_count + i] = b->nodes[b->node_ids[left.start + i]].first_index;
            leaf_count += n;
            sp--;
        } else {
            middle = split_node(b, out, node_count, left.node_id, left.start, left.end);
            node = &out[left.node_id];
            left.depth++;
            right.depth = left.depth;
            right.end = left.end;
            right.start = left.end = middle;
            left.node_id = node->first_index;
            right.node_id = node->first_index + 1;
            out[*node_count - 1].depth = left.depth;
            out[*node_count - 2].depth = right.depth;
            stack[sp] = left;
            stack[++sp] = right;
            if (left.depth > *height) *height = left.depth;
        }
    }
    out[1].depth = out[2].depth = 1;
    root->box = aunion(out[1].box, out[2].box);
}

/* BVHBuilder::buildMesh (bvh_builder.h:265-341) */
static void build_mesh(MeshD *mesh, const float *pos, const uint32_t *pidx, const float *nrm, const uint32_t *nidx,
                       const float *uv, const uint32_t *uidx, uint32_t leaf) {
    uint32_t T = mesh->tri_count;
    Builder b;
    builder_init(&b, T);
    for (uint32_t i = 0; i < T; i++) {
        b.nodes[i].first_index = b.node_ids[i] = i;
        v3 v1 = V3(pos[3 * pidx[3 * i]], pos[3 * pidx[3 * i] + 1], pos[3 * pidx[3 * i] + 2]);
        v3 v2 = V3(pos[3 * pidx[3 * i + 1]], pos[3 * pidx[3 * i + 1] + 1], pos[3 * pidx[3 * i + 1] + 2]);
        v3 v3_ = V3(pos[3 * pidx[3 * i + 2]], pos[3 * pidx[3 * i + 2] + 1], pos[3 * pidx[3 * i + 2] + 2]);
        v3 mn = vmin(vmin(v1, v2), v3_), mx = vmax(vmax(v1, v2), v3_);
        float d;
        d = mx.x - mn.x; if (d < 0) d = -d; if (d < EPS) { mn.x -= EPS; mx.x += EPS; }
        d = mx.y - mn.y; if (d < 0) d = -d; if (d < EPS) { mn.y -= EPS; mx.y += EPS; }
        d = mx.z - mn.z; if (d < 0) d = -d; if (d < EPS) { mn.z -= EPS; mx.z += EPS; }
        b.nodes[i].box.min = mn;
        b.nodes[i].box.max = mx;
    }
    mesh->nodes = (Node *)calloc(2 * (size_t)T + 1, sizeof(Node));
    bvh_build(&b, mesh->nodes, &mesh->node_count, &mesh->height, T, leaf);
    mesh->tris = (Tri *)calloc(T ? T : 1, sizeof(Tri));
    for (uint32_t i = 0; i < T; i++) {
        uint32_t t = b.leaf_ids[i];
        const uint32_t *ix = pidx + 3 * t;
        v3 v1 = V3(pos[3 * ix[0]], pos[3 * ix[0] + 1], pos[3 * ix[0] + 2]);
        v3 v2 = V3(pos[3 * ix[1]], pos[3 * ix[1] + 1], pos[3 * ix[1] + 2]);
        v3 v3_ = V3(pos[3 * ix[2]], pos[3 * ix[2] + 1], pos[3 * ix[2] + 2]);
        Tri *tr = &mesh->tris[i];
        tr->l2t.X = vsub(v3_, v1);
        tr->l2t.Y = vsub(v2, v1);
        tr->l2t.Z = vcross(tr->l2t.X, tr->l2t.Y);
        float area = vlen(tr->l2t.Z);
        tr->normal = tr->l2t.Z = vdivs(tr->l2t.Z, area);
        tr->position = v1;
        tr->l2t = m3_inverted(tr->l2t);
        tr->uv_coverage = 1.0f;
        if (mesh->normals_count) {
            const uint32_t *nx = nidx + 3 * t;
            tr->n1 = V3(nrm[3 * nx[0]], nrm[3 * nx[0] + 1], nrm[3 * nx[0] + 2]);
            tr->n2 = V3(nrm[3 * nx[1]], nrm[3 * nx[1] + 1], nrm[3 * nx[1] + 2]);
            tr->n3 = V3(nrm[3 * nx[2]], nrm[3 * nx[2] + 1], nrm[3 * nx[2] + 2]);
        }
        if (mesh->uvs_count) {
            const uint32_t *ux = uidx + 3 * t;
            tr->uv1.u = uv[2 * ux[0]]; tr->uv1.v = uv[2 * ux[0] + 1];
            tr->uv2.u = uv[2 * ux[1]]; tr->uv2.v = uv[2 * ux[1] + 1];
            tr->uv3.u = uv[2 * ux[2]]; tr->uv3.v = uv[2 * ux[2] + 1];
            float area_uv = (tr->uv2.u - tr->uv1.u) * (tr->uv3.v - tr->uv1.v) - (tr->uv3.u - tr->uv1.u) * (tr->uv2.v - tr->uv1.v);
            tr->uv_coverage = fabsf(area_uv / area);
        }
    }
    builder_free(&b);
}

/* ---------------------------------------------------------------- rays (src/slim/core/ray.h) */
typedef struct {
    v3 position, normal;
    v2 uv;
    float distance, uv_coverage, cone_width, NdotRd, scaling_factor;
    uint32_t id;
    int from_behind;
} Hit;

typedef struct {
    v3 origin, scaled_origin, direction, rcp;
    int sl, sb, sk;                 /* Sides: sign bits of x, y, z (left, bottom, back) */
} Ray;

static inline void ray_reset(Ray *r, v3 o, v3 d) {                                                        /* ray.h:40-49 */
    r->origin = o;
    r->direction = d;
    r->rcp = vsdiv(1.0f, d);
    r->scaled_origin = vmul(vneg(o), r->rcp);
    r->sl = signbit(d.x) ? 1 : 0;
    r->sb = signbit(d.y) ? 1 : 0;
    r->sk = signbit(d.z) ? 1 : 0;
}
static inline v3 ray_at(const Ray *r, float t) { return vscaleadd(r->direction, t, r->origin); }         /* ray.h:26 */

static inline void ray_localize(Ray *r, v3 o, v3 d, const Geo *g) {                                       /* ray.h:29-34 */
    v3 inv_scale = vsdiv(1.0f, g->scale);
    quat inv_rot = qconj(g->q);
    ray_reset(r, vmul(inv_scale, qrot(inv_rot, vsub(o, g->position))), vmul(inv_scale, qrot(inv_rot, d)));
}

static inline int hits_aabb(const Ray *r, const aabb *b, float *near_d, float *far_d) {                   /* ray.h:51-59 */
    v3 mn = V3(r->sl ? b->max.x : b->min.x, r->sb ? b->max.y : b->min.y, r->sk ? b->max.z : b->min.z);
    v3 mx = V3(r->sl ? b->min.x : b->max.x, r->sb ? b->min.y : b->max.y, r->sk ? b->min.z : b->max.z);
    mn = vmuladd(mn, r->rcp, r->scaled_origin);
    mx = vmuladd(mx, r->rcp, r->scaled_origin);
    float m = vmaximum(mn);
    *near_d = MAX2(0, m);
    *far_d = vminimum(mx);
    return *near_d <= *far_d;
}

static inline int hits_plane(const Ray *r, v3 po, v3 pn, Hit *hit) {                                      /* ray.h:61-82 */
    float NdotRd = vdot(pn, r->direction);
    if (NdotRd == 0) return 0;
    float NdotRoP = vdot(pn, vsub(po, r->origin));
    if (NdotRoP == 0) return 0;
    int facing = NdotRd < 0;
    hit->from_behind = NdotRoP > 0;
    if (hit->from_behind == facing) return 0;
    float t = NdotRoP / NdotRd;
    hit->distance = t;
    hit->position = ray_at(r, t);
    hit->normal = pn;
    return 1;
}

static inline int on_checkerboard(float u, float v) {                                                      /* base.h:277-279 */
    return ((uint8_t)(v * 4.0f)) % 2 == ((uint8_t)(u * 4.0f)) % 2;
}

static inline v2 uv_box_side(int side, float X, float Y, float Z) {                                       /* base.h:424-452 */
    float u, v;
    switch (side) {
        case SIDE_TOP: u = X; v = Z; break;
        case SIDE_BOTTOM: u = -X; v = -Z; break;
        case SIDE_LEFT: u = -Z; v = Y; break;
        case SIDE_RIGHT: u = Z; v = Y; break;
        case SIDE_FRONT: u = X; v = Y; break;
        default: u = -X; v = Y; break;
    }
    v2 r = {u * 0.5f + 0.5f, v * 0.5f + 0.5f};
    return r;
}

static inline v2 uv_sphere(float X, float Y, float Z) {                                                   /* base.h:454-483 */
    float u, v;
    float z_over_x = X ? (Z / X) : 2, y_over_x = X ? (Y / X) : 2;
    if (z_over_x <= 1 && z_over_x >= -1 && y_over_x <= 1 && y_over_x >= -1) {
        u = z_over_x;
        v = X > 0 ? y_over_x : -y_over_x;
    } else {
        float x_over_z = Z ? (X / Z) : 2, y_over_z = Z ? (Y / Z) : 2;
        if (x_over_z <= 1 && x_over_z >= -1 && y_over_z <= 1 && y_over_z >= -1) {
            u = -x_over_z;
            v = Z > 0 ? y_over_z : -y_over_z;
        } else {
            u = X / (Y > 0 ? Y : -Y);
            v = Z / Y;
        }
    }
    v2 r = {u * 0.5f + 0.5f, v * 0.5f + 0.5f};
    return r;
}

static int hits_quad(const Ray *r, Hit *hit, int transparent) {                                            /* ray.h:84-120 */
    if (r->direction.y == 0) return 0;
    if (r->origin.y == 0) return 0;
    hit->from_behind = r->origin.y < 0;
    if (hit->from_behind == (r->direction.y < 0)) return 0;
    float t = fabsf(r->origin.y * r->rcp.y);
    if (t > hit->distance) return 0;
    hit->position = ray_at(r, t);
    if (hit->position.x < -1 || hit->position.x > +1 || hit->position.z < -1 || hit->position.z > +1) return 0;
    hit->uv.u = hit->position.x * 0.5f + 0.5f;
    hit->uv.v = hit->position.z * 0.5f + 0.5f;
    if (transparent && on_checkerboard(hit->uv.u, hit->uv.v)) return 0;
    hit->distance = t;
    hit->normal = V3(0.0f, 1.0f, 0.0f);
    hit->uv_coverage = 0.25f;
    return 1;
}

static inline int box_side(int axis, int sl, int sb, int sk) {                                            /* base.h:542-548 */
    int sx = 3 * sl, sy = 3 * sb, sz = 3 * sk;
    return (axis == 1) * (1 << sx) | (axis == 2) * (2 << sy) | (axis == 4) * (4 << sz);
}

static int hits_box(const Ray *r, Hit *hit, int transparent) {                                             /* ray.h:122-175 */
    v3 sgn = V3(1.0f - 2.0f * (float)r->sl, 1.0f - 2.0f * (float)r->sb, 1.0f - 2.0f * (float)r->sk);
    v3 signed_rcp = vmul(sgn, r->rcp);
    v3 near_t = vsub(r->scaled_origin, signed_rcp);
    v3 far_t = vadd(r->scaled_origin, signed_rcp);
    int near_axis, far_axis, side;
    float far_hit_t = vminimum_axis(far_t, &far_axis);
    if (far_hit_t < 0) return SIDE_NONE;
    float near_hit_t = vmaximum_axis(near_t, &near_axis);
    if (near_hit_t > hit->distance || far_hit_t < (near_hit_t > 0 ? near_hit_t : 0)) return SIDE_NONE;
    float t = near_hit_t;
    hit->from_behind = near_hit_t < 0;
    if (hit->from_behind) {
        if (far_hit_t > hit->distance) return SIDE_NONE;
        t = far_hit_t;
        side = box_side(far_axis, !r->sl, !r->sb, !r->sk);
    } else
        side = box_side(near_axis, r->sl, r->sb, r->sk);
    hit->position = ray_at(r, t);
    hit->uv = uv_box_side(side, hit->position.x, hit->position.y, hit->position.z);
    if (transparent && on_checkerboard(hit->uv.u, hit->uv.v)) {
        if (hit->from_behind || far_hit_t > hit->distance) return SIDE_NONE;
        side = box_side(far_axis, !r->sl, !r->sb, !r->sk);
        hit->from_behind = 1;
        t = far_hit_t;
        hit->position = ray_at(r, t);
        hit->uv = uv_box_side(side, hit->position.x, hit->position.y, hit->position.z);
        if (on_checkerboard(hit->uv.u, hit->uv.v)) return SIDE_NONE;
    }
    hit->distance = t;
    hit->normal.x = (float)((int)(side == SIDE_RIGHT) - (int)(side == SIDE_LEFT));
    hit->normal.y = (float)((int)(side == SIDE_TOP) - (int)(side == SIDE_BOTTOM));
    hit->normal.z = (float)((int)(side == SIDE_FRONT) - (int)(side == SIDE_BACK));
    hit->uv_coverage = 0.25f;
    return side;
}

static int hits_sphere(const Ray *r, Hit *hit, int transparent, float *out_sq_dist) {                       /* ray.h:177-215 */
    float t_to_closest = -(vdot(r->origin, r->direction));
    if (t_to_closest <= 0) return 0;
    float dsq = vsqlen(r->direction);
    float sq_dist = vsqlen(r->origin) * dsq - t_to_closest * t_to_closest;
    if (sq_dist > 1.0f) return 0;
    float delta = sqrtf(dsq - sq_dist);
    dsq = 1.0f / dsq;
    float t = (t_to_closest - delta) * dsq;
    if (t > hit->distance) return 0;
    hit->normal = ray_at(r, t);
    hit->uv = uv_sphere(hit->normal.x, hit->normal.y, hit->normal.z);
    hit->from_behind = t <= 0 || (transparent && on_checkerboard(hit->uv.u, hit->uv.v));
    if (hit->from_behind) {
        t = (t_to_closest + delta) * dsq;
        if (t <= 0 || t > hit->distance) return 0;
        hit->normal = ray_at(r, t);
        hit->uv = uv_sphere(hit->normal.x, hit->normal.y, hit->normal.z);
        if (transparent && on_checkerboard(hit->uv.u, hit->uv.v)) return 0;
    }
    hit->distance = t;
    hit->uv_coverage = 1.0f / UNIT_SPHERE_AREA_OVER_SIX;
    if (out_sq_dist) *out_sq_dist = sq_dist * dsq;
    return 1;
}

static int hits_tet(const Ray *r, Hit *hit, int transparent) {                                             /* ray.h:217-311 */
    int found = 0;
    Hit cur;
    memset(&cur, 0, sizeof cur);
    cur.uv_coverage = SQRT3 / 4.0f;
    for (int t = 0; t < 4; t++) {
        v3 tp, fnrm;
        tp = fnrm = vs(t == 3 ? TET_MAX : -TET_MAX);
        switch (t) {
            case 0: fnrm.y = TET_MAX; break;
            case 1: fnrm.x = TET_MAX; break;
            case 2: fnrm.z = TET_MAX; break;
            case 3: tp.y = -TET_MAX; break;
        }
        cur.distance = hit->distance;
        if (!hits_plane(r, tp, fnrm, &cur)) continue;
        m3 tm;
        switch (t) {
            case 0: tm = M3(V3(TET_MAX, -TET_MIN, -TET_MAX), V3(TET_MIN, TET_MIN, TET_MAX), V3(-TET_MIN, TET_MAX, -TET_MAX)); break;
            case 1: tm = M3(V3(TET_MIN, TET_MIN, TET_MAX), V3(-TET_MIN, TET_MAX, -TET_MAX), V3(TET_MAX, -TET_MIN, -TET_MAX)); break;
            case 2: tm = M3(V3(-TET_MIN, TET_MAX, -TET_MAX), V3(TET_MAX, -TET_MIN, -TET_MAX), V3(TET_MIN, TET_MIN, TET_MAX)); break;
            default: tm = M3(V3(-TET_MAX, TET_MIN, TET_MAX), V3(TET_MIN, TET_MIN, TET_MAX), V3(TET_MIN, -TET_MAX, TET_MAX)); break;
        }
        tp = m3_mulv(tm, vsub(cur.position, tp));
        if (tp.x < 0 || tp.y < 0 || tp.y + tp.x > 1) continue;
        cur.uv.u = tp.x;
        cur.uv.v = tp.y;
        if (transparent && on_checkerboard(cur.uv.u, cur.uv.v)) continue;
        if (cur.distance < hit->distance) {
            *hit = cur;
            found = 1;
        }
    }
    return found;
}

/* ---------------------------------------------------------------- tracers (src/slim/scene/scene_tracer.h, mesh_tracer.h) */
typedef struct {
    Ray aux_ray;
    Hit aux_hit, triangle_hit;
    uint32_t *stack, *mesh_stack;
    float sph_b, sph_c, sph_tn, sph_tf, sph_tmax;   /* SphereTracer state (ray.h:314-346) */
    Counters *cnt;
} Tracer;

/* MeshTracer::hitTriangles (mesh_tracer.h:23-49) */
static int hit_triangles(Tracer *tr, const Tri *tris, uint32_t n, float closest, const Ray *r, Hit *hit, int any_hit) {
    int found = 0;
    closest = MIN2(closest, hit->distance);
    for (uint32_t i = 0; i < n; i++) {
        const Tri *t = &tris[i];
        tr->cnt->tris++;
        if (hits_plane(r, t->position, t->normal, &tr->triangle_hit)) {
            v3 UV = m3_mulv(t->l2t, vsub(tr->triangle_hit.position, t->position));
            if (UV.x < 0 || UV.y < 0 || (UV.x + UV.y) > 1 || tr->triangle_hit.distance >= closest) continue;
            closest = tr->triangle_hit.distance;
            *hit = tr->triangle_hit;
            hit->uv.u = UV.x;
            hit->uv.v = UV.y;
            hit->uv_coverage = t->uv_coverage;
            hit->id = i;
            found = 1;
            if (any_hit) break;
        }
    }
    return found;
}

/* MeshTracer::trace (mesh_tracer.h:51-134) */
static int mesh_trace(Tracer *tr, const MeshD *mesh, Ray *r, Hit *hit, int any_hit) {
    float ln, rn, lf, rf;
    int found = 0;
    tr->cnt->roots++;
    if (!(hits_aabb(r, &mesh->nodes[0].box, &ln, &lf) && ln < hit->distance)) return 0;
    if (mesh->nodes[0].leaf_count) return hit_triangles(tr, mesh->tris, mesh->tri_count, lf, r, hit, any_hit);
    const Node *left = mesh->nodes + mesh->nodes[0].first_index, *right;
    uint32_t top = 0;
    for (;;) {
        right = left + 1;
        tr->cnt->pairs++;
        int hl = hits_aabb(r, &left->box, &ln, &lf) && ln < hit->distance;
        int hr = hits_aabb(r, &right->box, &rn, &rf) && rn < hit->distance;
        if (hl) {
            if (left->leaf_count) {
                if (hit_triangles(tr, mesh->tris + left->first_index, left->leaf_count, lf, r, hit, any_hit)) {
                    hit->id += left->first_index;
                    found = 1;
                    if (any_hit) break;
                }
                left = NULL;
            }
        } else left = NULL;
        if (hr) {
            if (right->leaf_count) {
                if (hit_triangles(tr, mesh->tris + right->first_index, right->leaf_count, rf, r, hit, any_hit)) {
                    hit->id += right->first_index;
                    found = 1;
                    if (any_hit) break;
                }
                right = NULL;
            }
        } else right = NULL;
        if (left) {
            if (right) {
                if (!any_hit && ln > rn) { const Node *t = left; left = right; right = t; }
                tr->mesh_stack[top++] = right->first_index;
            }
            left = mesh->nodes + left->first_index;
        } else if (right) {
            left = mesh->nodes + right->first_index;
        } else {
            if (top == 0) break;
            left = mesh->nodes + tr->mesh_stack[--top];
        }
    }
    if (found && !any_hit && (mesh->normals_count | mesh->uvs_count)) {
        const Tri *t = &mesh->tris[hit->id];
        float a = hit->uv.u, b = hit->uv.v, c = 1 - a - b;
        tr->cnt->mesh_hits++;
        if (mesh->uvs_count) {
            hit->uv.u = t->uv3.u * a + (t->uv2.u * b + t->uv1.u * c);
            hit->uv.v = t->uv3.v * a + (t->uv2.v * b + t->uv1.v * c);
        }
        if (mesh->normals_count) {
            hit->normal.x = t->n3.x * a + (t->n2.x * b + t->n1.x * c);
            hit->normal.y = t->n3.y * a + (t->n2.y * b + t->n1.y * c);
            hit->normal.z = t->n3.z * a + (t->n2.z * b + t->n1.z * c);
        }
    }
    return found;
}

/* SceneTracer::hitGeometryInLocalSpace (scene_tracer.h:138-165) */
static int hit_geometry_local(Tracer *tr, const Scene *s, const Geo *g, const Ray *r, Hit *hit, int any_hit) {
    tr->cnt->insts++;
    ray_localize(&tr->aux_ray, r->origin, r->direction, g);
    aabb box;
    if (g->type == GT_MESH) box = s->meshes[g->id].box;
    else {
        box.max = vs(g->type == GT_TET ? TET_MAX : 1.0f);
        box.min = vs(-box.max.x);
        if (g->type == GT_QUAD) { box.min.y = -EPS; box.max.y = EPS; }
    }
    float n, f;
    if (!hits_aabb(&tr->aux_ray, &box, &n, &f)) return 0;
    int transparent = (g->flags & GEO_TRANSPARENT) != 0;
    switch (g->type) {
        case GT_QUAD: return hits_quad(&tr->aux_ray, hit, transparent);
        case GT_BOX: return hits_box(&tr->aux_ray, hit, transparent) != SIDE_NONE;
        case GT_SPHERE: return hits_sphere(&tr->aux_ray, hit, transparent, NULL);
        case GT_TET: return hits_tet(&tr->aux_ray, hit, transparent);
        case GT_MESH: return mesh_trace(tr, &s->meshes[g->id], &tr->aux_ray, hit, any_hit);
        default: return 0;
    }
}

/* SceneTracer::hitGeometries (scene_tracer.h:111-136); returns geometry index or -1 */
static int hit_geometries(Tracer *tr, const Scene *s, const uint32_t *ids, uint32_t n, float closest, const Ray *r, Hit *hit, int any_hit) {
    int hit_geo = -1;
    uint32_t vis = any_hit ? GEO_SHADOWING : GEO_VISIBLE;
    tr->aux_hit.distance = MIN2(closest + EPS, hit->distance);
    for (uint32_t i = 0; i < n; i++) {
        const Geo *g = &s->geos[ids[i]];
        if (!(g->flags & vis)) continue;
        if (hit_geometry_local(tr, s, g, r, &tr->aux_hit, any_hit)) {
            if (any_hit) return (int)ids[i];
            if (tr->aux_hit.distance < hit->distance) {
                hit_geo = (int)ids[i];
                *hit = tr->aux_hit;
                hit->NdotRd = -(vdot(hit->normal, tr->aux_ray.direction));
            }
        }
    }
    return hit_geo;
}

/* SceneTracer::trace (scene_tracer.h:26-98); returns geometry index or -1 */
static int scene_trace(Tracer *tr, const Scene *s, Ray *r, Hit *hit, int any_hit, float max_distance) {
    tr->cnt->rays++;
    ray_reset(r, vscaleadd(r->direction, TRACE_OFFSET, r->origin), r->direction);
    hit->distance = max_distance;
    float ln, rn, lf, rf;
    tr->cnt->roots++;
    if (!(hits_aabb(r, &s->snodes[0].box, &ln, &lf) && ln < hit->distance)) return -1;
    if (s->snodes[0].leaf_count) return hit_geometries(tr, s, s->leaf_ids, s->G, lf, r, hit, any_hit);
    const Node *left = s->snodes + s->snodes[0].first_index, *right;
    int closest = -1, g;
    uint32_t top = 0;
    for (;;) {
        right = left + 1;
        tr->cnt->pairs++;
        int hl = hits_aabb(r, &left->box, &ln, &lf) && ln < hit->distance;
        int hr = hits_aabb(r, &right->box, &rn, &rf) && rn < hit->distance;
        if (hl) {
            if (left->leaf_count) {
                g = hit_geometries(tr, s, s->leaf_ids + left->first_index, left->leaf_count, lf, r, hit, any_hit);
                if (g >= 0) { closest = g; if (any_hit) break; }
                left = NULL;
            }
        } else left = NULL;
        if (hr) {
            if (right->leaf_count) {
                g = hit_geometries(tr, s, s->leaf_ids + right->first_index, right->leaf_count, rf, r, hit, any_hit);
                if (g >= 0) { closest = g; if (any_hit) break; }
                right = NULL;
            }
        } else right = NULL;
        if (left) {
            if (right) {
                if (!any_hit && ln > rn) { const Node *t = left; left = right; right = t; }
                tr->stack[top++] = right->first_index;
            }
            left = s->snodes + left->first_index;
        } else if (right) {
            left = s->snodes + right->first_index;
        } else {
            if (top == 0) break;
            left = s->snodes + tr->stack[--top];
        }
    }
    return closest;
}

/* SphereTracer::hit (ray.h:317-336) via SceneTracer::hitLight (scene_tracer.h:100-109) */
static int hit_light(Tracer *tr, const Lite *l, const Ray *r, float *max_distance) {
    if (l->flags & LIGHT_DIRECTIONAL) return 0;
    float radius = l->intensity * LIGHT_INTENSITY_RADIUS_FACTOR;
    float oor = LIGHT_RADIUS_INTENSITY_FACTOR / l->intensity;
    tr->sph_tmax = *max_distance * oor;
    v3 rc = vscale(vsub(l->pos, r->origin), oor);
    tr->sph_b = vdot(r->direction, rc);
    tr->sph_c = vsqlen(rc) - 1;
    float h = tr->sph_b * tr->sph_b - tr->sph_c;
    if (h >= 0) {
        h = sqrtf(h);
        tr->sph_tn = tr->sph_b - h;
        tr->sph_tf = tr->sph_b + h;
        if (tr->sph_tf > 0 && tr->sph_tn < tr->sph_tmax) {
            *max_distance = tr->sph_tmax * radius;
            return 1;
        }
    }
    return 0;
}
static float integrate_density(const Tracer *tr) {                                                          /* ray.h:338-346 */
    float tn = MAX2(tr->sph_tn, 0), tf = MIN2(tr->sph_tf, tr->sph_tmax);
    float tn2 = tn * tn, tf2 = tf * tf, c = tr->sph_c, b = tr->sph_b;
    return (c * tn - b * tn2 + tn * tn2 / 3.0f - (c * tf - b * tf2 + tf * tf2 / 3.0f)) * (3.0f / 4.0f);
}

/* ---------------------------------------------------------------- textures (src/slim/core/texture.h) */
/* IEEE binary16 bits -> the float of the same value (every half is a float exactly). */
static float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t bits;
    if (e == 31) bits = sign | 0x7F800000u | (m << 13);                      /* infinity, NaN */
    else if (e) bits = sign | ((e + 112u) << 23) | (m << 13);                  /* normal */
    else if (!m) bits = sign;                                                 /* zero */
    else {                                                                    /* subnormal: m * 2^-24 */
        float f = (float)m * 5.9604644775390625e-8f;
        memcpy(&bits, &f, 4);
        bits |= sign;
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}
float oracle_half_to_float(uint16_t h) { return half_to_float(h); }

static color mip_sample(const Mip *m, float u, float v, Counters *cnt) {                                    /* texture.h:17-41 */
    cnt->taps++;
    if (u > 1) u -= (float)((uint64_t)u);
    if (v > 1) v -= (float)((uint64_t)v);
    const float U = u * (float)m->w + 0.5f, Vv = v * (float)m->h + 0.5f;
    const uint64_t x = (uint64_t)U, y = (uint64_t)Vv;
    const float r = U - (float)x, b = Vv - (float)y, l = 1 - r, t = 1 - b;
    if (m->fmt == 2) {   /* half texels: the same bilinear weights without the 1/255 scale, the same order */
        const float tl = t * l, tr = t * r, bl = b * l, br = b * r;
        const uint64_t row = (uint64_t)m->w + 2;
        const uint16_t *p = (const uint16_t *)m->quads + 4 * (y * row + x), *pb = p + 4 * row;
        color c;
        c.r = half_to_float(pb[4]) * br + (half_to_float(pb[0]) * bl + (half_to_float(p[4]) * tr + half_to_float(p[0]) * tl));
        c.g = half_to_float(pb[5]) * br + (half_to_float(pb[1]) * bl + (half_to_float(p[5]) * tr + half_to_float(p[1]) * tl));
        c.b = half_to_float(pb[6]) * br + (half_to_float(pb[2]) * bl + (half_to_float(p[6]) * tr + half_to_float(p[2]) * tl));
        return c;
    }
    const float tl = t * l * COLOR_COMPONENT_TO_FLOAT, tr = t * r * COLOR_COMPONENT_TO_FLOAT;
    const float bl = b * l * COLOR_COMPONENT_TO_FLOAT, br = b * r * COLOR_COMPONENT_TO_FLOAT;
    if (m->fmt == 1) {   /* RGBA8 texels: grid texel (x, y) is quad (x, y)'s top-left corner */
        const uint64_t row = (uint64_t)m->w + 2;
        const uint8_t *p = m->quads + 4 * (y * row + x), *pb = p + 4 * row;
        color c;
        c.r = (float)pb[4] * br + ((float)pb[0] * bl + ((float)p[4] * tr + (float)p[0] * tl));
        c.g = (float)pb[5] * br + ((float)pb[1] * bl + ((float)p[5] * tr + (float)p[1] * tl));
        c.b = (float)pb[6] * br + ((float)pb[2] * bl + ((float)p[6] * tr + (float)p[2] * tl));
        return c;
    }
    const uint8_t *q = m->quads + 12 * (y * (m->w + 1) + x);
    color c;
    c.r = (float)q[3] * br + ((float)q[2] * bl + ((float)q[1] * tr + (float)q[0] * tl));
    c.g = (float)q[7] * br + ((float)q[6] * bl + ((float)q[5] * tr + (float)q[4] * tl));
    c.b = (float)q[11] * br + ((float)q[10] * bl + ((float)q[9] * tr + (float)q[8] * tl));
    return c;
}
static uint32_t mip_level(float texel_area, uint32_t mip_count) {                                          /* texture.h:47-54 */
    uint32_t m = 0;
    while (texel_area > 1 && ++m < mip_count) texel_area *= 0.25f;
    if (m >= mip_count) m = mip_count - 1;
    return m;
}
static color tex_sample(const Tex *t, float u, float v, float coverage, Counters *cnt) {                    /* texture.h:68-70 */
    uint32_t m = (t->flags & IMG_MIPMAP) ? mip_level(coverage * (float)((uint64_t)t->width * t->height), t->mip_count) : 0;
    return mip_sample(&t->mips[m], u, v, cnt);
}
static color tex_sample_cube(const Tex *t, float X, float Y, float Z, Counters *cnt) {                      /* texture.h:72-116 */
    float u, v;
    int mip = 0;
    float right = signbit(X) ? 0.0f : 1.0f, left = signbit(X) ? 1.0f : 0.0f;
    float z_over_x = X ? (Z / X) : 2, y_over_x = X ? (Y / X) : 2;
    if (z_over_x <= 1 && z_over_x >= -1 && y_over_x <= 1 && y_over_x >= -1) {
        u = z_over_x * -0.125f + (0.5f * right + 0.125f);
        v = y_over_x * (left - 0.5f) + 0.5f;
    } else {
        float x_over_z = Z ? (X / Z) : 2, y_over_z = Z ? (Y / Z) : 2;
        if (x_over_z <= 1 && x_over_z >= -1 && y_over_z <= 1 && y_over_z >= -1) {
            float sz = signbit(Z) ? 1.0f : 0.0f;
            v = y_over_z * (sz - 0.5f) + +0.5f;
            u = x_over_z * 0.125f + (sz * 0.5f + 0.375f);
        } else {
            u = X / fabsf(Y) * 0.5f + 0.5f;
            v = Z / Y * 0.5f + 0.5f;
            mip = 2 - !signbit(Y);
        }
    }
    return mip_sample(&t->mips[mip], u, v, cnt);
}

/* ---------------------------------------------------------------- shading (src/slim/scene/material.h, renderer/surface_shader.h) */
static float ggx_D(float roughness, float NdotH) {                                                          /* material.h:27-36 */
    float a = roughness * roughness;
    float denom = NdotH * NdotH * (a - 1.0f) + 1.0f;
    return a / (PI_F * denom * denom);
}
static float ggx_G(float roughness, float NdotL, float NdotV) {                                              /* material.h:38-55 */
    float a = roughness * roughness, k = a * 0.5f, omk = 1.0f - k;
    float denom = NdotV * omk + k;
    float result = NdotV / fmaxf(denom, EPS);
    denom = NdotL * omk + k;
    result *= NdotL / fmaxf(denom, EPS);
    return result;
}
static color schlick(float HdotL, color F0) {                                                                /* material.h:57-59 */
    float p = powf(1.0f - HdotL, 5.0f);
    return cadd(F0, cscale(c1minus(F0), p));
}
static color cook_torrance(float roughness, float NdotL, float NdotV, float HdotL, float NdotH, color F0, color *F) { /* material.h:61-70 */
    *F = schlick(HdotL, F0);
    float D = ggx_D(roughness, NdotH), G = ggx_G(roughness, NdotL, NdotV);
    return cscale(*F, D * G / (4.0f * NdotL * NdotV));
}
static float tone_map(float x) {                                                                             /* base.h:777-786 */
    x = x - 0.004f;
    if (x < 0.0f) x = 0.0f;
    float x2s = x * x * 6.2f;
    return (x2s + x * 0.5f) / (x2s + x * 1.7f + 0.06f);
}

typedef struct {
    int geo;
    const Mat *mat;
    color F, albedo_from_map, Fs, Fd;
    Ray shadow_ray;
    Hit shadow_hit;
    v3 P, N, V, L, R, RF, H, quad_v[4];
    float Ld, Ld2, NdotL, NdotV, NdotH, HdotL, IOR;
    int refracted;
} Surface;

static color sample_slot(const Scene *s, const Mat *m, int slot, v2 uv, float coverage, Counters *cnt) {      /* surface_shader.h:9-11 */
    return (s->texs && m->texture_count > (uint32_t)slot) ? tex_sample(&s->texs[m->texture_ids[slot]], uv.u, uv.v, coverage, cnt) : cs(0);
}
static v3 rotate_normal(v3 normal, color sample, float magnitude) {                                          /* material.h:83-86 */
    v3 N = vnorm(vscaleadd(V3(sample.r, sample.b, sample.g), 2.0f, vs(-1.0f)));
    quat q = qaxisangle(vnorm(V3(N.z, 0, -N.x)), acosf(N.y) * magnitude);
    return qrot(q, normal);
}

static void radiance_fraction(Surface *sf) {                                                                  /* surface_shader.h:109-145 */
    const Mat *m = sf->mat;
    sf->Fs = cs(0);
    sf->Fd = cmul(m->albedo, sf->albedo_from_map);
    if (m->brdf == BRDF_CT) {
        sf->Fd = cscale(sf->Fd, (1.0f - m->metalness) * ONE_OVER_PI);
        if (sf->NdotV > 0.0f && m->roughness > 0.0f) {
            sf->H = vnorm(vadd(sf->L, sf->V));
            sf->NdotH = clamp01(vdot(sf->N, sf->H));
            sf->HdotL = clamp01(vdot(sf->H, sf->L));
            sf->Fs = cook_torrance(m->roughness, sf->NdotL, sf->NdotV, sf->HdotL, sf->NdotH, m->reflectivity, &sf->F);
            sf->Fd = cmul(sf->Fd, c1minus(sf->F));
        }
    } else {
        sf->Fd = cscale(sf->Fd, m->roughness * ONE_OVER_PI);
        if (m->brdf != BRDF_LAMBERT) {
            float spec, exponent;
            if (m->brdf == BRDF_PHONG) { exponent = 4.0f; spec = clamp01(vdot(sf->R, sf->L)); }
            else { exponent = 16.0f; spec = clamp01(vdot(sf->N, vnorm(vadd(sf->L, sf->V)))); }
            if (spec > 0.0f) sf->Fs = cscale(m->reflectivity, powf(spec, exponent) * (1.0f - m->roughness));
        }
    }
}

static int facing_light(Surface *sf, const Lite *l) {                                                         /* surface_shader.h:147-160 */
    if (l->flags & LIGHT_DIRECTIONAL) {
        sf->Ld = INFINITY;
        sf->Ld2 = 1.0f;
        sf->L = vneg(l->pos);
    } else {
        sf->L = vsub(l->pos, sf->P);
        sf->Ld2 = vsqlen(sf->L);
        sf->Ld = sqrtf(sf->Ld2);
        sf->L = vdivs(sf->L, sf->Ld);
    }
    sf->NdotL = clamp01(vdot(sf->L, sf->N));
    return sf->NdotL > 0.0f;
}

static void shade_from_light(Surface *sf, const Lite *l, const Scene *s, Tracer *tr, color *c) {           /* surface_shader.h:43-59 */
    if (facing_light(sf, l)) {
        int lit = 1;
        if (l->flags & LIGHT_SHADOWING) {
            sf->shadow_ray.origin = sf->P;
            sf->shadow_ray.direction = sf->L;
            lit = scene_trace(tr, s, &sf->shadow_ray, &sf->shadow_hit, 1, sf->Ld) < 0;
        }
        if (lit) {
            radiance_fraction(sf);
            *c = cmuladd(cadd(sf->Fs, sf->Fd), cscale(l->color, sf->NdotL * l->intensity / sf->Ld2), *c);
        }
    }
}

static void prepare_for_shading(Surface *sf, const Scene *s, const Ray *r, Hit *hit, Counters *cnt) {       /* surface_shader.h:61-107 */
    const Geo *g = &s->geos[sf->geo];
    v2 rep = s->mats[g->material_id].uv_repeat;
    hit->uv.u *= rep.u;
    hit->uv.v *= rep.v;
    if (hit->from_behind) hit->normal = vneg(hit->normal);
    hit->cone_width = hit->distance * hit->scaling_factor;
    hit->cone_width *= hit->cone_width;
    hit->cone_width *= hit->cone_width;
    hit->uv_coverage *= hit->cone_width / (rep.u * rep.v * hit->NdotRd * fabsf(vdot(vssub(1.0f, hit->normal), g->scale)));
    hit->position = ray_at(r, hit->distance);
    hit->normal = vnorm(qrot(g->q, vmul(g->scale, hit->normal)));
    sf->mat = &s->mats[g->material_id];
    const Mat *m = sf->mat;
    if (m->texture_count && (m->flags & MAT_NORMAL_MAP))
        hit->normal = rotate_normal(hit->normal, sample_slot(s, m, 1, hit->uv, hit->uv_coverage, cnt), m->normal_magnitude);
    sf->P = hit->position;
    sf->N = hit->normal;
    sf->R = sf->RF = vreflect(r->direction, sf->N);
    sf->V = vneg(r->direction);
    sf->NdotV = clamp01(vdot(sf->N, sf->V));
    sf->albedo_from_map = (m->texture_count && (m->flags & MAT_ALBEDO_MAP)) ? sample_slot(s, m, 0, hit->uv, hit->uv_coverage, cnt) : cs(1.0f);
    sf->refracted = (m->flags & MAT_REFRACTIVE) != 0;
    if (sf->refracted) {
        sf->IOR = hit->from_behind ? (m->ior / IOR_AIR) : (IOR_AIR / m->ior);
        float c = sf->IOR * sf->IOR * (1.0f - (sf->NdotV * sf->NdotV));
        sf->refracted = c < 1.0f;
        if (sf->refracted)
            sf->RF = vnorm(vscaleadd(sf->N, sf->IOR * sf->NdotV - sqrtf(1 - c), vscale(r->direction, sf->IOR)));
    }
}

static v3 area_light_vector(Surface *sf, const Geo *g) {                                                       /* surface_shader.h:162-186 */
    const float sx = g->scale.x, sz = g->scale.z;
    if (sx == 0 || sz == 0) return vs(0.0f);
    v3 U = qrot(g->q, V3(sx < 0 ? -sx : sx, 0.0f, 0.0f));
    v3 Vv = qrot(g->q, V3(0.0f, 0.0f, sz < 0 ? -sz : sz));
    sf->quad_v[0] = vsub(vsub(g->position, U), Vv);
    sf->quad_v[1] = vsub(vadd(g->position, U), Vv);
    sf->quad_v[2] = vadd(vadd(g->position, U), Vv);
    sf->quad_v[3] = vadd(vsub(g->position, U), Vv);
    v3 u1 = vnorm(vsub(sf->quad_v[0], sf->P)), u2 = vnorm(vsub(sf->quad_v[1], sf->P));
    v3 u3 = vnorm(vsub(sf->quad_v[2], sf->P)), u4 = vnorm(vsub(sf->quad_v[3], sf->P));
    return vadd(vadd(vadd(vscale(vcross(u1, u2), acosf(vdot(u1, u2)) * 0.5f),
                          vscale(vcross(u2, u3), acosf(vdot(u2, u3)) * 0.5f)),
                     vscale(vcross(u3, u4), acosf(vdot(u3, u4)) * 0.5f)),
                vscale(vcross(u4, u1), acosf(vdot(u4, u1)) * 0.5f));
}

static void shade_from_emissive_quads(Surface *sf, const Scene *s, color *c) {                                /* surface_shader.h:188-264 */
    for (uint32_t i = 0; i < s->G; i++) {
        const Geo *eq = &s->geos[i];
        if ((int)i == sf->geo || eq->type != GT_QUAD || !(s->mats[eq->material_id].flags & MAT_EMISSIVE)) continue;
        sf->L = vsub(eq->position, sf->P);
        sf->NdotL = vdot(sf->N, sf->L);
        if (sf->NdotL <= 0.0f) continue;
        sf->Ld2 = vsqlen(sf->L);
        sf->Ld = (float)sqrt((double)sf->Ld2);
        float Ld_rcp = 1.0f / sf->Ld;
        sf->L = vscale(sf->L, Ld_rcp);
        sf->NdotL *= Ld_rcp;
        v3 Ro = vscaleadd(sf->L, TRACE_OFFSET, sf->P);
        ray_localize(&sf->shadow_ray, Ro, sf->L, eq);
        sf->shadow_hit.distance = INFINITY;
        sf->shadow_ray.direction = vnorm(sf->shadow_ray.direction);
        (void)hits_quad(&sf->shadow_ray, &sf->shadow_hit, (eq->flags & GEO_TRANSPARENT) != 0);
        float emission_intensity = vdot(sf->N, area_light_vector(sf, eq));
        if (emission_intensity > 0) {
            int skip = 1;
            for (int k = 0; k < 4; k++)
                if (vdot(sf->N, vsub(sf->quad_v[k], sf->P)) >= 0) { skip = 0; break; }
            if (skip) continue;
            float shaded = 1.0f, sq_dist;
            for (uint32_t j = 0; j < s->G; j++) {
                const Geo *sg = &s->geos[j];
                if (j == i || (int)j == sf->geo) continue;
                ray_localize(&sf->shadow_ray, Ro, sf->L, sg);
                sf->shadow_hit.distance = INFINITY;
                sf->shadow_ray.direction = vnorm(sf->shadow_ray.direction);
                float d = 1.0f;
                if (sg->type == GT_SPHERE) {
                    if (hits_sphere(&sf->shadow_ray, &sf->shadow_hit, (sg->flags & GEO_TRANSPARENT) != 0, &sq_dist))
                        d -= (1.0f - sqrtf(sq_dist)) / (sf->shadow_hit.distance * emission_intensity * 3.0f);
                } else if (sg->type == GT_QUAD) {
                    if (hits_quad(&sf->shadow_ray, &sf->shadow_hit, (sg->flags & GEO_TRANSPARENT) != 0))
                        d -= 3.0f * (1.0f - MAX2(fabsf(sf->shadow_hit.position.x), fabsf(sf->shadow_hit.position.z))) /
                             (sf->shadow_hit.distance * emission_intensity);
                }
                shaded = MIN2(shaded, d);
            }
            if (shaded > 0.0f) {
                radiance_fraction(sf);
                *c = cmuladd(cadd(sf->Fd, sf->Fs), cscale(s->mats[eq->material_id].emission, emission_intensity * shaded * 7.0f), *c);
            }
        }
    }
}

static float depth_at(const Scene *s, v3 p) { return m3_mulv(s->proj.inv_rot, vsub(p, s->proj.camera_position)).z; } /* camera.h:66 */

/* Per-iteration record of one path, for the radiance cache's training targets (no counterpart in the reference):
 * the integrator below accumulates colour += throughput_i * current_i, throughput_{i+1} = throughput_i * next_i,
 * so the radiance leaving hit i is L_i = current_i + next_i * L_{i+1}.  Shaded hits give one sample each: the
 * 14-float hit record (position, shading normal, view direction, roughness, metalness, albedo times the albedo
 * map) and L_i. */
#define PATH_MAX_ITER 16
typedef struct {
    int n_iter;
    color cur[PATH_MAX_ITER], next[PATH_MAX_ITER];
    int shaded[PATH_MAX_ITER];
    float raw[PATH_MAX_ITER][14];
} PathRec;

/* renderPixelBeauty (src/slim/renderer/ray_tracer.h:19-108) */
static void render_beauty_rec(const Scene *s, Tracer *tr, int max_depth, v3 direction, float scaling, color *out,
                              float *depth, PathRec *rec) {
    Surface sf;
    memset(&sf, 0, sizeof sf);
    Ray ray;
    Hit hit;
    memset(&hit, 0, sizeof hit);
    hit.scaling_factor = scaling;
    color col = cs(0), cur, next_tp = cs(0), tp = cs(1.0f);
    *depth = INFINITY;
    ray_reset(&ray, s->proj.camera_position, vnorm(direction));
    uint32_t depth_left = (uint32_t)max_depth;
    if (rec) rec->n_iter = 0;
    while (depth_left) {
        cur = cs(0);
        int shaded = 0, bounced = 0;
        sf.geo = scene_trace(tr, s, &ray, &hit, 0, INFINITY);
        if (s->ray_cones) hit.scaling_factor = scaling;
        if (sf.geo >= 0) {
            const Geo *g = &s->geos[sf.geo];
            sf.mat = &s->mats[g->material_id];
            if ((sf.mat->flags & MAT_EMISSIVE) && g->type == GT_QUAD && !hit.from_behind) {
                depth_left = 0;
                cur = sf.mat->emission;
            } else {
                prepare_for_shading(&sf, s, &ray, &hit, tr->cnt);
                if (depth_left == (uint32_t)max_depth) *depth = depth_at(s, hit.position);
                shaded = 1;
                if (rec && rec->n_iter < PATH_MAX_ITER) {
                    float *r = rec->raw[rec->n_iter];
                    r[0] = sf.P.x; r[1] = sf.P.y; r[2] = sf.P.z; r[3] = sf.N.x; r[4] = sf.N.y; r[5] = sf.N.z;
                    r[6] = sf.V.x; r[7] = sf.V.y; r[8] = sf.V.z; r[9] = sf.mat->roughness; r[10] = sf.mat->metalness;
                    r[11] = sf.mat->albedo.r * sf.albedo_from_map.r;
                    r[12] = sf.mat->albedo.g * sf.albedo_from_map.g;
                    r[13] = sf.mat->albedo.b * sf.albedo_from_map.b;
                }
                for (uint32_t i = 0; i < s->L; i++) shade_from_light(&sf, &s->lights[i], s, tr, &cur);
                if (s->scene_flags & 1) shade_from_emissive_quads(&sf, s, &cur);
                if (s->sky_irradiance >= 0 && s->sky_radiance >= 0) {
                    sf.L = sf.N;
                    sf.NdotL = 1.0f;
                    color D = tex_sample_cube(&s->texs[s->sky_irradiance], sf.N.x, sf.N.y, sf.N.z, tr->cnt);
                    color S = tex_sample_cube(&s->texs[s->sky_radiance], sf.R.x, sf.R.y, sf.R.z, tr->cnt);
                    radiance_fraction(&sf);
                    cur = cmuladd(D, sf.Fd, cmuladd(sf.Fs, S, cur));
                }
                if ((sf.mat->flags & (MAT_REFLECTIVE | MAT_REFRACTIVE)) && --depth_left) {
                    sf.F = schlick(clamp01(vdot(sf.N, sf.R)), sf.mat->reflectivity);
                    next_tp = sf.refracted ? c1minus(sf.F) : sf.F;
                    ray_reset(&ray, hit.position, sf.RF);
                    bounced = 1;
                } else depth_left = 0;
            }
        } else {
            depth_left = 0;
            if (s->sky_color >= 0)
                cur = tex_sample_cube(&s->texs[s->sky_color], ray.direction.x, ray.direction.y, ray.direction.z, tr->cnt);
        }
        for (uint32_t i = 0; i < s->L; i++)
            if (hit_light(tr, &s->lights[i], &ray, &hit.distance)) {
                double p = pow((double)integrate_density(tr), 8.0) * 4;
                cur = cscaleadd(s->lights[i].color, (float)p, cur);
            }
        if (rec && rec->n_iter < PATH_MAX_ITER) {
            rec->cur[rec->n_iter] = cur;
            rec->next[rec->n_iter] = bounced ? next_tp : cs(0);
            rec->shaded[rec->n_iter] = shaded;
            rec->n_iter++;
        }
        col = cmuladd(cur, tp, col);
        tp = cmul(tp, next_tp);
    }
    out->r = tone_map(col.r);
    out->g = tone_map(col.g);
    out->b = tone_map(col.b);
}

static void render_beauty(const Scene *s, Tracer *tr, int max_depth, v3 direction, float scaling, color *out, float *depth) {
    render_beauty_rec(s, tr, max_depth, direction, scaling, out, depth, NULL);
}

static color mip_color(int level) {   /* RayTracingRenderer mip_level_colors (renderer.h:50-58), Color(ColorID) (base.h:805-858) */
    switch (level) {
        case 0: return C3(1.0f, 0.5f, 0.5f);    /* BrightRed */
        case 1: return C3(0.75f, 0.75f, 0.0f);  /* BrightYellow */
        case 2: return C3(0.5f, 1.0f, 0.5f);    /* BrightGreen */
        case 3: return C3(0.75f, 0.0f, 0.75f);  /* BrightMagenta */
        case 4: return C3(0.0f, 0.75f, 0.75f);  /* BrightCyan */
        case 5: return C3(0.5f, 0.5f, 1.0f);    /* BrightBlue */
        case 6: return cs(0.75f);               /* BrightGrey */
        case 7: return cs(0.5f);                /* Grey */
        default: return cs(0.25f);              /* DarkGrey */
    }
}

/* renderPixelDebugMode (src/slim/renderer/ray_tracer.h:110-142) */
static void render_debug(const Scene *s, Tracer *tr, int mode, v3 direction, float scaling, color *out, float *depth) {
    Surface sf;
    memset(&sf, 0, sizeof sf);
    Ray ray;
    Hit hit;
    memset(&hit, 0, sizeof hit);
    hit.scaling_factor = scaling;
    *out = cs(0);
    *depth = INFINITY;
    ray_reset(&ray, s->proj.camera_position, vnorm(direction));
    sf.geo = scene_trace(tr, s, &ray, &hit, 0, INFINITY);
    if (s->ray_cones) hit.scaling_factor = scaling;
    if (sf.geo >= 0) {
        prepare_for_shading(&sf, s, &ray, &hit, tr->cnt);
        *depth = depth_at(s, hit.position);
        switch (mode) {
            case RM_UVS: *out = C3(hit.uv.u, hit.uv.v, 1.0f); break;
            case RM_DEPTH: *out = cs(4.0f / hit.distance); break;
            case RM_NORMALS: *out = C3(hit.normal.x * 0.5f + 0.5f, hit.normal.y * 0.5f + 0.5f, hit.normal.z * 0.5f + 0.5f); break;
            case RM_NORMAL_MAP: {
                color n = sample_slot(s, sf.mat, 1, hit.uv, hit.uv_coverage, tr->cnt);
                *out = C3(n.r * 0.5f + 0.5f, n.g * 0.5f + 0.5f, n.b * 0.5f + 0.5f);
            } break;
            case RM_MIP_LEVEL:
                *out = s->NT ? mip_color((int)mip_level(hit.uv_coverage * (float)((uint64_t)s->texs[0].width * s->texs[0].height), s->texs[0].mip_count)) : cs(0.5f);
                break;
            default: break;
        }
    }
}

/* closed-form primary ray of d_render (src/slim/renderer/renderer_GPU.h:36-43) */
static void primary_ray(const Scene *s, int x, int y, v3 *dir, float *scaling) {
    const Projection *p = &s->proj;
    float a = (float)x * p->sample_size + p->c_start.u;
    float b = (float)(-y) * p->sample_size + p->c_start.v;
    *scaling = 1.0f / sqrtf(p->sq_dist + (a * a + b * b));
    *dir = vadd(vadd(p->start, vscale(p->down, (float)y)), vscale(p->right, (float)x));
}

/* ---------------------------------------------------------------- scene derivation */
static void update_aabbs(Scene *s) {                                                                          /* scene.h:165-179, transform.h:33-48 */
    for (uint32_t i = 0; i < s->G; i++) {
        const Geo *g = &s->geos[i];
        aabb b;
        if (g->type == GT_MESH) b = s->meshes[g->id].box;
        else {
            b.max = vs(g->type == GT_TET ? TET_MAX : 1.0f);
            b.min = vs(-b.max.x);
            if (g->type == GT_QUAD) b.min.y = b.max.y = 0.0f;
        }
        v3 mn = vs(INFINITY), mx = vs(-INFINITY);
        for (int i2 = 0; i2 < 2; i2++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) {
                    v3 p = V3(i2 ? b.max.x : b.min.x, j ? b.max.y : b.min.y, k ? b.max.z : b.min.z);
                    p = vadd(qrot(g->q, vmul(g->scale, p)), g->position);
                    mn = vmin(mn, p);
                    mx = vmax(mx, p);
                }
        s->geo_aabbs[i].min = mn;
        s->geo_aabbs[i].max = mx;
    }
}

static void update_scene_bvh(Scene *s, uint32_t max_leaf) {                                                   /* scene.h:181-191 */
    Builder b;
    builder_init(&b, s->G);
    for (uint32_t i = 0; i < s->G; i++) {
        b.nodes[i].box = s->geo_aabbs[i];
        b.nodes[i].first_index = b.node_ids[i] = i;
    }
    bvh_build(&b, s->snodes, &s->snode_count, &s->sheight, s->G, max_leaf);
    for (uint32_t i = 0; i < s->G; i++) s->leaf_ids[i] = b.leaf_ids[i];
    builder_free(&b);
}

static void projection_reset(Scene *s, const float *cam) {                                                     /* camera.h:69-83 */
    m3 o = m3_from_euler(cam[0], cam[1], cam[2]);
    float fw = (float)(uint16_t)s->width, fh = (float)(uint16_t)s->height;
    float hw = fw * 0.5f, hh = fh * 0.5f;
    Projection *p = &s->proj;
    p->sample_size = s->ssaa ? 0.5f : 1.0f;
    p->sq_dist = hh * cam[6];
    p->c_start.u = (p->sample_size * 0.5f) - hw;
    p->c_start.v = hh - (p->sample_size * 0.5f);
    p->inv_rot = m3_inverted(o);
    p->camera_position = V3(cam[3], cam[4], cam[5]);
    p->down = vscale(vneg(o.Y), p->sample_size);
    p->right = vscale(o.X, p->sample_size);
    p->start = vadd(vadd(vscale(o.X, p->c_start.u), vscale(o.Y, p->c_start.v)), vscale(o.Z, p->sq_dist));
    p->sq_dist *= p->sq_dist;
}

/* ---------------------------------------------------------------- exported API */
void *oracle_scene_load(const char *path, int ray_cones) {
    Scene *s = (Scene *)calloc(1, sizeof(Scene));
    if (spec_open(&s->sf, path)) { free(s); return NULL; }
    const int32_t *fr = (const int32_t *)spec_get(&s->sf, "frame", NULL);
    const float *cam = (const float *)spec_get(&s->sf, "camera", NULL);
    s->width = fr[0]; s->height = fr[1]; s->ssaa = fr[2]; s->max_depth = fr[3]; s->render_mode = fr[4];
    s->sky_color = (int)(signed char)fr[5]; s->sky_radiance = (int)(signed char)fr[6]; s->sky_irradiance = (int)(signed char)fr[7];
    s->G = (uint32_t)fr[8]; s->M = (uint32_t)fr[9]; s->L = (uint32_t)fr[10]; s->NM = (uint32_t)fr[11]; s->NT = (uint32_t)fr[12];
    s->ray_cones = ray_cones;
    const float *gf = (const float *)spec_get(&s->sf, "geo_f", NULL);
    const uint32_t *gu = (const uint32_t *)spec_get(&s->sf, "geo_u", NULL);
    /* geo_q: the orientations as quaternions (axis, amount), as an app holds them after composing rotations
     * (Transform::orientation *= rot, quat.h:56-65); without it they come from the Euler angles */
    const float *gq = (const float *)spec_get(&s->sf, "geo_q", NULL);
    s->geos = (Geo *)calloc(s->G, sizeof(Geo));
    for (uint32_t i = 0; i < s->G; i++) {
        const float *f = gf + 9 * i;
        if (gq) {
            s->geos[i].q.axis = V3(gq[4 * i], gq[4 * i + 1], gq[4 * i + 2]);
            s->geos[i].q.w = gq[4 * i + 3];
        } else {
            s->geos[i].q = quat_from_euler(f[0], f[1], f[2]);
        }
        s->geos[i].position = V3(f[3], f[4], f[5]);
        s->geos[i].scale = V3(f[6], f[7], f[8]);
        s->geos[i].type = gu[4 * i]; s->geos[i].material_id = gu[4 * i + 1];
        s->geos[i].id = gu[4 * i + 2]; s->geos[i].flags = gu[4 * i + 3] & 0xFFu;
    }
    const float *mf = (const float *)spec_get(&s->sf, "mat_f", NULL);
    const uint32_t *mu = (const uint32_t *)spec_get(&s->sf, "mat_u", NULL);
    s->mats = (Mat *)calloc(s->M ? s->M : 1, sizeof(Mat));
    for (uint32_t i = 0; i < s->M; i++) {
        const float *f = mf + 16 * i;
        const uint32_t *u = mu + 5 * i;
        Mat *m = &s->mats[i];
        m->albedo = C3(f[0], f[1], f[2]); m->roughness = f[3]; m->normal_magnitude = f[4];
        m->uv_repeat.u = f[5]; m->uv_repeat.v = f[6]; m->emission = C3(f[7], f[8], f[9]); m->ior = f[10];
        m->reflectivity = C3(f[11], f[12], f[13]); m->metalness = f[14];
        m->flags = u[0] & 0xFFu; m->texture_count = u[1] & 0xFFu; m->texture_ids[0] = u[2] & 0xFFu;
        m->texture_ids[1] = u[3] & 0xFFu; m->brdf = u[4];
    }
    if (s->L) {
        const float *lf = (const float *)spec_get(&s->sf, "light_f", NULL);
        const uint32_t *lu = (const uint32_t *)spec_get(&s->sf, "light_u", NULL);
        s->lights = (Lite *)calloc(s->L, sizeof(Lite));
        for (uint32_t i = 0; i < s->L; i++) {
            s->lights[i].color = C3(lf[7 * i], lf[7 * i + 1], lf[7 * i + 2]);
            s->lights[i].pos = V3(lf[7 * i + 3], lf[7 * i + 4], lf[7 * i + 5]);
            s->lights[i].intensity = lf[7 * i + 6];
            s->lights[i].flags = lu[i] & 0xFFu;
        }
    }
    s->meshes = (MeshD *)calloc(s->NM ? s->NM : 1, sizeof(MeshD));
    for (uint32_t k = 0; k < s->NM; k++) {
        char nm[32];
        snprintf(nm, sizeof nm, "mesh%u_info", k);
        const uint32_t *info = (const uint32_t *)spec_get(&s->sf, nm, NULL);
        MeshD *mesh = &s->meshes[k];
        mesh->tri_count = info[1]; mesh->normals_count = info[2]; mesh->uvs_count = info[3];
        snprintf(nm, sizeof nm, "mesh%u_pos", k);
        const float *pos = (const float *)spec_get(&s->sf, nm, NULL);
        snprintf(nm, sizeof nm, "mesh%u_pidx", k);
        const uint32_t *pidx = (const uint32_t *)spec_get(&s->sf, nm, NULL);
        snprintf(nm, sizeof nm, "mesh%u_nrm", k);
        const float *nrm = (const float *)spec_get(&s->sf, nm, NULL);
        snprintf(nm, sizeof nm, "mesh%u_nidx", k);
        const uint32_t *nidx = (const uint32_t *)spec_get(&s->sf, nm, NULL);
        snprintf(nm, sizeof nm, "mesh%u_uv", k);
        const float *uv = (const float *)spec_get(&s->sf, nm, NULL);
        snprintf(nm, sizeof nm, "mesh%u_uidx", k);
        const uint32_t *uidx = (const uint32_t *)spec_get(&s->sf, nm, NULL);
        mesh->box.min = vs(INFINITY);
        mesh->box.max = vs(-INFINITY);
        for (uint32_t v = 0; v < info[0]; v++) {
            v3 p = V3(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]);
            mesh->box.min = vmin(mesh->box.min, p);
            mesh->box.max = vmax(mesh->box.max, p);
        }
        build_mesh(mesh, pos, pidx, nrm, nidx, uv, uidx, info[4]);
    }
    s->texs = s->NT ? (Tex *)calloc(s->NT, sizeof(Tex)) : NULL;
    for (uint32_t k = 0; k < s->NT; k++) {
        char nm[32];
        snprintf(nm, sizeof nm, "tex%u_info", k);
        const uint32_t *info = (const uint32_t *)spec_get(&s->sf, nm, NULL);
        snprintf(nm, sizeof nm, "tex%u_dims", k);
        const uint32_t *dims = (const uint32_t *)spec_get(&s->sf, nm, NULL);
        Tex *t = &s->texs[k];
        t->width = info[0]; t->height = info[1]; t->flags = info[2]; t->mip_count = info[3];
        t->mips = (Mip *)calloc(t->mip_count, sizeof(Mip));
        for (uint32_t m = 0; m < t->mip_count; m++) {
            snprintf(nm, sizeof nm, "tex%u_mip%u", k, m);
            t->mips[m].w = dims[2 * m];
            t->mips[m].h = dims[2 * m + 1];
            t->mips[m].quads = (const uint8_t *)spec_get(&s->sf, nm, NULL);
            t->mips[m].fmt = (t->flags >> 24) & 3u;
        }
    }
    for (uint32_t i = 0; i < s->G; i++)
        if (s->geos[i].type == GT_QUAD && (s->mats[s->geos[i].material_id].flags & MAT_EMISSIVE)) s->scene_flags = 1;
    s->geo_aabbs = (aabb *)calloc(s->G, sizeof(aabb));
    s->snodes = (Node *)calloc(2 * (size_t)s->G + 1, sizeof(Node));
    s->leaf_ids = (uint32_t *)calloc(s->G, 4);
    update_aabbs(s);
    update_scene_bvh(s, 1);
    projection_reset(s, cam);
    return s;
}

void oracle_scene_free(void *h) {
    Scene *s = (Scene *)h;
    if (!s) return;
    for (uint32_t k = 0; k < s->NM; k++) { free(s->meshes[k].nodes); free(s->meshes[k].tris); }
    for (uint32_t k = 0; k < s->NT; k++) free(s->texs[k].mips);
    free(s->geos); free(s->mats); free(s->lights); free(s->meshes); free(s->texs);
    free(s->geo_aabbs); free(s->snodes); free(s->leaf_ids);
    spec_close(&s->sf);
    free(s);
}

static uint32_t max_mesh_height(const Scene *s) {
    uint32_t h = 0;
    for (uint32_t k = 0; k < s->NM; k++) h = s->meshes[k].height > h ? s->meshes[k].height : h;
    return h;
}

/* Renders rows [y0, y1) (sample rows when SSAA).  mode/max_depth < 0 take the spec's values.
 * colors: 3 floats per sample (pre-canvas, tone-mapped beauty or debug colour); depths: 1 float.
 * counters: 8 x u64 = rays, pairs, roots, instances, triangles, mesh closest-hit fetches, texel taps, pixels. */
int oracle_render(void *h, int y0, int y1, int mode, int max_depth, float *colors, float *depths, uint64_t *counters) {
    Scene *s = (Scene *)h;
    if (mode < 0) mode = s->render_mode;
    if (max_depth < 0) max_depth = s->max_depth;
    int Ws = s->width * (s->ssaa ? 2 : 1);
    Counters cnt;
    memset(&cnt, 0, sizeof cnt);
    Tracer *tr = (Tracer *)calloc(1, sizeof(Tracer));
    tr->stack = (uint32_t *)calloc(s->G + 2, 4);
    tr->mesh_stack = (uint32_t *)calloc(max_mesh_height(s) + 2, 4);
    tr->cnt = &cnt;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < Ws; x++) {
            size_t o = (size_t)(y - y0) * Ws + x;
            v3 dir;
            float scaling, depth;
            color c;
            primary_ray(s, x, y, &dir, &scaling);
            if (mode == RM_BEAUTY) render_beauty(s, tr, max_depth, dir, scaling, &c, &depth);
            else render_debug(s, tr, mode, dir, scaling, &c, &depth);
            cnt.pixels++;
            if (colors) { colors[3 * o] = c.r; colors[3 * o + 1] = c.g; colors[3 * o + 2] = c.b; }
            if (depths) depths[o] = depth;
        }
    if (counters) memcpy(counters, &cnt, sizeof cnt);
    free(tr->stack); free(tr->mesh_stack); free(tr);
    return 0;
}

/* Training samples of sample (x, y)'s beauty path (at most 16): raw (14 floats each), L (3 floats each) and the
 * path vertex (1 = primary hit) of every shaded hit, in path order.  Returns the number of samples. */
int oracle_path_samples(void *h, int x, int y, int max_depth, float *raw, float *L, int32_t *vertex) {
    Scene *s = (Scene *)h;
    if (max_depth < 0) max_depth = s->max_depth;
    Counters cnt;
    memset(&cnt, 0, sizeof cnt);
    Tracer *tr = (Tracer *)calloc(1, sizeof(Tracer));
    tr->stack = (uint32_t *)calloc(s->G + 2, 4);
    tr->mesh_stack = (uint32_t *)calloc(max_mesh_height(s) + 2, 4);
    tr->cnt = &cnt;
    PathRec *rec = (PathRec *)calloc(1, sizeof(PathRec));
    v3 dir;
    float scaling, depth;
    color c;
    primary_ray(s, x, y, &dir, &scaling);
    render_beauty_rec(s, tr, max_depth, dir, scaling, &c, &depth, rec);
    color Li = cs(0);
    int n = 0;
    for (int i = rec->n_iter - 1; i >= 0; i--) {
        Li = C3(rec->cur[i].r + rec->next[i].r * Li.r, rec->cur[i].g + rec->next[i].g * Li.g,
                rec->cur[i].b + rec->next[i].b * Li.b);
        rec->cur[i] = Li;   /* now L_i */
    }
    for (int i = 0; i < rec->n_iter; i++) {
        if (!rec->shaded[i]) continue;
        memcpy(raw + 14 * n, rec->raw[i], 14 * sizeof(float));
        L[3 * n] = rec->cur[i].r; L[3 * n + 1] = rec->cur[i].g; L[3 * n + 2] = rec->cur[i].b;
        vertex[n] = i + 1;
        n++;
    }
    free(rec); free(tr->stack); free(tr->mesh_stack); free(tr);
    return n;
}

/* Per-iteration path counts of beauty rows [y0, y1): counts[i] += the samples whose path runs iteration i of
 * renderPixelBeauty's loop (i = 0 for the camera ray; src/slim/renderer/ray_tracer.h:42-105), for i < n.
 * max_depth <= 16.  The wavefront frame's path counters are checked against these. */
int oracle_iteration_counts(void *h, int y0, int y1, int max_depth, uint64_t *counts, int n) {
    Scene *s = (Scene *)h;
    if (max_depth < 0) max_depth = s->max_depth;
    if (max_depth > PATH_MAX_ITER) return -1;
    int Ws = s->width * (s->ssaa ? 2 : 1);
    Counters cnt;
    memset(&cnt, 0, sizeof cnt);
    Tracer *tr = (Tracer *)calloc(1, sizeof(Tracer));
    tr->stack = (uint32_t *)calloc(s->G + 2, 4);
    tr->mesh_stack = (uint32_t *)calloc(max_mesh_height(s) + 2, 4);
    tr->cnt = &cnt;
    PathRec *rec = (PathRec *)calloc(1, sizeof(PathRec));
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < Ws; x++) {
            v3 dir;
            float scaling, depth;
            color c;
            primary_ray(s, x, y, &dir, &scaling);
            render_beauty_rec(s, tr, max_depth, dir, scaling, &c, &depth, rec);
            for (int i = 0; i < rec->n_iter && i < n; i++) counts[i]++;
        }
    free(rec); free(tr->stack); free(tr->mesh_stack); free(tr);
    return 0;
}

/* Window content of a canvas (Canvas::drawToWindow / getPixelContent, src/slim/draw/canvas.h:119-126, 187-190,
 * 254-265; Pixel::asContent, src/slim/core/base.h:1166-1171): per window pixel, under SSAA the mean of its 2x2
 * sample quad (((p0 + p1) + p2) + p3) * 0.25 - 0 when all four opacities are 0 - else the pixel itself (0 when
 * its opacity is 0); then 0xRRGGBB with each channel 255 * sqrt(c) truncated, 255 above 1.  pixels: RGBA floats
 * in the canvas layout (2x2 quads per pixel under SSAA). */
int oracle_canvas_content(const float *pixels, int width, int height, int ssaa, uint32_t *content) {
    const int step = ssaa ? 4 : 1;
    for (int i = 0; i < width * height; i++) {
        const float *p = pixels + (size_t)4 * step * i;
        float c[4];
        if (ssaa) {
            if (p[3] == 0.0f && p[7] == 0.0f && p[11] == 0.0f && p[15] == 0.0f) { content[i] = 0; continue; }
            for (int k = 0; k < 4; k++) c[k] = (((p[k] + p[4 + k]) + p[8 + k]) + p[12 + k]) * 0.25f;
        } else {
            if (p[3] == 0.0f) { content[i] = 0; continue; }
            for (int k = 0; k < 4; k++) c[k] = p[k];
        }
        uint32_t v = 0;
        for (int k = 0; k < 3; k++) {
            const uint32_t b = c[k] > 1.0f ? 255u : (uint32_t)(uint8_t)(255.0f * sqrtf(c[k]));
            v = (v << 8) | b;
        }
        content[i] = v;
    }
    return 0;
}

/* Primary closest hit per sample: geometry index (-1 miss), leaf-ordered triangle id (-1 if not a mesh), t. */
int oracle_primary_hits(void *h, int y0, int y1, int32_t *geo, int32_t *tri, float *t) {
    Scene *s = (Scene *)h;
    int Ws = s->width * (s->ssaa ? 2 : 1);
    Counters cnt;
    memset(&cnt, 0, sizeof cnt);
    Tracer *tr = (Tracer *)calloc(1, sizeof(Tracer));
    tr->stack = (uint32_t *)calloc(s->G + 2, 4);
    tr->mesh_stack = (uint32_t *)calloc(max_mesh_height(s) + 2, 4);
    tr->cnt = &cnt;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < Ws; x++) {
            size_t o = (size_t)(y - y0) * Ws + x;
            v3 dir;
            float scaling;
            primary_ray(s, x, y, &dir, &scaling);
            Ray ray;
            Hit hit;
            memset(&hit, 0, sizeof hit);
            ray_reset(&ray, s->proj.camera_position, vnorm(dir));
            int g = scene_trace(tr, s, &ray, &hit, 0, INFINITY);
            geo[o] = g;
            tri[o] = (g >= 0 && s->geos[g].type == GT_MESH) ? (int32_t)hit.id : -1;
            t[o] = g >= 0 ? hit.distance : INFINITY;
        }
    free(tr->stack); free(tr->mesh_stack); free(tr);
    return 0;
}

/* Derived data, under the same names the reference driver writes (oracle/ref_driver.cpp).
 * Returns the element count (floats or u32) and copies at most cap elements. */
static size_t put(void *out, size_t cap, const void *src, size_t n, size_t elem) {
    if (out) memcpy(out, src, (n < cap ? n : cap) * elem);
    return n;
}
static size_t put_nodes(const Node *nodes, uint32_t count, int part, void *out, size_t cap) {
    size_t per = part == 0 ? 6 : 3, n = per * count;
    if (!out) return n;
    for (uint32_t i = 0; i < count && (i + 1) * per <= cap; i++) {
        if (part == 0) {
            float v[6] = {nodes[i].box.min.x, nodes[i].box.min.y, nodes[i].box.min.z, nodes[i].box.max.x, nodes[i].box.max.y, nodes[i].box.max.z};
            memcpy((float *)out + 6 * i, v, sizeof v);
        } else {
            uint32_t u[3] = {nodes[i].first_index, nodes[i].leaf_count, nodes[i].depth};
            memcpy((uint32_t *)out + 3 * i, u, sizeof u);
        }
    }
    return n;
}

size_t oracle_get(void *h, const char *name, void *out, size_t cap) {
    Scene *s = (Scene *)h;
    if (!strcmp(name, "geo_quat")) {
        size_t n = 4 * (size_t)s->G;
        if (out) for (uint32_t i = 0; i < s->G && 4 * (i + 1) <= cap; i++) {
            float q[4] = {s->geos[i].q.axis.x, s->geos[i].q.axis.y, s->geos[i].q.axis.z, s->geos[i].q.w};
            memcpy((float *)out + 4 * i, q, sizeof q);
        }
        return n;
    }
    if (!strcmp(name, "geo_aabb")) return put(out, cap, s->geo_aabbs, 6 * (size_t)s->G, 4);
    if (!strcmp(name, "sbvh_f")) return put_nodes(s->snodes, s->snode_count, 0, out, cap);
    if (!strcmp(name, "sbvh_u")) return put_nodes(s->snodes, s->snode_count, 1, out, cap);
    if (!strcmp(name, "sbvh_meta")) { uint32_t m[2] = {s->snode_count, s->sheight}; return put(out, cap, m, 2, 4); }
    if (!strcmp(name, "sbvh_leaf")) return put(out, cap, s->leaf_ids, s->G, 4);
    if (!strcmp(name, "proj")) {
        const Projection *p = &s->proj;
        float v[25] = {p->inv_rot.X.x, p->inv_rot.X.y, p->inv_rot.X.z, p->inv_rot.Y.x, p->inv_rot.Y.y, p->inv_rot.Y.z,
                       p->inv_rot.Z.x, p->inv_rot.Z.y, p->inv_rot.Z.z, p->start.x, p->start.y, p->start.z,
                       p->right.x, p->right.y, p->right.z, p->down.x, p->down.y, p->down.z,
                       p->camera_position.x, p->camera_position.y, p->camera_position.z,
                       p->c_start.u, p->c_start.v, p->sq_dist, p->sample_size};
        return put(out, cap, v, 25, 4);
    }
    unsigned k;
    char rest[24];
    if (sscanf(name, "mesh%u_%23s", &k, rest) == 2 && k < s->NM) {
        const MeshD *m = &s->meshes[k];
        if (!strcmp(rest, "bvh_f")) return put_nodes(m->nodes, m->node_count, 0, out, cap);
        if (!strcmp(rest, "bvh_u")) return put_nodes(m->nodes, m->node_count, 1, out, cap);
        if (!strcmp(rest, "bvh_meta")) { uint32_t mm[2] = {m->node_count, m->height}; return put(out, cap, mm, 2, 4); }
        if (!strcmp(rest, "tri")) return put(out, cap, m->tris, 32 * (size_t)m->tri_count, 4);
        if (!strcmp(rest, "aabb")) return put(out, cap, &m->box, 6, 4);
    }
    return 0;
}

/* Rebuilds the scene BVH with a different leaf size; max_leaf >= G puts every geometry in the root leaf, which
 * turns SceneTracer::trace into a linear scan over all geometries (the brute-force check of traversal). */
void oracle_set_scene_leaf_size(void *h, uint32_t max_leaf) { update_scene_bvh((Scene *)h, max_leaf); }

/* ---------------------------------------------------------------- unit-level entry points for closed-form tests */
float oracle_tone_map(float x) { return tone_map(x); }
float oracle_ggx_D(float r, float NdotH) { return ggx_D(r, NdotH); }
float oracle_ggx_G(float r, float NdotL, float NdotV) { return ggx_G(r, NdotL, NdotV); }
void oracle_schlick(float HdotL, const float *F0, float *out) {
    color c = schlick(HdotL, C3(F0[0], F0[1], F0[2]));
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}
uint32_t oracle_mip_level(float area, uint32_t count) { return mip_level(area, count); }
void oracle_quat_from_euler(float x, float y, float z, float *out) {
    quat q = quat_from_euler(x, y, z);
    out[0] = q.axis.x; out[1] = q.axis.y; out[2] = q.axis.z; out[3] = q.w;
}
void oracle_mip_sample(uint32_t w, uint32_t h, const uint8_t *quads, float u, float v, float *out) {
    Counters cnt;
    Mip m = {w, h, quads, 0};
    color c = mip_sample(&m, u, v, &cnt);
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}
/* The same for a mip in texel storage fmt (0 quads, 1 RGBA8, 2 RGBA16F). */
void oracle_mip_sample_fmt(uint32_t w, uint32_t h, const uint8_t *data, uint32_t fmt, float u, float v, float *out) {
    Counters cnt;
    Mip m = {w, h, data, fmt};
    color c = mip_sample(&m, u, v, &cnt);
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}
/* Slab test of a ray (o, d) against a box; returns hit flag, writes near/far. */
int oracle_hits_aabb(const float *o, const float *d, const float *box, float *nf) {
    Ray r;
    ray_reset(&r, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]));
    aabb b = {V3(box[0], box[1], box[2]), V3(box[3], box[4], box[5])};
    return hits_aabb(&r, &b, &nf[0], &nf[1]);
}
/* Unit-shape intersection in object space.  out: t, position xyz, normal xyz, u, v, from_behind. */
int oracle_intersect(int type, const float *o, const float *d, int transparent, float max_t, float *out) {
    Ray r;
    ray_reset(&r, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]));
    Hit h;
    memset(&h, 0, sizeof h);
    h.distance = max_t;
    int ok = 0;
    switch (type) {
        case GT_QUAD: ok = hits_quad(&r, &h, transparent); break;
        case GT_BOX: ok = hits_box(&r, &h, transparent) != SIDE_NONE; break;
        case GT_SPHERE: ok = hits_sphere(&r, &h, transparent, NULL); if (ok) h.position = ray_at(&r, h.distance); break;
        case GT_TET: ok = hits_tet(&r, &h, transparent); break;
        default: return 0;
    }
    float v[10] = {h.distance, h.position.x, h.position.y, h.position.z, h.normal.x, h.normal.y, h.normal.z, h.uv.u, h.uv.v, (float)h.from_behind};
    memcpy(out, v, sizeof v);
    return ok;
}
/* SAH build over arbitrary boxes (N x 6 floats).  nodes_f: (2N) x 6, nodes_u: (2N) x 3, leaf_ids: N, meta: node_count, height. */
int oracle_build_bvh(const float *boxes, uint32_t N, uint32_t max_leaf, float *nodes_f, uint32_t *nodes_u, uint32_t *leaf_ids, uint32_t *meta) {
    Builder b;
    builder_init(&b, N);
    for (uint32_t i = 0; i < N; i++) {
        b.nodes[i].box.min = V3(boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]);
        b.nodes[i].box.max = V3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]);
        b.nodes[i].first_index = b.node_ids[i] = i;
    }
    Node *nodes = (Node *)calloc(2 * (size_t)N + 1, sizeof(Node));
    uint32_t count, height;
    bvh_build(&b, nodes, &count, &height, N, max_leaf);
    put_nodes(nodes, count, 0, nodes_f, 6 * (size_t)count);
    put_nodes(nodes, count, 1, nodes_u, 3 * (size_t)count);
    for (uint32_t i = 0; i < N; i++) leaf_ids[i] = b.leaf_ids[i];
    meta[0] = count;
    meta[1] = height;
    free(nodes);
    builder_free(&b);
    return 0;
}

/* ---------------------------------------------------------------- wireframe overlays (src/slim/draw)
 * drawBox / drawBVH / drawMesh -> drawEdge -> _drawLine -> Canvas::setPixel, drawn sequentially as the reference
 * draws them, into a canvas of RGBA floats and depths in the render's layout (2x2 quads per pixel under SSAA). */
typedef struct {
    float camera_euler[3], camera_position[3], focal_length;
    uint32_t width, height, ssaa;
    float near_distance, far_distance;
} OvView;                                                   /* the layout of slimrt_overlay_view */
typedef struct { float q[4], position[3], scale[3]; } OvTransform;
typedef struct {
    OvTransform t;
    float mn[3], mx[3], color[3], opacity;
    uint32_t line_width, sides;
} OvBox;                                                    /* the layout of slimrt_overlay_box */

typedef struct {
    m3 rot;                                                 /* Camera::orientation */
    v3 cam_pos;
    float focal, near_d, far_d;
    float f_width, f_height, h_width, h_height, width_over_height, height_over_width;   /* Dimensions */
    float scale_x, scale_y;                                 /* Frustum::Projection::scale */
    int width, height, ssaa;
    float *pixels, *depths;
} OvCanvas;

static OvCanvas ov_canvas(const OvView *v, float *pixels, float *depths) {
    OvCanvas c;
    c.rot = m3_from_euler(v->camera_euler[0], v->camera_euler[1], v->camera_euler[2]);
    c.cam_pos = V3(v->camera_position[0], v->camera_position[1], v->camera_position[2]);
    c.focal = v->focal_length;
    c.near_d = v->near_distance;
    c.far_d = v->far_distance;
    c.width = (uint16_t)v->width;                                                            /* base.h:1312-1323 */
    c.height = (uint16_t)v->height;
    c.f_width = (float)c.width;
    c.f_height = (float)c.height;
    c.h_width = c.f_width * 0.5f;
    c.h_height = c.f_height * 0.5f;
    c.width_over_height = c.f_width / c.f_height;
    c.height_over_width = c.f_height / c.f_width;
    c.scale_x = c.focal * c.height_over_width;                                               /* frustum.h:25-27 */
    c.scale_y = c.focal;
    c.ssaa = v->ssaa ? 1 : 0;
    c.pixels = pixels;
    c.depths = depths;
    return c;
}

static void ov_set_pixel(const OvCanvas *cv, int x, int y, const float col[3], float opacity, float depth) { /* canvas.h:128-185 */
    int w = cv->width, h = cv->height;
    if (cv->ssaa) { w <<= 1; h <<= 1; }
    if (x < 0 || y < 0 || x >= w || y >= h) return;
    int override_ = 0;
    if (opacity < 0) { opacity = -opacity; override_ = 1; }
    opacity = clamp01(opacity);
    float px[4] = {clamp01(col[0]), clamp01(col[1]), clamp01(col[2]), opacity};
    for (int k = 0; k < 3; k++) px[k] *= px[k] * px[3];
    const size_t off = cv->ssaa ? ((size_t)cv->width * (y >> 1) + (x >> 1)) * 4 + 2 * (y & 1) + (x & 1)
                                : (size_t)cv->width * y + x;
    float *out = cv->pixels + 4 * off, *out_depth = cv->depths + off;
    if (override_ || (out_depth[0] == INFINITY && out[0] == 0 && out[1] == 0 && out[2] == 0) ||
        (opacity == 1.0f && depth == INFINITY)) {
        memcpy(out, px, sizeof px);
        out_depth[0] = depth;
        return;
    }
    const float *bg, *fg;                                    /* _sortPixelsByDepth, canvas.h:267-276 */
    float old[4];
    memcpy(old, out, sizeof old);
    if (depth == INFINITY || depth < *out_depth) { *out_depth = depth; bg = old; fg = px; }
    else { bg = px; fg = old; }
    float res[4];
    if (fg[3] == 1) memcpy(res, fg, sizeof res);
    else for (int k = 0; k < 4; k++) res[k] = fg[k] + bg[k] * (1.0f - fg[3]);        /* alphaBlendOver, base.h:1162 */
    memcpy(out, res, sizeof res);
}

typedef struct { float first, last; } RangeF;
typedef struct { int first, last; } RangeI;
static void rangef_sub(RangeF *r, float a, float b) {                                                  /* base.h:305-317 */
    if (b < a) { float t = b; b = a; a = t; }
    if (r->last < a || b < r->first) { r->first = -1; r->last = -2; }
    else { r->first = r->first < a ? a : r->first; r->last = b < r->last ? b : r->last; }
}
static int rangei_has(RangeI r, int v) { return (r.first <= v) && (v <= r.last); }
#define OV_FRAC(x) ((x) - floorf(x))                                                                   /* base.h:219-220 */
#define OV_1MFRAC(x) (1 - OV_FRAC(x))

static void ov_draw_line(const OvCanvas *cv, float x1, float y1, float z1, float x2, float y2, float z2,
                         const float col[3], float opacity, uint8_t line_width) {                     /* line.h:70-273 */
    RangeF fx = {x1 <= x2 ? x1 : x2, x1 <= x2 ? x2 : x1};
    RangeF fy = {y1 <= y2 ? y1 : y2, y1 <= y2 ? y2 : y1};
    const float left = 0.0f, top = 0.0f, right = (float)(cv->width - 1), bottom = (float)(cv->height - 1);
    x1 += left; x2 += left; y1 += top; y2 += top;
    fx.first += left; fx.last += left;
    fy.first += top; fy.last += top;
    rangef_sub(&fx, left, right);
    rangef_sub(&fy, top, bottom);
    rangef_sub(&fx, 0, cv->f_width - 1.0f);
    rangef_sub(&fy, 0, cv->f_height - 1.0f);
    if (fx.last < fx.first || fy.last < fy.first) return;
    RangeI xr = {(int)fx.first, (int)ceilf(fx.last)}, yr = {(int)fy.first, (int)ceilf(fy.last)};
    if (xr.last == cv->width) xr.last--;
    if (yr.last == cv->height) yr.last--;
    if (cv->ssaa) {
        x1 += x1; x2 += x2; y1 += y1; y2 += y2;
        xr.first <<= 1; yr.first <<= 1; xr.last <<= 1; yr.last <<= 1;
        line_width <<= 1;
        line_width++;
    }
    float tmp, z_range, range_remap, gap, grad, first_offset, last_offset, z = 0, z_curr = 0, z_step = 0;
    float first_x, last_x, first_y, last_y;
    int x, y, start_x, end_x, start_y, end_y;
    const float dx = x2 - x1, dy = y2 - y1;
    const int has_depth = (z1 != 0.0f) || (z2 != 0.0f);
    if (fabsf(dx) > fabsf(dy)) {
        if (x2 < x1) { tmp = x2; x2 = x1; x1 = tmp; tmp = y2; y2 = y1; y1 = tmp; tmp = z2; z2 = z1; z1 = tmp; }
        grad = dy / dx;
        first_x = roundf(x1); last_x = roundf(x2);
        first_offset = first_x - x1; last_offset = last_x - x2;
        first_y = y1 + grad * first_offset; last_y = y2 + grad * last_offset;
        start_x = (int)first_x; start_y = (int)first_y; end_x = (int)last_x; end_y = (int)last_y;
        x = start_x; y = start_y;
        gap = OV_1MFRAC(x1 + 0.5f);
        if (rangei_has(xr, x)) {
            if (rangei_has(yr, y)) ov_set_pixel(cv, x, y, col, OV_1MFRAC(first_y) * gap * opacity, z1);
            for (uint8_t i = 0; i < line_width; i++) if (rangei_has(yr, ++y)) ov_set_pixel(cv, x, y, col, opacity, z1);
            if (rangei_has(yr, ++y)) ov_set_pixel(cv, x, y, col, OV_FRAC(first_y) * gap * opacity, z1);
        }
        x = end_x; y = end_y;
        gap = OV_FRAC(x2 + 0.5f);
        if (rangei_has(xr, x)) {
            if (rangei_has(yr, y)) ov_set_pixel(cv, x, y, col, OV_1MFRAC(last_y) * gap * opacity, z2);
            for (uint8_t i = 0; i < line_width; i++) if (rangei_has(yr, ++y)) ov_set_pixel(cv, x, y, col, opacity, z2);
            if (rangei_has(yr, ++y)) ov_set_pixel(cv, x, y, col, OV_FRAC(last_y) * gap * opacity, z2);
        }
        if (has_depth) {
            z1 = 1.0f / z1; z2 = 1.0f / z2;
            z_range = z2 - z1;
            range_remap = z_range / (x2 - x1);
            z1 += range_remap * (first_offset + 1.0f);
            z2 += range_remap * (last_offset - 1.0f);
            z_curr = z1;
            z_range = z2 - z1;
            z_step = z_range / (last_x - first_x - 1.0f);
        } else z = 0.0f;
        gap = first_y + grad;
        for (x = start_x + 1; x < end_x; x++) {
            if (rangei_has(xr, x)) {
                y = (int)gap;
                if (has_depth) z = 1.0f / z_curr;
                if (rangei_has(yr, y)) ov_set_pixel(cv, x, y, col, OV_1MFRAC(gap) * opacity, z);
                for (uint8_t i = 0; i < line_width; i++) if (rangei_has(yr, ++y)) ov_set_pixel(cv, x, y, col, opacity, z);
                if (rangei_has(yr, ++y)) ov_set_pixel(cv, x, y, col, OV_FRAC(gap) * opacity, z);
            }
            gap += grad;
            if (has_depth) z_curr += z_step;
        }
    } else {
        if (y2 < y1) { tmp = x2; x2 = x1; x1 = tmp; tmp = y2; y2 = y1; y1 = tmp; tmp = z2; z2 = z1; z1 = tmp; }
        grad = dx / dy;
        first_y = roundf(y1); last_y = roundf(y2);
        first_offset = y1 - first_y; last_offset = last_y - y2;
        first_x = x1 + grad * first_offset; last_x = x2 + grad * last_offset;
        start_y = (int)first_y; start_x = (int)first_x; end_y = (int)last_y; end_x = (int)last_x;
        x = start_x; y = start_y;
        gap = OV_1MFRAC(y1 + 0.5f);
        if (rangei_has(yr, y)) {
            if (rangei_has(xr, x)) ov_set_pixel(cv, x, y, col, OV_1MFRAC(first_x) * gap * opacity, z1);
            for (uint8_t i = 0; i < line_width; i++) if (rangei_has(xr, ++x)) ov_set_pixel(cv, x, y, col, opacity, z1);
            if (rangei_has(xr, ++x)) ov_set_pixel(cv, x, y, col, OV_FRAC(first_x) * gap * opacity, z1);
        }
        x = end_x; y = end_y;
        gap = OV_FRAC(y2 + 0.5f);
        if (rangei_has(yr, y)) {
            if (rangei_has(xr, x)) ov_set_pixel(cv, x, y, col, OV_1MFRAC(last_x) * gap * opacity, z2);
            for (uint8_t i = 0; i < line_width; i++) if (rangei_has(xr, ++x)) ov_set_pixel(cv, x, y, col, opacity, z2);
            if (rangei_has(xr, ++x)) ov_set_pixel(cv, x, y, col, OV_FRAC(last_x) * gap * opacity, z2);
        }
        if (has_depth) {
            z1 = 1.0f / z1; z2 = 1.0f / z2;
            z_range = z2 - z1;
            range_remap = z_range / (y2 - y1);
            z1 += range_remap * (first_offset + 1.0f);
            z2 += range_remap * (last_offset - 1.0f);
            z_range = z2 - z1;
            z_step = z_range / (last_y - first_y - 1.0f);
            z_curr = z1;
        } else z = 0.0f;
        gap = first_x + grad;
        for (y = start_y + 1; y < end_y; y++) {
            if (rangei_has(yr, y)) {
                if (has_depth) z = 1.0f / z_curr;
                x = (int)gap;
                if (rangei_has(xr, x)) ov_set_pixel(cv, x, y, col, OV_1MFRAC(gap) * opacity, z);
                for (uint8_t i = 0; i < line_width; i++) if (rangei_has(xr, ++x)) ov_set_pixel(cv, x, y, col, opacity, z);
                if (rangei_has(xr, ++x)) ov_set_pixel(cv, x, y, col, OV_FRAC(gap) * opacity, z);
            }
            gap += grad;
            if (has_depth) z_curr += z_step;
        }
    }
}

static v3 ov_lerp(v3 a, v3 b, float by) { return vscaleadd(vsub(b, a), by, a); }                     /* vec3.h:360-362 */

static int ov_clip(v3 *A, v3 *B, v3 N) {                                                   /* one frustum side */
    const float nA = vdot(N, *A), nB = vdot(N, *B);
    const int out = (nA < 0) | ((nB < 0) << 1);
    if (out) {
        if (out == 3) return 0;
        if (out & 1) *A = ov_lerp(*A, *B, nA / (nA - nB));
        else *B = ov_lerp(*B, *A, nB / (nB - nA));
    }
    return 1;
}

static void ov_draw_edge(const OvCanvas *cv, v3 from_world, v3 to_world, const float col[3], float opacity,
                         uint8_t line_width) {                                                 /* edge.h:7-17 */
    /* Camera::internPos (camera.h:47): transposed rotation times (p - position) */
    const m3 T = M3(V3(cv->rot.X.x, cv->rot.Y.x, cv->rot.Z.x), V3(cv->rot.X.y, cv->rot.Y.y, cv->rot.Z.y),
                    V3(cv->rot.X.z, cv->rot.Y.z, cv->rot.Z.z));
    v3 A = m3_mulv(T, vsub(from_world, cv->cam_pos)), B = m3_mulv(T, vsub(to_world, cv->cam_pos));
    /* Frustum::cullAndClipEdge (frustum.h:67-141) */
    float d = cv->near_d;
    int out = (A.z < d) | ((B.z < d) << 1);
    if (out) {
        if (out == 3) return;
        if (out & 1) A = ov_lerp(A, B, (d - A.z) / (B.z - A.z));
        else B = ov_lerp(B, A, (d - B.z) / (A.z - B.z));
    }
    d = cv->far_d;
    out = (A.z > d) | ((B.z > d) << 1);
    if (out) {
        if (out == 3) return;
        if (out & 1) A = ov_lerp(A, B, (A.z - d) / (A.z - B.z));
        else B = ov_lerp(B, A, (B.z - d) / (B.z - A.z));
    }
    if (!ov_clip(&A, &B, V3(cv->focal, 0, cv->width_over_height))) return;
    if (!ov_clip(&A, &B, V3(-cv->focal, 0, cv->width_over_height))) return;
    if (!ov_clip(&A, &B, V3(0, cv->focal, 1))) return;
    if (!ov_clip(&A, &B, V3(0, -cv->focal, 1))) return;
    /* Frustum::projectPoint (frustum.h:180-184) */
    float ax = ((cv->scale_x * A.x / A.z) + 1) * cv->h_width, ay = ((cv->scale_y * A.y / A.z) + 1) * cv->h_height;
    float bx = ((cv->scale_x * B.x / B.z) + 1) * cv->h_width, by = ((cv->scale_y * B.y / B.z) + 1) * cv->h_height;
    ay = cv->f_height - ay;
    by = cv->f_height - by;
    ov_draw_line(cv, ax, ay, A.z - EPS, bx, by, B.z - EPS, col, opacity, line_width);
}

static v3 ov_extern(const OvTransform *t, v3 p) {                                          /* transform.h:10 */
    const quat q = Q(V3(t->q[0], t->q[1], t->q[2]), t->q[3]);
    return vadd(qrot(q, vmul(p, V3(t->scale[0], t->scale[1], t->scale[2]))),
                V3(t->position[0], t->position[1], t->position[2]));
}

/* drawBox (box.h:7-34) */
void oracle_draw_boxes(const OvView *view, const OvBox *boxes, uint32_t count, float *pixels, float *depths) {
    const OvCanvas cv = ov_canvas(view, pixels, depths);
    for (uint32_t b = 0; b < count; b++) {
        const OvBox *bx = &boxes[b];
        const v3 mn = V3(bx->mn[0], bx->mn[1], bx->mn[2]), mx = V3(bx->mx[0], bx->mx[1], bx->mx[2]);
        const v3 corner[8] = {                                                                  /* box.h:19-29 */
            V3(mn.x, mx.y, mx.z), V3(mx.x, mx.y, mx.z), V3(mn.x, mn.y, mx.z), V3(mx.x, mn.y, mx.z),
            V3(mn.x, mx.y, mn.z), V3(mx.x, mx.y, mn.z), V3(mn.x, mn.y, mn.z), V3(mx.x, mn.y, mn.z)};
        v3 w[8];
        for (int i = 0; i < 8; i++) w[i] = ov_extern(&bx->t, corner[i]);
        /* BoxEdges::setFrom (box.h:52-77): front top, bottom, left, right; back ...; left bottom, left top;
           right bottom, right top */
        static const int E[12][2] = {{0, 1}, {2, 3}, {2, 0}, {3, 1}, {4, 5}, {6, 7},
                                     {6, 4}, {7, 5}, {2, 6}, {0, 4}, {3, 7}, {1, 5}};
        const uint8_t sides = (uint8_t)bx->sides;
        const uint8_t lw = (uint8_t)bx->line_width;
        if (sides == 63) {
            for (int e = 0; e < 12; e++) ov_draw_edge(&cv, w[E[e][0]], w[E[e][1]], bx->color, bx->opacity, lw);
        } else {                                                                                /* box.h:21-32 */
            enum { L = 1, Bo = 2, Bk = 4, R = 8, T = 16, F = 32 };
            static const int S[12][3] = {{0, F, T}, {1, F, Bo}, {2, F, L}, {3, F, R}, {4, Bk, T}, {5, Bk, Bo},
                                         {6, Bk, L}, {7, Bk, R}, {9, L, T}, {8, L, Bo}, {11, R, T}, {10, R, Bo}};
            for (int i = 0; i < 12; i++)
                if ((sides & S[i][1]) | (sides & S[i][2]))
                    ov_draw_edge(&cv, w[E[S[i][0]][0]], w[E[S[i][0]][1]], bx->color, bx->opacity, lw);
        }
    }
}

/* drawMesh without normals (mesh.h:7-22) */
void oracle_draw_mesh_wireframe(const OvView *view, const float *positions, const uint32_t *edges, uint32_t edge_count,
                                const OvTransform *t, const float *color3, float opacity, uint32_t line_width,
                                float *pixels, float *depths) {
    const OvCanvas cv = ov_canvas(view, pixels, depths);
    for (uint32_t i = 0; i < edge_count; i++) {
        const float *a = positions + 3 * (size_t)edges[2 * i], *b = positions + 3 * (size_t)edges[2 * i + 1];
        ov_draw_edge(&cv, ov_extern(t, V3(a[0], a[1], a[2])), ov_extern(t, V3(b[0], b[1], b[2])), color3, opacity,
                     (uint8_t)line_width);
    }
}
