// Per-pixel ray-tracing core: data layout in device memory and the trace / shade functions.
//
// Everything here is __host__ __device__ so that the CPU side of RayTracingRenderer::render's CPU/GPU
// toggle (src/slim/renderer/renderer.h:63-85) runs the same arithmetic as the gfx950 kernels.
// Floating-point operation order follows the reference statement by statement and the library is
// compiled with -ffp-contract=off, so GPU results reproduce the reference CPU renderer (which has no
// FMA on baseline x86-64: src/slim/core/base.h:84-99).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define SLIM_HD __host__ __device__ __forceinline__

namespace slimrt {

// ------------------------------------------------------------------ constants (src/slim/core/base.h:125-160)
constexpr float kEps = 0.0001f;
constexpr float kTraceOffset = 0.0001f;
constexpr float kPi = 3.14159265f;
constexpr float kOneOverPi = 1.0f / kPi;
constexpr float kSqrt3 = 1.73205080757f;
constexpr float kSphereCoverage = 1.0f / ((4.0f * kPi) / 6.0f);
constexpr float kTetMax = 0.577350259f;
constexpr float kTetMin = 0.288675159f;
constexpr float kColorToFloat = 0.00392156862f;
constexpr float kIorAir = 1.0003f;
constexpr float kLightRadiusFactor = 1.0f / 64.0f;   // LIGHT_INTENSITY_RADIUS_FACTOR (material.h:4-5)
constexpr float kLightInvRadiusFactor = 64.0f;

enum GeometryType : uint32_t { GT_None = 0, GT_Mesh, GT_Grid, GT_Box, GT_Curve, GT_Tet, GT_Quad, GT_Sphere };
enum BRDF : uint32_t { BRDF_Lambert = 0, BRDF_Blinn, BRDF_Phong, BRDF_CookTorrance };
enum RenderMode : uint32_t { RM_Normals = 0, RM_NormalMap, RM_Beauty, RM_Depth, RM_MipLevel, RM_UVs };
enum Side : int { S_None = 0, S_Left = 1, S_Bottom = 2, S_Back = 4, S_Right = 8, S_Top = 16, S_Front = 32 };
constexpr uint32_t MAT_REFLECTIVE = 1, MAT_REFRACTIVE = 2, MAT_EMISSIVE = 4, MAT_ALBEDO_MAP = 16, MAT_NORMAL_MAP = 32;
constexpr uint32_t GEO_VISIBLE = 1, GEO_SHADOWING = 2, GEO_TRANSPARENT = 4;
constexpr uint32_t LIGHT_DIRECTIONAL = 1, LIGHT_SHADOWING = 2;
constexpr uint32_t IMG_MIPMAP = 1u << 4;
// texel storage of a texture: bits 24-25 of its flags (no reference counterpart; the reference stores quads)
constexpr uint32_t kTexelsShift = 24, kTexelsMask = 3u << kTexelsShift;
constexpr uint32_t kTexelQuads = 0, kTexelsRGBA8 = 1, kTexelsRGBA16F = 2;

// ------------------------------------------------------------------ device data layout (all 16-byte aligned)
// BVH node, 32 bytes; siblings are adjacent so a pair is one 64-byte fetch (BVHNode, src/slim/scene/bvh.h:5-15).
struct alignas(16) Node {
    float minx, miny, minz;
    uint32_t first_index;        // child pair index, or first leaf item
    float maxx, maxy, maxz;
    uint32_t leaf_count_depth;   // leaf_count | depth << 16
};
// Geometry with its transform (Geometry/Transform, src/slim/core/transform.h:5-65), 64 bytes.
struct alignas(16) Geo {
    float qx, qy, qz, qw;        // orientation quaternion (axis, amount)
    float px, py, pz; uint32_t type;
    float sx, sy, sz; uint32_t material_id;
    float isx, isy, isz;         // 1 / scale, as Ray::localize computes it (ray.h:30)
    uint32_t id_flags;           // mesh id << 8 | flags
};
// Material (src/slim/scene/material.h:93-117), 64 bytes.
struct alignas(16) Mat {
    float ar, ag, ab, roughness;
    float er, eg, eb, ior;
    float fr, fg, fb, metalness;           // reflectivity (F0)
    float uv_ru, uv_rv, normal_magnitude;
    uint32_t packed;                       // flags | texture_count << 8 | brdf << 12 | tex0 << 16 | tex1 << 24
};
// Light (src/slim/scene/material.h:17-25), 32 bytes.
struct alignas(16) Light {
    float r, g, b, intensity;
    float x, y, z; uint32_t flags;
};
// Mesh record, 64 bytes: object-space AABB, node/triangle bases in the concatenated arrays, counts, and the
// bases of the mesh's 4-wide nodes, 8-wide nodes and 8-wide leaf records (16-byte units).
struct alignas(16) MeshRec {
    float minx, miny, minz; uint32_t node_base;
    float maxx, maxy, maxz; uint32_t tri_base;
    uint32_t node_count, tri_count, normals_count, uvs_count;
    uint32_t node4_base, node8_base, leaf8_base, pad;
};
// 4-wide mesh BVH node, 128 bytes: the children's boxes as structure of arrays, then per child either a node
// index (count 0) or a leaf's first triangle and triangle count; unused slots have child = ~0u.
struct alignas(16) Node4 {
    float minx[4], miny[4], minz[4];
    float maxx[4], maxy[4], maxz[4];
    uint32_t child[4], count[4];
};
// 8-wide mesh BVH node, 128 bytes, with quantised child boxes (the compressed wide node of Ylitie, Karras and Laine,
// "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017, with this build's own layout).
// Child box bounds are grid points o + q * 2^(e - 127) per axis, q in 0..255, chosen on the host so that every
// decoded box contains the child's exact box (each bound checked with the same fused multiply-add the walk uses);
// a slab test on the decoded box with the reference's arithmetic is monotone in the bounds, so it accepts every
// ray the exact box accepts.  Slots hold the children ordered for the walk's octant order (slot s is visited at
// position s ^ (7 ^ ray octant)).  Per slot: count 0 = inner child (child = mesh-relative node index), 1..254 = leaf
// of that many triangles (child = its record in the leaf array, 16-byte units), 255 = empty.
struct alignas(16) Node8 {
    float ox, oy, oz;
    uint32_t exps;               // bytes: biased exponents of the x, y, z grid steps; the fourth byte is unused
    uint32_t child[8];
    uint32_t count[2];           // one byte per slot
    uint32_t qlo[3][2], qhi[3][2];   // per axis, one byte per slot
    uint32_t pad[6];
};
// An 8-wide leaf record: the leaf's exact box (the sibling-pair node's), its first triangle (mesh-relative, leaf
// order) and triangle count, followed by its triangles' hit records (TriHit).  The walk repeats the reference's
// exact leaf test before testing the triangles, so quantised boxes change which leaves are fetched, not results.
struct alignas(16) LeafHead {
    float minx, miny, minz; uint32_t first;
    float maxx, maxy, maxz; uint32_t count;
};
// Triangle split into the 48 bytes every intersection test reads and the 64 bytes only the closest hit reads
// (Triangle is 128 bytes, src/slim/scene/mesh.h:19-24).  The test uses the first two columns of the
// local-to-tangent matrix (the barycentric coordinates), the position and the plane normal; the third column is
// read by nothing.
struct alignas(16) TriHit {
    float xx, xy, yx, yy;        // local_to_tangent rows X and Y, first two columns
    float zx, zy, px, py;        // row Z, first two columns | position.x, position.y
    float pz, nx, ny, nz;        // position.z | plane normal
};
struct alignas(16) TriAttr {
    float n1x, n1y, n1z, uv1u;
    float n2x, n2y, n2z, uv1v;
    float n3x, n3y, n3z, uv2u;
    float uv2v, uv3u, uv3v, uv_coverage;
};
struct alignas(16) TexRec { uint32_t width, height, flags, mip_count; uint32_t mip_base, pad0, pad1, pad2; };
// A mip's storage starts at byte 16 * base of the texture array.  fmt kTexelQuads: (w + 1) * (h + 1) texel quads
// of 12 bytes, R{TL,TR,BL,BR} G B (the reference's layout); kTexelsRGBA8 / kTexelsRGBA16F: (w + 2) * (h + 2)
// texels of 4 / 8 bytes, rows top to bottom, with a one-texel border holding what the quads' outer corners hold.
struct alignas(16) MipRec { uint32_t w, h; uint32_t base, fmt; };

struct SceneView {
    const Geo *geos;
    const Mat *mats;
    const Light *lights;
    const Node *snodes;          // scene BVH, node 0 = root
    const uint32_t *leaf_ids;    // scene leaf order -> geometry index
    const MeshRec *meshes;
    const Node *mnodes;          // all mesh BVH nodes
    const Node4 *mnodes4;        // all 4-wide mesh BVH nodes (collapsed from mnodes)
    const Node8 *mnodes8;        // all 8-wide mesh BVH nodes (collapsed from mnodes)
    const uint4 *leaf8;          // all 8-wide leaf records (LeafHead + TriHit records), 16-byte units
    const TriHit *tri_hit;
    const TriAttr *tri_attr;
    const TexRec *textures;
    const MipRec *mips;
    const uint8_t *quads;
    uint32_t G, M, L, NM, NT, flags;   // flags bit 0: scene has emissive quads
};

struct Projection {   // CameraRayProjection (src/slim/scene/camera.h:59-83)
    float inv_rot[9];
    float start[3], right[3], down[3], camera_position[3];
    float c_start[2];
    float sq_dist, sample_size;
};

struct RenderParams {
    Projection proj;
    uint32_t width, height;      // canvas dimensions (samples are 2x2 per pixel under SSAA)
    uint32_t ssaa;
    uint32_t max_depth, render_mode;
    int32_t sky_color, sky_radiance, sky_irradiance;
    uint32_t output;             // 0 canvas Pixel (clamp(c)^2, 1), 1 radiance (tone-mapped colour, 1)
    uint32_t flags;              // bit 0: keep the pixel's ray-cone scaling factor through hits (else the
                                 // reference data flow, which overwrites it on every hit); bits 1-5: wavefront
                                 // forms; bit 6: 4-wide mesh BVHs
    float mip_colors[9][3];
};

SLIM_HD bool keeps_ray_cones(const RenderParams &rp) { return (rp.flags & 1u) != 0; }
SLIM_HD bool wide_mesh_bvh(const RenderParams &rp) { return (rp.flags & 64u) != 0; }
// The meshes' node width a frame asks for: 8 (SLIMRT_BVH8), 4 (SLIMRT_BVH4) or the sibling-pair walk.
SLIM_HD uint32_t mesh_node_width(const RenderParams &rp) { return (rp.flags & 1024u) ? 8u : (rp.flags & 64u) ? 4u : 2u; }

// ------------------------------------------------------------------ small algebra
struct v3 { float x, y, z; };
struct v2 { float u, v; };
struct color { float r, g, b; };
struct quat { v3 a; float w; };

SLIM_HD v3 V3(float x, float y, float z) { return v3{x, y, z}; }
SLIM_HD v3 vs(float s) { return v3{s, s, s}; }
SLIM_HD v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SLIM_HD v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SLIM_HD v3 operator*(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
SLIM_HD v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
SLIM_HD v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
SLIM_HD v3 divs(v3 a, float s) { float f = 1.0f / s; return v3{a.x * f, a.y * f, a.z * f}; }
SLIM_HD float dot(v3 a, v3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
SLIM_HD v3 cross(v3 a, v3 b) { return v3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)}; }
SLIM_HD float sqlen(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
SLIM_HD v3 normalized(v3 a) { return divs(a, sqrtf(sqlen(a))); }
SLIM_HD v3 scale_add(v3 a, float f, v3 b) { return v3{a.x * f + b.x, a.y * f + b.y, a.z * f + b.z}; }
SLIM_HD v3 mul_add(v3 a, v3 f, v3 b) { return v3{a.x * f.x + b.x, a.y * f.y + b.y, a.z * f.z + b.z}; }
SLIM_HD v3 reflected(v3 d, v3 n) { return scale_add(n, -2 * dot(d, n), d); }
SLIM_HD float fmax2(float a, float b) { return a > b ? a : b; }   // Max (base.h:117)
SLIM_HD float fmin2(float a, float b) { return a < b ? a : b; }   // Min (base.h:118)
SLIM_HD float clamp01(float v) { float mn = v < 1.0f ? v : 1.0f; return mn > 0.0f ? mn : 0.0f; }
SLIM_HD uint32_t sign_bit(float x) { return __builtin_bit_cast(uint32_t, x) >> 31; }

SLIM_HD color C3(float r, float g, float b) { return color{r, g, b}; }
SLIM_HD color cs(float s) { return color{s, s, s}; }
SLIM_HD color operator+(color a, color b) { return color{a.r + b.r, a.g + b.g, a.b + b.b}; }
SLIM_HD color operator*(color a, color b) { return color{a.r * b.r, a.g * b.g, a.b * b.b}; }
SLIM_HD color operator*(color a, float s) { return color{a.r * s, a.g * s, a.b * s}; }
SLIM_HD color cmul_add(color a, color f, color b) { return color{a.r * f.r + b.r, a.g * f.g + b.g, a.b * f.b + b.b}; }
SLIM_HD color cscale_add(color a, float f, color b) { return color{a.r * f + b.r, a.g * f + b.g, a.b * f + b.b}; }
SLIM_HD color one_minus(color a) { return color{1.0f - a.r, 1.0f - a.g, 1.0f - a.b}; }

// quaternion rotation q * v (src/slim/math/quat.h:37-43)
SLIM_HD v3 qrot(quat q, v3 v) {
    v3 out = cross(q.a, v);
    v3 qqv = cross(q.a, out);
    out = scale_add(out, q.w, qqv);
    return scale_add(out, 2.0f, v);
}
SLIM_HD quat qnorm(quat q) {
    float len = sqrtf(q.a.x * q.a.x + q.a.y * q.a.y + q.a.z * q.a.z + q.w * q.w);
    float f = 1.0f / len;
    return quat{q.a * f, q.w * f};
}
SLIM_HD quat axis_angle(v3 axis, float radians) {   // quat.h:78-84
    radians *= 0.5f;
    return qnorm(quat{axis * sinf(radians), cosf(radians)});
}

// ------------------------------------------------------------------ rays (src/slim/core/ray.h)
struct Ray {
    v3 origin, scaled_origin, direction, rcp;
    uint32_t sl, sb, sk;   // sign bits of the direction: Sides left / bottom / back
};
struct Hit {
    v3 position, normal;
    v2 uv;
    float distance, uv_coverage, NdotRd, scaling_factor;
    uint32_t id;
    bool from_behind;
};

SLIM_HD void ray_reset(Ray &r, v3 o, v3 d) {   // ray.h:40-49
    r.origin = o;
    r.direction = d;
    r.rcp = v3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    r.scaled_origin = (-o) * r.rcp;
    r.sl = sign_bit(d.x);
    r.sb = sign_bit(d.y);
    r.sk = sign_bit(d.z);
}
SLIM_HD v3 ray_at(const Ray &r, float t) { return scale_add(r.direction, t, r.origin); }

SLIM_HD quat geo_quat(const Geo &g) { return quat{v3{g.qx, g.qy, g.qz}, g.qw}; }

SLIM_HD void ray_localize(Ray &r, v3 o, v3 d, const Geo &g) {   // ray.h:29-34
    v3 inv_scale{g.isx, g.isy, g.isz};
    quat inv_rot{v3{-g.qx, -g.qy, -g.qz}, g.qw};
    ray_reset(r, inv_scale * qrot(inv_rot, o - v3{g.px, g.py, g.pz}), inv_scale * qrot(inv_rot, d));
}

// Slab test with octant-swapped bounds (ray.h:51-59)
SLIM_HD bool hits_box_bounds(const Ray &r, float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                             float &near_d, float &far_d) {
    float ax = r.sl ? mxx : mnx, bx = r.sl ? mnx : mxx;
    float ay = r.sb ? mxy : mny, by = r.sb ? mny : mxy;
    float az = r.sk ? mxz : mnz, bz = r.sk ? mnz : mxz;
    float tx0 = ax * r.rcp.x + r.scaled_origin.x, ty0 = ay * r.rcp.y + r.scaled_origin.y, tz0 = az * r.rcp.z + r.scaled_origin.z;
    float tx1 = bx * r.rcp.x + r.scaled_origin.x, ty1 = by * r.rcp.y + r.scaled_origin.y, tz1 = bz * r.rcp.z + r.scaled_origin.z;
    float m = fmax2(tx0, fmax2(ty0, tz0));
    near_d = fmax2(0, m);
    far_d = fmin2(tx1, fmin2(ty1, tz1));
    return near_d <= far_d;
}
SLIM_HD bool hits_node(const Ray &r, const Node &n, float &near_d, float &far_d) {
    return hits_box_bounds(r, n.minx, n.miny, n.minz, n.maxx, n.maxy, n.maxz, near_d, far_d);
}

SLIM_HD bool hits_plane(const Ray &r, v3 po, v3 pn, Hit &hit) {   // ray.h:61-82
    float NdotRd = dot(pn, r.direction);
    if (NdotRd == 0) return false;
    float NdotRoP = dot(pn, po - r.origin);
    if (NdotRoP == 0) return false;
    bool facing = NdotRd < 0;
    hit.from_behind = NdotRoP > 0;
    if (hit.from_behind == facing) return false;
    float t = NdotRoP / NdotRd;
    hit.distance = t;
    hit.position = ray_at(r, t);
    hit.normal = pn;
    return true;
}

SLIM_HD bool on_checkerboard(float u, float v) {   // base.h:277-279
    return ((uint8_t)(v * 4.0f)) % 2 == ((uint8_t)(u * 4.0f)) % 2;
}

SLIM_HD v2 uv_box_side(int side, float X, float Y, float Z) {   // base.h:424-452
    float u, v;
    switch (side) {
        case S_Top: u = X; v = Z; break;
        case S_Bottom: u = -X; v = -Z; break;
        case S_Left: u = -Z; v = Y; break;
        case S_Right: u = Z; v = Y; break;
        case S_Front: u = X; v = Y; break;
        default: u = -X; v = Y; break;
    }
    return v2{u * 0.5f + 0.5f, v * 0.5f + 0.5f};
}

SLIM_HD v2 uv_sphere(float X, float Y, float Z) {   // base.h:454-483
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
    return v2{u * 0.5f + 0.5f, v * 0.5f + 0.5f};
}

SLIM_HD bool hits_quad(const Ray &r, Hit &hit, bool transparent) {   // ray.h:84-120
    if (r.direction.y == 0) return false;
    if (r.origin.y == 0) return false;
    hit.from_behind = r.origin.y < 0;
    if (hit.from_behind == (r.direction.y < 0)) return false;
    float t = fabsf(r.origin.y * r.rcp.y);
    if (t > hit.distance) return false;
    v3 p = ray_at(r, t);
    hit.position = p;
    if (p.x < -1 || p.x > +1 || p.z < -1 || p.z > +1) return false;
    hit.uv = v2{p.x * 0.5f + 0.5f, p.z * 0.5f + 0.5f};
    if (transparent && on_checkerboard(hit.uv.u, hit.uv.v)) return false;
    hit.distance = t;
    hit.normal = v3{0.0f, 1.0f, 0.0f};
    hit.uv_coverage = 0.25f;
    return true;
}

SLIM_HD int box_side(int axis, uint32_t sl, uint32_t sb, uint32_t sk) {   // OctantShifts::getBoxSide (base.h:542-548)
    return (axis == 1) * (1 << (3 * sl)) | (axis == 2) * (2 << (3 * sb)) | (axis == 4) * (4 << (3 * sk));
}

SLIM_HD bool hits_box(const Ray &r, Hit &hit, bool transparent) {   // ray.h:122-175
    v3 sgn{1.0f - 2.0f * (float)r.sl, 1.0f - 2.0f * (float)r.sb, 1.0f - 2.0f * (float)r.sk};
    v3 srcp = sgn * r.rcp;
    v3 nt = r.scaled_origin - srcp, ft = r.scaled_origin + srcp;
    int near_axis = 1, far_axis = 1, side;
    float far_hit_t = ft.x;
    if (ft.y < far_hit_t) { far_hit_t = ft.y; far_axis = 2; }
    if (ft.z < far_hit_t) { far_hit_t = ft.z; far_axis = 4; }
    if (far_hit_t < 0) return false;
    float near_hit_t = nt.x;
    if (nt.y > near_hit_t) { near_hit_t = nt.y; near_axis = 2; }
    if (nt.z > near_hit_t) { near_hit_t = nt.z; near_axis = 4; }
    if (near_hit_t > hit.distance || far_hit_t < (near_hit_t > 0 ? near_hit_t : 0)) return false;
    float t = near_hit_t;
    hit.from_behind = near_hit_t < 0;
    if (hit.from_behind) {
        if (far_hit_t > hit.distance) return false;
        t = far_hit_t;
        side = box_side(far_axis, !r.sl, !r.sb, !r.sk);
    } else
        side = box_side(near_axis, r.sl, r.sb, r.sk);
    hit.position = ray_at(r, t);
    hit.uv = uv_box_side(side, hit.position.x, hit.position.y, hit.position.z);
    if (transparent && on_checkerboard(hit.uv.u, hit.uv.v)) {
        if (hit.from_behind || far_hit_t > hit.distance) return false;
        side = box_side(far_axis, !r.sl, !r.sb, !r.sk);
        hit.from_behind = true;
        t = far_hit_t;
        hit.position = ray_at(r, t);
        hit.uv = uv_box_side(side, hit.position.x, hit.position.y, hit.position.z);
        if (on_checkerboard(hit.uv.u, hit.uv.v)) return false;
    }
    hit.distance = t;
    hit.normal.x = (float)((int)(side == S_Right) - (int)(side == S_Left));
    hit.normal.y = (float)((int)(side == S_Top) - (int)(side == S_Bottom));
    hit.normal.z = (float)((int)(side == S_Front) - (int)(side == S_Back));
    hit.uv_coverage = 0.25f;
    return true;
}

SLIM_HD bool hits_sphere(const Ray &r, Hit &hit, bool transparent, float *out_sq_dist) {   // ray.h:177-215
    float t_to_closest = -(dot(r.origin, r.direction));
    if (t_to_closest <= 0) return false;
    float dsq = sqlen(r.direction);
    float sq_dist = sqlen(r.origin) * dsq - t_to_closest * t_to_closest;
    if (sq_dist > 1.0f) return false;
    float delta = sqrtf(dsq - sq_dist);
    dsq = 1.0f / dsq;
    float t = (t_to_closest - delta) * dsq;
    if (t > hit.distance) return false;
    hit.normal = ray_at(r, t);
    hit.uv = uv_sphere(hit.normal.x, hit.normal.y, hit.normal.z);
    hit.from_behind = t <= 0 || (transparent && on_checkerboard(hit.uv.u, hit.uv.v));
    if (hit.from_behind) {
        t = (t_to_closest + delta) * dsq;
        if (t <= 0 || t > hit.distance) return false;
        hit.normal = ray_at(r, t);
        hit.uv = uv_sphere(hit.normal.x, hit.normal.y, hit.normal.z);
        if (transparent && on_checkerboard(hit.uv.u, hit.uv.v)) return false;
    }
    hit.distance = t;
    hit.uv_coverage = kSphereCoverage;
    if (out_sq_dist) *out_sq_dist = sq_dist * dsq;
    return true;
}

SLIM_HD bool hits_tet(const Ray &r, Hit &hit, bool transparent) {   // ray.h:217-311
    bool found = false;
    Hit cur = hit;
    cur.uv_coverage = kSqrt3 / 4.0f;
    for (int t = 0; t < 4; t++) {
        v3 tp = vs(t == 3 ? kTetMax : -kTetMax), fn = tp;
        if (t == 0) fn.y = kTetMax;
        else if (t == 1) fn.x = kTetMax;
        else if (t == 2) fn.z = kTetMax;
        else tp.y = -kTetMax;
        cur.distance = hit.distance;
        if (!hits_plane(r, tp, fn, cur)) continue;
        v3 X, Y, Z;
        if (t == 0) { X = V3(kTetMax, -kTetMin, -kTetMax); Y = V3(kTetMin, kTetMin, kTetMax); Z = V3(-kTetMin, kTetMax, -kTetMax); }
        else if (t == 1) { X = V3(kTetMin, kTetMin, kTetMax); Y = V3(-kTetMin, kTetMax, -kTetMax); Z = V3(kTetMax, -kTetMin, -kTetMax); }
        else if (t == 2) { X = V3(-kTetMin, kTetMax, -kTetMax); Y = V3(kTetMax, -kTetMin, -kTetMax); Z = V3(kTetMin, kTetMin, kTetMax); }
        else { X = V3(-kTetMax, kTetMin, kTetMax); Y = V3(kTetMin, kTetMin, kTetMax); Z = V3(kTetMin, -kTetMax, kTetMax); }
        v3 d = cur.position - tp;
        v3 q{X.x * d.x + Y.x * d.y + Z.x * d.z, X.y * d.x + Y.y * d.y + Z.y * d.z, X.z * d.x + Y.z * d.y + Z.z * d.z};
        if (q.x < 0 || q.y < 0 || q.y + q.x > 1) continue;
        cur.uv = v2{q.x, q.y};
        if (transparent && on_checkerboard(cur.uv.u, cur.uv.v)) continue;
        if (cur.distance < hit.distance) {
            hit = cur;
            found = true;
        }
    }
    return found;
}

// ------------------------------------------------------------------ textures (src/slim/core/texture.h)
SLIM_HD color mip_sample(const SceneView &s, const MipRec &m, float u, float v) {   // texture.h:17-41
    if (u > 1) u -= (float)((uint64_t)u);
    if (v > 1) v -= (float)((uint64_t)v);
    const float U = u * (float)m.w + 0.5f, V = v * (float)m.h + 0.5f;
    // For u, v in [0, 1] these are (u32)U and (u32)V; the clamp only defines the NaN / negative cases, where the
    // reference's conversion is undefined, so that no input can index outside the mip.
    const uint32_t x = (uint32_t)fminf(fmaxf(U, 0.0f), (float)m.w), y = (uint32_t)fminf(fmaxf(V, 0.0f), (float)m.h);
    const float rr = U - (float)x, b = V - (float)y, l = 1 - rr, t = 1 - b;
    if (m.fmt == kTexelsRGBA16F) {
        // half-float texels: the same bilinear weights without the 1/255 scale, the same order of terms
        const float tl = t * l, tr = t * rr, bl = b * l, br = b * rr;
        const uint64_t row = (uint64_t)m.w + 2;
        const uint8_t *src = s.quads + 16ull * m.base + 8 * ((uint64_t)y * row + x);
        _Float16 p[8], pb[8];   // TL, TR and BL, BR: one 16-byte load per row
        __builtin_memcpy(p, src, 16);
        __builtin_memcpy(pb, src + 8 * row, 16);
        color c;
        c.r = (float)pb[4] * br + ((float)pb[0] * bl + ((float)p[4] * tr + (float)p[0] * tl));
        c.g = (float)pb[5] * br + ((float)pb[1] * bl + ((float)p[5] * tr + (float)p[1] * tl));
        c.b = (float)pb[6] * br + ((float)pb[2] * bl + ((float)p[6] * tr + (float)p[2] * tl));
        return c;
    }
    const float tl = t * l * kColorToFloat, tr = t * rr * kColorToFloat;
    const float bl = b * l * kColorToFloat, br = b * rr * kColorToFloat;
    if (m.fmt == kTexelsRGBA8) {
        // texel (x, y) of the bordered grid is the quad's top-left corner; its right and lower neighbours the rest
        const uint64_t row = (uint64_t)m.w + 2;
        const uint8_t *src = s.quads + 16ull * m.base + 4 * ((uint64_t)y * row + x);
        uint32_t t2[2], b2[2];   // TL, TR and BL, BR: one 8-byte load per row
        __builtin_memcpy(t2, src, 8);
        __builtin_memcpy(b2, src + 4 * row, 8);
        const uint32_t TL = t2[0], TR = t2[1], BL = b2[0], BR = b2[1];
        color c;
        c.r = (float)(BR & 255) * br + ((float)(BL & 255) * bl + ((float)(TR & 255) * tr + (float)(TL & 255) * tl));
        c.g = (float)((BR >> 8) & 255) * br + ((float)((BL >> 8) & 255) * bl +
                                               ((float)((TR >> 8) & 255) * tr + (float)((TL >> 8) & 255) * tl));
        c.b = (float)((BR >> 16) & 255) * br + ((float)((BL >> 16) & 255) * bl +
                                                ((float)((TR >> 16) & 255) * tr + (float)((TL >> 16) & 255) * tl));
        return c;
    }
    const uint8_t *q = s.quads + 16ull * m.base + 12ull * ((uint64_t)y * (m.w + 1) + x);
    uint32_t R, G, B;
    __builtin_memcpy(&R, q, 4);
    __builtin_memcpy(&G, q + 4, 4);
    __builtin_memcpy(&B, q + 8, 4);
    color c;
    c.r = (float)(R >> 24) * br + ((float)((R >> 16) & 255) * bl + ((float)((R >> 8) & 255) * tr + (float)(R & 255) * tl));
    c.g = (float)(G >> 24) * br + ((float)((G >> 16) & 255) * bl + ((float)((G >> 8) & 255) * tr + (float)(G & 255) * tl));
    c.b = (float)(B >> 24) * br + ((float)((B >> 16) & 255) * bl + ((float)((B >> 8) & 255) * tr + (float)(B & 255) * tl));
    return c;
}
SLIM_HD uint32_t mip_level(float texel_area, uint32_t mip_count) {   // texture.h:47-54
    uint32_t m = 0;
    while (texel_area > 1 && ++m < mip_count) texel_area *= 0.25f;
    if (m >= mip_count) m = mip_count - 1;
    return m;
}
SLIM_HD color tex_sample(const SceneView &s, uint32_t tex, float u, float v, float coverage) {   // texture.h:68-70
    const TexRec &t = s.textures[tex];
    uint32_t m = (t.flags & IMG_MIPMAP) ? mip_level(coverage * (float)((uint64_t)t.width * t.height), t.mip_count) : 0;
    return mip_sample(s, s.mips[t.mip_base + m], u, v);
}
SLIM_HD color tex_sample_cube(const SceneView &s, uint32_t tex, float X, float Y, float Z) {   // texture.h:72-116
    float u, v;
    uint32_t mip = 0;
    const float right = sign_bit(X) ? 0.0f : 1.0f, left = sign_bit(X) ? 1.0f : 0.0f;
    float z_over_x = X ? (Z / X) : 2, y_over_x = X ? (Y / X) : 2;
    if (z_over_x <= 1 && z_over_x >= -1 && y_over_x <= 1 && y_over_x >= -1) {
        u = z_over_x * -0.125f + (0.5f * right + 0.125f);
        v = y_over_x * (left - 0.5f) + 0.5f;
    } else {
        float x_over_z = Z ? (X / Z) : 2, y_over_z = Z ? (Y / Z) : 2;
        if (x_over_z <= 1 && x_over_z >= -1 && y_over_z <= 1 && y_over_z >= -1) {
            const float sz = sign_bit(Z) ? 1.0f : 0.0f;
            v = y_over_z * (sz - 0.5f) + +0.5f;
            u = x_over_z * 0.125f + (sz * 0.5f + 0.375f);
        } else {
            u = X / fabsf(Y) * 0.5f + 0.5f;
            v = Z / Y * 0.5f + 0.5f;
            mip = 2 - !sign_bit(Y);
        }
    }
    const TexRec &t = s.textures[tex];
    return mip_sample(s, s.mips[t.mip_base + mip], u, v);
}

// ------------------------------------------------------------------ material helpers
SLIM_HD uint32_t mat_flags(const Mat &m) { return m.packed & 0xFFu; }
SLIM_HD uint32_t mat_tex_count(const Mat &m) { return (m.packed >> 8) & 0xFu; }
SLIM_HD uint32_t mat_brdf(const Mat &m) { return (m.packed >> 12) & 0xFu; }
SLIM_HD uint32_t mat_tex(const Mat &m, int slot) { return (m.packed >> (16 + 8 * slot)) & 0xFFu; }
SLIM_HD uint32_t geo_id(const Geo &g) { return g.id_flags >> 8; }
SLIM_HD uint32_t geo_flags(const Geo &g) { return g.id_flags & 0xFFu; }

}  // namespace slimrt
