// SlimTracin application-layer API on the MI355X runtime (C++17, header-only over include/slimrt.h).
//
// The types, their member order (so the reference's aggregate initialisers compile unchanged) and the render
// call follow the reference's application layer: Scene (src/slim/scene/scene.h), Geometry/Transform
// (src/slim/core/transform.h), Material/Light (src/slim/scene/material.h), Mesh/Triangle (src/slim/scene/mesh.h),
// Texture (src/slim/core/texture.h), Camera/CameraRayProjection (src/slim/scene/camera.h), Canvas
// (src/slim/draw/canvas.h), Viewport (src/slim/viewport/viewport.h) and RayTracingRenderer
// (src/slim/renderer/renderer.h).  render(viewport, update_scene, use_GPU) renders on the MI355X when use_GPU
// is true and on host threads otherwise; host-side derivation (orientations, projection, AABBs, SAH builds)
// goes through the library's slimrt_host_* entry points so both sides agree bit for bit.
//
// Integers are 32-bit (u32 = uint32_t), matching the reference's asset-file layout.  Errors throw
// std::runtime_error with the library's message; there is no silent fallback from GPU to CPU.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../slimrt.h"

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int16_t i16;
typedef int32_t i32;
typedef float f32;
typedef double f64;

#define EPS 0.0001f
#define DEG_TO_RAD 0.0174533f
#define IOR_AIR 1.0003f
#define IOR_GLASS 1.52f
#define MATERIAL_IS_REFLECTIVE ((u8)1)
#define MATERIAL_IS_REFRACTIVE ((u8)2)
#define MATERIAL_IS_EMISSIVE ((u8)4)
#define MATERIAL_HAS_ALBEDO_MAP ((u8)16)
#define MATERIAL_HAS_NORMAL_MAP ((u8)32)
#define GEOMETRY_IS_VISIBLE ((u8)1)
#define GEOMETRY_IS_SHADOWING ((u8)2)
#define GEOMETRY_IS_TRANSPARENT ((u8)4)
#define CAMERA_DEFAULT__FOCAL_LENGTH 2.0f
#define USE_GPU_BY_DEFAULT true
#define RAY_TRACER_DEFAULT_SETTINGS_MAX_DEPTH 3

enum GeometryType { GeometryType_None = 0, GeometryType_Mesh, GeometryType_Grid, GeometryType_Box, GeometryType_Curve,
                    GeometryType_Tet, GeometryType_Quad, GeometryType_Sphere, GeometryType_Light, GeometryType_Count };
enum BRDFType { BRDF_Lambert, BRDF_Blinn, BRDF_Phong, BRDF_CookTorrance };
enum RenderMode { RenderMode_Normals, RenderMode_NormalMap, RenderMode_Beauty, RenderMode_Depth, RenderMode_MipLevel,
                  RenderMode_UVs };
enum AntiAliasing { NoAA, MSAA, SSAA };
enum LightFlags { Light_None, Light_IsDirectional, Light_IsShadowing };
enum ColorID { Black, White, Grey, Red, Green, Blue, Cyan, Magenta, Yellow, DarkRed, DarkGreen, DarkBlue, DarkGrey,
               BrightRed, BrightGreen, BrightBlue, BrightGrey, BrightCyan, BrightMagenta, BrightYellow, DarkCyan,
               DarkMagenta, DarkYellow, F0_Default, F0_Water, F0_Glass_Low, F0_Glass_High, F0_Plastic,
               F0_PlasticHight, F0_Ruby, F0_Diamond, F0_Iron, F0_Copper, F0_Gold, F0_Aluminium, F0_Silver };

struct Color {
    f32 r = 0, g = 0, b = 0;
    Color() = default;
    Color(f32 v) : r{v}, g{v}, b{v} {}
    Color(f32 r, f32 g, f32 b) : r{r}, g{g}, b{b} {}
    // the arithmetic the example apps use (src/slim/core/base.h:1030-1040; multiply then add, no FMA)
    Color operator-(const Color &o) const { return {r - o.r, g - o.g, b - o.b}; }
    Color scaleAdd(f32 factor, const Color &add) const { return {r * factor + add.r, g * factor + add.g, b * factor + add.b}; }
    Color lerpTo(const Color &to, f32 by) const { return (to - *this).scaleAdd(by, *this); }
    Color(ColorID id) {   // the values of Color(ColorID) (src/slim/core/base.h:805-858)
        switch (id) {
            case White: r = g = b = 1.0f; break;
            case Grey: r = g = b = 0.5f; break;
            case DarkGrey: r = g = b = 0.25f; break;
            case BrightGrey: r = g = b = 0.75f; break;
            case Red: r = 1.0f; break;
            case Green: g = 1.0f; break;
            case Blue: b = 1.0f; break;
            case DarkRed: r = 0.5f; break;
            case DarkGreen: g = 0.5f; break;
            case DarkBlue: b = 0.5f; break;
            case DarkCyan: g = b = 0.5f; break;
            case DarkMagenta: r = b = 0.5f; break;
            case DarkYellow: r = g = 0.5f; break;
            case BrightRed: r = 1.0f; g = b = 0.5f; break;
            case BrightGreen: g = 1.0f; r = b = 0.5f; break;
            case BrightBlue: b = 1.0f; r = g = 0.5f; break;
            case Cyan: g = b = 1.0f; break;
            case Magenta: r = b = 1.0f; break;
            case Yellow: r = g = 1.0f; break;
            case BrightCyan: g = b = 0.75f; break;
            case BrightMagenta: r = b = 0.75f; break;
            case BrightYellow: r = g = 0.75f; break;
            case F0_Default: r = g = b = 0.04f; break;
            case F0_Water: r = g = b = 0.02f; break;
            case F0_Plastic: r = g = b = 0.03f; break;
            case F0_PlasticHight: r = g = b = 0.05f; break;
            case F0_Glass_Low: r = g = b = 0.03f; break;
            case F0_Glass_High: r = g = b = 0.08f; break;
            case F0_Ruby: r = g = b = 0.08f; break;
            case F0_Diamond: r = g = b = 0.17f; break;
            case F0_Iron: r = 0.56f; g = 0.57f; b = 0.58f; break;
            case F0_Copper: r = 0.95f; g = 0.64f; b = 0.54f; break;
            case F0_Gold: r = 1.0f; g = 0.71f; b = 0.29f; break;
            case F0_Aluminium: r = 0.91f; g = 0.92f; b = 0.92f; break;
            case F0_Silver: r = 0.95f; g = 0.93f; b = 0.88f; break;
            default: break;
        }
    }
};

inline f32 clampedValue(f32 value, f32 from, f32 to) {   // base.h:222-225
    f32 mn = value < to ? value : to;
    return mn > from ? mn : from;
}
inline f32 clampedValue(f32 value) { return clampedValue(value, 0.0f, 1.0f); }

struct Pixel {   // Pixel (src/slim/core/base.h:1091-1094): what Canvas::setPixel stores
    Color color;
    f32 opacity = 0;
};

struct vec2 {
    f32 x = 0, y = 0;
    vec2() = default;
    vec2(f32 v) : x{v}, y{v} {}
    vec2(f32 x, f32 y) : x{x}, y{y} {}
};
struct UV : vec2 {
    UV() : vec2{1.0f} {}
    UV(f32 v) : vec2{v} {}
    UV(f32 u, f32 v) : vec2{u, v} {}
};

struct vec3 {
    f32 x = 0, y = 0, z = 0;
    vec3() = default;
    vec3(f32 v) : x{v}, y{v}, z{v} {}
    vec3(f32 x, f32 y, f32 z) : x{x}, y{y}, z{z} {}
    vec3 operator+(const vec3 &o) const { return {x + o.x, y + o.y, z + o.z}; }
    vec3 operator-(const vec3 &o) const { return {x - o.x, y - o.y, z - o.z}; }
    vec3 operator*(f32 s) const { return {x * s, y * s, z * s}; }
};

struct AABB {
    vec3 min, max;
};

// Quaternion with the reference's composition and rotation formulas (src/slim/math/quat.h:37-84)
struct quat {
    vec3 axis{0.0f};
    f32 amount = 1.0f;
    quat() = default;
    quat(vec3 axis, f32 amount) : axis{axis}, amount{amount} {}
    void operator*=(const quat &r) {
        f32 x = axis.x, X = r.axis.x, y = axis.y, Y = r.axis.y, z = axis.z, Z = r.axis.z, w = amount, W = r.amount;
        axis.x = X * w + W * x + Z * y - Y * z;
        axis.y = Y * w - Z * x + W * y + X * z;
        axis.z = Z * w + Y * x - X * y + W * z;
        amount = W * w - X * x - Y * y - Z * z;
    }
    static quat AxisAngle(const vec3 &a, f32 radians) {
        radians *= 0.5f;
        quat q{a * sinf(radians), cosf(radians)};
        f32 f = 1.0f / sqrtf(q.axis.x * q.axis.x + q.axis.y * q.axis.y + q.axis.z * q.axis.z + q.amount * q.amount);
        return {q.axis * f, q.amount * f};
    }
    static quat RotationAroundX(f32 r) { return AxisAngle({1, 0, 0}, r); }
    static quat RotationAroundY(f32 r) { return AxisAngle({0, 1, 0}, r); }
    static quat RotationAroundZ(f32 r) { return AxisAngle({0, 0, 1}, r); }
    f32 length() const { return sqrtf(axis.x * axis.x + axis.y * axis.y + axis.z * axis.z + amount * amount); }
    quat operator*(f32 s) const { return {axis * s, amount * s}; }
    quat operator/(f32 s) const { return *this * (1.0f / s); }
    quat normalized() const { return *this / length(); }
    quat operator*(const quat &r) const {
        return {{amount * r.axis.x + axis.x * r.amount + axis.y * r.axis.z - axis.z * r.axis.y,
                 amount * r.axis.y - axis.x * r.axis.z + axis.y * r.amount + axis.z * r.axis.x,
                 amount * r.axis.z + axis.x * r.axis.y - axis.y * r.axis.x + axis.z * r.amount},
                amount * r.amount - axis.x * r.axis.x - axis.y * r.axis.y - axis.z * r.axis.z};
    }
};

// Euler-angle orientations (Orientation<T>, src/slim/core/base.h:554-621): rotation z, then x, then y.
struct OrientationUsingQuaternion : quat {
    OrientationUsingQuaternion(f32 x = 0, f32 y = 0, f32 z = 0) { setRotation(x, y, z); }
    OrientationUsingQuaternion &operator=(const quat &q) { axis = q.axis; amount = q.amount; return *this; }
    void setRotation(f32 x, f32 y, f32 z) {
        ex = x; ey = y; ez = z;
        float e[3] = {x, y, z}, q[4];
        slimrt_host_orientation(e, q);
        axis = vec3{q[0], q[1], q[2]};
        amount = q[3];
    }
    void rotate(f32 x, f32 y, f32 z) { setRotation(ex + x, ey + y, ex + z); }   // as base.h:558-560
    f32 ex = 0, ey = 0, ez = 0;
};
struct OrientationUsing3x3Matrix {
    OrientationUsing3x3Matrix(f32 x = 0, f32 y = 0, f32 z = 0) : x{x}, y{y}, z{z} {}
    void setRotation(f32 X, f32 Y, f32 Z) { x = X; y = Y; z = Z; }
    void rotate(f32 X, f32 Y) { x += X; y += Y; }
    f32 x, y, z;
};

struct Transform {
    OrientationUsingQuaternion orientation{};
    vec3 position{0.0f};
    vec3 scale{1.0f};
};

struct Geometry {
    Transform transform{};
    GeometryType type{GeometryType_None};
    u32 material_id = 0, id = 0;
    u8 flags = GEOMETRY_IS_VISIBLE | GEOMETRY_IS_SHADOWING;
    ColorID color{White};
};

struct Material {
    Color albedo = 1.0f;
    f32 roughness = 1.0f;
    u8 flags = 0;
    u8 texture_count = 0;
    u8 texture_ids[2] = {0, 0};
    BRDFType brdf{BRDF_CookTorrance};
    f32 normal_magnitude = 1.0f;
    UV uv_repeat = 1.0f;
    Color emission = 0.0f;
    f32 IOR = 1.0f;
    Color reflectivity{F0_Default};
    f32 metalness = 0.0f;
    bool isEmissive() const { return flags & MATERIAL_IS_EMISSIVE; }
};

struct Light {
    Color color{White};
    vec3 position_or_direction{};
    f32 intensity = 1.0f;
    u8 flags = Light_IsShadowing;
};

struct TexelQuad { u8 R[4], G[4], B[4]; };   // TL, TR, BL, BR per channel (src/slim/core/texture.h:5-11)
struct TextureMip { u32 width = 0, height = 0; TexelQuad *texel_quads = nullptr; };
struct Texture {
    u32 width = 0, height = 0, mip_count = 0, flags = 0;   // flags: ImageFlags bits (mipmap = 1 << 4, ...)
    TextureMip *mips = nullptr;
};

typedef slimrt_bvh_node BVHNode;
struct BVH { BVHNode *nodes = nullptr; u32 node_count = 0; u32 height = 0; };
typedef slimrt_triangle Triangle;
struct TriangleVertexIndices { u32 v1, v2, v3; };
struct EdgeVertexIndices { u32 from, to; };   // src/slim/scene/mesh.h:8-10

struct Mesh {
    AABB aabb;
    BVH bvh;
    Triangle *triangles = nullptr;
    vec3 *vertex_positions = nullptr;
    vec3 *vertex_normals = nullptr;
    vec2 *vertex_uvs = nullptr;
    TriangleVertexIndices *vertex_position_indices = nullptr;
    TriangleVertexIndices *vertex_normal_indices = nullptr;
    TriangleVertexIndices *vertex_uvs_indices = nullptr;
    EdgeVertexIndices *edge_vertex_indices = nullptr;
    u32 triangle_count = 0, vertex_count = 0, edge_count = 0, normals_count = 0, uvs_count = 0;
};

namespace slimtracin_detail {
inline void check(int rc, const slimrt_context *ctx, const char *what) {
    if (rc != SLIMRT_OK)
        throw std::runtime_error(std::string(what) + ": " + (ctx ? slimrt_last_error(ctx) : "invalid argument"));
}
// Storage for arrays the API allocates on the caller's behalf (meshes built in memory, loaded assets).
inline std::vector<std::vector<unsigned char>> &arena() {
    static std::vector<std::vector<unsigned char>> a;
    return a;
}
template <typename T> T *alloc(size_t n) {
    arena().emplace_back(n * sizeof(T) + 1);
    return reinterpret_cast<T *>(arena().back().data());
}
}  // namespace slimtracin_detail

// BVHBuilder::buildMesh (src/slim/scene/bvh_builder.h:265-341) through slimrt_host_build_mesh.
struct BVHBuilder {
    u32 max_leaf_size;
    explicit BVHBuilder(u32 max_leaf_size = 2) : max_leaf_size{max_leaf_size} {}
    void buildMesh(Mesh &m) const {
        using namespace slimtracin_detail;
        if (!m.bvh.nodes) m.bvh.nodes = alloc<BVHNode>(2 * (size_t)m.triangle_count);
        if (!m.triangles) m.triangles = alloc<Triangle>(m.triangle_count);
        float ab[6];
        check(slimrt_host_build_mesh((const float *)m.vertex_positions, m.vertex_count, (const u32 *)m.vertex_position_indices,
                                     (const float *)m.vertex_normals, m.normals_count, (const u32 *)m.vertex_normal_indices,
                                     (const float *)m.vertex_uvs, m.uvs_count, (const u32 *)m.vertex_uvs_indices,
                                     m.triangle_count, max_leaf_size, m.bvh.nodes, &m.bvh.node_count, &m.bvh.height,
                                     m.triangles, ab), nullptr, "buildMesh");
        m.aabb = AABB{{ab[0], ab[1], ab[2]}, {ab[3], ab[4], ab[5]}};
    }
};

// Readers for the reference's .mesh and .texture files (src/slim/serialization/mesh.h:56-122, texture.h:29-70;
// 4-byte u32 layout as shipped).
inline bool load(Mesh &m, const char *path) {
    using namespace slimtracin_detail;
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    u32 h[7];
    bool ok = fread(h, 4, 7, f) == 7;
    m = Mesh{};
    m.vertex_count = h[0]; m.triangle_count = h[1]; m.edge_count = h[2]; m.uvs_count = h[3]; m.normals_count = h[4];
    m.bvh.node_count = h[5]; m.bvh.height = h[6] & 0xFFu;
    float ab[6];
    ok = ok && fread(ab, 4, 6, f) == 6;
    m.aabb = AABB{{ab[0], ab[1], ab[2]}, {ab[3], ab[4], ab[5]}};
    m.triangles = alloc<Triangle>(m.triangle_count);
    m.vertex_positions = alloc<vec3>(m.vertex_count);
    m.vertex_position_indices = alloc<TriangleVertexIndices>(m.triangle_count);
    ok = ok && fread(m.triangles, sizeof(Triangle), m.triangle_count, f) == m.triangle_count;
    ok = ok && fread(m.vertex_positions, 12, m.vertex_count, f) == m.vertex_count;
    ok = ok && fread(m.vertex_position_indices, 12, m.triangle_count, f) == m.triangle_count;
    m.edge_vertex_indices = alloc<EdgeVertexIndices>(m.edge_count);
    ok = ok && fread(m.edge_vertex_indices, 8, m.edge_count, f) == m.edge_count;
    if (m.uvs_count) {
        m.vertex_uvs = alloc<vec2>(m.uvs_count);
        m.vertex_uvs_indices = alloc<TriangleVertexIndices>(m.triangle_count);
        ok = ok && fread(m.vertex_uvs, 8, m.uvs_count, f) == m.uvs_count;
        ok = ok && fread(m.vertex_uvs_indices, 12, m.triangle_count, f) == m.triangle_count;
    }
    if (m.normals_count) {
        m.vertex_normals = alloc<vec3>(m.normals_count);
        m.vertex_normal_indices = alloc<TriangleVertexIndices>(m.triangle_count);
        ok = ok && fread(m.vertex_normals, 12, m.normals_count, f) == m.normals_count;
        ok = ok && fread(m.vertex_normal_indices, 12, m.triangle_count, f) == m.triangle_count;
    }
    m.bvh.nodes = alloc<BVHNode>(m.bvh.node_count);
    ok = ok && fread(m.bvh.nodes, sizeof(BVHNode), m.bvh.node_count, f) == m.bvh.node_count;
    fclose(f);
    return ok;
}
inline bool load(Texture &t, const char *path) {
    using namespace slimtracin_detail;
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    u32 info[8];   // ImageInfo: width, height, size, stride, tile_width, tile_height, mip_count, flags
    bool ok = fread(info, 4, 8, f) == 8;
    t = Texture{};
    t.width = info[0]; t.height = info[1]; t.mip_count = info[6]; t.flags = info[7];
    t.mips = alloc<TextureMip>(t.mip_count);
    for (u32 i = 0; ok && i < t.mip_count; i++) {
        u32 wh[2];
        ok = fread(wh, 4, 2, f) == 2;
        t.mips[i].width = wh[0];
        t.mips[i].height = wh[1];
        size_t n = (size_t)(wh[0] + 1) * (wh[1] + 1);
        t.mips[i].texel_quads = alloc<TexelQuad>(n);
        ok = ok && fread(t.mips[i].texel_quads, 12, n, f) == n;
    }
    fclose(f);
    return ok;
}

// Writers in the same layout (Mesh / Texture / BVH save, src/slim/serialization/mesh.h:56-65 and 104-122,
// texture.h:61-68 with ImageInfo's header, base.h:1544-1546 and 1588-1595, bvh.h:18-48).
inline bool save(const Mesh &m, const char *path) {
    FILE *f = fopen(path, "wb");
    if (!f) return false;
    const u32 edges = m.edge_vertex_indices ? m.edge_count : 0;
    const u32 h[7] = {m.vertex_count, m.triangle_count, edges, m.uvs_count, m.normals_count, m.bvh.node_count, m.bvh.height};
    bool ok = fwrite(h, 4, 7, f) == 7;
    const float ab[6] = {m.aabb.min.x, m.aabb.min.y, m.aabb.min.z, m.aabb.max.x, m.aabb.max.y, m.aabb.max.z};
    ok = ok && fwrite(ab, 4, 6, f) == 6;
    ok = ok && fwrite(m.triangles, sizeof(Triangle), m.triangle_count, f) == m.triangle_count;
    ok = ok && fwrite(m.vertex_positions, 12, m.vertex_count, f) == m.vertex_count;
    ok = ok && fwrite(m.vertex_position_indices, 12, m.triangle_count, f) == m.triangle_count;
    ok = ok && fwrite(m.edge_vertex_indices, 8, edges, f) == edges;
    if (m.uvs_count) {
        ok = ok && fwrite(m.vertex_uvs, 8, m.uvs_count, f) == m.uvs_count;
        ok = ok && fwrite(m.vertex_uvs_indices, 12, m.triangle_count, f) == m.triangle_count;
    }
    if (m.normals_count) {
        ok = ok && fwrite(m.vertex_normals, 12, m.normals_count, f) == m.normals_count;
        ok = ok && fwrite(m.vertex_normal_indices, 12, m.triangle_count, f) == m.triangle_count;
    }
    ok = ok && fwrite(m.bvh.nodes, sizeof(BVHNode), m.bvh.node_count, f) == m.bvh.node_count;
    return fclose(f) == 0 && ok;
}
inline bool save(const Texture &t, const char *path) {
    if (t.flags & SLIMRT_TEXELS_MASK) return false;   // the .texture format holds texel quads only
    FILE *f = fopen(path, "wb");
    if (!f) return false;
    const u32 info[8] = {t.width, t.height, t.width * t.height, t.width, 0, 0, t.mip_count, t.flags};
    bool ok = fwrite(info, 4, 8, f) == 8;
    for (u32 i = 0; ok && i < t.mip_count; i++) {
        const u32 wh[2] = {t.mips[i].width, t.mips[i].height};
        const size_t n = (size_t)(wh[0] + 1) * (wh[1] + 1);
        ok = fwrite(wh, 4, 2, f) == 2 && fwrite(t.mips[i].texel_quads, 12, n, f) == n;
    }
    return fclose(f) == 0 && ok;
}
inline bool load(BVH &b, const char *path) {
    using namespace slimtracin_detail;
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    u32 h[2];
    bool ok = fread(h, 4, 2, f) == 2;
    b = BVH{};
    if (ok) {
        b.node_count = h[0];
        b.height = h[1];
        b.nodes = alloc<BVHNode>(b.node_count);
        ok = fread(b.nodes, sizeof(BVHNode), b.node_count, f) == b.node_count;
    }
    fclose(f);
    return ok;
}
inline bool save(const BVH &b, const char *path) {
    FILE *f = fopen(path, "wb");
    if (!f) return false;
    const u32 h[2] = {b.node_count, b.height};
    bool ok = fwrite(h, 4, 2, f) == 2 && fwrite(b.nodes, sizeof(BVHNode), b.node_count, f) == b.node_count;
    return fclose(f) == 0 && ok;
}

// The reference converter (src/bmp2texture.cpp) in memory: a Texture from an RGB image (width * height * 3 bytes,
// rows top to bottom) and ImageFlags bits, through slimrt_host_bake_texture.
inline Texture bakeTexture(const u8 *rgb, u32 width, u32 height, u32 flags) {
    using namespace slimtracin_detail;
    size_t total = 0;
    u32 mip_count = 0, dims[32], out_flags = 0;
    check(slimrt_host_bake_texture(rgb, width, height, flags, nullptr, 0, &total, &mip_count, dims, &out_flags),
          nullptr, "bakeTexture");
    u8 *quads = alloc<u8>(total * 12);
    check(slimrt_host_bake_texture(rgb, width, height, flags, quads, total, &total, &mip_count, dims, &out_flags),
          nullptr, "bakeTexture");
    Texture t;
    t.width = width;
    t.height = height;
    t.flags = out_flags;
    t.mip_count = mip_count;
    t.mips = alloc<TextureMip>(mip_count);
    for (u32 i = 0; i < mip_count; i++) {
        t.mips[i].width = dims[2 * i];
        t.mips[i].height = dims[2 * i + 1];
        t.mips[i].texel_quads = reinterpret_cast<TexelQuad *>(quads);
        quads += 12 * (size_t)(dims[2 * i] + 1) * (dims[2 * i + 1] + 1);
    }
    return t;
}

// Bytes of one mip in the texture's storage: the reference's texel quads, or a bordered grid of RGBA8 / RGBA16F
// texels (SLIMRT_TEXELS_* in the flags; include/slimrt.h).
inline size_t mipBytes(const Texture &t, u32 mip) {
    const size_t w = t.mips[mip].width, h = t.mips[mip].height;
    switch (t.flags & SLIMRT_TEXELS_MASK) {
        case SLIMRT_TEXELS_RGBA8: return 4 * (w + 2) * (h + 2);
        case SLIMRT_TEXELS_RGBA16F: return 8 * (w + 2) * (h + 2);
        default: return 12 * (w + 1) * (h + 1);
    }
}

// A texel-quad texture in texel storage (no reference counterpart): RGBA8 samples exactly as the quads whenever
// no quad's corners differ from the texel grid (*mismatches, summed over the mips, may be null); RGBA16F holds each
// component as the nearest half to u8 / 255.  The texels point into memory the API keeps, as for loaded assets.
inline Texture toTexelStorage(const Texture &t, u32 storage, u64 *mismatches = nullptr) {
    using namespace slimtracin_detail;
    if ((t.flags & SLIMRT_TEXELS_MASK) || (storage != SLIMRT_TEXELS_RGBA8 && storage != SLIMRT_TEXELS_RGBA16F))
        throw std::runtime_error("toTexelStorage: a texel-quad texture and SLIMRT_TEXELS_RGBA8 or _RGBA16F expected");
    Texture out = t;
    out.flags = t.flags | storage;
    out.mips = alloc<TextureMip>(t.mip_count);
    if (mismatches) *mismatches = 0;
    for (u32 i = 0; i < t.mip_count; i++) {
        out.mips[i].width = t.mips[i].width;
        out.mips[i].height = t.mips[i].height;
        u8 *texels = alloc<u8>(mipBytes(out, i));
        uint64_t bad = 0;
        check(slimrt_host_texels_from_quads((const uint8_t *)t.mips[i].texel_quads, t.mips[i].width, t.mips[i].height,
                                            storage, texels, &bad), nullptr, "toTexelStorage");
        if (mismatches) *mismatches += bad;
        out.mips[i].texel_quads = reinterpret_cast<TexelQuad *>(texels);
    }
    return out;
}

struct Camera {
    OrientationUsing3x3Matrix orientation{};
    vec3 position{0};
    f32 focal_length{CAMERA_DEFAULT__FOCAL_LENGTH};
};

struct Dimensions {
    u32 width_times_height = 0;
    u16 width = 0, height = 0;
    void update(u16 w, u16 h) { width = w; height = h; width_times_height = (u32)w * h; }
};

// Canvas (src/slim/draw/canvas.h:11-39): pixels (x4 under SSAA) and depths, in caller-owned buffers or, when
// none are given (as the reference's platform layer hands its canvases memory), in the canvas's own storage,
// sized at each render to the current dimensions.
struct Canvas {
    Dimensions dimensions;
    AntiAliasing antialias = NoAA;
    Pixel *pixels = nullptr;
    f32 *depths = nullptr;
    std::vector<Pixel> own_pixels;
    std::vector<f32> own_depths;
    Canvas() = default;
    Canvas(u16 width, u16 height, AntiAliasing aa = NoAA, Pixel *pixels = nullptr, f32 *depths = nullptr)
        : antialias{aa}, pixels{pixels}, depths{depths} { dimensions.update(width, height); }
    size_t sampleCount() const { return (size_t)dimensions.width_times_height * (antialias == SSAA ? 4 : 1); }
    void ensureStorage() {
        if (pixels && pixels != own_pixels.data()) return;   // the caller's buffers
        own_pixels.resize(sampleCount());
        own_depths.resize(sampleCount());
        pixels = own_pixels.data();
        depths = own_depths.data();
    }
};

struct Viewport {
    Canvas &canvas;
    Camera *camera;
    Viewport(Canvas &canvas, Camera *camera) : canvas{canvas}, camera{camera} {}
};

// CameraRayProjection::reset (src/slim/scene/camera.h:69-83) through slimrt_host_projection
struct CameraRayProjection {
    slimrt_render_params params{};
    void reset(const Camera &camera, const Dimensions &dim, bool antialias) {
        float e[3] = {camera.orientation.x, camera.orientation.y, camera.orientation.z};
        float p[3] = {camera.position.x, camera.position.y, camera.position.z};
        slimtracin_detail::check(slimrt_host_projection(e, p, camera.focal_length, dim.width, dim.height,
                                                        antialias ? 1u : 0u, &params), nullptr, "projection.reset");
    }
};

struct SceneCounts {
    u32 geometries, cameras, lights, materials, textures, meshes, grids, boxes, tets, quads, curves;
    SceneCounts(u32 geometries = 0, u32 cameras = 0, u32 lights = 0, u32 materials = 0, u32 textures = 0,
                u32 meshes = 0, u32 grids = 0, u32 boxes = 0, u32 tets = 0, u32 quads = 0, u32 curves = 0)
        : geometries{geometries}, cameras{cameras}, lights{lights}, materials{materials}, textures{textures},
          meshes{meshes}, grids{grids}, boxes{boxes}, tets{tets}, quads{quads}, curves{curves} {}
};

#define SCENE_HAD_EMISSIVE_QUADS 1

struct Scene {
    SceneCounts counts;
    u16 flags = 0, mesh_stack_size = 0;
    Geometry *geometries;
    Camera *cameras;
    Light *lights;
    Material *materials;
    Texture *textures;
    Mesh *meshes;
    std::vector<float> aabbs;
    std::vector<BVHNode> bvh_nodes;
    std::vector<u32> bvh_leaf_geometry_indices;
    BVH bvh;

    // Scene::Scene (scene.h:141-246) for in-memory arrays; texture/mesh files are loaded beforehand with load().
    Scene(SceneCounts counts, Geometry *geometries = nullptr, Camera *cameras = nullptr, Light *lights = nullptr,
          Material *materials = nullptr, Texture *textures = nullptr, Mesh *meshes = nullptr)
        : counts{counts}, geometries{geometries}, cameras{cameras}, lights{lights}, materials{materials},
          textures{textures}, meshes{meshes} {
        for (u32 i = 0; i < counts.meshes; i++)
            if (meshes[i].bvh.height > mesh_stack_size) mesh_stack_size = (u16)meshes[i].bvh.height;
        if (counts.meshes) mesh_stack_size += 2;
        for (u32 i = 0; i < counts.geometries; i++)
            if (geometries[i].type == GeometryType_Quad && materials[geometries[i].material_id].isEmissive())
                flags = SCENE_HAD_EMISSIVE_QUADS;
        updateAABBs();
        updateBVH();
    }

    std::vector<slimrt_geometry> geometryRecords() const {
        std::vector<slimrt_geometry> g(counts.geometries);
        for (u32 i = 0; i < counts.geometries; i++) {
            const Geometry &s = geometries[i];
            g[i] = slimrt_geometry{{s.transform.orientation.axis.x, s.transform.orientation.axis.y,
                                    s.transform.orientation.axis.z, s.transform.orientation.amount},
                                   {s.transform.position.x, s.transform.position.y, s.transform.position.z},
                                   {s.transform.scale.x, s.transform.scale.y, s.transform.scale.z},
                                   (u32)s.type, s.material_id, s.id, s.flags};
        }
        return g;
    }
    std::vector<slimrt_mesh> meshRecords() const {
        std::vector<slimrt_mesh> m(counts.meshes);
        for (u32 i = 0; i < counts.meshes; i++) {
            const Mesh &s = meshes[i];
            m[i] = slimrt_mesh{{s.aabb.min.x, s.aabb.min.y, s.aabb.min.z}, {s.aabb.max.x, s.aabb.max.y, s.aabb.max.z},
                               s.bvh.nodes, s.bvh.node_count, s.bvh.height, s.triangles, s.triangle_count,
                               s.normals_count, s.uvs_count};
        }
        return m;
    }
    void updateAABBs() {   // scene.h:176-179
        auto g = geometryRecords();
        auto m = meshRecords();
        aabbs.resize(6 * (size_t)counts.geometries);
        slimtracin_detail::check(slimrt_host_world_aabbs(g.data(), counts.geometries, m.data(), counts.meshes, aabbs.data()),
                                 nullptr, "updateAABBs");
    }
    void updateBVH(u16 max_leaf_size = 1) {   // scene.h:181-191
        bvh_nodes.resize(2 * (size_t)counts.geometries);
        bvh_leaf_geometry_indices.resize(counts.geometries);
        slimtracin_detail::check(slimrt_host_build_bvh(aabbs.data(), counts.geometries, max_leaf_size, bvh_nodes.data(),
                                                       &bvh.node_count, &bvh.height, bvh_leaf_geometry_indices.data()),
                                 nullptr, "updateBVH");
        bvh.nodes = bvh_nodes.data();
    }
};

// Traversal stacks live on the device (LDS) or in the host renderer; kept for source compatibility.
struct SceneTracer {
    SceneTracer(u32 stack_size = 0, u32 mesh_stack_size = 0) { (void)stack_size; (void)mesh_stack_size; }
};

struct RayTracerSettings {
    u8 max_depth;
    char skybox_color_texture_id, skybox_radiance_texture_id, skybox_irradiance_texture_id;
    RenderMode render_mode;
    ColorID mip_level_colors[9];
};

namespace slimtracin_detail {
struct Contexts {
    slimrt_context *gpu = nullptr, *cpu = nullptr;
    std::string gpu_error;
    bool gpu_tried = false;
};
inline Contexts &contexts() {
    static Contexts c;
    return c;
}
inline std::vector<slimrt_material> materialRecords(const Scene &s) {
    std::vector<slimrt_material> m(s.counts.materials);
    for (u32 i = 0; i < s.counts.materials; i++) {
        const Material &a = s.materials[i];
        m[i] = slimrt_material{{a.albedo.r, a.albedo.g, a.albedo.b}, a.roughness,
                               {a.emission.r, a.emission.g, a.emission.b}, a.IOR,
                               {a.reflectivity.r, a.reflectivity.g, a.reflectivity.b}, a.metalness,
                               {a.uv_repeat.x, a.uv_repeat.y}, a.normal_magnitude, a.flags, a.texture_count,
                               {a.texture_ids[0], a.texture_ids[1]}, (u32)a.brdf};
    }
    return m;
}
inline std::vector<slimrt_light> lightRecords(const Scene &s) {
    std::vector<slimrt_light> l(s.counts.lights);
    for (u32 i = 0; i < s.counts.lights; i++) {
        const Light &a = s.lights[i];
        l[i] = slimrt_light{{a.color.r, a.color.g, a.color.b}, a.intensity,
                            {a.position_or_direction.x, a.position_or_direction.y, a.position_or_direction.z}, a.flags};
    }
    return l;
}
inline void uploadAll(slimrt_context *ctx, const Scene &s) {
    auto g = s.geometryRecords();
    auto m = s.meshRecords();
    auto mats = materialRecords(s);
    auto lights = lightRecords(s);
    std::vector<std::vector<slimrt_texture_mip>> mips(s.counts.textures);
    std::vector<slimrt_texture> tex(s.counts.textures);
    for (u32 i = 0; i < s.counts.textures; i++) {
        const Texture &t = s.textures[i];
        for (u32 k = 0; k < t.mip_count; k++)
            mips[i].push_back(slimrt_texture_mip{t.mips[k].width, t.mips[k].height, (const uint8_t *)t.mips[k].texel_quads});
        tex[i] = slimrt_texture{t.width, t.height, t.flags, t.mip_count, mips[i].data()};
    }
    slimrt_scene_desc d{g.data(), s.counts.geometries, mats.data(), s.counts.materials, lights.data(), s.counts.lights,
                        m.data(), s.counts.meshes, tex.data(), s.counts.textures, s.bvh.nodes, s.bvh.node_count,
                        s.bvh.height, s.bvh_leaf_geometry_indices.data()};
    check(slimrt_upload_scene(ctx, &d), ctx, "upload scene");
}
inline slimrt_context *gpu() {
    Contexts &c = contexts();
    if (!c.gpu) throw std::runtime_error("MI355X path unavailable: " + c.gpu_error);
    return c.gpu;
}
}  // namespace slimtracin_detail

// renderer_GPU.h's free functions (renderer_GPU.h:70-181), with the MI355X runtime behind them.
inline void initDataOnGPU(const Scene &scene) {
    using namespace slimtracin_detail;
    Contexts &c = contexts();
    if (!c.gpu && !c.gpu_tried) {
        c.gpu_tried = true;
        if (slimrt_create(&c.gpu, SLIMRT_DEVICE_GPU) != SLIMRT_OK) {
            c.gpu = nullptr;
            c.gpu_error = "no HIP device";
            return;   // the CPU side still works; a GPU render will throw
        }
    }
    if (c.gpu) uploadAll(c.gpu, scene);
}
inline void uploadGeometries(const Scene &s) {
    auto g = s.geometryRecords();
    slimtracin_detail::check(slimrt_upload_geometries(slimtracin_detail::gpu(), g.data(), s.counts.geometries),
                             slimtracin_detail::gpu(), "uploadGeometries");
}
inline void uploadMaterials(const Scene &s) {
    auto m = slimtracin_detail::materialRecords(s);
    slimtracin_detail::check(slimrt_upload_materials(slimtracin_detail::gpu(), m.data(), s.counts.materials),
                             slimtracin_detail::gpu(), "uploadMaterials");
}
inline void uploadLights(const Scene &s) {
    auto l = slimtracin_detail::lightRecords(s);
    slimtracin_detail::check(slimrt_upload_lights(slimtracin_detail::gpu(), l.data(), s.counts.lights),
                             slimtracin_detail::gpu(), "uploadLights");
}
inline void uploadCameras(const Scene &) {}   // cameras are read on the host only
inline void uploadSceneBVH(const Scene &s) {
    slimtracin_detail::check(slimrt_upload_scene_bvh(slimtracin_detail::gpu(), s.bvh.nodes, s.bvh.node_count,
                                                     s.bvh_leaf_geometry_indices.data(), s.bvh.height),
                             slimtracin_detail::gpu(), "uploadSceneBVH");
}

struct RayTracingRenderer {
    Scene &scene;
    SceneTracer &scene_tracer;
    CameraRayProjection &projection;
    RayTracerSettings settings;

    explicit RayTracingRenderer(Scene &scene, SceneTracer &scene_tracer, CameraRayProjection &projection,
                                u8 max_depth = RAY_TRACER_DEFAULT_SETTINGS_MAX_DEPTH, char skybox_color_texture_id = -1,
                                char skybox_radiance_texture_id = -1, char skybox_irradiance_texture_id = -1,
                                RenderMode render_mode = RenderMode_Beauty)
        : scene{scene}, scene_tracer{scene_tracer}, projection{projection} {
        settings.skybox_color_texture_id = skybox_color_texture_id;
        settings.skybox_radiance_texture_id = skybox_radiance_texture_id;
        settings.skybox_irradiance_texture_id = skybox_irradiance_texture_id;
        settings.max_depth = max_depth;
        settings.render_mode = render_mode;
        const ColorID mips[9] = {BrightRed, BrightYellow, BrightGreen, BrightMagenta, BrightCyan, BrightBlue,
                                 BrightGrey, Grey, DarkGrey};
        for (int i = 0; i < 9; i++) settings.mip_level_colors[i] = mips[i];
        initDataOnGPU(scene);
    }

    // RayTracingRenderer::render (renderer.h:63-85): use_GPU selects the MI355X kernels, otherwise host threads.
    void render(const Viewport &viewport, bool update_scene = true, bool use_GPU = false) {
        using namespace slimtracin_detail;
        Canvas &canvas = viewport.canvas;
        canvas.ensureStorage();
        if (update_scene) {
            scene.updateAABBs();
            scene.updateBVH();
            if (use_GPU) {
                uploadLights(scene);
                uploadCameras(scene);
                uploadGeometries(scene);
                uploadSceneBVH(scene);
            }
        }
        slimrt_render_params p = projection.params;
        p.width = canvas.dimensions.width;
        p.height = canvas.dimensions.height;
        p.ssaa = canvas.antialias == SSAA;
        p.max_depth = settings.max_depth;
        p.render_mode = (u32)settings.render_mode;
        p.skybox_color = settings.skybox_color_texture_id;
        p.skybox_radiance = settings.skybox_radiance_texture_id;
        p.skybox_irradiance = settings.skybox_irradiance_texture_id;
        p.output = SLIMRT_OUTPUT_CANVAS;
        p.flags = 0;
        for (int i = 0; i < 9; i++) {
            Color c{settings.mip_level_colors[i]};
            p.mip_level_colors[i][0] = c.r; p.mip_level_colors[i][1] = c.g; p.mip_level_colors[i][2] = c.b;
        }
        slimrt_context *ctx;
        if (use_GPU) {
            ctx = gpu();
        } else {
            Contexts &c = contexts();
            if (!c.cpu) check(slimrt_create(&c.cpu, SLIMRT_DEVICE_CPU), nullptr, "create CPU context");
            ctx = c.cpu;
            uploadAll(ctx, scene);   // the host context reads the scene's current arrays
        }
        check(slimrt_render(ctx, &p, (slimrt_pixel *)canvas.pixels, canvas.depths), ctx, "render");
        if (frameHook()) frameHook()(scene, viewport, p, use_GPU);
    }

    // Called after every render with what was rendered (the scene, the viewport and its canvas, the frame's
    // parameters); frame recorders for offline checks install one (include/slimtracin/app.hpp).
    using FrameHook = void (*)(const Scene &, const Viewport &, const slimrt_render_params &, bool use_GPU);
    static FrameHook &frameHook() {
        static FrameHook hook = nullptr;
        return hook;
    }
};

// ---------------------------------------------------------------------------------------- wireframe overlays
// drawBox / drawBVH / drawMesh (src/slim/draw/box.h:7-34, bvh.h:8-28, mesh.h:7-22) into the viewport's canvas
// after a render, drawn by the library exactly as the reference draws them (slimrt_draw_boxes,
// slimrt_draw_mesh_wireframe).  The canvas here is host memory, so the host context draws; a canvas in HBM is
// drawn by the same entry points on a GPU context.  The viewport uses the reference's default frustum.
#define BOX__ALL_SIDES 63
enum BoxSide { BoxSide_None = 0, BoxSide_Left = 1, BoxSide_Bottom = 2, BoxSide_Back = 4, BoxSide_Right = 8,
               BoxSide_Top = 16, BoxSide_Front = 32 };
struct Box {   // Box with its vertices given by their AABB (BoxVertices(min, max), src/slim/scene/box.h:17-29)
    AABB vertices{vec3{-1.0f}, vec3{1.0f}};
};

namespace slimtracin_detail {
inline slimrt_context *cpu() {
    Contexts &c = contexts();
    if (!c.cpu) check(slimrt_create(&c.cpu, SLIMRT_DEVICE_CPU), nullptr, "create CPU context");
    return c.cpu;
}
inline slimrt_overlay_view overlayView(const Viewport &v) {
    const Camera &cam = *v.camera;
    return slimrt_overlay_view{{cam.orientation.x, cam.orientation.y, cam.orientation.z},
                               {cam.position.x, cam.position.y, cam.position.z}, cam.focal_length,
                               v.canvas.dimensions.width, v.canvas.dimensions.height,
                               v.canvas.antialias == SSAA ? 1u : 0u, 0.001f, 1000.0f};
}
inline slimrt_overlay_transform overlayTransform(const Transform &t) {
    return slimrt_overlay_transform{{t.orientation.axis.x, t.orientation.axis.y, t.orientation.axis.z, t.orientation.amount},
                                    {t.position.x, t.position.y, t.position.z}, {t.scale.x, t.scale.y, t.scale.z}};
}
inline void drawBoxes(const Viewport &viewport, const slimrt_overlay_box *boxes, u32 count) {
    const slimrt_overlay_view view = overlayView(viewport);
    check(slimrt_draw_boxes(cpu(), &view, boxes, count, viewport.canvas.pixels, viewport.canvas.depths), cpu(),
          "draw boxes");
}
}  // namespace slimtracin_detail

inline void drawBox(const Box &box, const Transform &transform, const Viewport &viewport, const Color &color = White,
                    f32 opacity = 1.0f, u8 line_width = 1, u8 sides = BOX__ALL_SIDES) {
    const AABB &a = box.vertices;
    const slimrt_overlay_box b{slimtracin_detail::overlayTransform(transform),
                               {a.min.x, a.min.y, a.min.z}, {a.max.x, a.max.y, a.max.z},
                               {color.r, color.g, color.b}, opacity, line_width, sides};
    slimtracin_detail::drawBoxes(viewport, &b, 1);
}

inline void drawBVH(const BVH &bvh, const Transform &transform, const Viewport &viewport, u16 min_depth = 0,
                    u16 max_depth = 5, f32 opacity = 0.25f, u8 line_width = 1) {
    std::vector<slimrt_overlay_box> boxes(bvh.node_count);
    const slimrt_overlay_transform t = slimtracin_detail::overlayTransform(transform);
    u32 n = 0;
    slimtracin_detail::check(slimrt_host_bvh_boxes(bvh.nodes, bvh.node_count, &t, min_depth, max_depth, opacity,
                                                   line_width, boxes.data(), &n), nullptr, "drawBVH");
    slimtracin_detail::drawBoxes(viewport, boxes.data(), n);
}

inline void drawMesh(const Mesh &mesh, const Transform &transform, bool draw_normals, const Viewport &viewport,
                     const Color &color = White, f32 opacity = 1.0f, u8 line_width = 1) {
    using namespace slimtracin_detail;
    const slimrt_overlay_view view = overlayView(viewport);
    const slimrt_overlay_transform t = overlayTransform(transform);
    const float col[3] = {color.r, color.g, color.b};
    check(slimrt_draw_mesh_wireframe(cpu(), &view, (const float *)mesh.vertex_positions, mesh.vertex_count,
                                     (const uint32_t *)mesh.edge_vertex_indices, mesh.edge_count, &t, col, opacity,
                                     line_width, viewport.canvas.pixels, viewport.canvas.depths),
          cpu(), "drawMesh");
    if (draw_normals && mesh.normals_count && mesh.vertex_normals && mesh.vertex_normal_indices) {
        // one segment per triangle corner, position to position + normal * 0.1 (mesh.h:24-33), in Red at half opacity
        std::vector<float> ends;
        std::vector<uint32_t> segs;
        for (u32 i = 0; i < mesh.triangle_count; i++) {
            const u32 *pi = &mesh.vertex_position_indices[i].v1, *ni = &mesh.vertex_normal_indices[i].v1;
            for (int k = 0; k < 3; k++) {
                const vec3 p = mesh.vertex_positions[pi[k]], to = mesh.vertex_normals[ni[k]] * 0.1f + p;
                segs.push_back((uint32_t)(ends.size() / 3));
                ends.insert(ends.end(), {p.x, p.y, p.z});
                segs.push_back((uint32_t)(ends.size() / 3));
                ends.insert(ends.end(), {to.x, to.y, to.z});
            }
        }
        const float red[3] = {1.0f, 0.0f, 0.0f};
        check(slimrt_draw_mesh_wireframe(cpu(), &view, ends.data(), (uint32_t)(ends.size() / 3), segs.data(),
                                         (uint32_t)(segs.size() / 2), &t, red, opacity * 0.5f, line_width,
                                         viewport.canvas.pixels, viewport.canvas.depths),
              cpu(), "drawMesh normals");
    }
}
