// The reference's seven example apps (src/examples/01_Lights.cpp .. 07_Meshes.cpp) on the SlimTracin
// application API over the MI355X runtime, run by the headless platform loop (include/slimtracin/app.hpp).
//
// Each app keeps its reference's scene setup - geometries, transforms, materials, lights, camera, bounce depth,
// skybox textures - and its per-frame animation (OnUpdate).  Left out: the HUD, picking and interactive camera
// navigation and key bindings, which need a window's input.  Assets: the textures and meshes are loaded from
// --assets DIR under the reference's file names (src/examples/*.texture, *.mesh); a texture whose .texture file
// is absent is baked from its .bmp source by the reference converter's procedure (loadBitmap + bakeTexture).
// What is absent in both forms is replaced by a stand-in, reported on stderr: floor and dog maps baked from
// procedural images, a skybox slot by the first cube map found (or a baked gradient), dragon.mesh by monkey.mesh
// and dog.mesh by cube.mesh.
//
//   slim_examples <01..07> [--frames N] [--size WxH] [--cpu|--gpu] [--ssaa] [--assets DIR] [--ppm PREFIX]
//                          [--record FRAME PREFIX]
//
// --ppm writes every frame's window content to PREFIX<frame>.ppm; --record writes frame FRAME's scene spec and
// canvas (PREFIX.spec, PREFIX.canvas) for checks against the CPU oracle.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "slimtracin/app.hpp"

using slimtracin_app::HeadlessOptions;

// ---------------------------------------------------------------------------------------------- assets
enum TextureID {   // src/examples/textures.h:6-21
    Floor_Albedo, Floor_Normal, Dog_Albedo, Dog_Normal,
    Cathedral_SkyboxColor, Cathedral_SkyboxRadiance, Cathedral_SkyboxIrradiance,
    Bolonga_SkyboxColor, Bolonga_SkyboxRadiance, Bolonga_SkyboxIrradiance,
    TextureCount
};
static const char *kTextureFiles[TextureCount] = {
    "floor_albedo.texture", "floor_normal.texture", "dog_albedo.texture", "dog_normal.texture",
    "cathedral_color.texture", "cathedral_radiance.texture", "cathedral_irradiance.texture",
    "bolonga_color.texture", "bolonga_radiance.texture", "bolonga_irradiance.texture"};
static Texture textures[TextureCount];

#define IMG_LINEAR (1u << 1)
#define IMG_MIPMAP (1u << 4)
#define IMG_WRAP (1u << 6)
#define IMG_NORMAL (1u << 7)
#define IMG_CUBEMAP (1u << 8)

// Procedural stand-ins, baked by the reference converter's procedure (bakeTexture): a two-tone checker with a
// soft ring pattern for albedo, the normals of a ridged height field for a normal map, a vertical gradient for a
// cube map (4-face strip + top + bottom, width = 6 * height).
static Texture bakedAlbedo(u32 size, Color a, Color b, u32 squares) {
    std::vector<u8> rgb(3 * (size_t)size * size);
    for (u32 y = 0; y < size; y++)
        for (u32 x = 0; x < size; x++) {
            const bool odd = ((x * squares / size) + (y * squares / size)) & 1u;
            const f32 fx = (f32)x / (f32)size - 0.5f, fy = (f32)y / (f32)size - 0.5f;
            const f32 ring = 0.85f + 0.15f * cosf(40.0f * sqrtf(fx * fx + fy * fy));
            const Color c = odd ? a : b;
            u8 *p = &rgb[3 * ((size_t)y * size + x)];
            p[0] = (u8)(255.0f * clampedValue(c.r * ring));
            p[1] = (u8)(255.0f * clampedValue(c.g * ring));
            p[2] = (u8)(255.0f * clampedValue(c.b * ring));
        }
    return bakeTexture(rgb.data(), size, size, IMG_MIPMAP | IMG_WRAP);
}
static Texture bakedNormalMap(u32 size, f32 frequency) {
    std::vector<u8> rgb(3 * (size_t)size * size);
    const f32 k = 6.2831853f * frequency / (f32)size;
    for (u32 y = 0; y < size; y++)
        for (u32 x = 0; x < size; x++) {
            const f32 dx = 0.35f * cosf(k * (f32)x) * sinf(0.5f * k * (f32)y);   // height sin(kx) sin(ky / 2)
            const f32 dy = 0.175f * sinf(k * (f32)x) * cosf(0.5f * k * (f32)y);
            const f32 inv = 1.0f / sqrtf(dx * dx + dy * dy + 1.0f);
            u8 *p = &rgb[3 * ((size_t)y * size + x)];   // r = x, g = y, b = up (rotated by the sampler)
            p[0] = (u8)(255.0f * (0.5f - 0.5f * dx * inv));
            p[1] = (u8)(255.0f * (0.5f - 0.5f * dy * inv));
            p[2] = (u8)(255.0f * (0.5f + 0.5f * inv));
        }
    return bakeTexture(rgb.data(), size, size, IMG_MIPMAP | IMG_WRAP | IMG_NORMAL | IMG_LINEAR);
}
static Texture bakedSky(u32 face) {
    const u32 w = 6 * face;
    std::vector<u8> rgb(3 * (size_t)w * face);
    for (u32 y = 0; y < face; y++)
        for (u32 x = 0; x < w; x++) {
            const bool top = x >= 4 * face && x < 5 * face, bottom = x >= 5 * face;
            const f32 t = top ? 1.0f : bottom ? 0.0f : 1.0f - (f32)y / (f32)(face - 1);
            u8 *p = &rgb[3 * ((size_t)y * w + x)];
            p[0] = (u8)(255.0f * (0.35f + 0.3f * t));
            p[1] = (u8)(255.0f * (0.35f + 0.4f * t));
            p[2] = (u8)(255.0f * (0.4f + 0.55f * t));
        }
    return bakeTexture(rgb.data(), w, face, IMG_CUBEMAP | IMG_LINEAR);
}

// The converter flags each texture's .bmp source is baked with when its .texture file is absent: the cube maps as
// the shipped irradiance maps were (cube map, linear), albedo maps mipmapped and wrapped, normal maps also
// renormalised and linear.
static const u32 kBitmapFlags[TextureCount] = {
    IMG_MIPMAP | IMG_WRAP, IMG_MIPMAP | IMG_WRAP | IMG_NORMAL | IMG_LINEAR,
    IMG_MIPMAP | IMG_WRAP, IMG_MIPMAP | IMG_WRAP | IMG_NORMAL | IMG_LINEAR,
    IMG_CUBEMAP | IMG_LINEAR, IMG_CUBEMAP | IMG_LINEAR, IMG_CUBEMAP | IMG_LINEAR,
    IMG_CUBEMAP | IMG_LINEAR, IMG_CUBEMAP | IMG_LINEAR, IMG_CUBEMAP | IMG_LINEAR};

static bool loadBaked(Texture &t, const std::string &dir, int i) {
    std::string bmp = dir + "/" + kTextureFiles[i];
    bmp.replace(bmp.size() - 8, 8, ".bmp");
    std::vector<u8> rgb;
    u32 w = 0, h = 0;
    if (!loadBitmap(bmp.c_str(), rgb, w, h)) return false;
    t = bakeTexture(rgb.data(), w, h, kBitmapFlags[i]);
    return true;
}

static void loadTextures(const std::string &dir) {
    bool loaded[TextureCount] = {};
    for (int i = 0; i < TextureCount; i++)
        loaded[i] = !dir.empty() && (load(textures[i], (dir + "/" + kTextureFiles[i]).c_str()) ||
                                     loadBaked(textures[i], dir, i));
    int cube = -1;
    for (int i = Cathedral_SkyboxColor; i < TextureCount && cube < 0; i++)
        if (loaded[i] && (textures[i].flags & IMG_CUBEMAP)) cube = i;
    for (int i = 0; i < TextureCount; i++) {
        if (loaded[i]) continue;
        switch (i) {
            case Floor_Albedo: textures[i] = bakedAlbedo(256, Color{0.75f, 0.7f, 0.6f}, Color{0.3f, 0.28f, 0.25f}, 8); break;
            case Floor_Normal: textures[i] = bakedNormalMap(256, 8.0f); break;
            case Dog_Albedo: textures[i] = bakedAlbedo(128, Color{0.8f, 0.55f, 0.3f}, Color{0.45f, 0.3f, 0.2f}, 4); break;
            case Dog_Normal: textures[i] = bakedNormalMap(128, 4.0f); break;
            default: textures[i] = cube >= 0 ? textures[cube] : bakedSky(64); break;
        }
        fprintf(stderr, "slim_examples: %s not found; stand-in: %s\n", kTextureFiles[i],
                i < Cathedral_SkyboxColor ? "procedural" : cube >= 0 ? kTextureFiles[cube] : "baked gradient");
    }
}

static std::string g_assets;
static bool loadMesh(Mesh &m, const char *name, const char *stand_in) {
    if (!g_assets.empty() && load(m, (g_assets + "/" + name).c_str())) return true;
    fprintf(stderr, "slim_examples: %s not found; stand-in: %s\n", name, stand_in);
    return !g_assets.empty() && load(m, (g_assets + "/" + stand_in).c_str());
}

// ------------------------------------------------------------------------------------------------- apps
// What every example app does per frame and on resize (01_Lights.cpp:55, 119-122): reset the projection for the
// canvas, render into the canvas on the GPU or the CPU, present the canvas to the window.
struct ExampleBase : SlimApp {
    bool use_gpu = USE_GPU_BY_DEFAULT;
    bool antialias = false;
    Canvas canvas;
    void OnWindowResize(u16 width, u16 height) override { canvas.dimensions.update(width, height); }
};

#define EXAMPLE_RENDER(renderer)                                   \
    void OnRender() override {                                     \
        renderer.render(viewport, true, use_gpu);                  \
        drawToWindow(canvas);                                      \
    }

namespace ex01 {   // src/examples/01_Lights.cpp:14-66
struct App : ExampleBase {
    Camera camera{{-25 * DEG_TO_RAD, 0, 0}, {0, 25, -45}}, *cameras{&camera};
    CameraRayProjection projection;
    Viewport viewport{canvas, &camera};
    Light key_light{{1.0f, 1.0f, 0.65f}, {20, 7, -5}, 1.1f * 130.0f};
    Light fill_light{{0.65f, 0.65f, 1.0f}, {-20, 6, -5}, 1.2f * 130.0f};
    Light rim_light{{1.0f, 0.25f, 0.25f}, {2, 5, 10}, 0.9f * 130.0f};
    Light lights[3]{key_light, fill_light, rim_light};
    u8 flags{MATERIAL_HAS_NORMAL_MAP | MATERIAL_HAS_ALBEDO_MAP};
    Material floor_material{0.8f, 0.2f, flags, 2, {Floor_Albedo, Floor_Normal}};
    Material *materials{&floor_material};
    Geometry floor{{{}, {}, {40, 1, 40}}, GeometryType_Quad};
    Geometry *geometries{&floor};
    Scene scene{{1, 1, 3, 1, TextureCount}, geometries, cameras, lights, materials, textures};
    SceneTracer scene_tracer{scene.counts.geometries, scene.mesh_stack_size};
    RayTracingRenderer renderer{scene, scene_tracer, projection, 1, Cathedral_SkyboxColor, Cathedral_SkyboxRadiance,
                                Cathedral_SkyboxIrradiance};
    void OnUpdate(f32) override { projection.reset(camera, canvas.dimensions, canvas.antialias == SSAA); }
    EXAMPLE_RENDER(renderer)
};
}  // namespace ex01

namespace ex02 {   // src/examples/02_Geometry.cpp:15-105
struct App : ExampleBase {
    Camera camera{{-25 * DEG_TO_RAD, 0, 0}, {-4, 15, -17}}, *cameras{&camera};
    CameraRayProjection projection;
    Viewport viewport{canvas, &camera};
    Light lights[6]{{{1.0f, 1.0f, 0.65f}, {15, 10, -5}, 1.1f * 50.0f},  {{0.65f, 0.65f, 1.0f}, {-16, 10, -5}, 1.2f * 50.0f},
                    {{1.0f, 0.25f, 0.25f}, {5, 8, 15}, 0.9f * 150.0f}, {{1.0f, 1.0f, 0.65f}, {3, 6, 2}, 15.0f},
                    {{0.65f, 0.65f, 1.0f}, {-9, 8, -2}, 15.0f},        {{0.75f, 1.0f, 0.75f}, {-3, 6, 14}, 15.0f}};
    enum MaterialID { Floor, Rough, MaterialCount };
    Material materials[MaterialCount]{{0.8f, 0.2f, MATERIAL_HAS_NORMAL_MAP | MATERIAL_HAS_ALBEDO_MAP, 2, {Floor_Albedo, Floor_Normal}},
                                      {0.8f, 0.7f}};
    u8 flags = GEOMETRY_IS_VISIBLE | GEOMETRY_IS_SHADOWING | GEOMETRY_IS_TRANSPARENT;
    Geometry geometries[4]{{{{}, {}, {40, 1, 40}}, GeometryType_Quad, Floor},
                           {{{}, {-9, 8, -2}, {2, 2, 3}}, GeometryType_Box, Rough, 0, flags},
                           {{{}, {-3, 6, 14}, {4, 3, 4}}, GeometryType_Tet, Rough, 0, flags},
                           {{{}, {3, 6, 2}, {4, 3, 3}}, GeometryType_Sphere, Rough, 0, flags}};
    Scene scene{{4, 1, 6, MaterialCount, TextureCount}, geometries, cameras, lights, materials, textures};
    SceneTracer scene_tracer{scene.counts.geometries, scene.mesh_stack_size};
    RayTracingRenderer renderer{scene, scene_tracer, projection, 1, Cathedral_SkyboxColor, Cathedral_SkyboxRadiance,
                                Cathedral_SkyboxIrradiance};
    void OnUpdate(f32 delta_time) override {
        projection.reset(camera, canvas.dimensions, canvas.antialias == SSAA);
        static vec3 axis{OrientationUsingQuaternion{DEG_TO_RAD, DEG_TO_RAD * 2, DEG_TO_RAD * 3}.axis};
        quat rot = quat::AxisAngle(axis, delta_time * 10.0f);
        for (u32 i = 0; i < scene.counts.geometries; i++) {
            Geometry &geo = geometries[i];
            if (geo.type != GeometryType_Quad && geo.type != GeometryType_Mesh) geo.transform.orientation *= rot;
        }
    }
    EXAMPLE_RENDER(renderer)
};
}  // namespace ex02

namespace ex03 {   // src/examples/03_BlinnPhong.cpp:48-110
struct App : ExampleBase {
    Camera camera{{-25 * DEG_TO_RAD, 0, 0}, {-4, 15, -17}}, *cameras{&camera};
    CameraRayProjection projection;
    Viewport viewport{canvas, &camera};
    Light lights[3]{{{1.0f, 1.0f, 0.65f}, {20, 20, -5}, 1.1f * 200.0f},
                    {{0.65f, 0.65f, 1.0f}, {-20, 20, -5}, 1.2f * 200.0f},
                    {{1.0f, 0.25f, 0.25f}, {5, 8, 15}, 0.9f * 200.0f}};
    enum MaterialID { Floor, Lambert, Phong, Blinn, MaterialCount };
    Material materials[MaterialCount]{
        {0.8f, 0.2f, MATERIAL_HAS_NORMAL_MAP | MATERIAL_HAS_ALBEDO_MAP, 2, {Floor_Albedo, Floor_Normal}},
        {1.0f, 0.7f, 0, 0, {}, BRDF_Lambert, 1.0f, 0.25f},
        {1.0f, 0.5f, MATERIAL_HAS_NORMAL_MAP, 2, {0, 1}, BRDF_Phong, 0.75f, 0.25f, 0.0f, 1.0f, {1.0f, 1.0f, 0.4f}},
        {{1.0f, 0.3f, 1.0f}, 0.5f, MATERIAL_HAS_NORMAL_MAP, 2, {0, 1}, BRDF_Blinn, 0.75f, 0.25f, 0.0f, 1.0f, {1.0f, 0.4f, 1.0f}}};
    Geometry geometries[4]{{{{}, {}, {40, 1, 40}}, GeometryType_Quad, Floor},
                           {{{}, {-9, 8, -2}, {2, 2, 3}}, GeometryType_Box, Lambert},
                           {{{}, {-3, 6, 14}, {4, 3, 4}}, GeometryType_Tet, Phong},
                           {{{}, {3, 6, 2}, {4, 3, 3}}, GeometryType_Sphere, Blinn}};
    Scene scene{{4, 1, 3, MaterialCount, TextureCount}, geometries, cameras, lights, materials, textures};
    SceneTracer scene_tracer{scene.counts.geometries, scene.mesh_stack_size};
    RayTracingRenderer renderer{scene, scene_tracer, projection, 1, Cathedral_SkyboxColor, Cathedral_SkyboxRadiance,
                                Cathedral_SkyboxIrradiance};
    void OnUpdate(f32 delta_time) override {
        projection.reset(camera, canvas.dimensions, canvas.antialias == SSAA);
        static vec3 axis{OrientationUsingQuaternion{0.02f, 0.04f, 0.06f}.axis};
        quat rot = quat::AxisAngle(axis, delta_time * 10.0f);
        for (u32 i = 0; i < scene.counts.geometries; i++) {
            Geometry &geo = geometries[i];
            if (geo.type != GeometryType_Quad && geo.type != GeometryType_Mesh)
                geo.transform.orientation = (geo.transform.orientation * rot).normalized();
        }
    }
    EXAMPLE_RENDER(renderer)
};
}  // namespace ex03

namespace ex04 {   // src/examples/04_GlassMirror.cpp:46-110
struct App : ExampleBase {
    Camera camera{{-25 * DEG_TO_RAD, 0, 0}, {-4, 15, -17}}, *cameras{&camera};
    CameraRayProjection projection;
    Viewport viewport{canvas, &camera};
    Light key_light{{1.0f, 1.0f, 0.65f}, {-9, 8, -2}, 100.0f};
    Light *lights{&key_light};
    enum MaterialID { Floor, Mirror, Glass, MaterialCount };
    Material materials[MaterialCount]{
        {0.8f, 0.2f, MATERIAL_HAS_NORMAL_MAP | MATERIAL_HAS_ALBEDO_MAP, 2, {Floor_Albedo, Floor_Normal}},
        {0.0f, 0.02f, MATERIAL_IS_REFLECTIVE | MATERIAL_HAS_NORMAL_MAP, 2, {0, 1}, BRDF_CookTorrance, 0.02f, 0.2f, 0.0f,
         IOR_AIR, F0_Aluminium},
        {0.05f, 0.25f, MATERIAL_IS_REFRACTIVE | MATERIAL_HAS_NORMAL_MAP, 2, {0, 1}, BRDF_CookTorrance, 0.08f, 0.05f, 0.0f,
         IOR_GLASS, F0_Glass_Low}};
    OrientationUsingQuaternion wall_rot{75 * DEG_TO_RAD, 0, 90.0f * DEG_TO_RAD};
    Geometry geometries[5]{{{{}, {}, {40, 1, 40}}, GeometryType_Quad, Floor},
                           {{wall_rot, {-6, 8, 15}, {6, 1, 10}}, GeometryType_Quad, Mirror},
                           {{{}, {-9, 8, -2}, {2, 2, 3}}, GeometryType_Box, Glass, 0, GEOMETRY_IS_VISIBLE},
                           {{{}, {-3, 5, -3}, {4, 3, 4}}, GeometryType_Tet, Glass, 0, GEOMETRY_IS_VISIBLE},
                           {{{}, {3, 7, 2}, {4, 3, 3}}, GeometryType_Sphere, Glass, 0, GEOMETRY_IS_VISIBLE}};
    Scene scene{{5, 1, 1, MaterialCount, TextureCount}, geometries, cameras, lights, materials, textures};
    SceneTracer scene_tracer{scene.counts.geometries, scene.mesh_stack_size};
    RayTracingRenderer renderer{scene, scene_tracer, projection, 5, Cathedral_SkyboxColor, Cathedral_SkyboxRadiance,
                                Cathedral_SkyboxIrradiance};
    void OnUpdate(f32 delta_time) override {
        projection.reset(camera, canvas.dimensions, canvas.antialias == SSAA);
        static vec3 axis{OrientationUsingQuaternion{0.02f, 0.04f, 0.06f}.axis};
        quat rot = quat::AxisAngle(axis, delta_time * 10.0f);
        for (u32 i = 0; i < scene.counts.geometries; i++) {
            Geometry &geo = geometries[i];
            if (geo.type != GeometryType_Quad && geo.type != GeometryType_Mesh)
                geo.transform.orientation = (geo.transform.orientation * rot).normalized();
        }
    }
    EXAMPLE_RENDER(renderer)
};
}  // namespace ex04

namespace ex05 {   // src/examples/05_PBR.cpp:13-86
#define GRID_SPACING 2.5f
#define GRID_SIZE 7
#define GRID_SIZE_F ((float)GRID_SIZE)
#define OBJECT_COUNT (GRID_SIZE * GRID_SIZE)
struct App : ExampleBase {
    Camera camera{{}, {-2.0f, -1.0f, -20.0f}};
    CameraRayProjection projection;
    Viewport viewport{canvas, &camera};
    Light lights[4]{{White, {+10, +10, -10}, 200}, {White, {-10, +10, -10}, 200},
                    {White, {+10, -20, -10}, 200}, {White, {-10, -20, -10}, 200}};
    Material materials[GRID_SIZE][GRID_SIZE];
    Geometry geometries[GRID_SIZE][GRID_SIZE];
    Scene scene{{OBJECT_COUNT, 1, 4, OBJECT_COUNT, TextureCount}, &geometries[0][0], &camera, lights,
                &materials[0][0], textures};
    SceneTracer scene_tracer{scene.counts.geometries, scene.mesh_stack_size};
    RayTracingRenderer renderer{scene, scene_tracer, projection, 1, Cathedral_SkyboxColor, Cathedral_SkyboxRadiance,
                                Cathedral_SkyboxIrradiance};
    App() {
        Color plastic{0.04f};
        for (u8 row = 0; row < GRID_SIZE; row++) {
            for (u8 col = 0; col < GRID_SIZE; col++) {
                Geometry &geo{geometries[row][col]};
                Material &mat{materials[row][col]};
                geo.type = GeometryType_Sphere;
                geo.transform.position.x = ((f32)col - (GRID_SIZE_F / 2.0f)) * GRID_SPACING;
                geo.transform.position.y = ((f32)row - (GRID_SIZE_F / 2.0f)) * GRID_SPACING;
                geo.material_id = GRID_SIZE * row + col;
                mat.brdf = BRDF_CookTorrance;
                mat.metalness = (f32)row / GRID_SIZE_F;
                mat.roughness = (f32)col / GRID_SIZE_F;
                mat.roughness = clampedValue(mat.roughness, 0.05f, 1.0f);
                mat.albedo = Color{0.5f, 0.0f, 0.0f};
                mat.reflectivity = plastic.lerpTo(mat.albedo, mat.metalness);
            }
        }
        // the renderer uploaded the scene when it was constructed, before these materials were set: main
        // uploads them again once the device is chosen (05_PBR.cpp:78, uploadMaterials)
    }
    void OnUpdate(f32) override { projection.reset(camera, canvas.dimensions, canvas.antialias == SSAA); }
    EXAMPLE_RENDER(renderer)
};
}  // namespace ex05

namespace ex06 {   // src/examples/06_AreaLights.cpp:48-110
struct App : ExampleBase {
    Camera camera{{-25 * DEG_TO_RAD, 0, 0}, {-4, 15, -17}}, *cameras{&camera};
    CameraRayProjection projection;
    Viewport viewport{canvas, &camera};
    enum MaterialID { Floor, Rough, Mirror, Emissive1, Emissive2, MaterialCount };
    Material materials[MaterialCount]{
        {0.8f, 0.2f, MATERIAL_HAS_NORMAL_MAP | MATERIAL_HAS_ALBEDO_MAP, 2, {Floor_Albedo, Floor_Normal}},
        {0.8f, 0.2f, MATERIAL_HAS_NORMAL_MAP, 2, {0, Floor_Normal}, BRDF_CookTorrance, 0.25f, 0.25f},
        {0.1f, 0.02f, MATERIAL_IS_REFLECTIVE | MATERIAL_HAS_NORMAL_MAP, 2, {0, 1}, BRDF_CookTorrance, 0.1f, 0.3f, 0.0f,
         IOR_AIR, F0_Aluminium},
        {0.0f, 0.0f, MATERIAL_IS_EMISSIVE, 0, {}, BRDF_CookTorrance, 1.0f, 1.0f, {0.9f, 0.5f, 0.5f}},
        {0.0f, 0.0f, MATERIAL_IS_EMISSIVE, 0, {}, BRDF_CookTorrance, 1.0f, 1.0f, {0.5f, 0.5f, 0.9f}}};
    Geometry geometries[5]{
        {{{}, {}, {40, 1, 40}}, GeometryType_Quad, Floor},
        {{{0.0f, 0.0f, -90.0f * DEG_TO_RAD}, {-12, 10, 5}, {4, 1, 8}}, GeometryType_Quad, Emissive1},
        {{{0.0f, 0.0f, +90.0f * DEG_TO_RAD}, {+15, 5, 5}, {4, 1, 8}}, GeometryType_Quad, Emissive2},
        {{{0.0f, 0.0f, -15.0f * DEG_TO_RAD}, {-7.5f, 4.5f, 2}, {2, 4, 3}}, GeometryType_Sphere, Rough},
        {{{0, 0, 90 * DEG_TO_RAD}, {6, 6, 1}, {5, 1, 5}}, GeometryType_Quad, Mirror}};
    Scene scene{{5, 1, 0, MaterialCount, TextureCount}, geometries, cameras, nullptr, materials, textures};
    SceneTracer scene_tracer{scene.counts.geometries, scene.mesh_stack_size};
    RayTracingRenderer renderer{scene, scene_tracer, projection, 3, Cathedral_SkyboxColor, Cathedral_SkyboxRadiance,
                                Cathedral_SkyboxIrradiance};
    void OnUpdate(f32) override { projection.reset(camera, canvas.dimensions, canvas.antialias == SSAA); }
    EXAMPLE_RENDER(renderer)
};
}  // namespace ex06

namespace ex07 {   // src/examples/07_Meshes.cpp:34-140
enum MesheID { Monkey, Dragon, Dog, MeshCount };
static Mesh meshes[MeshCount];
struct App : ExampleBase {
    Camera camera{{-25 * DEG_TO_RAD, 0, 0}, {0, 7, -11}}, *cameras{&camera};
    CameraRayProjection projection;
    Viewport viewport{canvas, &camera};
    Light lights[3]{{{1.0f, 0.25f, 0.25f}, {6, 5, 2}, 0.9f * 40.0f},
                    {{0.25f, 1.0f, 0.25f}, {-3.3f, 0.6f, -3.0f}, 20.0f},
                    {{0.25f, 0.25f, 1.0f}, {-0.6f, 1.75f, -3.15f}, 20.0f}};
    enum MaterialID { Floor, DogMaterial, Rough, Phong, Blinn, Mirror, Glass, MaterialCount };
    u8 flags{MATERIAL_HAS_NORMAL_MAP | MATERIAL_HAS_ALBEDO_MAP};
    Material materials[MaterialCount]{
        {0.8f, 0.2f, flags, 2, {Floor_Albedo, Floor_Normal}},
        {1.0f, 0.2f, flags, 2, {Dog_Albedo, Dog_Normal}},
        {0.8f, 0.7f},
        {1.0f, 0.5f, 0, 0, {}, BRDF_Phong, 1.0f, 1.0, 0.0f, IOR_AIR, {1.0f, 1.0f, 0.4f}},
        {{1.0f, 0.3f, 1.0f}, 0.5f, 0, 0, {}, BRDF_Blinn, 1.0f, 1.0, 0.0f, IOR_AIR, {1.0f, 0.4f, 1.0f}},
        {0.0f, 0.02f, MATERIAL_IS_REFLECTIVE, 0, {}, BRDF_CookTorrance, 1.0f, 1.0, 0.0f, IOR_AIR, F0_Aluminium},
        {0.05f, 0.25f, MATERIAL_IS_REFRACTIVE, 0, {}, BRDF_CookTorrance, 1.0f, 1.0f, 0.0f, IOR_GLASS, F0_Glass_Low}};
    OrientationUsingQuaternion rot{0, -45 * DEG_TO_RAD, 0};
    Geometry geometries[5]{{{{}, {}, {40, 1, 40}}, GeometryType_Quad, Floor},
                           {{{}, {4, 1, -4}}, GeometryType_Sphere, Mirror},
                           {{rot, {6, 4.5, 2}, 0.4f}, GeometryType_Mesh, Glass, Monkey},
                           {{{}, {-2, 2, -3}}, GeometryType_Mesh, Glass, Dragon},
                           {{rot, {4, 2.1f, 3}, 0.8f}, GeometryType_Mesh, DogMaterial, Dog}};
    Geometry &dragon{geometries[3]};
    Scene scene{{5, 1, 3, MaterialCount, TextureCount, MeshCount}, geometries, cameras, lights, materials, textures,
                meshes};
    SceneTracer scene_tracer{scene.counts.geometries, scene.mesh_stack_size};
    RayTracingRenderer renderer{scene, scene_tracer, projection, 5, Cathedral_SkyboxColor, Cathedral_SkyboxRadiance,
                                Cathedral_SkyboxIrradiance};
    void OnUpdate(f32 delta_time) override {
        projection.reset(camera, canvas.dimensions, canvas.antialias == SSAA);
        quat r = quat::RotationAroundY(delta_time * 0.25f);
        for (u32 i = 1; i < scene.counts.geometries; i++) {
            Geometry &geo{geometries[i]};
            if (geo.material_id == Glass)
                geo.flags &= ~GEOMETRY_IS_SHADOWING;
            else
                geo.flags |= GEOMETRY_IS_SHADOWING;
            if (&geo != &dragon) {
                r.amount = -r.amount;
                geo.transform.orientation *= r;
            }
        }
    }
    EXAMPLE_RENDER(renderer)
};
}  // namespace ex07

// ------------------------------------------------------------------------------------------------- main
template <class A>
static int run(const HeadlessOptions &o, bool gpu, bool ssaa) {
    A app;
    app.use_gpu = gpu;
    app.canvas.antialias = ssaa ? SSAA : NoAA;
    if constexpr (std::is_same<A, ex05::App>::value) {
        if (gpu) uploadMaterials(app.scene);
    }
    return slimtracin_app::runHeadless(app, o);
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <01..07> [--frames N] [--size WxH] [--cpu|--gpu] [--ssaa] [--assets DIR] "
                        "[--ppm PREFIX] [--record FRAME PREFIX]\n", argv[0]);
        return 2;
    }
    const std::string which = argv[1];
    HeadlessOptions o;
    bool gpu = true, ssaa = false;
    for (int i = 2; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--frames" && i + 1 < argc) o.frames = (u32)atoi(argv[++i]);
        else if (a == "--size" && i + 1 < argc) {
            unsigned w = 0, h = 0;
            if (sscanf(argv[++i], "%ux%u", &w, &h) != 2 || !w || !h || w > 65535 || h > 65535) return 2;
            o.width = (u16)w;
            o.height = (u16)h;
        } else if (a == "--cpu") gpu = false;
        else if (a == "--gpu") gpu = true;
        else if (a == "--ssaa") ssaa = true;
        else if (a == "--assets" && i + 1 < argc) g_assets = argv[++i];
        else if (a == "--ppm" && i + 1 < argc) o.ppm_prefix = argv[++i];
        else if (a == "--record" && i + 2 < argc) {
            slimtracin_app::recorder().target = (uint32_t)atoi(argv[++i]);
            slimtracin_app::recorder().prefix = argv[++i];
        } else {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    try {
        loadTextures(g_assets);
        if (which == "07") {
            if (!loadMesh(ex07::meshes[ex07::Monkey], "monkey.mesh", "monkey.mesh") ||
                !loadMesh(ex07::meshes[ex07::Dragon], "dragon.mesh", "monkey.mesh") ||
                !loadMesh(ex07::meshes[ex07::Dog], "dog.mesh", "cube.mesh")) {
                fprintf(stderr, "slim_examples: 07 needs monkey.mesh and cube.mesh in --assets\n");
                return 2;
            }
        }
        int rc;
        if (which == "01") rc = run<ex01::App>(o, gpu, ssaa);
        else if (which == "02") rc = run<ex02::App>(o, gpu, ssaa);
        else if (which == "03") rc = run<ex03::App>(o, gpu, ssaa);
        else if (which == "04") rc = run<ex04::App>(o, gpu, ssaa);
        else if (which == "05") rc = run<ex05::App>(o, gpu, ssaa);
        else if (which == "06") rc = run<ex06::App>(o, gpu, ssaa);
        else if (which == "07") rc = run<ex07::App>(o, gpu, ssaa);
        else {
            fprintf(stderr, "no example %s\n", which.c_str());
            return 2;
        }
        if (rc == 0 && slimtracin_app::recorder().target != ~0u && !slimtracin_app::recorder().written) {
            fprintf(stderr, "slim_examples: frame %u was not rendered\n", slimtracin_app::recorder().target);
            return 3;
        }
        return rc;
    } catch (const std::exception &e) {
        fprintf(stderr, "slim_examples: %s\n", e.what());
        return 1;
    }
}
