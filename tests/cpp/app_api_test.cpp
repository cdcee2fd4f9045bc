// Drives the SlimTracin application-layer API (include/slimtracin/slimtracin.hpp) the way the reference's
// example apps do (src/examples/07_Meshes.cpp:34-140): geometries, materials, lights, meshes built with
// BVHBuilder, a Camera, CameraRayProjection::reset, a Canvas + Viewport and RayTracingRenderer::render.
//
//   app_api_test spec <scene.spec> <out.bin> <cpu|gpu>   render a scene spec through the API, write the canvas
//   app_api_test texels8|texels16 <scene.spec> <out.bin> <cpu|gpu>  as spec, with every texel-quad texture turned
//       into RGBA8 / RGBA16F texel storage by toTexelStorage
//   app_api_test demo <monkey.mesh> <out.bin> <cpu|gpu>  a hand-built scene with a loaded .mesh asset
//   app_api_test overlay<LW> <scene.spec> <out.bin> <cpu|gpu>  spec, then the app's wireframe overlays with line
//       width LW (drawSceneBVH of src/examples/03_BlinnPhong.cpp:194-209, drawBVH + drawMesh per mesh geometry,
//       a selection box and two of its sides); writes the canvas before and after them
//
// Output: raw float32 canvas pixels (RGBA per sample) followed by the depths.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "slimtracin/slimtracin.hpp"
#include "specfile.h"

static int write_canvas(const char *path, const std::vector<Pixel> &px, const std::vector<f32> &depths) {
    FILE *f = fopen(path, "wb");
    if (!f) return 1;
    fwrite(px.data(), sizeof(Pixel), px.size(), f);
    fwrite(depths.data(), sizeof(f32), depths.size(), f);
    fclose(f);
    return 0;
}

static int run_spec(const char *spec_path, const char *out, bool gpu, int overlay_lw = -1, u32 texel_storage = 0) {
    SpecFile sf;
    if (spec_open(&sf, spec_path)) return 2;
    const int32_t *fr = (const int32_t *)spec_get(&sf, "frame", nullptr);
    const float *cam = (const float *)spec_get(&sf, "camera", nullptr);
    const u32 G = fr[8], M = fr[9], L = fr[10], NM = fr[11], NT = fr[12];
    const float *gf = (const float *)spec_get(&sf, "geo_f", nullptr);
    const uint32_t *gu = (const uint32_t *)spec_get(&sf, "geo_u", nullptr);
    std::vector<Geometry> geometries(G);
    for (u32 i = 0; i < G; i++) {
        const float *f = gf + 9 * i;
        geometries[i] = Geometry{{{f[0], f[1], f[2]}, {f[3], f[4], f[5]}, {f[6], f[7], f[8]}},
                                 (GeometryType)gu[4 * i], gu[4 * i + 1], gu[4 * i + 2], (u8)gu[4 * i + 3]};
    }
    const float *mf = (const float *)spec_get(&sf, "mat_f", nullptr);
    const uint32_t *mu = (const uint32_t *)spec_get(&sf, "mat_u", nullptr);
    std::vector<Material> materials(M);
    for (u32 i = 0; i < M; i++) {
        const float *f = mf + 16 * i;
        const uint32_t *u = mu + 5 * i;
        materials[i] = Material{{f[0], f[1], f[2]}, f[3], (u8)u[0], (u8)u[1], {(u8)u[2], (u8)u[3]}, (BRDFType)u[4],
                                f[4], {f[5], f[6]}, {f[7], f[8], f[9]}, f[10], {f[11], f[12], f[13]}, f[14]};
    }
    std::vector<Light> lights(L);
    const float *lf = (const float *)spec_get(&sf, "light_f", nullptr);
    const uint32_t *lu = (const uint32_t *)spec_get(&sf, "light_u", nullptr);
    for (u32 i = 0; i < L; i++)
        lights[i] = Light{{lf[7 * i], lf[7 * i + 1], lf[7 * i + 2]}, {lf[7 * i + 3], lf[7 * i + 4], lf[7 * i + 5]},
                          lf[7 * i + 6], (u8)lu[i]};
    std::vector<Mesh> meshes(NM);
    for (u32 k = 0; k < NM; k++) {
        char nm[32];
        snprintf(nm, sizeof nm, "mesh%u_info", k);
        const uint32_t *info = (const uint32_t *)spec_get(&sf, nm, nullptr);
        Mesh &m = meshes[k];
        m.vertex_count = info[0]; m.triangle_count = info[1]; m.normals_count = info[2]; m.uvs_count = info[3];
        snprintf(nm, sizeof nm, "mesh%u_pos", k);  m.vertex_positions = (vec3 *)spec_get(&sf, nm, nullptr);
        snprintf(nm, sizeof nm, "mesh%u_pidx", k); m.vertex_position_indices = (TriangleVertexIndices *)spec_get(&sf, nm, nullptr);
        snprintf(nm, sizeof nm, "mesh%u_nrm", k);  m.vertex_normals = (vec3 *)spec_get(&sf, nm, nullptr);
        snprintf(nm, sizeof nm, "mesh%u_nidx", k); m.vertex_normal_indices = (TriangleVertexIndices *)spec_get(&sf, nm, nullptr);
        snprintf(nm, sizeof nm, "mesh%u_uv", k);   m.vertex_uvs = (vec2 *)spec_get(&sf, nm, nullptr);
        snprintf(nm, sizeof nm, "mesh%u_uidx", k); m.vertex_uvs_indices = (TriangleVertexIndices *)spec_get(&sf, nm, nullptr);
        uint32_t edge_words = 0;
        snprintf(nm, sizeof nm, "mesh%u_edges", k);
        m.edge_vertex_indices = (EdgeVertexIndices *)spec_get(&sf, nm, &edge_words);
        m.edge_count = m.edge_vertex_indices ? edge_words / 2 : 0;
        BVHBuilder{info[4]}.buildMesh(m);
    }
    std::vector<Texture> textures(NT);
    std::vector<std::vector<TextureMip>> mips(NT);
    for (u32 k = 0; k < NT; k++) {
        char nm[32];
        snprintf(nm, sizeof nm, "tex%u_info", k);
        const uint32_t *info = (const uint32_t *)spec_get(&sf, nm, nullptr);
        snprintf(nm, sizeof nm, "tex%u_dims", k);
        const uint32_t *dims = (const uint32_t *)spec_get(&sf, nm, nullptr);
        mips[k].resize(info[3]);
        for (u32 m = 0; m < info[3]; m++) {
            snprintf(nm, sizeof nm, "tex%u_mip%u", k, m);
            mips[k][m] = TextureMip{dims[2 * m], dims[2 * m + 1], (TexelQuad *)spec_get(&sf, nm, nullptr)};
        }
        textures[k] = Texture{info[0], info[1], info[3], info[2], mips[k].data()};
        if (texel_storage && !(textures[k].flags & SLIMRT_TEXELS_MASK)) {
            u64 mismatches = 0;
            textures[k] = toTexelStorage(textures[k], texel_storage, &mismatches);
            fprintf(stderr, "texture %u in texel storage: %llu mismatching quads\n", k, (unsigned long long)mismatches);
        }
    }

    Camera camera{{cam[0], cam[1], cam[2]}, {cam[3], cam[4], cam[5]}, cam[6]};
    const bool ssaa = fr[2] != 0;
    const size_t samples = (size_t)fr[0] * fr[1] * (ssaa ? 4 : 1);
    std::vector<Pixel> pixels(samples);
    std::vector<f32> depths(samples);
    Canvas canvas{(u16)fr[0], (u16)fr[1], ssaa ? SSAA : NoAA, pixels.data(), depths.data()};
    Viewport viewport{canvas, &camera};
    Scene scene{{G, 1, L, M, NT, NM}, geometries.data(), &camera, lights.data(), materials.data(),
                NT ? textures.data() : nullptr, NM ? meshes.data() : nullptr};
    SceneTracer scene_tracer{scene.counts.geometries, scene.mesh_stack_size};
    CameraRayProjection projection;
    RayTracingRenderer renderer{scene, scene_tracer, projection, (u8)fr[3], (char)fr[5], (char)fr[6], (char)fr[7],
                                (RenderMode)fr[4]};
    projection.reset(camera, canvas.dimensions, canvas.antialias == SSAA);
    renderer.render(viewport, true, gpu);
    if (overlay_lw < 0) return write_canvas(out, pixels, depths);
    const std::vector<Pixel> before = pixels;
    const std::vector<f32> depths_before = depths;
    const u8 lw = (u8)overlay_lw;
    Box box;
    for (u32 i = 0; i < G; i++) {
        const Geometry &geo = geometries[i];
        const f32 m = geo.type == GeometryType_Tet ? 0.577350259f : 1.0f;   // TET_MAX
        box.vertices = AABB{vec3{-m}, vec3{m}};
        if (geo.type == GeometryType_Quad) {
            box.vertices.min.y = -EPS;
            box.vertices.max.y = EPS;
        }
        drawBox(box, geo.transform, viewport, BrightYellow, 0.5f, lw);
    }
    drawBVH(scene.bvh, {}, viewport, 1, (u16)scene.bvh.height, 0.25f, lw);
    for (u32 i = 0; i < G; i++) {
        const Geometry &geo = geometries[i];
        if (geo.type != GeometryType_Mesh) continue;
        drawBVH(meshes[geo.id].bvh, geo.transform, viewport, 0, 3, 0.25f, lw);
        drawMesh(meshes[geo.id], geo.transform, false, viewport, White, 0.5f, lw);
    }
    if (G > 1) {
        drawBox(Box{}, geometries[1].transform, viewport, Yellow, 0.5f, 0);
        drawBox(Box{}, geometries[1].transform, viewport, Red, 0.5f, 1, BoxSide_Left | BoxSide_Top);
    }
    FILE *f = fopen(out, "wb");
    if (!f) return 1;
    fwrite(before.data(), sizeof(Pixel), before.size(), f);
    fwrite(depths_before.data(), sizeof(f32), depths_before.size(), f);
    fwrite(pixels.data(), sizeof(Pixel), pixels.size(), f);
    fwrite(depths.data(), sizeof(f32), depths.size(), f);
    fclose(f);
    return 0;
}

static int run_demo(const char *mesh_path, const char *out, bool gpu) {
    // A small scene in the style of src/examples/07_Meshes.cpp: a floor, a mirror sphere, a glass monkey.
    Camera camera{{-25 * DEG_TO_RAD, 0, 0}, {0, 7, -11}};
    Light lights[2] = {Light{{1.0f, 0.25f, 0.25f}, {6, 5, 2}, 0.9f * 40.0f},
                       Light{{0.25f, 1.0f, 0.25f}, {-3.3f, 0.6f, -3.0f}, 20.0f}};
    Material materials[3] = {Material{0.8f, 0.2f},
                             Material{0.0f, 0.02f, MATERIAL_IS_REFLECTIVE, 0, {}, BRDF_CookTorrance, 1.0f, 1.0f,
                                      0.0f, IOR_AIR, F0_Aluminium},
                             Material{0.05f, 0.25f, MATERIAL_IS_REFRACTIVE, 0, {}, BRDF_CookTorrance, 1.0f, 1.0f,
                                      0.0f, IOR_GLASS, F0_Glass_Low}};
    Mesh meshes[1];
    if (!load(meshes[0], mesh_path)) return 3;
    OrientationUsingQuaternion rot{0, -45 * DEG_TO_RAD, 0};
    Geometry geometries[3] = {Geometry{{{}, {}, {40, 1, 40}}, GeometryType_Quad, 0},
                              Geometry{{{}, {4, 1, -4}}, GeometryType_Sphere, 1},
                              Geometry{{rot, {0, 2.5f, 0}, 1.2f}, GeometryType_Mesh, 2, 0}};
    geometries[2].flags &= ~GEOMETRY_IS_SHADOWING;
    std::vector<Pixel> pixels(320 * 180);
    std::vector<f32> depths(320 * 180);
    Canvas canvas{320, 180, NoAA, pixels.data(), depths.data()};
    Viewport viewport{canvas, &camera};
    Scene scene{{3, 1, 2, 3, 0, 1}, geometries, &camera, lights, materials, nullptr, meshes};
    SceneTracer scene_tracer{scene.counts.geometries, scene.mesh_stack_size};
    CameraRayProjection projection;
    RayTracingRenderer renderer{scene, scene_tracer, projection, 4};
    projection.reset(camera, canvas.dimensions, false);
    renderer.render(viewport, true, gpu);
    // animate as the demo does, then render again (per-frame scene update path)
    quat spin = quat::RotationAroundY(0.25f);
    geometries[2].transform.orientation *= spin;
    renderer.render(viewport, true, gpu);
    return write_canvas(out, pixels, depths);
}

// Asset files through the C++ API: load and save again (mesh, texture, bvh by the input's extension); for a
// texture also the converter's re-bake of its own faces (<out>.baked), as the Python tests do with assets.py.
static int run_assets(const char *in, const char *out) {
    const std::string path(in);
    auto ends = [&](const char *e) { return path.size() > strlen(e) && path.compare(path.size() - strlen(e), strlen(e), e) == 0; };
    if (ends(".mesh")) {
        Mesh m;
        return load(m, in) && save(m, out) ? 0 : 3;
    }
    if (ends(".bvh")) {
        BVH b;
        return load(b, in) && save(b, out) ? 0 : 3;
    }
    Texture t;
    if (!load(t, in) || !save(t, out)) return 3;
    if (!(t.flags & (1u << 8))) return 0;
    // the cube's input image: texel (x, y) of each mip is quad (x + 1, y + 1).TL, stored as byte - 1
    const u32 n = t.height, W = 6 * n;
    std::vector<u8> rgb((size_t)W * n * 3);
    u32 x0 = 0;
    for (u32 k = 0; k < 3; k++) {
        const TextureMip &m = t.mips[k];
        for (u32 y = 0; y < m.height; y++)
            for (u32 x = 0; x < m.width; x++) {
                const TexelQuad &q = m.texel_quads[(size_t)(y + 1) * (m.width + 1) + x + 1];
                const u8 c[3] = {q.R[0], q.G[0], q.B[0]};
                for (int ch = 0; ch < 3; ch++) rgb[((size_t)y * W + x0 + x) * 3 + ch] = c[ch] ? (u8)(c[ch] + 1) : 0;
            }
        x0 += m.width;
    }
    Texture baked = bakeTexture(rgb.data(), W, n, t.flags);
    return save(baked, (std::string(out) + ".baked").c_str()) ? 0 : 3;
}

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s spec|demo|assets <input> <out> cpu|gpu\n", argv[0]);
        return 2;
    }
    const bool gpu = !strcmp(argv[4], "gpu");
    try {
        if (!strcmp(argv[1], "spec")) return run_spec(argv[2], argv[3], gpu);
        if (!strcmp(argv[1], "texels8")) return run_spec(argv[2], argv[3], gpu, -1, SLIMRT_TEXELS_RGBA8);
        if (!strcmp(argv[1], "texels16")) return run_spec(argv[2], argv[3], gpu, -1, SLIMRT_TEXELS_RGBA16F);
        if (!strncmp(argv[1], "overlay", 7)) return run_spec(argv[2], argv[3], gpu, atoi(argv[1] + 7));
        if (!strcmp(argv[1], "demo")) return run_demo(argv[2], argv[3], gpu);
        if (!strcmp(argv[1], "assets")) return run_assets(argv[2], argv[3]);
    } catch (const std::exception &e) {
        fprintf(stderr, "app_api_test: %s\n", e.what());
        return 1;
    }
    return 2;
}
