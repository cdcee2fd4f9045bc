// Driver that runs the reference SlimTracin CPU renderer, compiled from its own headers where they lie,
// on a scene-spec file, and writes what it computed.  Test infrastructure: it validates the CPU oracle
// (oracle/slim_oracle.c) and is timed as the reference CPU baseline.  Nothing here is product code.
//
// The reference is used unmodified.  Two build-time facts (see oracle/Makefile):
//  * MAX_TRIANGLES_PER_MESH_BVH_NODE (src/slim/core/base.h:159) is redefined to a variable after base.h,
//    so BVHBuilder::buildMesh (src/slim/scene/bvh_builder.h:305) builds with the spec's leaf size.
//  * The platform layer (os::, timers::) is never called: every allocator is handed memory up front.
//
// Scene setup mirrors what Scene::Scene does for in-memory arrays (src/slim/scene/scene.h:58-163),
// without the file loading.  Per pixel the closed-form primary ray of d_render
// (src/slim/renderer/renderer_GPU.h:36-43) is used and renderPixel (src/slim/renderer/ray_tracer.h:144)
// produces the pre-canvas colour and depth.  The tracer lives in zero-initialised storage, as it does
// in the reference apps (`new ExampleApp()` value-initialises the app, src/examples/01_Lights.cpp:138).
#include <stdlib.h>
#include <new>
#include <stdio.h>
#include <string.h>
#include <time.h>
#include <thread>
#include <vector>
#include <string>
#include <algorithm>

#include "slim/core/base.h"
static unsigned short g_mesh_leaf_size = 2;
#undef MAX_TRIANGLES_PER_MESH_BVH_NODE
#define MAX_TRIANGLES_PER_MESH_BVH_NODE g_mesh_leaf_size

#include "slim/scene/selection.h"
#include "slim/renderer/renderer.h"
#include "slim/draw/bvh.h"
#include "slim/draw/mesh.h"

#include "specfile.h"

struct OutFile {
    std::vector<std::string> names;
    std::vector<uint32_t> dtypes;
    std::vector<std::vector<uint8_t>> blobs;
    std::vector<uint32_t> counts;
    template <typename T> void add(const char *name, uint32_t dtype, const T *p, size_t n) {
        names.push_back(name);
        dtypes.push_back(dtype);
        counts.push_back((uint32_t)n);
        blobs.emplace_back((const uint8_t *)p, (const uint8_t *)p + n * sizeof(T));
    }
    bool write(const char *path) {
        FILE *f = fopen(path, "wb");
        if (!f) return false;
        fwrite("SLMSPEC1", 1, 8, f);
        uint32_t hdr[2] = {(uint32_t)names.size(), 0};
        fwrite(hdr, 4, 2, f);
        for (size_t i = 0; i < names.size(); i++) {
            char nm[24] = {0};
            strncpy(nm, names[i].c_str(), 23);
            fwrite(nm, 1, 24, f);
            fwrite(&dtypes[i], 4, 1, f);
            fwrite(&counts[i], 4, 1, f);
            fwrite(blobs[i].data(), 1, blobs[i].size(), f);
            size_t pad = (8 - blobs[i].size() % 8) % 8;
            static const char zeros[8] = {0};
            fwrite(zeros, 1, pad, f);
        }
        fclose(f);
        return true;
    }
};

struct Harness {   // value-initialised (zeroed) like the reference's app objects
    SceneTracer tracer{(u32 *)nullptr, (u32 *)nullptr};
    SurfaceShader surface;
    Ray ray;
    RayHit hit;
};

static double now_ms() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static memory::MonotonicAllocator arena(u64 bytes) {
    memory::MonotonicAllocator a;
    a.address = (u8 *)calloc(1, bytes);
    a.capacity = bytes;
    a.occupied = 0;
    return a;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s spec out [--threads N] [--rows y0 y1] [--reps R] [--no-pixels] [--overlay LW]\n",
                argv[0]);
        return 2;
    }
    int threads = 1, reps = 1, y0 = 0, y1 = -1, overlay_lw = -1;
    bool pixels = true;
    for (int i = 3; i < argc; i++) {
        if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--reps") && i + 1 < argc) reps = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--rows") && i + 2 < argc) { y0 = atoi(argv[++i]); y1 = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "--no-pixels")) pixels = false;
        else if (!strcmp(argv[i], "--overlay") && i + 1 < argc) overlay_lw = atoi(argv[++i]);
    }
    SpecFile sf;
    if (spec_open(&sf, argv[1])) { fprintf(stderr, "cannot read spec %s\n", argv[1]); return 1; }
    const int32_t *frame = (const int32_t *)spec_get(&sf, "frame", nullptr);
    const float *cam = (const float *)spec_get(&sf, "camera", nullptr);
    const int W = frame[0], H = frame[1], ssaa = frame[2];
    const u32 G = (u32)frame[8], M = (u32)frame[9], L = (u32)frame[10], NM = (u32)frame[11], NT = (u32)frame[12];
    if (y1 < 0) y1 = H * (ssaa ? 2 : 1);

    // --- geometries, materials, lights --------------------------------------------------------
    const float *geo_f = (const float *)spec_get(&sf, "geo_f", nullptr);
    const uint32_t *geo_u = (const uint32_t *)spec_get(&sf, "geo_u", nullptr);
    std::vector<Geometry> geometries(G);
    for (u32 i = 0; i < G; i++) {
        Geometry &g = geometries[i];
        const float *f = geo_f + 9 * i;
        g.transform.orientation = OrientationUsingQuaternion{f[0], f[1], f[2]};
        g.transform.position = vec3{f[3], f[4], f[5]};
        g.transform.scale = vec3{f[6], f[7], f[8]};
        g.type = (GeometryType)geo_u[4 * i];
        g.material_id = geo_u[4 * i + 1];
        g.id = geo_u[4 * i + 2];
        g.flags = (u8)geo_u[4 * i + 3];
    }
    const float *mat_f = (const float *)spec_get(&sf, "mat_f", nullptr);
    const uint32_t *mat_u = (const uint32_t *)spec_get(&sf, "mat_u", nullptr);
    std::vector<Material> materials(M);
    for (u32 i = 0; i < M; i++) {
        Material &m = materials[i];
        const float *f = mat_f + 16 * i;
        const uint32_t *u = mat_u + 5 * i;
        m.albedo = Color{f[0], f[1], f[2]};
        m.roughness = f[3];
        m.normal_magnitude = f[4];
        m.uv_repeat.u = f[5];
        m.uv_repeat.v = f[6];
        m.emission = Color{f[7], f[8], f[9]};
        m.IOR = f[10];
        m.reflectivity = Color{f[11], f[12], f[13]};
        m.metalness = f[14];
        m.flags = (u8)u[0];
        m.texture_count = (u8)u[1];
        m.texture_ids[0] = (u8)u[2];
        m.texture_ids[1] = (u8)u[3];
        m.brdf = (BRDFType)u[4];
    }
    std::vector<Light> lights(L);
    if (L) {
        const float *lf = (const float *)spec_get(&sf, "light_f", nullptr);
        const uint32_t *lu = (const uint32_t *)spec_get(&sf, "light_u", nullptr);
        for (u32 i = 0; i < L; i++) {
            lights[i].color = Color{lf[7 * i], lf[7 * i + 1], lf[7 * i + 2]};
            lights[i].position_or_direction = vec3{lf[7 * i + 3], lf[7 * i + 4], lf[7 * i + 5]};
            lights[i].intensity = lf[7 * i + 6];
            lights[i].flags = (u8)lu[i];
        }
    }

    // --- meshes: positions/normals/uvs + BVHBuilder::buildMesh ------------------------------
    std::vector<Mesh> meshes(NM);
    u32 max_tris = 0;
    u16 mesh_stack_size = 0;
    std::vector<std::vector<BVHNode>> mesh_nodes(NM);
    std::vector<std::vector<Triangle>> mesh_tris(NM);
    std::vector<std::vector<vec3>> mesh_pos(NM), mesh_nrm(NM);
    std::vector<std::vector<vec2>> mesh_uv(NM);
    std::vector<std::vector<TriangleVertexIndices>> mesh_pidx(NM), mesh_nidx(NM), mesh_uidx(NM);
    std::vector<std::vector<EdgeVertexIndices>> mesh_edges(NM);
    for (u32 k = 0; k < NM; k++) {
        char nm[32];
        snprintf(nm, sizeof nm, "mesh%u_info", k);
        const uint32_t *info = (const uint32_t *)spec_get(&sf, nm, nullptr);
        u32 V = info[0], T = info[1], NN = info[2], NU = info[3];
        Mesh &mesh = meshes[k];
        mesh = Mesh{};
        mesh.vertex_count = V;
        mesh.triangle_count = T;
        mesh.normals_count = NN;
        mesh.uvs_count = NU;
        snprintf(nm, sizeof nm, "mesh%u_pos", k);
        const float *pos = (const float *)spec_get(&sf, nm, nullptr);
        snprintf(nm, sizeof nm, "mesh%u_pidx", k);
        const uint32_t *pidx = (const uint32_t *)spec_get(&sf, nm, nullptr);
        mesh_pos[k].resize(V);
        mesh.aabb.min = INFINITY;
        mesh.aabb.max = -INFINITY;
        for (u32 v = 0; v < V; v++) {
            mesh_pos[k][v] = vec3{pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]};
            mesh.aabb.min = minimum(mesh.aabb.min, mesh_pos[k][v]);
            mesh.aabb.max = maximum(mesh.aabb.max, mesh_pos[k][v]);
        }
        mesh_pidx[k].resize(T);
        for (u32 t = 0; t < T; t++)
            for (int j = 0; j < 3; j++) mesh_pidx[k][t].ids[j] = pidx[3 * t + j];
        mesh.vertex_positions = mesh_pos[k].data();
        mesh.vertex_position_indices = mesh_pidx[k].data();
        if (NN) {
            snprintf(nm, sizeof nm, "mesh%u_nrm", k);
            const float *nrm = (const float *)spec_get(&sf, nm, nullptr);
            snprintf(nm, sizeof nm, "mesh%u_nidx", k);
            const uint32_t *nidx = (const uint32_t *)spec_get(&sf, nm, nullptr);
            mesh_nrm[k].resize(NN);
            for (u32 v = 0; v < NN; v++) mesh_nrm[k][v] = vec3{nrm[3 * v], nrm[3 * v + 1], nrm[3 * v + 2]};
            mesh_nidx[k].resize(T);
            for (u32 t = 0; t < T; t++)
                for (int j = 0; j < 3; j++) mesh_nidx[k][t].ids[j] = nidx[3 * t + j];
            mesh.vertex_normals = mesh_nrm[k].data();
            mesh.vertex_normal_indices = mesh_nidx[k].data();
        }
        if (NU) {
            snprintf(nm, sizeof nm, "mesh%u_uv", k);
            const float *uv = (const float *)spec_get(&sf, nm, nullptr);
            snprintf(nm, sizeof nm, "mesh%u_uidx", k);
            const uint32_t *uidx = (const uint32_t *)spec_get(&sf, nm, nullptr);
            mesh_uv[k].resize(NU);
            for (u32 v = 0; v < NU; v++) mesh_uv[k][v] = vec2{uv[2 * v], uv[2 * v + 1]};
            mesh_uidx[k].resize(T);
            for (u32 t = 0; t < T; t++)
                for (int j = 0; j < 3; j++) mesh_uidx[k][t].ids[j] = uidx[3 * t + j];
            mesh.vertex_uvs = mesh_uv[k].data();
            mesh.vertex_uvs_indices = mesh_uidx[k].data();
        }
        snprintf(nm, sizeof nm, "mesh%u_edges", k);
        uint32_t n_edge_words = 0;
        if (const uint32_t *ev = (const uint32_t *)spec_get(&sf, nm, &n_edge_words)) {   // for the overlays
            mesh_edges[k].resize(n_edge_words / 2);
            for (u32 e = 0; e < n_edge_words / 2; e++) mesh_edges[k][e] = EdgeVertexIndices{ev[2 * e], ev[2 * e + 1]};
            mesh.edge_count = n_edge_words / 2;
            mesh.edge_vertex_indices = mesh_edges[k].data();
        }
        mesh_nodes[k].resize(2 * T);
        mesh_tris[k].resize(T);
        mesh.bvh.nodes = mesh_nodes[k].data();
        mesh.triangles = mesh_tris[k].data();
        g_mesh_leaf_size = (unsigned short)info[4];
        memory::MonotonicAllocator a = arena(BVHBuilder::getSizeInBytes(T) + 4096);
        BVHBuilder builder{T, &a};
        builder.buildMesh(mesh);
        free((void *)(a.address - a.occupied));
        max_tris = Max(max_tris, T);
        mesh_stack_size = Max(mesh_stack_size, (u16)mesh.bvh.height);
    }
    if (NM) mesh_stack_size += 2;

    // --- textures ------------------------------------------------------------------------------
    std::vector<Texture> textures(NT);
    std::vector<std::vector<TextureMip>> tex_mips(NT);
    for (u32 k = 0; k < NT; k++) {
        char nm[32];
        snprintf(nm, sizeof nm, "tex%u_info", k);
        const uint32_t *info = (const uint32_t *)spec_get(&sf, nm, nullptr);
        snprintf(nm, sizeof nm, "tex%u_dims", k);
        const uint32_t *dims = (const uint32_t *)spec_get(&sf, nm, nullptr);
        Texture &t = textures[k];
        t.width = info[0];
        t.height = info[1];
        t.flags.flags = info[2];
        t.mip_count = info[3];
        tex_mips[k].resize(t.mip_count);
        for (u32 m = 0; m < t.mip_count; m++) {
            snprintf(nm, sizeof nm, "tex%u_mip%u", k, m);
            tex_mips[k][m].width = dims[2 * m];
            tex_mips[k][m].height = dims[2 * m + 1];
            tex_mips[k][m].texel_quads = (TexelQuad *)spec_get(&sf, nm, nullptr);
        }
        t.mips = tex_mips[k].data();
    }

    // --- scene (SceneData filled as Scene::Scene does for in-memory arrays) --------------------
    SceneData data{};
    data.counts = SceneCountsData{G, 1, L, M, NT, NM, 0, 0, 0, 0, 0};
    data.flags = 0;
    data.mesh_stack_size = mesh_stack_size;
    data.geometries = geometries.data();
    data.lights = L ? lights.data() : nullptr;
    data.materials = materials.data();
    data.textures = NT ? textures.data() : nullptr;
    data.meshes = NM ? meshes.data() : nullptr;
    std::vector<BVHNode> scene_nodes(2 * G);
    std::vector<u32> leaf_ids(G);
    std::vector<AABB> aabbs(G);
    data.bvh.nodes = scene_nodes.data();
    data.bvh.node_count = G * 2;
    data.bvh.height = (u8)G;
    data.bvh_leaf_geometry_indices = leaf_ids.data();
    data.aabbs = aabbs.data();
    u32 max_leaf = Max(max_tris, G);
    memory::MonotonicAllocator sa = arena(BVHBuilder::getSizeInBytes(G) + sizeof(BVHBuilder) + 4096);
    BVHBuilder scene_builder{G, &sa};
    data.bvh_builder = &scene_builder;
    (void)max_leaf;
    for (u32 i = 0; i < G; i++)
        if (geometries[i].type == GeometryType_Quad && materials[geometries[i].material_id].isEmissive())
            data.flags = SCENE_HAD_EMISSIVE_QUADS;
    Scene &scene = *(Scene *)&data;   // as renderer_GPU.h:42 does with its SceneData
    scene.updateAABBs();
    scene.updateBVH();

    // --- camera + projection ---------------------------------------------------------------------
    Camera camera;
    camera.orientation.setRotation(cam[0], cam[1], cam[2]);
    camera.position = vec3{cam[3], cam[4], cam[5]};
    camera.focal_length = cam[6];
    Dimensions dim;
    dim.update((u16)W, (u16)H);
    CameraRayProjection projection;
    projection.reset(camera, dim, ssaa != 0);

    RayTracerSettings settings{};
    settings.max_depth = (u8)frame[3];
    settings.render_mode = (RenderMode)frame[4];
    settings.skybox_color_texture_id = (char)frame[5];
    settings.skybox_radiance_texture_id = (char)frame[6];
    settings.skybox_irradiance_texture_id = (char)frame[7];
    settings.mip_level_colors[0] = BrightRed;
    settings.mip_level_colors[1] = BrightYellow;
    settings.mip_level_colors[2] = BrightGreen;
    settings.mip_level_colors[3] = BrightMagenta;
    settings.mip_level_colors[4] = BrightCyan;
    settings.mip_level_colors[5] = BrightBlue;
    settings.mip_level_colors[6] = BrightGrey;
    settings.mip_level_colors[7] = Grey;
    settings.mip_level_colors[8] = DarkGrey;

    const int Ws = W * (ssaa ? 2 : 1);
    const int rows = y1 - y0;
    std::vector<float> colors((size_t)rows * Ws * 3), depths((size_t)rows * Ws);
    std::vector<int32_t> hit_geo((size_t)rows * Ws), hit_tri((size_t)rows * Ws);
    std::vector<float> hit_t((size_t)rows * Ws);

    auto render_rows = [&](int ya, int yb, bool record_hits) {
        Harness *h = new Harness();
        std::vector<u32> sstack(G + 2), mstack(mesh_stack_size + 2);
        h->tracer.stack = sstack.data();
        h->tracer.mesh_tracer.stack = mstack.data();
        for (int y = ya; y < yb; y++) {
            for (int x = 0; x < Ws; x++) {
                size_t o = (size_t)(y - y0) * Ws + x;
                Color color;
                f32 depth;
                h->ray.pixel_coords.x = x;
                h->ray.pixel_coords.y = y;
                h->hit.scaling_factor = 1.0f / sqrtf(projection.squared_distance_to_projection_plane +
                    vec2{(i32)x, -(i32)y}.scaleAdd(projection.sample_size, projection.C_start).squaredLength());
                vec3 dir = projection.getRayDirectionAt(x, y);
                if (pixels) {
                    renderPixel(settings, projection, scene, h->tracer, h->surface, h->ray, h->hit, dir, color, depth);
                    colors[3 * o] = color.r;
                    colors[3 * o + 1] = color.g;
                    colors[3 * o + 2] = color.b;
                    depths[o] = depth;
                }
                if (record_hits) {
                    h->ray.reset(projection.camera_position, dir.normalized());
                    Geometry *g = h->tracer.trace(h->ray, h->hit, scene);
                    hit_geo[o] = g ? (int32_t)(g - scene.geometries) : -1;
                    hit_tri[o] = (g && g->type == GeometryType_Mesh) ? (int32_t)h->hit.id : -1;
                    hit_t[o] = g ? h->hit.distance : INFINITY;
                }
            }
        }
        delete h;
    };
    auto run = [&](bool record_hits) {
        if (threads <= 1) { render_rows(y0, y1, record_hits); return; }
        std::vector<std::thread> pool;
        int per = (rows + threads - 1) / threads;
        for (int t = 0; t < threads; t++) {
            int a = y0 + t * per, b = std::min(y1, a + per);
            if (a < b) pool.emplace_back(render_rows, a, b, record_hits);
        }
        for (auto &th : pool) th.join();
    };

    run(true);
    double best = 0;
    if (pixels) {
        best = 1e30;
        for (int r = 0; r < reps; r++) {
            double t0 = now_ms();
            run(false);
            best = std::min(best, now_ms() - t0);
        }
    }

    OutFile out;
    std::vector<uint32_t> content;
    std::vector<float> cv_before, cv_after, cv_depth_before, cv_depth_after;
    if (pixels && y0 == 0 && y1 == H * (ssaa ? 2 : 1)) {
        // The canvas as the app keeps it (Canvas::setPixel with opacity -1, src/slim/renderer/renderer.h:97) and the
        // window content Canvas::drawToWindow makes of it (getPixelContent, src/slim/draw/canvas.h:119-126, 187-190).
        memory::canvas_memory_capacity = 0;   // the canvas gets the driver's buffers (canvas.h:21-39)
        Canvas canvas{(u16)W, (u16)H, ssaa ? SSAA : NoAA};
        std::vector<Pixel> canvas_pixels((size_t)Ws * rows);
        std::vector<f32> canvas_depths((size_t)Ws * rows);
        canvas.pixels = canvas_pixels.data();
        canvas.depths = canvas_depths.data();
        for (int y = 0; y < rows; y++)
            for (int x = 0; x < Ws; x++) {
                const size_t o = (size_t)y * Ws + x;
                canvas.setPixel(x, y, Color{colors[3 * o], colors[3 * o + 1], colors[3 * o + 2]}, -1, depths[o]);
            }
        content.resize((size_t)W * H);
        const int step = ssaa ? 4 : 1;
        for (size_t i = 0; i < content.size(); i++) content[i] = canvas.getPixelContent(canvas_pixels.data() + i * step);
        if (overlay_lw >= 0) {
            // The canvas before and after the app's wireframe overlays: the scene-BVH view of
            // examples/03_BlinnPhong.cpp:194-209 (a box per geometry, then the scene BVH's nodes from depth 1),
            // each mesh geometry's BVH (depths 0-3) and wireframe, and a selection box with two highlighted sides
            // (draw/selection.h:24-33), all with line width LW where the apps use 1.
            for (const Pixel &p : canvas_pixels) { cv_before.push_back(p.color.r); cv_before.push_back(p.color.g);
                                                   cv_before.push_back(p.color.b); cv_before.push_back(p.opacity); }
            cv_depth_before = canvas_depths;
            Viewport viewport{canvas, &camera};
            const u8 lw = (u8)overlay_lw;
            Box box;
            AABB aabb;
            for (u32 i = 0; i < scene.counts.geometries; i++) {
                Geometry &geo{scene.geometries[i]};
                aabb.max = geo.type == GeometryType_Tet ? TET_MAX : 1.0f;
                aabb.min = -aabb.max.x;
                if (geo.type == GeometryType_Quad) {
                    aabb.min.y = -EPS;
                    aabb.max.y = EPS;
                }
                box.vertices = aabb;
                drawBox(box, geo.transform, viewport, BrightYellow, 0.5f, lw);
            }
            drawBVH(scene.bvh, {}, viewport, 1, scene.bvh.height, 0.25f, lw);
            for (u32 i = 0; i < scene.counts.geometries; i++) {
                Geometry &geo{scene.geometries[i]};
                if (geo.type != GeometryType_Mesh) continue;
                drawBVH(scene.meshes[geo.id].bvh, geo.transform, viewport, 0, 3, 0.25f, lw);
                drawMesh(scene.meshes[geo.id], geo.transform, false, viewport, White, 0.5f, lw);
            }
            if (scene.counts.geometries > 1) {
                Box unit;
                drawBox(unit, scene.geometries[1].transform, viewport, Yellow, 0.5f, 0);
                drawBox(unit, scene.geometries[1].transform, viewport, Red, 0.5f, 1, BoxSide_Left | BoxSide_Top);
            }
            for (const Pixel &p : canvas_pixels) { cv_after.push_back(p.color.r); cv_after.push_back(p.color.g);
                                                  cv_after.push_back(p.color.b); cv_after.push_back(p.opacity); }
            cv_depth_after = canvas_depths;
        }
    }
    if (pixels) {
        out.add("color", 0, colors.data(), colors.size());
        out.add("depth", 0, depths.data(), depths.size());
    }
    if (!content.empty()) out.add("content", 1, content.data(), content.size());
    if (!cv_after.empty()) {
        out.add("ov_in_pixels", 0, cv_before.data(), cv_before.size());
        out.add("ov_in_depths", 0, cv_depth_before.data(), cv_depth_before.size());
        out.add("ov_pixels", 0, cv_after.data(), cv_after.size());
        out.add("ov_depths", 0, cv_depth_after.data(), cv_depth_after.size());
    }
    out.add("hit_geo", 2, hit_geo.data(), hit_geo.size());
    out.add("hit_tri", 2, hit_tri.data(), hit_tri.size());
    out.add("hit_t", 0, hit_t.data(), hit_t.size());
    std::vector<float> quats(4 * G), gaabb(6 * G);
    for (u32 i = 0; i < G; i++) {
        const quat &q = geometries[i].transform.orientation;
        quats[4 * i] = q.axis.x; quats[4 * i + 1] = q.axis.y; quats[4 * i + 2] = q.axis.z; quats[4 * i + 3] = q.amount;
        const AABB &b = aabbs[i];
        float v[6] = {b.min.x, b.min.y, b.min.z, b.max.x, b.max.y, b.max.z};
        memcpy(&gaabb[6 * i], v, sizeof v);
    }
    out.add("geo_quat", 0, quats.data(), quats.size());
    out.add("geo_aabb", 0, gaabb.data(), gaabb.size());
    auto dump_bvh = [&](const char *prefix, const BVH &bvh) {
        std::vector<float> f(6 * bvh.node_count);
        std::vector<uint32_t> u(3 * bvh.node_count);
        for (u32 i = 0; i < bvh.node_count; i++) {
            const BVHNode &n = bvh.nodes[i];
            float v[6] = {n.aabb.min.x, n.aabb.min.y, n.aabb.min.z, n.aabb.max.x, n.aabb.max.y, n.aabb.max.z};
            memcpy(&f[6 * i], v, sizeof v);
            u[3 * i] = (uint32_t)n.first_index;
            u[3 * i + 1] = n.leaf_count;
            u[3 * i + 2] = n.depth;
        }
        uint32_t meta[2] = {(uint32_t)bvh.node_count, bvh.height};
        std::string p(prefix);
        out.add((p + "_f").c_str(), 0, f.data(), f.size());
        out.add((p + "_u").c_str(), 1, u.data(), u.size());
        out.add((p + "_meta").c_str(), 1, meta, 2);
    };
    dump_bvh("sbvh", scene.bvh);
    std::vector<uint32_t> leaves(G);
    for (u32 i = 0; i < G; i++) leaves[i] = (uint32_t)leaf_ids[i];
    out.add("sbvh_leaf", 1, leaves.data(), leaves.size());
    for (u32 k = 0; k < NM; k++) {
        char nm[32];
        snprintf(nm, sizeof nm, "mesh%u_bvh", k);
        dump_bvh(nm, meshes[k].bvh);
        snprintf(nm, sizeof nm, "mesh%u_tri", k);
        out.add(nm, 0, (const float *)mesh_tris[k].data(), mesh_tris[k].size() * 32);
        float ab[6] = {meshes[k].aabb.min.x, meshes[k].aabb.min.y, meshes[k].aabb.min.z,
                       meshes[k].aabb.max.x, meshes[k].aabb.max.y, meshes[k].aabb.max.z};
        snprintf(nm, sizeof nm, "mesh%u_aabb", k);
        out.add(nm, 0, ab, 6);
    }
    float proj[25];
    const mat3 &R = projection.inverted_camera_rotation;
    float pv[25] = {R.X.x, R.X.y, R.X.z, R.Y.x, R.Y.y, R.Y.z, R.Z.x, R.Z.y, R.Z.z,
                    projection.start.x, projection.start.y, projection.start.z,
                    projection.right.x, projection.right.y, projection.right.z,
                    projection.down.x, projection.down.y, projection.down.z,
                    projection.camera_position.x, projection.camera_position.y, projection.camera_position.z,
                    projection.C_start.x, projection.C_start.y,
                    projection.squared_distance_to_projection_plane, projection.sample_size};
    memcpy(proj, pv, sizeof proj);
    out.add("proj", 0, proj, 25);
    float timing[3] = {(float)best, (float)reps, (float)threads};
    out.add("timing", 0, timing, 3);
    uint32_t range[2] = {(uint32_t)y0, (uint32_t)y1};
    out.add("rows", 1, range, 2);
    if (!out.write(argv[2])) { fprintf(stderr, "cannot write %s\n", argv[2]); return 1; }
    printf("slimref: %dx%d rows %d..%d threads %d best %.3f ms\n", Ws, H, y0, y1, threads, best);
    spec_close(&sf);
    return 0;
}
