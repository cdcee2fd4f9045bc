// SlimApp and a headless Linux platform loop for the SlimTracin application API (C++17, header-only).
//
// The reference runs its example apps in a Win32 window: WinMain creates the window and loops
// OnWindowRedraw -> OnUpdate(delta_time) + OnRender while the app runs (src/slim/platforms/win32.h:164-258,
// src/slim/app.h:6-33), and OnRender ends with Canvas::drawToWindow (src/slim/draw/canvas.h:119-126).  This
// build has no display, so the loop here is headless: the window is a size and a content buffer, each frame's
// window content can be written to a PPM image, and the update step advances by a fixed delta time so that a
// run is reproducible.  A frame recorder writes what a chosen frame rendered (the scene as it was, camera,
// settings, and the canvas) as a scene-spec container, so the frame can be checked offline.
#pragma once

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "slimtracin.hpp"

// The window the canvas is presented to (Win32's window::width / height / content in the reference).
namespace window {
inline u16 width = 0, height = 0;
inline std::vector<u32> &content() {
    static std::vector<u32> c;
    return c;
}
}  // namespace window

// Canvas::drawToWindow (canvas.h:119-126): the canvas's window content, 0xRRGGBB per window pixel (the mean of
// each 2x2 sample quad under SSAA), through slimrt_canvas_content on the host context.
inline void drawToWindow(const Canvas &canvas) {
    using namespace slimtracin_detail;
    window::content().resize(canvas.dimensions.width_times_height);
    check(slimrt_canvas_content(cpu(), canvas.pixels, canvas.dimensions.width, canvas.dimensions.height,
                                canvas.antialias == SSAA ? 1u : 0u, window::content().data()),
          cpu(), "drawToWindow");
}

// loadBitmap (src/slim/platforms/win32_bitmap.h:13-33, 75-91, 122-160): an uncompressed 24- or 32-bit BMP as RGB
// bytes, rows top to bottom (a file stored bottom-up is flipped, unless `flip` asks to keep its stored order
// reversed as the converter's -f does), blue-green-red(-alpha) components reordered to red-green-blue and alpha
// dropped (*has_alpha tells whether the file had it).  Rows are read at the file's 4-byte-aligned stride; the reference reads them unpadded, which is the
// same for widths whose rows need no padding (every image it ships).  False when the file is not such a BMP.
inline bool loadBitmap(const char *path, std::vector<u8> &rgb, u32 &width, u32 &height, bool flip = false,
                       bool *has_alpha = nullptr) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    unsigned char h[54];
    bool ok = fread(h, 1, 54, f) == 54 && h[0] == 'B' && h[1] == 'M';
    uint32_t offset = 0, compression = 0;
    int32_t w = 0, hh = 0;
    uint16_t bpp = 0;
    if (ok) {
        std::memcpy(&offset, h + 10, 4);
        std::memcpy(&w, h + 18, 4);
        std::memcpy(&hh, h + 22, 4);
        std::memcpy(&bpp, h + 28, 2);
        std::memcpy(&compression, h + 30, 4);
        ok = (bpp == 24 || bpp == 32) && (compression == 0 || (compression == 3 && bpp == 32)) && w > 0 && hh != 0 &&
             w <= 65536 && hh <= 65536 && hh >= -65536;
    }
    const bool bottom_up = hh > 0;
    width = (u32)w;
    height = (u32)(bottom_up ? hh : -hh);
    const size_t comps = bpp / 8, stride = ((size_t)width * comps + 3) & ~(size_t)3;
    std::vector<unsigned char> raw(ok ? stride * height : 0);
    ok = ok && fseek(f, (long)offset, SEEK_SET) == 0 && fread(raw.data(), 1, raw.size(), f) == raw.size();
    fclose(f);
    if (!ok) return false;
    if (has_alpha) *has_alpha = bpp == 32;
    rgb.resize((size_t)width * height * 3);
    const bool reverse = bottom_up != flip;
    for (u32 y = 0; y < height; y++) {
        const unsigned char *src = raw.data() + stride * (reverse ? height - 1 - y : y);
        u8 *dst = rgb.data() + (size_t)y * width * 3;
        for (u32 x = 0; x < width; x++, src += comps, dst += 3) {
            dst[0] = src[2];
            dst[1] = src[1];
            dst[2] = src[0];
        }
    }
    return true;
}

// SlimApp (src/slim/app.h:6-33) without the mouse and HUD hooks the headless loop has no source for.
struct SlimApp {
    bool is_running{true};
    virtual ~SlimApp() = default;
    virtual void OnWindowResize(u16 width, u16 height) {}
    virtual void OnKeyChanged(u8 key, bool pressed) {}
    virtual void OnRender() {}
    virtual void OnUpdate(f32 delta_time) {}
    virtual void OnWindowRedraw(f32 delta_time) {
        OnUpdate(delta_time);
        OnRender();
    }
};

namespace slimtracin_app {

// ------------------------------------------------------------------------------------------- frame recording
// The container format of scene specs (magic "SLMSPEC1", section count, then named typed sections, each
// payload padded to 8 bytes), as the CPU oracle and the reference driver read it.
struct SpecWriter {
    std::vector<unsigned char> body;
    uint32_t count = 0;
    void section(const char *name, uint32_t dtype, const void *data, size_t elems, size_t elem_size) {
        unsigned char head[32] = {};
        const size_t len = std::strlen(name);
        std::memcpy(head, name, len < 23 ? len : 23);   // NUL-padded to 24 bytes
        const uint32_t n = (uint32_t)elems;
        std::memcpy(head + 24, &dtype, 4);
        std::memcpy(head + 28, &n, 4);
        body.insert(body.end(), head, head + 32);
        const unsigned char *p = (const unsigned char *)data;
        body.insert(body.end(), p, p + elems * elem_size);
        body.resize((body.size() + 7) & ~(size_t)7, 0);
        count++;
    }
    void f32s(const char *name, const std::vector<float> &v) { section(name, 0, v.data(), v.size(), 4); }
    void u32s(const char *name, const std::vector<uint32_t> &v) { section(name, 1, v.data(), v.size(), 4); }
    void i32s(const char *name, const std::vector<int32_t> &v) { section(name, 2, v.data(), v.size(), 4); }
    bool write(const std::string &path) const {
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) return false;
        const uint32_t hdr[2] = {count, 0};
        bool ok = fwrite("SLMSPEC1", 1, 8, f) == 8 && fwrite(hdr, 4, 2, f) == 2 &&
                  fwrite(body.data(), 1, body.size(), f) == body.size();
        return fclose(f) == 0 && ok;
    }
};

// The largest leaf of a mesh BVH: building the mesh's triangles with that leaf size gives the same BVH (a node
// is a leaf exactly when it holds at most the builder's leaf size, and no node holds more than the largest leaf
// without being split, whatever leaf size above it the mesh was built with).
inline uint32_t leafSize(const Mesh &m) {
    uint32_t mx = 1;
    for (u32 i = 0; i < m.bvh.node_count; i++) mx = m.bvh.nodes[i].leaf_count > mx ? m.bvh.nodes[i].leaf_count : mx;
    return mx;
}

// The frame as the renderer saw it: the scene's arrays (geometry orientations as quaternions, section "geo_q",
// beside zero Euler angles), the camera, the frame parameters, and the textures' texel quads.
inline SpecWriter sceneSpec(const Scene &s, const Viewport &v, const slimrt_render_params &p) {
    SpecWriter w;
    const SceneCounts &c = s.counts;
    w.i32s("frame", {(int32_t)p.width, (int32_t)p.height, (int32_t)p.ssaa, (int32_t)p.max_depth,
                     (int32_t)p.render_mode, p.skybox_color, p.skybox_radiance, p.skybox_irradiance,
                     (int32_t)c.geometries, (int32_t)c.materials, (int32_t)c.lights, (int32_t)c.meshes,
                     (int32_t)c.textures});
    const Camera &cam = *v.camera;
    w.f32s("camera", {cam.orientation.x, cam.orientation.y, cam.orientation.z, cam.position.x, cam.position.y,
                      cam.position.z, cam.focal_length});
    std::vector<float> gf, gq;
    std::vector<uint32_t> gu;
    for (u32 i = 0; i < c.geometries; i++) {
        const Geometry &g = s.geometries[i];
        const Transform &t = g.transform;
        gf.insert(gf.end(), {0.0f, 0.0f, 0.0f, t.position.x, t.position.y, t.position.z, t.scale.x, t.scale.y, t.scale.z});
        gq.insert(gq.end(), {t.orientation.axis.x, t.orientation.axis.y, t.orientation.axis.z, t.orientation.amount});
        gu.insert(gu.end(), {(uint32_t)g.type, g.material_id, g.id, g.flags});
    }
    w.f32s("geo_f", gf);
    w.u32s("geo_u", gu);
    w.f32s("geo_q", gq);
    std::vector<float> mf;
    std::vector<uint32_t> mu;
    for (u32 i = 0; i < c.materials; i++) {
        const Material &m = s.materials[i];
        mf.insert(mf.end(), {m.albedo.r, m.albedo.g, m.albedo.b, m.roughness, m.normal_magnitude, m.uv_repeat.x,
                             m.uv_repeat.y, m.emission.r, m.emission.g, m.emission.b, m.IOR, m.reflectivity.r,
                             m.reflectivity.g, m.reflectivity.b, m.metalness, 0.0f});
        mu.insert(mu.end(), {m.flags, m.texture_count, m.texture_ids[0], m.texture_ids[1], (uint32_t)m.brdf});
    }
    w.f32s("mat_f", mf);
    w.u32s("mat_u", mu);
    if (c.lights) {
        std::vector<float> lf;
        std::vector<uint32_t> lu;
        for (u32 i = 0; i < c.lights; i++) {
            const Light &l = s.lights[i];
            lf.insert(lf.end(), {l.color.r, l.color.g, l.color.b, l.position_or_direction.x, l.position_or_direction.y,
                                 l.position_or_direction.z, l.intensity});
            lu.push_back(l.flags);
        }
        w.f32s("light_f", lf);
        w.u32s("light_u", lu);
    }
    char nm[32];
    for (u32 k = 0; k < c.meshes; k++) {
        const Mesh &m = s.meshes[k];
        snprintf(nm, sizeof nm, "mesh%u_info", k);
        w.u32s(nm, {m.vertex_count, m.triangle_count, m.normals_count, m.uvs_count, leafSize(m)});
        snprintf(nm, sizeof nm, "mesh%u_pos", k);
        w.section(nm, 0, m.vertex_positions, 3 * (size_t)m.vertex_count, 4);
        snprintf(nm, sizeof nm, "mesh%u_pidx", k);
        w.section(nm, 1, m.vertex_position_indices, 3 * (size_t)m.triangle_count, 4);
        if (m.normals_count) {
            snprintf(nm, sizeof nm, "mesh%u_nrm", k);
            w.section(nm, 0, m.vertex_normals, 3 * (size_t)m.normals_count, 4);
            snprintf(nm, sizeof nm, "mesh%u_nidx", k);
            w.section(nm, 1, m.vertex_normal_indices, 3 * (size_t)m.triangle_count, 4);
        }
        if (m.uvs_count) {
            snprintf(nm, sizeof nm, "mesh%u_uv", k);
            w.section(nm, 0, m.vertex_uvs, 2 * (size_t)m.uvs_count, 4);
            snprintf(nm, sizeof nm, "mesh%u_uidx", k);
            w.section(nm, 1, m.vertex_uvs_indices, 3 * (size_t)m.triangle_count, 4);
        }
    }
    for (u32 k = 0; k < c.textures; k++) {
        const Texture &t = s.textures[k];
        snprintf(nm, sizeof nm, "tex%u_info", k);
        w.u32s(nm, {t.width, t.height, t.flags, t.mip_count});
        std::vector<uint32_t> dims;
        for (u32 i = 0; i < t.mip_count; i++) dims.insert(dims.end(), {t.mips[i].width, t.mips[i].height});
        snprintf(nm, sizeof nm, "tex%u_dims", k);
        w.u32s(nm, dims);
        for (u32 i = 0; i < t.mip_count; i++) {
            snprintf(nm, sizeof nm, "tex%u_mip%u", k, i);
            w.section(nm, 3, t.mips[i].texel_quads, mipBytes(t, i), 1);
        }
    }
    return w;
}

// Records frame `target` of a run: <prefix>.spec (the scene spec above) and <prefix>.canvas (the canvas's
// pixels, RGBA float32 per sample, then its depths).
struct Recorder {
    std::string prefix;
    uint32_t frame = 0, target = ~0u;
    bool written = false;
};
inline Recorder &recorder() {
    static Recorder r;
    return r;
}
inline void recordFrame(const Scene &s, const Viewport &v, const slimrt_render_params &p, bool) {
    Recorder &r = recorder();
    if (r.frame != r.target || r.prefix.empty()) return;
    if (!sceneSpec(s, v, p).write(r.prefix + ".spec")) throw std::runtime_error("cannot write " + r.prefix + ".spec");
    const Canvas &cv = v.canvas;
    FILE *f = fopen((r.prefix + ".canvas").c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + r.prefix + ".canvas");
    const size_t n = cv.sampleCount();
    bool ok = fwrite(cv.pixels, sizeof(Pixel), n, f) == n && fwrite(cv.depths, sizeof(f32), n, f) == n;
    if (fclose(f) != 0 || !ok) throw std::runtime_error("cannot write " + r.prefix + ".canvas");
    r.written = true;
}

// ------------------------------------------------------------------------------------------- the loop
struct HeadlessOptions {
    u16 width = 320, height = 180;
    u32 frames = 1;
    f32 delta_time = 1.0f / 60.0f;     // fixed: a run is reproducible
    std::string ppm_prefix;            // when set, frame f's window content goes to <prefix><f>.ppm
};

inline bool writePPM(const std::string &path, u16 width, u16 height, const std::vector<u32> &content) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    fprintf(f, "P6\n%u %u\n255\n", width, height);
    std::vector<unsigned char> row(3 * (size_t)width);
    bool ok = true;
    for (u32 y = 0; y < height && ok; y++) {
        for (u32 x = 0; x < width; x++) {
            const u32 c = content[(size_t)y * width + x];
            row[3 * x] = (unsigned char)(c >> 16);
            row[3 * x + 1] = (unsigned char)(c >> 8);
            row[3 * x + 2] = (unsigned char)c;
        }
        ok = fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    return fclose(f) == 0 && ok;
}

// WinMain's loop without a window (win32.h:164-258): size the window, then redraw frame after frame.
inline int runHeadless(SlimApp &app, const HeadlessOptions &o) {
    window::width = o.width;
    window::height = o.height;
    app.OnWindowResize(o.width, o.height);
    RayTracingRenderer::frameHook() = recordFrame;
    for (u32 f = 0; f < o.frames && app.is_running; f++) {
        recorder().frame = f;
        window::content().clear();
        app.OnWindowRedraw(o.delta_time);
        if (!o.ppm_prefix.empty() && window::content().size() == (size_t)o.width * o.height &&
            !writePPM(o.ppm_prefix + std::to_string(f) + ".ppm", o.width, o.height, window::content()))
            return 1;
    }
    return 0;
}

}  // namespace slimtracin_app
