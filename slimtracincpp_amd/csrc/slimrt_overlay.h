// Wireframe overlays drawn into a canvas: boxes (drawBox, src/slim/draw/box.h:7-34), the node boxes of a BVH
// (drawBVH, draw/bvh.h:8-28) and the edges of a mesh (drawMesh, draw/mesh.h:7-22).  Every edge goes through
// drawEdge (draw/edge.h:7-17): to view space, near/far and frustum-side clipping (Frustum::cullAndClipEdge,
// viewport/frustum.h:67-141), projection to the canvas (projectPoint, frustum.h:180-188), then the antialiased
// line of _drawLine (draw/line.h:70-273) whose samples are blended into the canvas by Canvas::setPixel
// (draw/canvas.h:128-185).
//
// Everything here is __host__ __device__: the CPU context draws sequentially with these functions, and the
// device path walks the same lines to emit its samples in drawing order (slimrt_overlay.hip).
#pragma once

#include "slimrt_core.h"

namespace slimrt {

// A canvas and the view it is drawn from (Viewport, viewport/viewport.h:7-61, with its default full-canvas bounds
// and Frustum).
struct OverlayView {
    v3 cam_x, cam_y, cam_z;     // the camera's rotation matrix (OrientationUsing3x3Matrix) by rows
    v3 cam_position;
    float focal_length;
    float near_distance, far_distance;
    float width_over_height;    // Dimensions (base.h:1300-1323)
    float f_width, f_height, h_width, h_height;
    float proj_x, proj_y;       // Frustum::Projection::scale.x, .y (frustum.h:25-27)
    int32_t width, height;      // canvas pixels
    uint32_t ssaa;
};

// One line of _drawLine in canvas coordinates, with its colour, opacity and width.
struct OverlayLine {
    float x1, y1, z1, x2, y2, z2;
    color col;
    float opacity;
    uint32_t line_width;
};

// Camera::internPos (scene/camera.h:47): the transposed rotation times (pos - position), mat3 * vec3
// (math/mat3.h:254-260).
SLIM_HD v3 camera_intern(const OverlayView &v, v3 p) {
    const v3 d = p - v.cam_position;
    return v3{v.cam_x.x * d.x + v.cam_x.y * d.y + v.cam_x.z * d.z, v.cam_y.x * d.x + v.cam_y.y * d.y + v.cam_y.z * d.z,
              v.cam_z.x * d.x + v.cam_z.y * d.y + v.cam_z.z * d.z};
}

// Transform::externPos (core/transform.h:10): scale, rotate, translate.
SLIM_HD v3 transform_extern(quat q, v3 position, v3 scale, v3 p) { return qrot(q, p * scale) + position; }

// vec3::lerpTo (math/vec3.h:360-362)
SLIM_HD v3 lerp_to(v3 a, v3 b, float by) { return scale_add(b - a, by, a); }

// Clips one side: keeps the part where both plane values are >= 0 (the "out" cases of cullAndClipEdge).
SLIM_HD bool clip_side(v3 &A, v3 &B, float nA, float nB) {
    const uint32_t out = (nA < 0 ? 1u : 0u) | (nB < 0 ? 2u : 0u);
    if (out == 3) return false;
    if (out & 1) A = lerp_to(A, B, nA / (nA - nB));
    else if (out & 2) B = lerp_to(B, A, nB / (nB - nA));
    return true;
}

// Frustum::cullAndClipEdge (viewport/frustum.h:67-141) on a view-space edge.
SLIM_HD bool cull_and_clip(const OverlayView &v, v3 &A, v3 &B) {
    float d = v.near_distance;
    uint32_t out = (A.z < d ? 1u : 0u) | (B.z < d ? 2u : 0u);
    if (out) {
        if (out == 3) return false;
        if (out & 1) A = lerp_to(A, B, (d - A.z) / (B.z - A.z));
        else B = lerp_to(B, A, (d - B.z) / (A.z - B.z));
    }
    d = v.far_distance;
    out = (A.z > d ? 1u : 0u) | (B.z > d ? 2u : 0u);
    if (out) {
        if (out == 3) return false;
        if (out & 1) A = lerp_to(A, B, (A.z - d) / (A.z - B.z));
        else B = lerp_to(B, A, (B.z - d) / (B.z - A.z));
    }
    v3 N{v.focal_length, 0.0f, v.width_over_height};   // left plane, facing right
    if (!clip_side(A, B, dot(N, A), dot(N, B))) return false;
    N.x = -N.x;                                        // right plane
    if (!clip_side(A, B, dot(N, A), dot(N, B))) return false;
    N = v3{0.0f, v.focal_length, 1.0f};                // bottom plane, facing up
    if (!clip_side(A, B, dot(N, A), dot(N, B))) return false;
    N.y = -N.y;                                        // top plane
    return clip_side(A, B, dot(N, A), dot(N, B));
}

// drawEdge (draw/edge.h:7-17) up to the line: world-space end points to a canvas line (z: view depth - EPS);
// false when the edge is culled.
SLIM_HD bool edge_line(const OverlayView &v, v3 from_world, v3 to_world, OverlayLine &l) {
    v3 A = camera_intern(v, from_world), B = camera_intern(v, to_world);
    if (!cull_and_clip(v, A, B)) return false;
    // Frustum::projectPoint (frustum.h:180-184)
    l.x1 = ((v.proj_x * A.x / A.z) + 1) * v.h_width;
    l.y1 = v.f_height - ((v.proj_y * A.y / A.z) + 1) * v.h_height;
    l.x2 = ((v.proj_x * B.x / B.z) + 1) * v.h_width;
    l.y2 = v.f_height - ((v.proj_y * B.y / B.z) + 1) * v.h_height;
    l.z1 = A.z - kEps;
    l.z2 = B.z - kEps;
    // A non-finite end point (degenerate transforms) would reach float-to-int conversions that C++ leaves
    // undefined in _drawLine; such lines are not drawn.
    return isfinite(l.x1) && isfinite(l.y1) && isfinite(l.x2) && isfinite(l.y2) && isfinite(l.z1) && isfinite(l.z2);
}

// RangeOf (core/base.h:282-318): contains, emptiness and sub() as the reference has them.
template <typename T>
struct RangeT {
    T first, last;
    SLIM_HD bool contains(int32_t x) const { return (first <= x) && (x <= last); }
    SLIM_HD bool empty() const { return last < first; }
    SLIM_HD void sub(T a, T b) {
        if (b < a) { T t = b; b = a; a = t; }
        if (last < a || b < first) {
            first = -1;
            last = -2;
        } else {
            first = first < a ? a : first;
            last = b < last ? b : last;
        }
    }
};

// (i32) of a float as the reference's x86-64 build computes it (cvttss2si): truncation, and INT32_MIN for NaN or
// values out of range, so that degenerate lines (NaN gradients) draw nothing on the GPU either.
SLIM_HD int32_t to_i32(float f) { return (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : INT32_MIN; }

SLIM_HD float fraction_of(float x) { return x - floorf(x); }              // fractionOf (base.h:219)
SLIM_HD float one_minus_fraction_of(float x) { return 1 - fraction_of(x); }

// _drawLine (draw/line.h:70-273) with the viewport's full-canvas bounds: calls put(x, y, opacity, depth) for every
// Canvas::setPixel the reference makes, in its order (first end point, last end point, then the span between).
// Coordinates are canvas samples (doubled under SSAA).
template <class Put>
SLIM_HD void walk_line(const OverlayView &v, OverlayLine l, Put &put) {
    float x1 = l.x1, y1 = l.y1, z1 = l.z1, x2 = l.x2, y2 = l.y2, z2 = l.z2;
    const float opacity = l.opacity;
    uint8_t line_width = (uint8_t)l.line_width;
    RangeT<float> fx{x1 <= x2 ? x1 : x2, x1 <= x2 ? x2 : x1};
    RangeT<float> fy{y1 <= y2 ? y1 : y2, y1 <= y2 ? y2 : y1};
    // viewport bounds: left = top = 0, right = width - 1, bottom = height - 1
    const float left = 0.0f, top = 0.0f, right = (float)(v.width - 1), bottom = (float)(v.height - 1);
    x1 += left; x2 += left; y1 += top; y2 += top;
    fx.first += left; fx.last += left;
    fy.first += top; fy.last += top;
    fx.sub(left, right);
    fy.sub(top, bottom);
    fx.sub(0, v.f_width - 1.0f);
    fy.sub(0, v.f_height - 1.0f);
    if (fx.empty() || fy.empty()) return;
    RangeT<int32_t> xr{to_i32(fx.first), to_i32(ceilf(fx.last))};
    RangeT<int32_t> yr{to_i32(fy.first), to_i32(ceilf(fy.last))};
    if (xr.last == v.width) xr.last--;
    if (yr.last == v.height) yr.last--;
    if (v.ssaa) {
        x1 += x1; x2 += x2; y1 += y1; y2 += y2;
        xr.first <<= 1; yr.first <<= 1; xr.last <<= 1; yr.last <<= 1;
        line_width = (uint8_t)(line_width << 1);
        line_width++;
    }
    const float dx = x2 - x1, dy = y2 - y1;
    const bool has_depth = (z1 != 0.0f) || (z2 != 0.0f);   // the canvas always has depths here
    float tmp, z_range, range_remap, z = 0.0f, z_curr = 0.0f, z_step = 0.0f;
    int32_t x, y;
    if (fabsf(dx) > fabsf(dy)) {   // shallow
        if (x2 < x1) {
            tmp = x2; x2 = x1; x1 = tmp;
            tmp = y2; y2 = y1; y1 = tmp;
            tmp = z2; z2 = z1; z1 = tmp;
        }
        const float grad = dy / dx;
        const float first_x = roundf(x1), last_x = roundf(x2);
        const float first_offset = first_x - x1, last_offset = last_x - x2;
        const float first_y = y1 + grad * first_offset, last_y = y2 + grad * last_offset;
        const int32_t start_x = to_i32(first_x), start_y = to_i32(first_y);
        const int32_t end_x = to_i32(last_x), end_y = to_i32(last_y);
        x = start_x;
        y = start_y;
        float gap = one_minus_fraction_of(x1 + 0.5f);
        if (xr.contains(x)) {
            if (yr.contains(y)) put(x, y, one_minus_fraction_of(first_y) * gap * opacity, z1);
            for (uint8_t i = 0; i < line_width; i++) if (yr.contains(++y)) put(x, y, opacity, z1);
            if (yr.contains(++y)) put(x, y, fraction_of(first_y) * gap * opacity, z1);
        }
        x = end_x;
        y = end_y;
        gap = fraction_of(x2 + 0.5f);
        if (xr.contains(x)) {
            if (yr.contains(y)) put(x, y, one_minus_fraction_of(last_y) * gap * opacity, z2);
            for (uint8_t i = 0; i < line_width; i++) if (yr.contains(++y)) put(x, y, opacity, z2);
            if (yr.contains(++y)) put(x, y, fraction_of(last_y) * gap * opacity, z2);
        }
        if (has_depth) {   // one over depth: start and step
            z1 = 1.0f / z1;
            z2 = 1.0f / z2;
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
            if (xr.contains(x)) {
                y = to_i32(gap);
                if (has_depth) z = 1.0f / z_curr;
                if (yr.contains(y)) put(x, y, one_minus_fraction_of(gap) * opacity, z);
                for (uint8_t i = 0; i < line_width; i++) if (yr.contains(++y)) put(x, y, opacity, z);
                if (yr.contains(++y)) put(x, y, fraction_of(gap) * opacity, z);
            }
            gap += grad;
            if (has_depth) z_curr += z_step;
        }
    } else {   // steep
        if (y2 < y1) {
            tmp = x2; x2 = x1; x1 = tmp;
            tmp = y2; y2 = y1; y1 = tmp;
            tmp = z2; z2 = z1; z1 = tmp;
        }
        const float grad = dx / dy;
        const float first_y = roundf(y1), last_y = roundf(y2);
        const float first_offset = y1 - first_y, last_offset = last_y - y2;
        const float first_x = x1 + grad * first_offset, last_x = x2 + grad * last_offset;
        const int32_t start_y = to_i32(first_y), start_x = to_i32(first_x);
        const int32_t end_y = to_i32(last_y), end_x = to_i32(last_x);
        x = start_x;
        y = start_y;
        float gap = one_minus_fraction_of(y1 + 0.5f);
        if (yr.contains(y)) {
            if (xr.contains(x)) put(x, y, one_minus_fraction_of(first_x) * gap * opacity, z1);
            for (uint8_t i = 0; i < line_width; i++) if (xr.contains(++x)) put(x, y, opacity, z1);
            if (xr.contains(++x)) put(x, y, fraction_of(first_x) * gap * opacity, z1);
        }
        x = end_x;
        y = end_y;
        gap = fraction_of(y2 + 0.5f);
        if (yr.contains(y)) {
            if (xr.contains(x)) put(x, y, one_minus_fraction_of(last_x) * gap * opacity, z2);
            for (uint8_t i = 0; i < line_width; i++) if (xr.contains(++x)) put(x, y, opacity, z2);
            if (xr.contains(++x)) put(x, y, fraction_of(last_x) * gap * opacity, z2);
        }
        if (has_depth) {
            z1 = 1.0f / z1;
            z2 = 1.0f / z2;
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
            if (yr.contains(y)) {
                if (has_depth) z = 1.0f / z_curr;
                x = to_i32(gap);
                if (xr.contains(x)) put(x, y, one_minus_fraction_of(gap) * opacity, z);
                for (uint8_t i = 0; i < line_width; i++) if (xr.contains(++x)) put(x, y, opacity, z);
                if (xr.contains(++x)) put(x, y, fraction_of(gap) * opacity, z);
            }
            gap += grad;
            if (has_depth) z_curr += z_step;
        }
    }
}

// The canvas offset of sample (x, y), or -1 when Canvas::setPixel would drop it (canvas.h:128-135, 147).
SLIM_HD int64_t overlay_offset(const OverlayView &v, int32_t x, int32_t y) {
    const int32_t w = v.ssaa ? v.width << 1 : v.width, h = v.ssaa ? v.height << 1 : v.height;
    if (x < 0 || y < 0 || x >= w || y >= h) return -1;
    return v.ssaa ? ((int64_t)v.width * (y >> 1) + (x >> 1)) * 4 + 2 * (y & 1) + (x & 1) : (int64_t)v.width * y + x;
}

// Canvas::setPixel (canvas.h:128-185, no MSAA) on one canvas sample: the new sample is clamp(c)^2 * opacity with
// that opacity; it replaces the old one when forced (negative opacity), when the old one is empty, or when it is
// opaque without depth; otherwise the nearer of the two is blended over the other (Pixel::alphaBlendOver,
// base.h:1162-1164; _sortPixelsByDepth, canvas.h:267-276).
SLIM_HD void blend_sample(float4 &out, float &out_depth, color c, float opacity, float depth) {
    bool override_ = false;
    if (opacity < 0) {
        opacity = -opacity;
        override_ = true;
    }
    opacity = clamp01(opacity);
    color pc{clamp01(c.r), clamp01(c.g), clamp01(c.b)};
    pc = pc * (pc * opacity);
    if (override_ || (out_depth == INFINITY && out.x == 0 && out.y == 0 && out.z == 0) ||
        (opacity == 1.0f && depth == INFINITY)) {
        out = make_float4(pc.r, pc.g, pc.b, opacity);
        out_depth = depth;
        return;
    }
    float4 fg, bg;
    const float4 mine = make_float4(pc.r, pc.g, pc.b, opacity);
    if (depth == INFINITY || depth < out_depth) {
        out_depth = depth;
        bg = out;
        fg = mine;
    } else {
        bg = mine;
        fg = out;
    }
    if (fg.w == 1) {
        out = fg;
    } else {
        const float k = 1.0f - fg.w;
        out = make_float4(fg.x + bg.x * k, fg.y + bg.y * k, fg.z + bg.z * k, fg.w + bg.w * k);
    }
}

// The vertex pairs of a Box's edges (BoxEdges::setFrom, scene/box.h:52-77) over its corners (BoxVertices,
// box.h:17-29: front_top_left, front_top_right, front_bottom_left, front_bottom_right, then the back four):
// front top, bottom, left, right; back top, bottom, left, right; left bottom, left top; right bottom, right top.
// (Switches rather than tables, so that host and device code share them.)
SLIM_HD void box_edge_verts(int e, int &a, int &b) {
    switch (e) {
        case 0: a = 0; b = 1; break;
        case 1: a = 2; b = 3; break;
        case 2: a = 2; b = 0; break;
        case 3: a = 3; b = 1; break;
        case 4: a = 4; b = 5; break;
        case 5: a = 6; b = 7; break;
        case 6: a = 6; b = 4; break;
        case 7: a = 7; b = 5; break;
        case 8: a = 2; b = 6; break;
        case 9: a = 0; b = 4; break;
        case 10: a = 3; b = 7; break;
        default: a = 1; b = 5; break;
    }
}
constexpr uint32_t kBoxAllSides = S_Top | S_Bottom | S_Left | S_Right | S_Front | S_Back;

// drawBox's subset branch (draw/box.h:21-32): its i-th test draws edge `edge` when either of two sides is chosen
// (front top, front bottom, front left, front right, back ..., left top, left bottom, right top, right bottom).
SLIM_HD void box_side_test(int i, int &edge, uint32_t &sides) {
    switch (i) {
        case 0: edge = 0; sides = S_Front | S_Top; break;
        case 1: edge = 1; sides = S_Front | S_Bottom; break;
        case 2: edge = 2; sides = S_Front | S_Left; break;
        case 3: edge = 3; sides = S_Front | S_Right; break;
        case 4: edge = 4; sides = S_Back | S_Top; break;
        case 5: edge = 5; sides = S_Back | S_Bottom; break;
        case 6: edge = 6; sides = S_Back | S_Left; break;
        case 7: edge = 7; sides = S_Back | S_Right; break;
        case 8: edge = 9; sides = S_Left | S_Top; break;
        case 9: edge = 8; sides = S_Left | S_Bottom; break;
        case 10: edge = 11; sides = S_Right | S_Top; break;
        default: edge = 10; sides = S_Right | S_Bottom; break;
    }
}

// The edges drawBox draws for `sides` (taken as the reference's u8), in its order: out[0..n).
SLIM_HD int box_edges(uint32_t sides, int out[12]) {
    sides &= 0xFFu;
    int n = 0;
    for (int i = 0; i < 12; i++) {
        if (sides == kBoxAllSides) {
            out[n++] = i;
            continue;
        }
        int e;
        uint32_t m;
        box_side_test(i, e, m);
        if (sides & m) out[n++] = e;
    }
    return n;
}

// Corner i of a box with local bounds mn..mx (BoxVertices(min, max), box.h:19-29).
SLIM_HD v3 box_corner(v3 mn, v3 mx, int i) {
    return v3{(i & 1) ? mx.x : mn.x, (i & 2) ? mn.y : mx.y, (i & 4) ? mn.z : mx.z};
}

}  // namespace slimrt
