// Texel-quad baking of RGB images: what the reference's converter (src/bmp2texture.cpp) writes into .texture files.
//
// A mip of w x h texels becomes (w + 1) x (h + 1) quads; quad (x, y) holds the four texels around its corner
// point, TL = texel (x - 1, y - 1), TR = (x, y - 1), BL = (x - 1, y), BR = (x, y), so one 12-byte fetch gives a
// bilinear footprint.  The quads on the border take their missing corners from the texel itself (clamp), from the
// opposite edge (wrap), or, for the three images of a cube map (4-face strip, top, bottom), from the adjacent face
// along the seam, with the eight cube corners set to the midpoint of the two texels they join.  Mips are 2 x 2 box
// filters down to 4 texels.  Colours are floats u8 / 255 (gamma 2.2 unless linear; normal maps renormalised) and
// are quantised by truncation of x * 255 at the end.
//
// The border writes happen per texel, row by row, in the converter's order (TextureMipLoader::load,
// bmp2texture.cpp:36-198), because later writes overwrite earlier ones at the corners and the files record that.
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>

#include "../../include/slimrt.h"

namespace {

struct rgb { float r, g, b; };

// IEEE binary32 -> binary16 bits, rounded to nearest, ties to even; overflow to infinity, NaN kept a (quiet) NaN.
uint16_t half_bits(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u, mag = x & 0x7FFFFFFFu;
    if (mag > 0x7F800000u) return (uint16_t)(sign | 0x7E00u | ((mag >> 13) & 0x3FFu));   // NaN
    if (mag >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);   // rounds to >= 65520: infinity
    if (mag < 0x38800000u) {   // below the smallest normal half (2^-14): subnormal or zero
        if (mag < 0x33000000u) return (uint16_t)sign;   // <= 2^-25: rounds to zero
        const uint32_t e = mag >> 23, m = (mag & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126 - e;   // the value is m * 2^(e - 150); a half subnormal counts 2^-24
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1), halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((mag >> 13) - (112u << 10));   // rebias the exponent 127 -> 15
    const uint32_t rem = mag & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}
struct quad { rgb tl, tr, bl, br; };

constexpr float kByteToFloat = 0.00392156862f;   // COLOR_COMPONENT_TO_FLOAT (src/slim/core/base.h:139)
constexpr float kFloatToByte = 255.0f;           // FLOAT_TO_COLOR_COMPONENT (base.h:140)
constexpr uint32_t kLinear = 1u << 1, kMipmap = 1u << 4, kWrap = 1u << 6, kNormal = 1u << 7, kCubemap = 1u << 8;
constexpr uint32_t kTile = 1u << 2, kChannel = 1u << 3;

// Pixel::lerpTo at 0.5 (base.h:1030-1032, 1104-1109): (to - from) * 0.5 + from, two roundings.
rgb midpoint(rgb from, rgb to) {
    return {(to.r - from.r) * 0.5f + from.r, (to.g - from.g) * 0.5f + from.g, (to.b - from.b) * 0.5f + from.b};
}

struct Mip {
    uint32_t w = 0, h = 0;
    std::vector<rgb> texels;   // w * h, row-major
    std::vector<quad> quads;   // (w + 1) * (h + 1), row-major
    void init(uint32_t W, uint32_t H) {
        w = W;
        h = H;
        texels.assign((size_t)w * h, rgb{0, 0, 0});
        quads.assign((size_t)(w + 1) * (h + 1), quad{});
    }
    quad &at(uint32_t x, uint32_t y) { return quads[(size_t)y * (w + 1) + x]; }

    // TextureMipLoader::load without the cube seams (bmp2texture.cpp:36-95).
    void build(bool wrap) {
        const uint32_t last_x = w - 1, last_y = h - 1;
        for (uint32_t y = 0; y < h; y++) {
            const bool top = y == 0, bottom = y == last_y;
            for (uint32_t x = 0; x < w; x++) {
                const rgb c = texels[(size_t)y * w + x];
                const bool left = x == 0, right = x == last_x;
                at(x, y).br = c;
                at(x + 1, y).bl = c;
                at(x, y + 1).tr = c;
                at(x + 1, y + 1).tl = c;
                if (left) {
                    at(x, y).bl = c;
                    at(x, y + 1).tl = c;
                } else if (right) {
                    at(x + 1, y).br = c;
                    at(x + 1, y + 1).tr = c;
                }
                if (wrap) {
                    if (left) {
                        at(w, y).br = c;
                        at(w, y + 1).tr = c;
                    } else if (right) {
                        at(0, y).bl = c;
                        at(0, y + 1).tl = c;
                    }
                    if (top) {
                        at(x, h).br = c;
                        at(x + 1, h).bl = c;
                        if (left) at(w, h).br = c;
                        else if (right) at(0, h).bl = c;
                    } else if (bottom) {
                        at(x, 0).tr = c;
                        at(x + 1, 0).tl = c;
                        if (left) at(w, 0).tr = c;
                        else if (right) at(0, 0).tl = c;
                    }
                } else if (top) {
                    at(x, y).tr = c;
                    at(x + 1, y).tl = c;
                    if (left) at(x, y).tl = c;
                    else if (right) at(x + 1, y).tr = c;
                } else if (bottom) {
                    at(x, y + 1).br = c;
                    at(x + 1, y + 1).bl = c;
                    if (left) at(x, y + 1).bl = c;
                    else if (right) at(x + 1, y + 1).br = c;
                }
            }
        }
    }
};

// The strip's seams with the top and bottom faces (CubeMapLoaderMode_Main, bmp2texture.cpp:97-128).  The strip is
// 4 faces of n x n: left, front, right, back.  Kept as the converter does it: the back face's seams read the top
// face from texel n down to 1 and the bottom face from texel n * n (one past its end, which the shipped files hold
// as black) down to n * (n - 1) + 1; the last two corners are addressed with a row of (n + 1) * 4 quads where the
// strip's rows hold 4n + 1, so the top-right corner lands on quad (2, 1) and the bottom-right one past the array
// (dropped here).
void stitch_strip(Mip &strip, const std::vector<rgb> &top, const std::vector<rgb> &bot, uint32_t n) {
    const uint32_t last = n - 1, last_row = n * last, last_texel = n * n - 1, qw = (n + 1) * 4, H = strip.h;
    auto bottom_texel = [&](size_t k) { return k < bot.size() ? bot[k] : rgb{0.0f, 0.0f, 0.0f}; };
    const size_t first = 0, final = (size_t)H * (strip.w + 1);   // linear indices of rows 0 and H
    std::vector<quad> &q = strip.quads;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t L = i, F = n + i, R = 2 * n + i, B = 3 * n + i;   // face offsets + i along the strip
        q[first + 1 + L].tl = q[first + L].tr = top[n * i];
        q[first + 1 + F].tl = q[first + F].tr = top[last_row + i];
        q[first + 1 + R].tl = q[first + R].tr = top[last_texel - n * i];
        q[first + 1 + B].tl = q[first + B].tr = top[n - i];
        q[final + 1 + L].bl = q[final + L].br = bottom_texel(n * (last - i));
        q[final + 1 + F].bl = q[final + F].br = bottom_texel(i);
        q[final + 1 + R].bl = q[final + R].br = bottom_texel(n * i + last);
        q[final + 1 + B].bl = q[final + B].br = bottom_texel(last_row + (n - i));
    }
    q[first].tl = midpoint(top[0], strip.texels[strip.w - 1]);
    q[first + qw - 1].tr = midpoint(top[0], strip.texels[0]);
    q[final].bl = midpoint(bot[last_row], strip.texels[(size_t)strip.w * last]);
    if (final + qw - 1 < q.size()) q[final + qw - 1].br = midpoint(bot[last_row], strip.texels[0]);
}

// The top face's seams with the strip's first row (CubeMapLoaderMode_Top, bmp2texture.cpp:129-160).
void stitch_top(Mip &face, const std::vector<rgb> &strip, uint32_t n) {
    const uint32_t last = n - 1;
    const rgb *left = &strip[0], *front = &strip[n], *right = &strip[2 * n], *back = &strip[3 * n];
    quad *first_row = &face.at(0, 0), *final_row = &face.at(0, n);
    first_row[0].tl = midpoint(left[0], back[last]);
    first_row[n].tr = midpoint(back[0], right[last]);
    final_row[0].bl = midpoint(front[0], left[last]);
    final_row[n].br = midpoint(right[0], front[last]);
    for (uint32_t i = 0; i < n; i++) {
        first_row[1 + i].tl = first_row[i].tr = back[last - i];
        final_row[1 + i].bl = final_row[i].br = front[i];
        face.at(0, i).bl = face.at(0, i + 1).tl = left[i];
        face.at(n, i).br = face.at(n, i + 1).tr = right[last - i];
    }
}

// The bottom face's seams with the strip's last row (CubeMapLoaderMode_Bottom, bmp2texture.cpp:161-198).
void stitch_bottom(Mip &face, const std::vector<rgb> &strip, uint32_t n) {
    const uint32_t last = n - 1, row = 4 * n * last;
    const rgb *left = &strip[row], *front = &strip[row + n], *right = &strip[row + 2 * n], *back = &strip[row + 3 * n];
    quad *first_row = &face.at(0, 0), *final_row = &face.at(0, n);
    first_row[0].tl = midpoint(front[0], left[last]);
    first_row[n].tr = midpoint(right[0], front[last]);
    final_row[0].bl = midpoint(left[0], back[last]);
    final_row[n].br = midpoint(back[0], right[last]);
    for (uint32_t i = 0; i < n; i++) {
        first_row[1 + i].tl = first_row[i].tr = front[i];
        final_row[1 + i].bl = final_row[i].br = back[last - i];
        face.at(0, i).bl = face.at(0, i + 1).tl = left[last - i];
        face.at(n, i).br = face.at(n, i + 1).tr = right[i];
    }
}

// loadMips (bmp2texture.cpp:201-237): texel (x, y) of the next mip = the average of quad (2x + 1, 2y + 1).
void halve(Mip &from, Mip &to, bool wrap) {
    to.init(from.w / 2, from.h / 2);
    for (uint32_t y = 0; y < to.h; y++)
        for (uint32_t x = 0; x < to.w; x++) {
            const quad &q = from.at(2 * x + 1, 2 * y + 1);
            to.texels[(size_t)y * to.w + x] = {0.25f * (q.tl.r + q.tr.r + q.bl.r + q.br.r),
                                               0.25f * (q.tl.g + q.tr.g + q.bl.g + q.br.g),
                                               0.25f * (q.tl.b + q.tr.b + q.bl.b + q.br.b)};
        }
    to.build(wrap);
}

uint8_t to_byte(float v) { return (uint8_t)(v * kFloatToByte); }

void quantize(const Mip &m, uint8_t *out) {
    for (const quad &q : m.quads) {
        const rgb c[4] = {q.tl, q.tr, q.bl, q.br};
        for (int k = 0; k < 4; k++) {
            out[k] = to_byte(c[k].r);
            out[4 + k] = to_byte(c[k].g);
            out[8 + k] = to_byte(c[k].b);
        }
        out += 12;
    }
}

// componentsToPixel (src/slim/platforms/win32_bitmap.h:24-33) and the normal-map renormalisation
// (bmp2texture.cpp:302-318).
rgb texel_from_bytes(const uint8_t *c, uint32_t flags) {
    rgb p{(float)c[0] * kByteToFloat, (float)c[1] * kByteToFloat, (float)c[2] * kByteToFloat};
    if (!(flags & kLinear)) {
        p.r = powf(p.r, 2.2f);
        p.g = powf(p.g, 2.2f);
        p.b = powf(p.b, 2.2f);
    }
    return p;
}

rgb renormalize(rgb p) {
    float r = p.r * 2.0f - 1.0f, g = p.g * 2.0f - 1.0f;
    const float l_rcp = 1.0f / sqrtf(r * r + g * g + 1.0f);
    r *= l_rcp;
    g *= l_rcp;
    const float b = sqrtf(1.0f - r * r - g * g);
    return {r * 0.5f + 0.5f, g * 0.5f + 0.5f, b * 0.5f + 0.5f};
}

}  // namespace

extern "C" {

int slimrt_host_bake_texture(const uint8_t *rgb_bytes, uint32_t width, uint32_t height, uint32_t flags,
                             uint8_t *quads_out, size_t quad_cap, size_t *quad_count_out, uint32_t *mip_count_out,
                             uint32_t *mip_dims_out, uint32_t *flags_out) {
    if (!rgb_bytes || !width || !height || !quad_count_out || !mip_count_out || !mip_dims_out || !flags_out)
        return SLIMRT_E_ARG;
    if (width > 65536 || height > 65536) return SLIMRT_E_ARG;
    std::vector<Mip> mips;
    flags &= ~kChannel;   // the converter clears the channel flag
    if (flags & kCubemap) {   // one image: 4-face strip, top face, bottom face, each of height x height
        if (width != 6 * height) return SLIMRT_E_ARG;
        flags &= ~(kWrap | kNormal | kMipmap | kTile);
        const uint32_t n = height;
        mips.resize(3);
        mips[0].init(4 * n, n);
        mips[1].init(n, n);
        mips[2].init(n, n);
        for (uint32_t y = 0; y < n; y++)
            for (uint32_t x = 0; x < width; x++) {
                const rgb p = texel_from_bytes(rgb_bytes + 3 * ((size_t)y * width + x), flags);
                if (x < 4 * n) mips[0].texels[(size_t)y * 4 * n + x] = p;
                else if (x < 5 * n) mips[1].texels[(size_t)y * n + (x - 4 * n)] = p;
                else mips[2].texels[(size_t)y * n + (x - 5 * n)] = p;
            }
        for (Mip &m : mips) m.build(false);
        stitch_strip(mips[0], mips[1].texels, mips[2].texels, n);
        stitch_top(mips[1], mips[0].texels, n);
        stitch_bottom(mips[2], mips[0].texels, n);
    } else {
        const bool wrap = (flags & kWrap) != 0;
        mips.resize(1);
        mips[0].init(width, height);
        for (size_t i = 0; i < (size_t)width * height; i++) {
            rgb p = texel_from_bytes(rgb_bytes + 3 * i, flags);
            mips[0].texels[i] = (flags & kNormal) ? renormalize(p) : p;
        }
        mips[0].build(wrap);
        if (flags & kMipmap)
            while (mips.back().w > 4 && mips.back().h > 4) {
                Mip next;
                halve(mips.back(), next, wrap);
                mips.push_back(std::move(next));
            }
    }
    size_t total = 0;
    for (const Mip &m : mips) total += m.quads.size();
    *quad_count_out = total;
    *mip_count_out = (uint32_t)mips.size();
    *flags_out = flags;
    if (mips.size() > 16) return SLIMRT_E_OVERFLOW;
    for (size_t i = 0; i < mips.size(); i++) {
        mip_dims_out[2 * i] = mips[i].w;
        mip_dims_out[2 * i + 1] = mips[i].h;
    }
    if (!quads_out) return SLIMRT_OK;   // size query
    if (quad_cap < total) return SLIMRT_E_OVERFLOW;
    for (const Mip &m : mips) {
        quantize(m, quads_out);
        quads_out += 12 * m.quads.size();
    }
    return SLIMRT_OK;
}

// The bordered texel grid of a mip's quads (SLIMRT_TEXELS_* storage, include/slimrt.h): grid texel (x, y) is quad
// (x, y)'s top-left corner for x <= w, y <= h; the last column and row come from the last quads' right and bottom
// corners.  Quads whose four corners are not the grid's (a converter that wrote a corner of one quad differently
// from the same texel in its neighbours) are counted: with none, RGBA8 texels sample exactly as the quads do.
int slimrt_host_texels_from_quads(const uint8_t *quads, uint32_t width, uint32_t height, uint32_t format,
                                  void *texels_out, uint64_t *mismatches_out) {
    if (!quads || !texels_out || width == 0 || height == 0) return SLIMRT_E_ARG;
    if (format != SLIMRT_TEXELS_RGBA8 && format != SLIMRT_TEXELS_RGBA16F) return SLIMRT_E_ARG;
    const size_t W = (size_t)width + 2, H = (size_t)height + 2, QW = (size_t)width + 1;
    std::vector<uint8_t> grid(W * H * 3);
    auto corner = [&](size_t qx, size_t qy, int k, int ch) { return quads[12 * (qy * QW + qx) + 4 * ch + k]; };
    for (size_t y = 0; y < H; y++)
        for (size_t x = 0; x < W; x++) {
            const size_t qx = x <= width ? x : width, qy = y <= height ? y : height;
            const int k = (x > width ? 1 : 0) + (y > height ? 2 : 0);   // TL, TR, BL, BR
            for (int ch = 0; ch < 3; ch++) grid[3 * (y * W + x) + ch] = corner(qx, qy, k, ch);
        }
    uint64_t bad = 0;
    for (size_t qy = 0; qy <= height; qy++)
        for (size_t qx = 0; qx <= width; qx++)
            for (int k = 0; k < 4; k++) {
                const size_t gx = qx + (k & 1), gy = qy + (k >> 1);
                bool same = true;
                for (int ch = 0; ch < 3; ch++) same = same && grid[3 * (gy * W + gx) + ch] == corner(qx, qy, k, ch);
                if (!same) { bad++; break; }
            }
    if (mismatches_out) *mismatches_out = bad;
    if (format == SLIMRT_TEXELS_RGBA8) {
        uint8_t *o = (uint8_t *)texels_out;
        for (size_t i = 0; i < W * H; i++) {
            o[4 * i] = grid[3 * i]; o[4 * i + 1] = grid[3 * i + 1]; o[4 * i + 2] = grid[3 * i + 2]; o[4 * i + 3] = 255;
        }
    } else {
        uint16_t *o = (uint16_t *)texels_out;
        for (size_t i = 0; i < W * H; i++) {
            for (int ch = 0; ch < 3; ch++) o[4 * i + ch] = half_bits((float)grid[3 * i + ch] / 255.0f);
            o[4 * i + 3] = half_bits(1.0f);
        }
    }
    return SLIMRT_OK;
}

uint16_t slimrt_host_float_to_half(float f) { return half_bits(f); }

}  // extern "C"
