// The reference's texture converter CLI (src/bmp2texture.cpp:239-361, README "bmp2texture src.bmp trg.texture"):
// a .bmp image baked into the renderer's .texture file - texel quads per mip, gamma 2.2 unless linear, optional
// box-filtered mips, wrapped or clamped borders, renormalised normal maps, stitched cube maps - through
// loadBitmap, bakeTexture (slimrt_host_bake_texture) and save.
//
//   bmp2texture src.bmp trg.texture [-c] [-f] [-l] [-m] [-n] [-w]
//     -c cube map (4-face strip, top face, bottom face; width = 6 * height)   -f flip vertically
//     -l linear (no gamma correction)   -m mipmaps   -n normal map   -w wrap
//
// -t (the reference's tiled component order) is refused: a texture read in tiled order is scrambled, and no
// texture the reference ships uses it.
#include <cstdio>
#include <vector>

#include "slimtracin/app.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s src.bmp trg.texture [-c] [-f] [-l] [-m] [-n] [-w]\n", argv[0]);
        return 2;
    }
    u32 flags = 0;   // ImageFlags bits (src/slim/core/base.h:1264-1277)
    bool flip = false;
    for (int i = 3; i < argc; i++) {
        const char *a = argv[i];
        if (a[0] != '-' || !a[1] || a[2]) {
            fprintf(stderr, "bmp2texture: unknown option %s\n", a);
            return 2;
        }
        switch (a[1]) {
            case 'l': flags |= 1u << 1; break;
            case 'm': flags |= 1u << 4; break;
            case 'f': flags |= 1u << 5; flip = true; break;
            case 'w': flags |= 1u << 6; break;
            case 'n': flags |= 1u << 7; break;
            case 'c': flags |= 1u << 8; break;
            case 't': fprintf(stderr, "bmp2texture: -t (tiled component order) is not supported\n"); return 2;
            default: fprintf(stderr, "bmp2texture: unknown option %s\n", a); return 2;
        }
    }
    std::vector<u8> rgb;
    u32 width = 0, height = 0;
    bool alpha = false;
    if (!loadBitmap(argv[1], rgb, width, height, flip, &alpha)) {
        fprintf(stderr, "bmp2texture: %s is not an uncompressed 24- or 32-bit BMP\n", argv[1]);
        return 1;
    }
    if (alpha) flags |= 1u;   // loadBitmap sets the alpha flag for 32-bit images; the converter keeps it
    try {
        Texture t = bakeTexture(rgb.data(), width, height, flags);
        if (!save(t, argv[2])) {
            fprintf(stderr, "bmp2texture: cannot write %s\n", argv[2]);
            return 1;
        }
    } catch (const std::exception &e) {
        fprintf(stderr, "bmp2texture: %s\n", e.what());
        return 1;
    }
    return 0;
}
