// The multi-view trace kernel (slimrt_views.hip launches it): one-wave workgroups with a WX x WY tile (64
// samples) of one view each, one workgroup per tile or a persistent grid fed by a device counter.  Its
// instantiations are split over two files by light terms (slimrt_views_all.hip: kLightsAll,
// slimrt_views_none.hip: kLightsNone), which compile in parallel.
#ifndef SLIMRT_VIEWS_TRACE_H
#define SLIMRT_VIEWS_TRACE_H

#include "slimrt_context.h"

namespace slimrt_views_detail {

struct ViewRec {
    RenderParams rp;
    float4 *pixels;
    float *depths;
    uint32_t tiles_x, tile_base, tiles, pad;
};

// One launch of the views kernel: its arguments and the form (tile shape 0..3, node width, persistent grid).
struct ViewsLaunch {
    SceneView sv;
    const ViewRec *recs;
    uint32_t *counter, *status;
    hipStream_t stream;
    size_t lds;
    uint32_t blocks, nviews, total_tiles, grab, scene_depth, mesh_depth, tile_shape, width;
    bool persistent;
};

hipError_t launch_views_all(const ViewsLaunch &l);    // scenes with lights or emissive quads
hipError_t launch_views_none(const ViewsLaunch &l);   // scenes with neither

}  // namespace slimrt_views_detail

namespace {

using slimrt_views_detail::ViewRec;
using slimrt_views_detail::ViewsLaunch;
using slimrt_views_detail::launch_views_all;
using slimrt_views_detail::launch_views_none;



// The tile of one view traced by one workgroup, exactly as the single-view render kernel traces it.  The view
// records are read through the constant address space (invariant loads).
template <int WX, int WY, uint32_t LIGHTS, class TR>
__device__ __forceinline__ void trace_view_tile(const SceneView &s, const ViewRec *views, uint32_t nviews, uint32_t tile,
                                                TR &tr, uint32_t lane) {
    typedef const __attribute__((address_space(4))) ViewRec cview;
    typedef const __attribute__((address_space(4))) uint32_t cword;
    cview *cv = (cview *)views;
    uint32_t lo = 0, hi = nviews;   // the view whose tile range holds this tile
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cv[mid].tile_base <= tile) lo = mid; else hi = mid;
    }
    RenderParams rp;
    {
        cword *src = (cword *)&cv[lo].rp;
        uint32_t *dst = (uint32_t *)&rp;
#pragma unroll
        for (uint32_t i = 0; i < sizeof(RenderParams) / 4; i++) dst[i] = src[i];
    }
    const uint32_t tiles_x = cv[lo].tiles_x, local = tile - cv[lo].tile_base;
    float4 *const pixels = cv[lo].pixels;
    float *const depths = cv[lo].depths;
    const uint32_t x = (local % tiles_x) * WX + (lane % WX);
    const uint32_t y = (local / tiles_x) * WY + (lane / WX);
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;
    if (x < ws && y < hs) {
        float px[4], depth;
        render_sample<TR, LIGHTS>(s, rp, x, y, tr, px, depth);
        const uint32_t o = canvas_offset(rp, x, y);
        pixels[o] = make_float4(px[0], px[1], px[2], px[3]);
        depths[o] = depth;
        tr.status |= nan_status(rp, px, depth);
    }
}

// Multi-view trace: one-wave workgroups with a WX x WY tile (64 samples), as the frame kernel (render_kernel).
// PERSISTENT: a grid sized to the chip takes tiles from the work counter (grab at a time) until it passes the
// total; otherwise one workgroup per tile, tile i on workgroup i (workgroups go round-robin to the 8 XCDs, so
// each XCD gets an even share).
// The wide walks at three waves per SIMD, as the frame kernel (render_kernel).  LIGHTS: kLightsAll for scenes with
// lights or emissive quads, else kLightsNone (no light terms compiled in, no VGPR spills).
template <int WX, int WY, bool PERSISTENT, uint32_t NW, uint32_t LIGHTS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NW >= 4 ? 3 : 1)))
void views_kernel(SceneView s, const ViewRec *__restrict__ views, uint32_t nviews, uint32_t total_tiles, uint32_t grab,
                  uint32_t *__restrict__ counter, uint32_t scene_depth, uint32_t mesh_depth,
                  uint32_t *__restrict__ status) {
    static_assert(WX * WY == 64, "a wave tile holds 64 samples");
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t item;
    const uint32_t lane = threadIdx.x;
    TracerT<NW> tr;
    bind_lds_stacks<NW, 64>(tr, lds, lane, scene_depth, mesh_depth);
    tr.cnt = nullptr;
    if constexpr (PERSISTENT) {
        for (;;) {
            __syncthreads();   // every lane has read the previous item
            if (lane == 0) item = atomicAdd(counter, grab);
            __syncthreads();
            const uint32_t first = __builtin_amdgcn_readfirstlane(item);
            if (first >= total_tiles) break;
            const uint32_t last = min(first + grab, total_tiles);
            for (uint32_t tile = first; tile < last; tile++) trace_view_tile<WX, WY, LIGHTS>(s, views, nviews, tile, tr, lane);
        }
    } else {
        const uint32_t tile = blockIdx.x;
        if (tile >= total_tiles) return;
        trace_view_tile<WX, WY, LIGHTS>(s, views, nviews, tile, tr, lane);
    }
    if (tr.status) atomicOr(status, tr.status);
}


template <uint32_t LIGHTS>
hipError_t launch_views(const ViewsLaunch &l) {
#define VIEWS_LAUNCH1(WX, WY, P, W)                                                                           \
    hipLaunchKernelGGL((views_kernel<WX, WY, P, W, LIGHTS>), dim3(l.blocks), dim3(64), l.lds, l.stream, l.sv,   \
                       l.recs, l.nviews, l.total_tiles, l.grab, l.counter, l.scene_depth, l.mesh_depth, l.status)
#define VIEWS_LAUNCH(WX, WY)                                                                                  \
    if (l.persistent) {                                                                                       \
        if (l.width == 8) VIEWS_LAUNCH1(WX, WY, true, 8u);                                                    \
        else if (l.width == 4) VIEWS_LAUNCH1(WX, WY, true, 4u);                                               \
        else VIEWS_LAUNCH1(WX, WY, true, 2u);                                                                 \
    } else {                                                                                                  \
        if (l.width == 8) VIEWS_LAUNCH1(WX, WY, false, 8u);                                                   \
        else if (l.width == 4) VIEWS_LAUNCH1(WX, WY, false, 4u);                                              \
        else VIEWS_LAUNCH1(WX, WY, false, 2u);                                                                \
    }
    switch (l.tile_shape) {
        case 0: VIEWS_LAUNCH(8, 8); break;
        case 1: VIEWS_LAUNCH(16, 4); break;
        case 2: VIEWS_LAUNCH(32, 2); break;
        default: VIEWS_LAUNCH(4, 16); break;
    }
#undef VIEWS_LAUNCH
#undef VIEWS_LAUNCH1
    return hipGetLastError();
}

}  // namespace

#endif
