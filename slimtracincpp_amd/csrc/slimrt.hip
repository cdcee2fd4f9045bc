// slimrt: gfx950 kernels and the C ABI around them (include/slimrt.h).
//
// Device layout in HBM (one flat array per kind, concatenated over meshes / textures):
//   geometries 64 B | materials 64 B | lights 32 B | scene BVH nodes 32 B (+ leaf ids) | mesh records 48 B |
//   mesh BVH nodes 32 B | triangle hit records 64 B | triangle attribute records 64 B | texture and mip records |
//   texel quads 12 B.
// The render kernel is a per-sample megakernel: each workgroup is one wave owning a 32x2 strip of samples
// (one-wave workgroups finish the frame 6-9% sooner than 4-wave 16x16 ones, whose waves wait for the slowest of
// the four before the workgroup's slot frees; 32x2 strips beat 8x8 squares by 4% and 16x4 by 2% on the mesh +
// PBR frame: a wave's canvas writes are two 512-byte rows).  In image order tile i is workgroup i: workgroups
// are dealt round-robin to the 8 XCDs, so every XCD gets an even share of each image row band (band-per-XCD
// dealing left whole XCDs on the costly bands and ran 18-20% slower); frames bound by their longest tiles
// dispatch those first instead (launch_tile_order).  Traversal stacks live in LDS, one column per thread, sized
// from the uploaded BVH heights.
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <thread>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <cstdio>
#include <cmath>

#include "slimrt_context.h"
#include "../../include/slimrt_nrc.h"

using namespace slimrt;

static_assert(sizeof(slimrt_bvh_node) == 32, "node layout");
static_assert(sizeof(Node) == 32, "node layout");
static_assert(sizeof(slimrt_triangle) == 128, "triangle layout");
static_assert(sizeof(slimrt_geometry) == 56 && sizeof(slimrt_material) == 80 && sizeof(slimrt_light) == 32, "record layout");
static_assert(sizeof(slimrt_mesh) == 64 && sizeof(slimrt_texture) == 24 && sizeof(slimrt_texture_mip) == 16, "record layout");
static_assert(sizeof(RenderParams) == sizeof(slimrt_render_params), "render params layout");

namespace {


// Dispatch order of a per-sample frame, costliest tiles first.  A frame's duration is set by its slowest waves:
// the waves whose samples see the glass and mirror instances run 50-70x longer than the median wave, and in
// image order they start a fifth of the way into the frame, so the frame ends with most of the chip idle while
// they finish.  Dispatched first, they overlap the rest of the frame.  The cost of a tile is its duration two
// frames earlier (a scene changes little from one frame to the next): the order for frame f + 1 is computed while
// frame f renders, from the durations frame f - 1 recorded, by one wave of frame f's kernel (order_wave, its first
// workgroup), so a frame is one launch on one stream.
//
// A tile's cost class is its duration's binary exponent and the two bits below it (four classes an octave; rank
// 0 = costliest).  The costly tiles - those within a factor 64 of the costliest class (narrowed 4 classes at a
// time while more than kHeavyMax), none if even the costliest class alone has more - go first, by rank and within
// a rank in image order (a counting sort over the at most 25 ranks of the window); the others follow in image
// order.  The frames that allocate nothing compute the same order in three launches right before the frame (their
// scratch holds one set of durations, so the order cannot be computed while the next durations are written), over
// waves that each own a contiguous segment of tiles (so image order within a class is segment order):
// order_count_kernel ranks the tiles and counts them per class and wave, with each wave's costliest rank, longest
// and summed durations; order_plan_kernel (one workgroup) picks the order, the window and every wave's first slot
// per class and for its other tiles; order_place_kernel writes each tile to its slot (about 25 us in all for a
// 1080p frame, on the frame's path).  Deterministic: both forms give the same order.
constexpr uint32_t kCostClasses = 128;
constexpr uint32_t kWindow = 24;              // costly window: ranks top .. top + 24 (a factor 64 in duration)
constexpr uint32_t kHeavyMax = 8192;          // at most this many costly tiles
constexpr uint32_t kOrderBlock = 256, kOrderMaxWaves = 256;   // order workgroups (4 waves) and waves at most
// per-wave words of the order scratch: class counts, then [costliest rank, longest, summed lo, summed hi], then
// the plan: first slot per window rank and the first slot of the other tiles
constexpr uint32_t kOrderStats = kCostClasses, kOrderSlots = kCostClasses + 4, kOrderWaveWords = kCostClasses + 4 + kWindow + 2;
constexpr uint32_t kOrderScratchWords = 8 + kOrderMaxWaves * kOrderWaveWords;   // [0..8): top, window, flag
constexpr uint32_t kMaxOrderedTiles = 1u << 30;
constexpr uint32_t kSplitShare = 32, kSplitMax = 2048;   // costliest tiles rendered as two half waves: 1 in 32
// With 8-wide nodes the first 1 in 1024 tiles of the order (at most 128) are rendered as four quarter waves: the
// 8-wide walk leaves less work in the frame (its ideal length, summed tile durations over the resident waves, fell
// 464 -> 421 us on the 1080p mesh + PBR frame) but not in the frame's longest paths, so the longest waves must be
// narrower to approach it (0.565 -> 0.518 ms).  With 4-wide nodes quarter waves only add waves (0.553 -> 0.56 ms
// for the costliest 32), and eight or sixteen parts are slower with either width.
constexpr uint32_t kQuarterShare = 1024, kQuarterMax = 128, kTopParts = 4;
uint32_t quarter_tiles(uint32_t tiles, uint32_t width) {
    return width == 8 ? std::min(tiles / kQuarterShare, kQuarterMax) : 0u;
}

__device__ __forceinline__ uint32_t cost_rank(uint32_t ticks) {
    if (ticks < 4u) return kCostClasses - 1u;
    const uint32_t e = 31u - (uint32_t)__builtin_clz(ticks);                       // >= 2
    return kCostClasses - 2u - ((e - 2u) * 4u + ((ticks >> (e - 2u)) & 3u));      // 7 .. 126, costliest lowest
}

// Waves over n tiles and the 64-tile rounds of each wave's segment.
__host__ __device__ inline void order_shape(uint32_t n, uint32_t &waves, uint32_t &rounds) {
    waves = n < 64u ? 1u : (n + 1023u) / 1024u * 4u;   // a workgroup of 4 waves per 1024 tiles...
    if (waves > kOrderMaxWaves) waves = kOrderMaxWaves;  // ...at most 64 workgroups
    rounds = (n + waves * 64u - 1u) / (waves * 64u);
}

__global__ __launch_bounds__(kOrderBlock) void order_count_kernel(const uint32_t *__restrict__ cost, uint32_t n,
                                                                  uint32_t *__restrict__ scratch) {
    __shared__ uint32_t hist[kOrderBlock / 64][kCostClasses];
    uint32_t waves, rounds;
    order_shape(n, waves, rounds);
    const uint32_t w_local = threadIdx.x >> 6, lane = threadIdx.x & 63u, wave = blockIdx.x * (kOrderBlock / 64) + w_local;
    for (uint32_t k = lane; k < kCostClasses; k += 64) hist[w_local][k] = 0;
    __syncthreads();
    const bool active = wave < waves;
    uint32_t lo = kCostClasses, longest = 0;
    uint64_t sum = 0;
    const uint32_t seg = wave * rounds * 64u;
    for (uint32_t r = 0; r < rounds && active; r++) {
        const uint32_t i = seg + r * 64u + lane;
        if (i < n) {
            const uint32_t v = cost[i], rk = cost_rank(v);
            atomicAdd(&hist[w_local][rk], 1u);
            lo = min(lo, rk);
            longest = max(longest, v);
            sum += v;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, off, 64));
        longest = max(longest, (uint32_t)__shfl_xor((int)longest, off, 64));
        sum += (uint64_t)__shfl_xor((long long)sum, off, 64);
    }
    __syncthreads();   // the histogram's atomics are done
    if (!active) return;
    uint32_t *out = scratch + 8 + wave * kOrderWaveWords;
    for (uint32_t k = lane; k < kCostClasses; k += 64) out[k] = hist[w_local][k];
    if (lane == 0) {
        out[kOrderStats] = lo;
        out[kOrderStats + 1] = longest;
        out[kOrderStats + 2] = (uint32_t)sum;
        out[kOrderStats + 3] = (uint32_t)(sum >> 32);
    }
}

// Sum over the workgroup (kOrderBlock threads), broadcast; red holds kOrderBlock / 64 words.
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *red) {
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t s = 0;
    for (uint32_t w = 0; w < kOrderBlock / 64; w++) s += red[w];
    return s;
}

// One workgroup, a thread per wave of the count.  cost_tiles (may be null): the tile count the durations were
// recorded for; any other count gives the image order (the durations belong to another frame size).  Writes
// order[n], the longest-first flag, and the plan: every wave's first slot per window rank (ranks in turn, waves in
// turn within a rank) and for its other tiles (after every costly tile, waves in turn).
__global__ __launch_bounds__(kOrderBlock) void order_plan_kernel(uint32_t n, uint32_t resident,
                                                                 const uint32_t *__restrict__ cost_tiles,
                                                                 uint32_t *__restrict__ scratch,
                                                                 uint32_t *__restrict__ order) {
    __shared__ uint32_t red[kOrderBlock / 64], wave_tot[kOrderBlock / 64][kWindow + 2];
    __shared__ uint32_t top_s, longest_s;
    __shared__ unsigned long long summed_s;
    uint32_t waves, rounds;
    order_shape(n, waves, rounds);
    const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
    const bool mine = t < waves;   // order_shape keeps waves <= kOrderBlock
    const uint32_t *st = scratch + 8 + t * kOrderWaveWords;
    // the costliest rank, the longest and the summed durations
    uint32_t lo = mine ? st[kOrderStats] : kCostClasses, longest = mine ? st[kOrderStats + 1] : 0u;
    unsigned long long sum = mine ? ((unsigned long long)st[kOrderStats + 2] | ((unsigned long long)st[kOrderStats + 3] << 32)) : 0ull;
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, off, 64));
        longest = max(longest, (uint32_t)__shfl_xor((int)longest, off, 64));
        sum += (unsigned long long)__shfl_xor((long long)sum, off, 64);
    }
    if (t == 0) { top_s = kCostClasses; longest_s = 0; summed_s = 0; }
    __syncthreads();
    if (lane == 0) {
        atomicMin(&top_s, lo);
        atomicMax(&longest_s, longest);
        atomicAdd(&summed_s, sum);
    }
    __syncthreads();
    const uint32_t top = top_s;
    // Longest first only pays when the frame is bound by its longest tiles: when the longest tile lasts at least
    // 3/4 of the frame's ideal length (summed durations over the resident waves).  Otherwise (large frames whose
    // work fills the chip for longer than any one tile) the image order keeps neighbouring tiles together, which
    // keeps what they share in L2 (4K combined workload: 3.81 ms in image order, 4.14 ms in cost order).
    const bool image = (cost_tiles && *cost_tiles != n) || 4ull * (unsigned long long)longest_s * resident <= 3ull * summed_s;
    // this wave's tiles per window rank, the window's totals and width
    uint32_t c[kWindow + 1], total[kWindow + 1];
#pragma unroll
    for (uint32_t k = 0; k <= kWindow; k++) c[k] = (mine && top + k < kCostClasses) ? st[top + k] : 0u;
#pragma unroll
    for (uint32_t k = 0; k <= kWindow; k++) total[k] = block_sum(c[k], red);
    int window = -1;   // the widest of 16, 12, 8, 4, 0 ranks with <= kHeavyMax tiles (the same in every thread)
    uint32_t heavy = 0;
    for (int w = (int)kWindow; w >= 0; w -= 4) {
        uint32_t h = 0;
        for (int k = 0; k <= w; k++) h += total[k];
        if (h <= kHeavyMax) {
            window = w;
            heavy = h;
            break;
        }
    }
    if (t == 0) {
        scratch[0] = top;
        scratch[1] = (uint32_t)window;
        scratch[2] = image ? 0u : 1u;
        order[n] = image ? 0u : 1u;
    }
    if (image) return;   // uniform, after the last barrier
    // exclusive scans over the waves of the window ranks' counts and of the other tiles' counts
    const uint32_t first = t * rounds * 64u, count = (mine && first < n) ? min(n - first, rounds * 64u) : 0u;
    uint32_t v[kWindow + 2], hv = 0;
#pragma unroll
    for (uint32_t k = 0; k <= kWindow; k++) {
        v[k] = (int)k <= window ? c[k] : 0u;
        hv += v[k];
    }
    v[kWindow + 1] = count - hv;
#pragma unroll
    for (uint32_t k = 0; k <= kWindow + 1; k++) {   // inclusive within the wave
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t u = (uint32_t)__shfl_up((int)v[k], off, 64);
            if (lane >= off) v[k] += u;
        }
        if (lane == 63) wave_tot[wid][k] = v[k];
    }
    __syncthreads();
    if (!mine) return;
    uint32_t *slot = scratch + 8 + t * kOrderWaveWords + kOrderSlots;
    uint32_t base = 0;
#pragma unroll
    for (uint32_t k = 0; k <= kWindow + 1; k++) {
        uint32_t before = 0;
        for (uint32_t w = 0; w < wid; w++) before += wave_tot[w][k];
        const uint32_t own = (k <= kWindow) ? ((int)k <= window ? c[k] : 0u) : count - hv;
        const uint32_t excl = before + v[k] - own;
        if (k <= kWindow) {
            slot[k] = base + excl;
            base += (int)k <= window ? total[k] : 0u;
        } else {
            slot[k] = heavy + excl;
        }
    }
}

__global__ __launch_bounds__(kOrderBlock) void order_place_kernel(const uint32_t *__restrict__ cost, uint32_t n,
                                                                  const uint32_t *__restrict__ scratch,
                                                                  uint32_t *__restrict__ order) {
    uint32_t waves, rounds;
    order_shape(n, waves, rounds);
    const uint32_t lane = threadIdx.x & 63u, wave = blockIdx.x * (kOrderBlock / 64) + (threadIdx.x >> 6);
    if (wave >= waves) return;
    const uint32_t seg = wave * rounds * 64u;
    if (!scratch[2]) {   // image order
        for (uint32_t r = 0; r < rounds; r++) {
            const uint32_t i = seg + r * 64u + lane;
            if (i < n) order[i] = i;
        }
        return;
    }
    const uint32_t top = scratch[0];
    const int window = (int)scratch[1];
    const uint32_t *sl = scratch + 8 + wave * kOrderWaveWords + kOrderSlots;
    uint32_t slot[kWindow + 1];
#pragma unroll
    for (uint32_t k = 0; k <= kWindow; k++) slot[k] = sl[k];
    uint32_t lb = sl[kWindow + 1];
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t r = 0; r < rounds; r++) {
        const uint32_t i = seg + r * 64u + lane;
        const uint32_t d = i < n ? cost_rank(cost[i]) - top : 0xFFu;
        const bool is_heavy = (int)d <= window, is_light = i < n && !is_heavy;
        uint32_t pos = 0;
#pragma unroll
        for (uint32_t k = 0; k <= kWindow; k++) {
            if ((int)k <= window) {   // uniform
                const uint64_t m = __ballot(d == k);
                if (d == k) pos = slot[k] + (uint32_t)__popcll(m & below);
                slot[k] += (uint32_t)__popcll(m);
            }
        }
        if (is_heavy) order[pos] = i;
        const uint64_t ml = __ballot(is_light);
        if (is_light) order[lb + (uint32_t)__popcll(ml & below)] = i;
        lb += (uint32_t)__popcll(ml);
    }
}

// The order of launch_tile_order's three launches, computed by the first workgroups of later frames (one wave
// each) in two stages over segments of kOrderSegTiles tiles: a counting wave per segment (class counts by LDS
// atomics, the costliest rank, the longest and summed durations, kSegWords words per segment) in one frame, then a
// placing wave per segment in the next (the plan from every segment's counts, then the segment's tiles: those of
// each window rank by ballots, the others by one ballot).  Segments in image order keep image order within a class.
// The durations are loaded 64 x kOrderWaveBatch at a time (the waves wait on memory: one wave for all 32,400 tiles
// of a 1080p frame took ~0.3 ms, longer than a frame of primary hits).
constexpr uint32_t kOrderWaveBatch = 32, kOrderSegTiles = 4096, kOrderMaxSegs = 64;
constexpr uint32_t kSegLo = kCostClasses, kSegLongest = kCostClasses + 1, kSegSum = kCostClasses + 2;
constexpr uint32_t kSegWords = kCostClasses + 4;
// LDS of a placing wave: the window ranks' counts per segment, their totals and prefixes; then the staged ranks
constexpr size_t kOrderWaveLds = ((kWindow + 1) * kOrderMaxSegs + 2 * (kWindow + 1)) * sizeof(uint32_t);
static_assert(kOrderWaveLds >= 64 * kOrderWaveBatch && kOrderWaveLds >= kCostClasses * sizeof(uint32_t), "");

// Segments of the order: segs waves of seg tiles each (a multiple of 64; the last may be shorter).
__host__ __device__ inline void order_segments(uint32_t n, uint32_t &segs, uint32_t &seg) {
    segs = (n + kOrderSegTiles - 1) / kOrderSegTiles;
    segs = segs < 1u ? 1u : (segs > kOrderMaxSegs ? kOrderMaxSegs : segs);
    seg = ((n + segs - 1) / segs + 63u) / 64u * 64u;
    segs = (n + seg - 1) / seg;
}

// The work of a frame's first workgroups: count_waves counting waves over count_cost (into count_hist), then
// place_waves placing waves over place_cost (counted into place_hist one frame earlier) into place_order.
struct OrderWork {
    const uint32_t *count_cost;
    uint32_t *count_hist;
    const uint32_t *place_cost;
    const uint32_t *place_hist;
    uint32_t *place_order;
    uint32_t count_waves, place_waves, seg, resident;
};

__device__ __forceinline__ void order_count_wave(const uint32_t *__restrict__ cost, uint32_t n, uint32_t seg,
                                                 uint32_t w, uint32_t *__restrict__ out, uint32_t *lds) {
    const uint32_t lane = threadIdx.x, first = w * seg, last = min(n, first + seg);
    uint32_t *hist = lds;
    for (uint32_t k = lane; k < kCostClasses; k += 64) hist[k] = 0;
    __syncthreads();
    uint32_t lo = kCostClasses, longest = 0;
    uint64_t sum = 0;
    for (uint32_t base = first; base < last; base += 64u * kOrderWaveBatch) {
        uint32_t v[kOrderWaveBatch];
#pragma unroll
        for (uint32_t j = 0; j < kOrderWaveBatch; j++) {
            const uint32_t i = base + j * 64u + lane;
            v[j] = i < last ? cost[i] : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < kOrderWaveBatch; j++) {
            if (base + j * 64u + lane < last) {
                const uint32_t rk = cost_rank(v[j]);
                atomicAdd(&hist[rk], 1u);
                lo = min(lo, rk);
                longest = max(longest, v[j]);
                sum += v[j];
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, off, 64));
        longest = max(longest, (uint32_t)__shfl_xor((int)longest, off, 64));
        sum += (uint64_t)__shfl_xor((long long)sum, off, 64);
    }
    __syncthreads();   // the class counts are complete
    for (uint32_t k = lane; k < kCostClasses; k += 64) out[k] = hist[k];
    if (lane == 0) {
        out[kSegLo] = lo;
        out[kSegLongest] = longest;
        out[kSegSum] = (uint32_t)sum;
        out[kSegSum + 1] = (uint32_t)(sum >> 32);
    }
}

__device__ __forceinline__ void order_place_wave(const uint32_t *__restrict__ cost, uint32_t n, uint32_t seg,
                                                 uint32_t segs, uint32_t w, uint32_t resident,
                                                 const uint32_t *__restrict__ hist, uint32_t *__restrict__ order,
                                                 uint32_t *lds) {
    const uint32_t lane = threadIdx.x, first = w * seg, last = min(n, first + seg);
    // the frame's costliest rank, longest and summed durations, from every segment's (lane j: segment j)
    const uint32_t *hj = hist + (size_t)lane * kSegWords;
    const bool mine = lane < segs;
    uint32_t top = mine ? hj[kSegLo] : kCostClasses, longest = mine ? hj[kSegLongest] : 0u;
    uint64_t sum = mine ? ((uint64_t)hj[kSegSum] | ((uint64_t)hj[kSegSum + 1] << 32)) : 0ull;
    for (int off = 32; off > 0; off >>= 1) {
        top = min(top, (uint32_t)__shfl_xor((int)top, off, 64));
        longest = max(longest, (uint32_t)__shfl_xor((int)longest, off, 64));
        sum += (uint64_t)__shfl_xor((long long)sum, off, 64);
    }
    // as order_plan_kernel: longest first only when the longest tile lasts at least 3/4 of the frame's ideal length
    if (4ull * (unsigned long long)longest * resident <= 3ull * sum) {
        for (uint32_t i = first + lane; i < last; i += 64u) order[i] = i;
        if (w == 0 && lane == 0) order[n] = 0u;
        return;
    }
    // the window ranks' counts per segment (lds[k * 64 + j]); lane k sums rank k over all segments and over those
    // before this one
    uint32_t *grid = lds, *tot = lds + (kWindow + 1) * kOrderMaxSegs, *pre = tot + (kWindow + 1);
#pragma unroll
    for (uint32_t k = 0; k <= kWindow; k++) grid[k * 64u + lane] = (mine && top + k < kCostClasses) ? hj[top + k] : 0u;
    __syncthreads();
    if (lane <= kWindow) {
        uint32_t t = 0, b = 0;
        for (uint32_t j = 0; j < segs; j++) {
            const uint32_t c = grid[lane * 64u + j];
            if (j < w) b += c;
            t += c;
        }
        tot[lane] = t;
        pre[lane] = b;
    }
    __syncthreads();
    uint32_t total[kWindow + 1], before[kWindow + 1];
#pragma unroll
    for (uint32_t k = 0; k <= kWindow; k++) {
        total[k] = tot[k];
        before[k] = pre[k];
    }
    int window = -1;   // the widest of 16, 12, 8, 4, 0 ranks with <= kHeavyMax tiles
    uint32_t heavy = 0;
    for (int wd = (int)kWindow; wd >= 0; wd -= 4) {
        uint32_t h = 0;
        for (int k = 0; k <= wd; k++) h += total[k];
        if (h <= kHeavyMax) {
            window = wd;
            heavy = h;
            break;
        }
    }
    // first slots of this segment: per window rank after the costlier ranks and this rank's tiles in earlier
    // segments; the other tiles after every window tile and the other tiles of earlier segments (full segments)
    uint32_t slot[kWindow + 1];
    uint32_t acc = 0, heavy_before = 0;
#pragma unroll
    for (uint32_t k = 0; k <= kWindow; k++) {
        const bool in = (int)k <= window;
        slot[k] = acc + (in ? before[k] : 0u);
        acc += in ? total[k] : 0u;
        heavy_before += in ? before[k] : 0u;
    }
    uint32_t lb = heavy + first - heavy_before;
    const uint64_t below = (1ull << lane) - 1ull;
    uint8_t *stage = reinterpret_cast<uint8_t *>(lds);
    for (uint32_t base = first; base < last; base += 64u * kOrderWaveBatch) {
        uint32_t v[kOrderWaveBatch];
#pragma unroll
        for (uint32_t j = 0; j < kOrderWaveBatch; j++) {
            const uint32_t i = base + j * 64u + lane;
            v[j] = i < last ? cost[i] : 0u;
        }
        __syncthreads();   // the previous batch's ranks (or the plan's counts) have been read
        // window rank, 0xFE for the other tiles, 0xFF past the segment
#pragma unroll
        for (uint32_t j = 0; j < kOrderWaveBatch; j++) {
            const uint32_t i = base + j * 64u + lane, r = cost_rank(v[j]) - top;
            stage[j * 64u + lane] = (uint8_t)(i >= last ? 0xFFu : (window >= 0 && r <= (uint32_t)window ? r : 0xFEu));
        }
        __syncthreads();
        const uint32_t rounds = min(kOrderWaveBatch, (last - base + 63u) / 64u);
        for (uint32_t j = 0; j < rounds; j++) {
            const uint32_t i = base + j * 64u + lane, d = stage[j * 64u + lane];
            const bool is_heavy = d < 0xFEu, is_light = d == 0xFEu;
            if (__ballot(is_heavy)) {   // uniform; ballots only (a shuffle waits on LDS, and they would chain)
                uint32_t pos = 0;
#pragma unroll
                for (uint32_t k = 0; k <= kWindow; k++) {
                    const uint64_t m = (int)k <= window ? __ballot(d == k) : 0ull;
                    if (m) {   // uniform
                        if (d == k) pos = slot[k] + (uint32_t)__popcll(m & below);
                        slot[k] += (uint32_t)__popcll(m);
                    }
                }
                if (is_heavy && pos < n) order[pos] = i;   // always below n: the counts are of these durations
            }
            const uint64_t ml = __ballot(is_light);
            const uint32_t q = lb + (uint32_t)__popcll(ml & below);
            if (is_light && q < n) order[q] = i;
            lb += (uint32_t)__popcll(ml);
        }
    }
    if (w == 0 && lane == 0) order[n] = 1u;
}

// The three launches of a frame's order on one stream.
void launch_tile_order(hipStream_t stream, const uint32_t *cost, uint32_t *order, uint32_t n, uint32_t resident,
                       const uint32_t *cost_tiles, uint32_t *order_scratch) {
    uint32_t waves, rounds;
    order_shape(n, waves, rounds);
    const uint32_t blocks = (waves + kOrderBlock / 64 - 1) / (kOrderBlock / 64);
    hipLaunchKernelGGL(order_count_kernel, dim3(blocks), dim3(kOrderBlock), 0, stream, cost, n, order_scratch);
    hipLaunchKernelGGL(order_plan_kernel, dim3(1), dim3(kOrderBlock), 0, stream, n, resident, cost_tiles,
                       order_scratch, order);
    hipLaunchKernelGGL(order_place_kernel, dim3(blocks), dim3(kOrderBlock), 0, stream, cost, n,
                       (const uint32_t *)order_scratch, order);
}

// The 4- and 8-wide frames at three waves per SIMD (168 VGPRs; only the instantiations with every light term
// spill, 32 and 35 VGPRs): the walk's dependent fetches want the waves more than the registers (6% faster than two
// on the mesh + PBR frame).  The 2-wide frame would spill 64 dwords a
// lane for no gain, and the counting pass keeps its registers: both stay at the compiler's two.  With an order
// (order_place_wave or launch_tile_order) workgroup k renders tile order[k] and records the tile's duration in
// cost (the first ow.count_waves + ow.place_waves workgroups work on later frames' orders, and the tiles' workgroups
// follow them); when order[tiles]
// is set, the first `split` tiles of the order (the costliest) are each rendered by two workgroups, lanes 0-31
// and lanes 32-63: a costly tile's 64 paths diverge, and half as many lanes per wave shorten the frame's longest
// waves (0.77 -> 0.65 ms on the mesh + PBR frame); the first `quarter` of them (8-wide nodes only, quarter_tiles)
// by top_parts workgroups of 64 / top_parts lanes each.
// LIGHTS selects the light terms (kLightsNone for scenes without lights or emissive quads: no VGPR spills at three
// waves per SIMD; kLightsQuads for emissive quads without point lights: no shadow-ray walks).
template <bool kCount, uint32_t NW, uint32_t LIGHTS = kLightsAll>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NW >= 4 && !kCount ? 3 : 1)))
void render_kernel(SceneView s, RenderParams rp, float4 *__restrict__ pixels, float *__restrict__ depths,
                   uint32_t scene_depth, uint32_t mesh_depth, uint32_t tiles_x, uint32_t tiles,
                   uint32_t *__restrict__ status, unsigned long long *__restrict__ counters,
                   const uint32_t *__restrict__ order, uint32_t *__restrict__ cost, uint32_t split,
                   uint32_t quarter, uint32_t top_parts, uint32_t *__restrict__ cost_tiles,
                   OrderWork ow) {
    extern __shared__ uint32_t lds[];
    uint32_t b = blockIdx.x;
    // the first workgroups count and place for later frames' orders; the tiles follow
    if (b < ow.count_waves) {
        order_count_wave(ow.count_cost, tiles, ow.seg, b, ow.count_hist + (size_t)b * kSegWords, lds);
        return;
    }
    b -= ow.count_waves;
    if (b < ow.place_waves) {
        order_place_wave(ow.place_cost, tiles, ow.seg, ow.place_waves, b, ow.resident, ow.place_hist, ow.place_order, lds);
        return;
    }
    b -= ow.place_waves;
    if (cost_tiles && b == 0 && threadIdx.x == 0) *cost_tiles = tiles;   // what cost will describe
    uint32_t tile = b, parts = 1, part = 0;   // the wave renders lanes [part, part + 1) * 64 / parts
    if (order) {
        // order[tiles] = 0: the order computation chose image order, no split
        const uint32_t sp = order[tiles] ? split : 0u, q = order[tiles] ? min(quarter, split) : 0u;
        const uint32_t P = top_parts;
        if (b >= tiles + sp + (P - 2) * q) return;
        if (b < P * q) {   // the first q tiles of the order as P part waves
            tile = order[b / P];
            parts = P;
            part = b % P;
        } else if (b < (P - 2) * q + 2 * sp) {   // the next sp - q as two half waves
            tile = order[q + ((b - P * q) >> 1)];
            parts = 2;
            part = (b - P * q) & 1u;
        } else tile = order[b - (P - 2) * q - sp];
    }
    if (tile >= tiles) return;
    const uint64_t start = cost ? (uint64_t)wall_clock64() : 0u;
    const uint32_t lane = threadIdx.x;
    const uint32_t x = (tile % tiles_x) * kWaveTileX + (lane % kWaveTileX);
    const uint32_t y = (tile / tiles_x) * kWaveTileY + (lane / kWaveTileX);
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;

    Counters cnt{};
    TracerT<NW> tr;
    bind_lds_stacks<NW, 64>(tr, lds, lane, scene_depth, mesh_depth);
    tr.cnt = kCount ? &cnt : nullptr;
    if (lane * parts / 64u == part && x < ws && y < hs) {
        float px[4], depth;
        render_sample<TracerT<NW>, LIGHTS>(s, rp, x, y, tr, px, depth);
        const uint32_t o = canvas_offset(rp, x, y);
        pixels[o] = make_float4(px[0], px[1], px[2], px[3]);
        depths[o] = depth;
        tr.status |= nan_status(rp, px, depth);
        if (tr.status) atomicOr(status, tr.status);
    }
    if (cost && lane == 0) {
        // A split tile records its part's duration times the parts (the part that ends last writes last, so a
        // costly half is not hidden by a cheap one): a part runs about its share of the tile's diverging paths, and
        // the scaled figure keeps the tile in the costliest classes two frames later (its own part's duration
        // would rank it lower, so it would run unsplit and late, then split again).
        const uint64_t d = ((uint64_t)wall_clock64() - start) * parts;
        cost[tile] = d > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)d;
    }
    if (kCount) {
        const uint64_t *c = &cnt.rays;
        for (int i = 0; i < 8; i++) {
            uint64_t v = c[i];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0 && v) atomicAdd(counters + i, (unsigned long long)v);
        }
    }
}

// ---------------------------------------------------------------- rendering with the neural radiance cache
// The frame kernel has the render kernel's tiling; one in train_one_in waves (chosen by a hash of the wave's
// position and the frame) traces full paths for training, so training and cached pixels never share a wave.
template <uint32_t NW>
__global__ __launch_bounds__(64) void nrc_frame_kernel(SceneView s, RenderParams rp, NrcFrame nf,
                                                       float4 *__restrict__ pixels, float *__restrict__ depths,
                                                       uint32_t scene_depth, uint32_t mesh_depth, uint32_t tiles_x,
                                                       uint32_t tiles, uint32_t *__restrict__ status) {
    extern __shared__ uint32_t lds[];
    const uint32_t tile = blockIdx.x;
    if (tile >= tiles) return;
    const uint32_t t = threadIdx.x, lane = t;
    const uint32_t x = (tile % tiles_x) * kWaveTileX + (lane % kWaveTileX);
    const uint32_t y = (tile / tiles_x) * kWaveTileY + (lane / kWaveTileX);
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;
    uint32_t h = tile * 0x9E3779B1u + nf.frame * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    const bool training = nf.train_one_in && (h % nf.train_one_in) == 0;

    TracerT<NW> tr;
    bind_lds_stacks<NW, 64>(tr, lds, t, scene_depth, mesh_depth);
    tr.cnt = nullptr;
    if (x < ws && y < hs) {
        v3 dir;
        float scaling, depth;
        primary_ray(rp.proj, x, y, dir, scaling);
        const uint32_t o = canvas_offset(rp, x, y);
        color col;
        if (render_beauty_nrc(s, rp, dir, scaling, tr, nf, training, o, col, depth)) {
            float px[4];
            output_value(rp, color{tone_map(col.r), tone_map(col.g), tone_map(col.b)}, px);
            pixels[o] = make_float4(px[0], px[1], px[2], px[3]);
            tr.status |= nan_status(rp, px, depth);
        }
        depths[o] = depth;
        if (depth != depth) tr.status |= kStatusNaN;
        if (tr.status) atomicOr(status, tr.status);
        if (training) (void)wave_append(true, nf.counters + NC_TRAIN_PIXELS);
    }
}

// Pixel = colour before the cache vertex + throughput * max(cache, 0), tone-mapped (ray_tracer.h:107).
__global__ void nrc_composite_kernel(RenderParams rp, const uint32_t *__restrict__ q_count, uint32_t q_cap,
                                     const uint32_t *__restrict__ q_pixel, const float *__restrict__ q_prefix,
                                     const float *__restrict__ q_tp, const float *__restrict__ radiance,
                                     float4 *__restrict__ pixels, uint32_t *__restrict__ status) {
    const uint32_t n = min(*q_count, q_cap);
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const color L{fmaxf(radiance[3 * q], 0.0f), fmaxf(radiance[3 * q + 1], 0.0f), fmaxf(radiance[3 * q + 2], 0.0f)};
        const color tp{q_tp[3 * q], q_tp[3 * q + 1], q_tp[3 * q + 2]};
        const color col = cmul_add(L, tp, color{q_prefix[3 * q], q_prefix[3 * q + 1], q_prefix[3 * q + 2]});
        float px[4];
        output_value(rp, color{tone_map(col.r), tone_map(col.g), tone_map(col.b)}, px);
        pixels[q_pixel[q]] = make_float4(px[0], px[1], px[2], px[3]);
        if (nan_status(rp, px, 0.0f)) atomicOr(status, kStatusNaN);
    }
}

// Canvas::drawToWindow: one thread per window pixel (its 2x2 sample quad is 64 contiguous bytes under SSAA).
__global__ __launch_bounds__(kBlock) void content_kernel(const float4 *__restrict__ pixels, uint32_t n, uint32_t ssaa,
                                                         uint32_t *__restrict__ content) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) content[i] = pixel_content(pixels + (size_t)i * (ssaa ? 4 : 1), ssaa != 0);
}

// One pick on the device: thread 0 traces, with the render kernels' LDS stacks.
__global__ __launch_bounds__(kBlock) void pick_kernel(SceneView s, RenderParams rp, int32_t x, int32_t y,
                                                      uint32_t scene_depth, uint32_t mesh_depth,
                                                      int32_t *__restrict__ out, uint32_t *__restrict__ status) {
    extern __shared__ uint32_t lds[];
    const uint32_t t = threadIdx.x;
    if (t != 0) return;
    TracerT<2> tr;
    bind_lds_stacks(tr, lds, t, scene_depth, mesh_depth);
    tr.cnt = nullptr;
    int32_t oi[3];
    float of[4];
    pick(s, rp, x, y, tr, oi, of);
    for (int k = 0; k < 3; k++) out[k] = oi[k];
    for (int k = 0; k < 4; k++) out[3 + k] = __builtin_bit_cast(int32_t, of[k]);
    if (tr.status) atomicOr(status, tr.status);
}

// Clears the frame counters and step losses (a kernel rather than a memset, so the frame can be captured).
__global__ void nrc_reset_kernel(uint32_t *__restrict__ counters, float *__restrict__ losses, uint32_t nloss) {
    const uint32_t i = threadIdx.x;
    if (i < NC_COUNT) counters[i] = 0u;
    if (i < nloss) losses[i] = 0.0f;
}

__global__ void nrc_gate_clear_kernel(float *__restrict__ cells, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) cells[i] = 0.0f;
}

// Folds the frame's per-cell sums into the moving averages (weight 1/16) and clears the sums.
__global__ void nrc_gate_fold_kernel(float *__restrict__ cells, uint32_t ncells) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ncells; i += gridDim.x * blockDim.x) {
        float *c = cells + (size_t)i * kGateFloats;
        const float cnt = c[2];
        if (cnt > 0.0f) {
            const float m1 = c[0] / cnt, m2 = c[1] / cnt;
            if (c[5] == 0.0f) {
                c[3] = m1;
                c[4] = m2;
            } else {
                c[3] += (m1 - c[3]) * (1.0f / 16.0f);
                c[4] += (m2 - c[4]) * (1.0f / 16.0f);
            }
            c[5] = fminf(c[5] + 1.0f, 1.0e6f);
            c[0] = 0.0f;
            c[1] = 0.0f;
            c[2] = 0.0f;
        }
    }
}

}  // namespace

namespace {

// ---------------------------------------------------------------- scratch of the frames that allocate nothing
// slimrt_render_device_noalloc keeps the tile durations its frames record and the dispatch order it computes from
// them in one file-scope device array, for frames of up to kNoAllocTiles tiles (32 x 2 samples each).  Words:
// [0] the tile count the recorded durations belong to, [1, 64) padding, then the durations, then the order
// (kNoAllocTiles + 1 words: the order and the longest-first flag), then the order computation's per-wave counts
// and plan.  Zero-initialised, so a device's first such
// frame takes the image order.  Every GPU holds its own copy until the process ends, shared by every context on
// that GPU.  In host code the array's name is not a device address: the address is looked up once per device
// (hipGetSymbolAddress, with that device current) and only that address reaches kernels.
constexpr uint32_t kNoAllocTiles = 131072;   // 8,388,608 samples: 3840 x 2160, or 1920 x 1080 with SSAA
constexpr uint32_t kNoAllocCost = 64, kNoAllocOrder = kNoAllocCost + kNoAllocTiles;
constexpr uint32_t kNoAllocOrderScratch = (kNoAllocOrder + kNoAllocTiles + 1 + 63) / 64 * 64;
constexpr uint32_t kNoAllocWords = kNoAllocOrderScratch + kOrderScratchWords;   // 1,126,656 bytes
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint32_t *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be the current device.
int noalloc_scratch_address(int device, uint32_t **out) {
    if (device < 0 || device >= kMaxDevices) return SLIMRT_E_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return SLIMRT_E_DEVICE;
        g_scratch[device] = static_cast<uint32_t *>(ptr);
    }
    *out = g_scratch[device];
    return SLIMRT_OK;
}

// Buffers and the event of cost-ordered frames (created on first use, grown with the tile count).
int prepare_tile_order(slimrt_context *c, uint32_t tiles) {
    if (!c->tile_rendered) HIP_TRY(c, hipEventCreateWithFlags(&c->tile_rendered, hipEventDisableTiming));
    if (c->tile_cap >= tiles) return SLIMRT_OK;
    if (c->tile_frame) HIP_TRY(c, hipEventSynchronize(c->tile_rendered));   // the last such frame is done
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int p = 0; p < 3; p++) {
        if (c->tile_cost[p]) HIP_TRY(c, ctx_free(c, c->tile_cost[p]));
        c->tile_cost[p] = nullptr;
        c->cost_tiles[p] = 0;
    }
    for (int p = 0; p < 2; p++) {
        for (uint32_t **b : {&c->tile_order[p], &c->tile_hist[p]}) {
            if (*b) HIP_TRY(c, ctx_free(c, *b));
            *b = nullptr;
        }
        c->hist_tiles[p] = c->order_tiles[p] = 0;
        c->order_used[p] = false;
    }
    c->tile_cap = 0;
    for (int p = 0; p < 3; p++) HIP_TRY(c, ctx_malloc(c, &c->tile_cost[p], tiles * sizeof(uint32_t)));
    for (int p = 0; p < 2; p++) {
        HIP_TRY(c, ctx_malloc(c, &c->tile_order[p], (tiles + 1) * sizeof(uint32_t)));   // + the longest-first flag
        HIP_TRY(c, ctx_malloc(c, &c->tile_hist[p], kOrderMaxSegs * kSegWords * sizeof(uint32_t)));
    }
    c->tile_cap = tiles;
    return SLIMRT_OK;
}

// Waves of the frame kernel the device holds at once (its workgroups are single waves), for this node width and
// stack size; cached per configuration.
int resident_waves(slimrt_context *c, uint32_t width, uint32_t lights, size_t lds, uint32_t &out) {
    const uint64_t key = ((uint64_t)lds << 6) | ((uint64_t)lights << 4) | width;
    if (c->resident_key != key || !c->resident) {
        int per_cu = 0, cus = 0;
        auto occupancy = [&](auto kernel) { return hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, lds); };
#define OCCUPANCY(NW)                                                                                           \
    HIP_TRY(c, lights == kLightsAll ? occupancy(render_kernel<false, NW, kLightsAll>)                              \
               : lights == kLightsQuads ? occupancy(render_kernel<false, NW, kLightsQuads>)                        \
                                        : occupancy(render_kernel<false, NW, kLightsNone>))
        if (width == 8) OCCUPANCY(8);
        else if (width == 4) OCCUPANCY(4);
        else OCCUPANCY(2);
#undef OCCUPANCY
        HIP_TRY(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        c->resident = (uint32_t)std::max(1, per_cu * cus);
        c->resident_key = key;
    }
    out = c->resident;
    return SLIMRT_OK;
}

int launch(slimrt_context *c, const slimrt_render_params *p, float4 *pixels, float *depths, bool count) {
    RenderParams rp;
    std::memcpy(&rp, p, sizeof rp);
    rp.ssaa = rp.ssaa ? 1u : 0u;
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;
    const uint32_t width = mesh_node_width(rp);
    const uint32_t scene_depth = c->scene_height + 1, mesh_depth = mesh_stack_depth(c, width);
    // sized for the 256-thread workgroups of the wavefront kernels (the frame kernel needs a quarter)
    const size_t lds = stack_lds_bytes(scene_depth, mesh_depth, width);
    if (lds > 160 * 1024) return fail(c, SLIMRT_E_ARG, "BVH too deep for the LDS traversal stacks");
    if (check_width(c, width)) return SLIMRT_E_ARG;
    SceneView v = device_view(c);
    if (c->timed) HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    if (!count && (rp.flags & SLIMRT_WAVEFRONT) && rp.render_mode == RM_Beauty && rp.output != SLIMRT_OUTPUT_HITS) {
        if (width == 8) return fail(c, SLIMRT_E_ARG, "the wavefront frame forms take node width 2 or 4, not SLIMRT_BVH8");
        int rc = launch_wavefront(c, rp, pixels, depths, (rp.flags & SLIMRT_WAVEFRONT_PER_BOUNCE) != 0,
                                  (rp.flags & SLIMRT_WAVEFRONT_SORT) != 0, 1u << ((rp.flags >> 4) & 3u));
        if (rc) return rc;
    } else {
        const uint32_t wtiles_x = (ws + kWaveTileX - 1) / kWaveTileX;
        const uint32_t wtiles = wtiles_x * ((hs + kWaveTileY - 1) / kWaveTileY);
        const size_t wlds = stack_lds_bytes(scene_depth, mesh_depth, width, 64);
        const bool by_cost = !count && !(rp.flags & SLIMRT_TILE_ORDER_FIXED) && wtiles < kMaxOrderedTiles;
        // Frame g renders in order tile_order[g % 2] and writes its durations to tile_cost[g % 3]; its first
        // workgroups count frame g - 1's durations into tile_hist[(g + 1) % 2] and place frame g - 2's (counted by
        // frame g - 1) into tile_order[(g + 1) % 2], the order of frame g + 1.
        const uint64_t g = c->tile_frame;
        const uint32_t cs = (uint32_t)(g % 3), op = (uint32_t)(g & 1u), on = op ^ 1u;
        const uint32_t c_prev = (uint32_t)((g + 2) % 3), c_prev2 = (uint32_t)((g + 1) % 3);
        uint32_t *order = nullptr, *cost = nullptr;
        OrderWork ow{};
        size_t klds = wlds;
        if (by_cost) {
            if (int rc = prepare_tile_order(c, wtiles)) return rc;
            // the last cost-ordered frame ran on another stream: this frame reads what it wrote
            if (g && c->order_stream != c->stream) HIP_TRY(c, hipStreamWaitEvent(c->stream, c->tile_rendered, 0));
            c->order_used[op] = c->order_tiles[op] == wtiles;
            order = c->order_used[op] ? c->tile_order[op] : nullptr;
            cost = c->tile_cost[cs];
            uint32_t segs = 0, seg = 0;
            order_segments(wtiles, segs, seg);
            ow.seg = seg;
            c->order_tiles[on] = 0;
            if (c->hist_tiles[op] == wtiles && c->cost_tiles[c_prev2] == wtiles) {   // frame g - 2, counted
                if (int rc = resident_waves(c, width, scene_lights(v), wlds, ow.resident)) return rc;
                ow.place_cost = c->tile_cost[c_prev2];
                ow.place_hist = c->tile_hist[op];
                ow.place_order = c->tile_order[on];
                ow.place_waves = segs;
                c->order_tiles[on] = wtiles;
            }
            c->hist_tiles[on] = 0;
            if (c->cost_tiles[c_prev] == wtiles) {   // frame g - 1
                ow.count_cost = c->tile_cost[c_prev];
                ow.count_hist = c->tile_hist[on];
                ow.count_waves = segs;
                c->hist_tiles[on] = wtiles;
            }
            if (ow.count_waves || ow.place_waves) klds = std::max(wlds, kOrderWaveLds);
        }
        const uint32_t split = order ? std::min(wtiles / kSplitShare, kSplitMax) : 0u;
        const uint32_t quarter = std::min(quarter_tiles(wtiles, width), split), tp = kTopParts;
        const uint32_t blocks = ow.count_waves + ow.place_waves + wtiles + split + (tp - 2) * quarter;
#define RENDER_LAUNCH(COUNT, WIDE, LIGHTS)                                                                      \
    hipLaunchKernelGGL((render_kernel<COUNT, WIDE, LIGHTS>), dim3(blocks), dim3(64), klds, c->stream, v, rp, pixels,    \
                       depths, scene_depth, mesh_depth, wtiles_x, wtiles, c->d_status, c->d_counters, order, cost,  \
                       split, quarter, tp, nullptr, ow)
        const uint32_t lights = scene_lights(v);
        if (count) {
            if (width == 8) RENDER_LAUNCH(true, 8, kLightsAll); else if (width == 4) RENDER_LAUNCH(true, 4, kLightsAll);
            else RENDER_LAUNCH(true, 2, kLightsAll);
        } else if (lights == kLightsAll) {
            if (width == 8) RENDER_LAUNCH(false, 8, kLightsAll); else if (width == 4) RENDER_LAUNCH(false, 4, kLightsAll);
            else RENDER_LAUNCH(false, 2, kLightsAll);
        } else if (lights == kLightsQuads) {
            if (width == 8) RENDER_LAUNCH(false, 8, kLightsQuads); else if (width == 4) RENDER_LAUNCH(false, 4, kLightsQuads);
            else RENDER_LAUNCH(false, 2, kLightsQuads);
        } else {
            if (width == 8) RENDER_LAUNCH(false, 8, kLightsNone); else if (width == 4) RENDER_LAUNCH(false, 4, kLightsNone);
            else RENDER_LAUNCH(false, 2, kLightsNone);
        }
#undef RENDER_LAUNCH
        HIP_TRY(c, hipGetLastError());
        if (by_cost) {
            HIP_TRY(c, hipEventRecord(c->tile_rendered, c->stream));
            c->order_stream = c->stream;
            c->cost_tiles[cs] = wtiles;
            c->tile_frame++;
        }
    }
    HIP_TRY(c, hipGetLastError());
    if (c->timed) HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    return SLIMRT_OK;
}

template <uint32_t NW>
void cpu_rows(const SceneView &v, const RenderParams &rp, unsigned ti, unsigned nthreads, uint32_t sd, uint32_t md,
              Counters *cnt, slimrt_pixel *pixels, float *depths, uint32_t &status);

int cpu_render(slimrt_context *c, const slimrt_render_params *p, slimrt_pixel *pixels, float *depths, uint64_t *counters) {
    RenderParams rp;
    std::memcpy(&rp, p, sizeof rp);
    rp.ssaa = rp.ssaa ? 1u : 0u;
    const SceneView v = host_view(c);
    const uint32_t width = mesh_node_width(rp);   // the sibling-pair walk unless a wide one is asked for
    if (check_width(c, width)) return SLIMRT_E_ARG;
    const uint32_t sd = c->scene_height + 1, md = mesh_stack_depth(c, width);
    unsigned nthreads = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    std::vector<Counters> per(nthreads);
    std::vector<uint32_t> overflow(nthreads, 0);
    auto work = [&](unsigned ti) {
        if (width == 8) cpu_rows<8>(v, rp, ti, nthreads, sd, md, counters ? &per[ti] : nullptr, pixels, depths, overflow[ti]);
        else if (width == 4) cpu_rows<4>(v, rp, ti, nthreads, sd, md, counters ? &per[ti] : nullptr, pixels, depths, overflow[ti]);
        else cpu_rows<2>(v, rp, ti, nthreads, sd, md, counters ? &per[ti] : nullptr, pixels, depths, overflow[ti]);
    };
    std::vector<std::thread> pool;
    for (unsigned i = 0; i < nthreads; i++) pool.emplace_back(work, i);
    for (auto &th : pool) th.join();
    if (counters) {
        std::memset(counters, 0, 8 * sizeof(uint64_t));
        for (auto &k : per) {
            const uint64_t *s = &k.rays;
            for (int i = 0; i < 8; i++) counters[i] += s[i];
        }
    }
    for (uint32_t o : overflow) {
        c->host_status |= o & kStatusNaN;
        if (o & kStatusStack) return fail(c, SLIMRT_E_OVERFLOW, "traversal stack overflow");
    }
    return SLIMRT_OK;
}

// Rows ti, ti + nthreads, ... of a CPU-context frame with mesh node width NW.
template <uint32_t NW>
void cpu_rows(const SceneView &v, const RenderParams &rp, unsigned ti, unsigned nthreads, uint32_t sd, uint32_t md,
              Counters *cnt, slimrt_pixel *pixels, float *depths, uint32_t &status) {
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;
    using Entry = typename TracerT<NW>::MeshEntry;
    std::vector<uint32_t> scene_stack(sd);
    std::vector<Entry> mesh_stack(md);
    TracerT<NW> tr;
    tr.scene_stack = Stack{scene_stack.data(), 1};
    tr.mesh_stack = StackT<Entry>{mesh_stack.data(), 1};
    tr.scene_depth = sd;
    tr.mesh_depth = md;
    tr.status = 0;
    if (cnt) *cnt = Counters{};
    tr.cnt = cnt;
    for (uint32_t y = ti; y < hs; y += nthreads)
        for (uint32_t x = 0; x < ws; x++) {
            float px[4], depth;
            render_sample(v, rp, x, y, tr, px, depth);
            if (pixels) {
                const uint32_t o = canvas_offset(rp, x, y);
                pixels[o] = slimrt_pixel{px[0], px[1], px[2], px[3]};
                depths[o] = depth;
                tr.status |= nan_status(rp, px, depth);
            }
        }
    status = tr.status;
}

// Collapses one mesh's sibling-pair BVH (nodes, mesh-relative indices) into 4-wide nodes appended to out: a
// node's children are its two children with, while fewer than four, the inner child of largest surface area
// replaced by its own two children.  Leaves keep their triangle ranges.  Returns the node index; height4 receives
// the collapsed tree's height, stack the most mesh stack entries the 4-wide walk can hold below this node: a
// visit pushes its k inner children and pops one, so stack = max(k, k - 1 + the largest stack of its inner
// children).
uint32_t collapse4(const Node *n2, uint32_t i2, uint32_t depth, std::vector<Node4> &out, uint32_t &height4,
                   uint32_t &stack) {
    auto is_leaf = [&](uint32_t k) { return (n2[k].leaf_count_depth & 0xFFFFu) != 0; };
    auto area = [&](uint32_t k) {
        const float ex = n2[k].maxx - n2[k].minx, ey = n2[k].maxy - n2[k].miny, ez = n2[k].maxz - n2[k].minz;
        return ex * ey + ey * ez + ez * ex;
    };
    uint32_t kids[4], nk = 2;
    kids[0] = n2[i2].first_index;
    kids[1] = n2[i2].first_index + 1;
    while (nk < 4) {
        int pick = -1;
        for (uint32_t j = 0; j < nk; j++)
            if (!is_leaf(kids[j]) && (pick < 0 || area(kids[j]) > area(kids[pick]))) pick = (int)j;
        if (pick < 0) break;
        const uint32_t k = kids[pick];
        kids[pick] = n2[k].first_index;
        kids[nk++] = n2[k].first_index + 1;
    }
    const uint32_t idx = (uint32_t)out.size();
    out.push_back(Node4{});
    height4 = std::max(height4, depth);
    Node4 node;
    uint32_t inner = 0, below = 0;
    for (uint32_t j = 0; j < 4; j++) {
        if (j >= nk) {
            node.minx[j] = node.miny[j] = node.minz[j] = INFINITY;
            node.maxx[j] = node.maxy[j] = node.maxz[j] = -INFINITY;
            node.child[j] = ~0u;
            node.count[j] = 0;
            continue;
        }
        const Node &c = n2[kids[j]];
        node.minx[j] = c.minx; node.miny[j] = c.miny; node.minz[j] = c.minz;
        node.maxx[j] = c.maxx; node.maxy[j] = c.maxy; node.maxz[j] = c.maxz;
        if (is_leaf(kids[j])) {
            node.child[j] = c.first_index;
            node.count[j] = c.leaf_count_depth & 0xFFFFu;
        } else {
            uint32_t child_stack = 0;
            node.child[j] = collapse4(n2, kids[j], depth + 1, out, height4, child_stack);
            node.count[j] = 0;
            inner++;
            below = std::max(below, child_stack);
        }
    }
    stack = inner ? std::max(inner, inner - 1 + below) : 0;
    out[idx] = node;
    return idx;
}

// The 4-wide tree of a validated sibling-pair mesh BVH (children after parents), root first; returns its height
// and sets stack to the mesh stack entries its walk needs.  A root that is a leaf becomes one node holding it.
uint32_t collapse_mesh_bvh4(const Node *n2, std::vector<Node4> &out, uint32_t &stack) {
    uint32_t h4 = 1;
    stack = 0;
    out.clear();
    if (n2[0].leaf_count_depth & 0xFFFFu) {
        Node4 leaf;
        for (uint32_t j = 0; j < 4; j++) {
            leaf.minx[j] = leaf.miny[j] = leaf.minz[j] = INFINITY;
            leaf.maxx[j] = leaf.maxy[j] = leaf.maxz[j] = -INFINITY;
            leaf.child[j] = ~0u;
            leaf.count[j] = 0;
        }
        leaf.minx[0] = n2[0].minx; leaf.miny[0] = n2[0].miny; leaf.minz[0] = n2[0].minz;
        leaf.maxx[0] = n2[0].maxx; leaf.maxy[0] = n2[0].maxy; leaf.maxz[0] = n2[0].maxz;
        leaf.child[0] = n2[0].first_index;
        leaf.count[0] = n2[0].leaf_count_depth & 0xFFFFu;
        out.push_back(leaf);
    } else {
        collapse4(n2, 0, 1, out, h4, stack);
    }
    return h4;
}

// The 8-wide tree of a validated sibling-pair mesh BVH (Node8, LeafHead).  collapse8 gathers a node's children: its
// two children with, while fewer than eight, the inner child of largest surface area replaced by its own two;
// leaves keep their triangle ranges.  emit8 writes the node: child boxes quantised on a per-axis grid o + q * 2^e
// (q in 0..255) spanning the node's box with two steps to spare on each side, each bound moved outward until the
// decoded value, computed with the fused multiply-add the walk uses, lies at least two ulps of the exact bound
// beyond it (if a bound cannot be reached the node's steps double and it is quantised again); slots filled greedily by how well each child's
// offset from the node's centre points along the octant diagonal of the slot (slot bit set: negative axis), so a
// ray whose direction lies in octant o meets the slots in the order s ^ (7 ^ o) nearly front to back.  Node
// indices are assigned depth first, root 0; leaves go to leaf_out (LeafHead, then the leaf's TriHit records, in
// 16-byte units).  Both return the node index; stack receives the mesh stack entries the walk can hold below the
// node: a visit pushes its k hit inner children and pops one, so stack = max(k, k - 1 + the largest below).
uint32_t emit8(const Node *n2, const TriHit *tris, const uint32_t *kids, uint32_t nk, const Node &par,
               std::vector<Node8> &out, std::vector<uint4> &leaf_out, uint32_t &stack);

double ulp(float x) {   // the spacing of floats at |x|
    const float a = std::fabs(x);
    return (double)std::nextafter(a, INFINITY) - (double)a;
}

uint32_t collapse8(const Node *n2, const TriHit *tris, uint32_t i2, std::vector<Node8> &out, std::vector<uint4> &leaf_out,
                   uint32_t &stack) {
    auto is_leaf = [&](uint32_t k) { return (n2[k].leaf_count_depth & 0xFFFFu) != 0; };
    auto area = [&](uint32_t k) {
        const float ex = n2[k].maxx - n2[k].minx, ey = n2[k].maxy - n2[k].miny, ez = n2[k].maxz - n2[k].minz;
        return ex * ey + ey * ez + ez * ex;
    };
    uint32_t kids[8], nk = 2;
    kids[0] = n2[i2].first_index;
    kids[1] = n2[i2].first_index + 1;
    while (nk < 8) {
        int pick = -1;
        for (uint32_t j = 0; j < nk; j++)
            if (!is_leaf(kids[j]) && (pick < 0 || area(kids[j]) > area(kids[pick]))) pick = (int)j;
        if (pick < 0) break;
        const uint32_t k = kids[pick];
        kids[pick] = n2[k].first_index;
        kids[nk++] = n2[k].first_index + 1;
    }
    return emit8(n2, tris, kids, nk, n2[i2], out, leaf_out, stack);
}

uint32_t emit8(const Node *n2, const TriHit *tris, const uint32_t *kids, uint32_t nk, const Node &par,
               std::vector<Node8> &out, std::vector<uint4> &leaf_out, uint32_t &stack) {
    // slot assignment: the (child, slot) pairs by decreasing score, each child and slot taken once
    const float cx = 0.5f * (par.minx + par.maxx), cy = 0.5f * (par.miny + par.maxy), cz = 0.5f * (par.minz + par.maxz);
    struct Pair { float score; uint32_t child, slot; };
    std::vector<Pair> pairs;
    for (uint32_t j = 0; j < nk; j++) {
        const Node &c = n2[kids[j]];
        const float dx = 0.5f * (c.minx + c.maxx) - cx, dy = 0.5f * (c.miny + c.maxy) - cy, dz = 0.5f * (c.minz + c.maxz) - cz;
        for (uint32_t sl = 0; sl < 8; sl++)
            pairs.push_back({((sl & 1u) ? -dx : dx) + ((sl & 2u) ? -dy : dy) + ((sl & 4u) ? -dz : dz), j, sl});
    }
    std::stable_sort(pairs.begin(), pairs.end(), [](const Pair &a, const Pair &b) { return a.score > b.score; });
    int slot_of[8], child_in[8];
    for (int j = 0; j < 8; j++) slot_of[j] = child_in[j] = -1;
    for (const Pair &pr : pairs)
        if (slot_of[pr.child] < 0 && child_in[pr.slot] < 0) {
            slot_of[pr.child] = (int)pr.slot;
            child_in[pr.slot] = (int)pr.child;
        }
    auto set_byte = [](uint32_t w[2], uint32_t k, uint32_t v) {
        w[k >> 2] = (w[k >> 2] & ~(0xFFu << (8 * (k & 3u)))) | (v << (8 * (k & 3u)));
    };
    // the grid: per axis a power-of-two step with the node's extent in at most 251 steps, origin two steps below
    const float pmin[3] = {par.minx, par.miny, par.minz}, pmax[3] = {par.maxx, par.maxy, par.maxz};
    int e[3];
    for (int a = 0; a < 3; a++) {
        const double ext = (double)pmax[a] - (double)pmin[a];
        e[a] = -126;
        while (e[a] < 127 && std::ldexp(251.0, e[a]) < ext) e[a]++;
    }
    Node8 node;
    for (;;) {
        node = Node8{};
        float o[3], step[3];
        for (int a = 0; a < 3; a++) {
            step[a] = std::ldexp(1.0f, e[a]);
            o[a] = pmin[a] - 2.0f * step[a];
            node.exps |= (uint32_t)(e[a] + 127) << (8 * a);
        }
        node.ox = o[0]; node.oy = o[1]; node.oz = o[2];
        bool ok = true;
        for (uint32_t sl = 0; sl < 8; sl++) {
            if (child_in[sl] < 0) {
                set_byte(node.count, sl, 0xFFu);
                for (int a = 0; a < 3; a++) { set_byte(node.qlo[a], sl, 255u); set_byte(node.qhi[a], sl, 0u); }
                continue;
            }
            const Node &c = n2[kids[child_in[sl]]];
            const float cmin[3] = {c.minx, c.miny, c.minz}, cmax[3] = {c.maxx, c.maxy, c.maxz};
            for (int a = 0; a < 3; a++) {
                const double lo = std::floor(((double)cmin[a] - (double)o[a]) / (double)step[a]);
                const double hi = std::ceil(((double)cmax[a] - (double)o[a]) / (double)step[a]);
                int ql = (int)std::min(std::max(lo, 0.0), 255.0), qh = (int)std::min(std::max(hi, 0.0), 255.0);
                // two ulps of the bound to spare: the walk's fused slab test on the decoded bound then stays at or
                // beyond the reference's rounded slab distances on the exact bound, for every ray
                const double lim_lo = (double)cmin[a] - 2.0 * ulp(cmin[a]), lim_hi = (double)cmax[a] + 2.0 * ulp(cmax[a]);
                while (ql > 0 && (double)std::fmaf((float)ql, step[a], o[a]) > lim_lo) ql--;
                while (qh < 255 && (double)std::fmaf((float)qh, step[a], o[a]) < lim_hi) qh++;
                ok = ok && (double)std::fmaf((float)ql, step[a], o[a]) <= lim_lo &&
                     (double)std::fmaf((float)qh, step[a], o[a]) >= lim_hi;
                set_byte(node.qlo[a], sl, (uint32_t)ql);
                set_byte(node.qhi[a], sl, (uint32_t)qh);
            }
        }
        if (ok) break;
        for (int a = 0; a < 3; a++) e[a] = std::min(e[a] + 1, 127);
    }
    const uint32_t idx = (uint32_t)out.size();
    out.push_back(Node8{});
    uint32_t inner = 0, below = 0;
    for (uint32_t sl = 0; sl < 8; sl++) {
        if (child_in[sl] < 0) continue;
        const uint32_t k = kids[child_in[sl]];
        const Node &c = n2[k];
        const uint32_t count = c.leaf_count_depth & 0xFFFFu;
        if (count) {
            set_byte(node.count, sl, count);
            node.child[sl] = (uint32_t)leaf_out.size();
            const LeafHead h{c.minx, c.miny, c.minz, c.first_index, c.maxx, c.maxy, c.maxz, count};
            uint4 u[2];
            std::memcpy(u, &h, sizeof h);
            leaf_out.insert(leaf_out.end(), u, u + 2);
            for (uint32_t t = 0; t < count; t++) {
                uint4 w[3];
                std::memcpy(w, &tris[c.first_index + t], sizeof(TriHit));
                leaf_out.insert(leaf_out.end(), w, w + 3);
            }
        } else {
            uint32_t child_stack = 0;
            set_byte(node.count, sl, 0u);
            node.child[sl] = collapse8(n2, tris, k, out, leaf_out, child_stack);
            inner++;
            below = std::max(below, child_stack);
        }
    }
    stack = inner ? std::max(inner, inner - 1 + below) : 0;
    out[idx] = node;
    return idx;
}

// The 8-wide tree of a validated sibling-pair mesh BVH (children after parents) whose leaves hold at most 254
// triangles, root first, and its leaf records; returns the mesh stack entries its walk needs.  A root that is a leaf
// becomes a node holding it.  The leaf records end with a record's worth of padding (the walk reads a leaf's first
// two triangle records together).
uint32_t collapse_mesh_bvh8(const Node *n2, const TriHit *tris, std::vector<Node8> &out, std::vector<uint4> &leaf_out) {
    uint32_t stack = 0;
    out.clear();
    leaf_out.clear();
    if (n2[0].leaf_count_depth & 0xFFFFu) {
        const uint32_t root = 0;
        emit8(n2, tris, &root, 1, n2[0], out, leaf_out, stack);
    } else {
        collapse8(n2, tris, 0, out, leaf_out, stack);
    }
    for (int i = 0; i < 6; i++) leaf_out.push_back(uint4{0, 0, 0, 0});
    return stack;
}

int ensure_outputs(slimrt_context *c, size_t n) {
    if (c->d_pixels.cap < n) {
        dev_free(c, c->d_pixels);
        dev_free(c, c->d_depths);
        HIP_TRY(c, ctx_malloc(c, &c->d_pixels.ptr, n * sizeof(float4)));
        c->d_pixels.cap = n;
        HIP_TRY(c, ctx_malloc(c, &c->d_depths.ptr, n * sizeof(float)));
        c->d_depths.cap = n;
    }
    return SLIMRT_OK;
}

constexpr uint32_t kMaxTrainSteps = 64;

void free_nrc(slimrt_context *c) {
    for (void *p : {(void *)c->q_raw, (void *)c->q_prefix, (void *)c->q_tp, (void *)c->q_out, (void *)c->q_pixel,
                    (void *)c->t_raw, (void *)c->t_target, (void *)c->t_next, (void *)c->t_link,
                    (void *)c->nrc_counters, (void *)c->nrc_losses, (void *)c->cells})
        if (p) (void)ctx_free(c, p);
    c->q_raw = c->q_prefix = c->q_tp = c->q_out = c->t_raw = c->t_target = c->t_next = c->nrc_losses = c->cells = nullptr;
    c->q_pixel = c->nrc_counters = c->t_link = nullptr;
    c->q_cap = c->t_cap = c->cells_log2 = 0;
}

int ensure_nrc(slimrt_context *c, uint32_t q_cap, uint32_t t_cap, uint32_t cells_log2, bool &cells_new) {
    cells_new = false;
    if (!c->nrc_counters) {
        HIP_TRY(c, ctx_malloc(c, &c->nrc_counters, NC_COUNT * sizeof(uint32_t)));
        HIP_TRY(c, ctx_malloc(c, &c->nrc_losses, kMaxTrainSteps * sizeof(float)));
    }
    if (c->q_cap < q_cap) {
        for (void *p : {(void *)c->q_raw, (void *)c->q_prefix, (void *)c->q_tp, (void *)c->q_out, (void *)c->q_pixel})
            if (p) HIP_TRY(c, ctx_free(c, p));
        c->q_raw = c->q_prefix = c->q_tp = c->q_out = nullptr;
        c->q_pixel = nullptr;
        c->q_cap = 0;
        HIP_TRY(c, ctx_malloc(c, &c->q_raw, (size_t)q_cap * kNrcRaw * sizeof(float)));
        HIP_TRY(c, ctx_malloc(c, &c->q_prefix, (size_t)q_cap * 3 * sizeof(float)));
        HIP_TRY(c, ctx_malloc(c, &c->q_tp, (size_t)q_cap * 3 * sizeof(float)));
        HIP_TRY(c, ctx_malloc(c, &c->q_out, (size_t)q_cap * 3 * sizeof(float)));
        HIP_TRY(c, ctx_malloc(c, &c->q_pixel, (size_t)q_cap * sizeof(uint32_t)));
        c->q_cap = q_cap;
    }
    if (c->t_cap < t_cap) {
        for (void *p : {(void *)c->t_raw, (void *)c->t_target, (void *)c->t_next, (void *)c->t_link})
            if (p) HIP_TRY(c, ctx_free(c, p));
        c->t_raw = c->t_target = c->t_next = nullptr;
        c->t_link = nullptr;
        c->t_cap = 0;
        HIP_TRY(c, ctx_malloc(c, &c->t_raw, (size_t)t_cap * kNrcRaw * sizeof(float)));
        HIP_TRY(c, ctx_malloc(c, &c->t_target, (size_t)t_cap * 3 * sizeof(float)));
        HIP_TRY(c, ctx_malloc(c, &c->t_next, (size_t)t_cap * 3 * sizeof(float)));
        HIP_TRY(c, ctx_malloc(c, &c->t_link, (size_t)t_cap * sizeof(uint32_t)));
        c->t_cap = t_cap;
    }
    if (c->cells_log2 != cells_log2) {
        if (c->cells) HIP_TRY(c, ctx_free(c, c->cells));
        c->cells = nullptr;
        c->cells_log2 = 0;
        HIP_TRY(c, ctx_malloc(c, &c->cells, ((size_t)kGateFloats << cells_log2) * sizeof(float)));
        c->cells_log2 = cells_log2;
        cells_new = true;
    }
    return SLIMRT_OK;
}

}  // namespace

extern "C" {

int slimrt_create(slimrt_context **out, uint32_t flags) {
    if (!out) return SLIMRT_E_ARG;
    *out = nullptr;
    auto *c = new slimrt_context();
    c->cpu = (flags & SLIMRT_DEVICE_CPU) != 0;
    if (!c->cpu) {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) {
            std::string msg = std::string("no HIP device available: ") + (e != hipSuccess ? hipGetErrorString(e) : "0 devices");
            delete c;
            fprintf(stderr, "slimrt_create: %s\n", msg.c_str());
            return SLIMRT_E_DEVICE;
        }
        (void)hipGetDevice(&c->device);
        if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
            ctx_malloc(c, &c->d_status, sizeof(uint32_t)) != hipSuccess ||
            ctx_malloc(c, &c->d_counters, 8 * sizeof(unsigned long long)) != hipSuccess ||
            hipMemset(c->d_status, 0, sizeof(uint32_t)) != hipSuccess ||
            noalloc_scratch_address(c->device, &c->noalloc_scratch) != SLIMRT_OK) {
            c->stream = c->own_stream;
            (void)slimrt_destroy(c);
            return SLIMRT_E_DEVICE;
        }
        c->stream = c->own_stream;
    }
    *out = c;
    return SLIMRT_OK;
}

int slimrt_destroy(slimrt_context *c) {
    if (!c) return SLIMRT_E_ARG;
    if (!c->cpu) {
        // the caller's stream, the stream of the last cost-ordered frame (it reads and writes the tile buffers) and
        // the side stream of banded wavefront frames all finish before anything is freed
        (void)hipStreamSynchronize(c->stream);
        if (c->tile_frame) (void)hipEventSynchronize(c->tile_rendered);
        if (c->side_stream) (void)hipStreamSynchronize(c->side_stream);
        dev_free(c, c->d_geos); dev_free(c, c->d_mats); dev_free(c, c->d_lights); dev_free(c, c->d_snodes); dev_free(c, c->d_mnodes);
        dev_free(c, c->d_mnodes4);
        dev_free(c, c->d_mnodes8);
        dev_free(c, c->d_leaf8);
        dev_free(c, c->d_leaf); dev_free(c, c->d_meshes); dev_free(c, c->d_tri_hit); dev_free(c, c->d_tri_attr); dev_free(c, c->d_tex);
        dev_free(c, c->d_mips); dev_free(c, c->d_quads); dev_free(c, c->d_pixels); dev_free(c, c->d_depths);
        free_nrc(c);
        dev_free(c, c->d_geo_stage);
        dev_free(c, c->ov_lines);
        dev_free(c, c->ov_samples);
        for (void *q : {c->wf_queue_a, c->wf_queue_b, c->wf_accum, c->wf_hits, (void *)c->wf_counters, (void *)c->wf_stats,
                        (void *)c->d_pick})
            if (q) (void)ctx_free(c, q);
        for (hipEvent_t e : c->wf_band)
            if (e) (void)hipEventDestroy(e);
        if (c->wf_fork) (void)hipEventDestroy(c->wf_fork);
        if (c->wf_join) (void)hipEventDestroy(c->wf_join);
        if (c->d_status) (void)ctx_free(c, c->d_status);
        if (c->d_counters) (void)ctx_free(c, c->d_counters);
        for (int p = 0; p < 3; p++)
            if (c->tile_cost[p]) (void)ctx_free(c, c->tile_cost[p]);
        for (int p = 0; p < 2; p++) {
            if (c->tile_order[p]) (void)ctx_free(c, c->tile_order[p]);
            if (c->tile_hist[p]) (void)ctx_free(c, c->tile_hist[p]);
        }
        if (c->tile_rendered) (void)hipEventDestroy(c->tile_rendered);
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        if (c->side_stream) (void)hipStreamDestroy(c->side_stream);
        if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    }
    delete c;
    return SLIMRT_OK;
}

const char *slimrt_last_error(const slimrt_context *c) { return c ? c->err.c_str() : "null context"; }

int slimrt_set_stream(slimrt_context *c, void *stream) {
    if (!c) return SLIMRT_E_ARG;
    if (c->cpu) return SLIMRT_OK;
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    return SLIMRT_OK;
}

int slimrt_upload_geometries(slimrt_context *c, const slimrt_geometry *g, uint32_t count) {
    if (!c || (!g && count)) return SLIMRT_E_ARG;
    if (c->loaded && count != c->geos.size()) return fail(c, SLIMRT_E_ARG, "geometry count differs from the scene's");
    c->geos.resize(count);
    for (uint32_t i = 0; i < count; i++) {
        if (g[i].type == GT_Mesh && g[i].id >= c->meshes.size() && c->loaded)
            return fail(c, SLIMRT_E_ARG, "geometry references a missing mesh");
        c->geos[i] = to_geo(g[i]);
    }
    update_scene_flags(c);
    return c->cpu ? SLIMRT_OK : dev_upload(c, c->d_geos, c->geos.data(), c->geos.size());
}

int slimrt_upload_materials(slimrt_context *c, const slimrt_material *m, uint32_t count) {
    if (!c || (!m && count)) return SLIMRT_E_ARG;
    c->mats.resize(count);
    for (uint32_t i = 0; i < count; i++) c->mats[i] = to_mat(m[i]);
    update_scene_flags(c);
    return c->cpu ? SLIMRT_OK : dev_upload(c, c->d_mats, c->mats.data(), c->mats.size());
}

int slimrt_upload_lights(slimrt_context *c, const slimrt_light *l, uint32_t count) {
    if (!c || (!l && count)) return SLIMRT_E_ARG;
    c->lights.resize(count);
    for (uint32_t i = 0; i < count; i++) c->lights[i] = to_light(l[i]);
    return c->cpu ? SLIMRT_OK : dev_upload(c, c->d_lights, c->lights.data(), c->lights.size());
}

int slimrt_upload_scene_bvh(slimrt_context *c, const slimrt_bvh_node *nodes, uint32_t node_count, const uint32_t *leaf_ids,
                            uint32_t height) {
    if (!c || !nodes || !leaf_ids || node_count == 0) return SLIMRT_E_ARG;
    const uint32_t G = (uint32_t)c->geos.size();
    c->snodes.resize(node_count);
    for (uint32_t i = 0; i < node_count; i++) {
        c->snodes[i] = to_node(nodes[i]);
        const uint32_t lc = nodes[i].leaf_count;
        // children always follow their parent (BVHBuilder::splitNode appends them), which also rules out cycles
        const bool bad = lc ? (nodes[i].first_index + lc > G)
                            : (nodes[i].first_index <= i || nodes[i].first_index + 1 >= node_count);
        if (bad) return fail(c, SLIMRT_E_ARG, "scene BVH node index out of range");
    }
    c->leaf_ids.assign(leaf_ids, leaf_ids + G);
    for (uint32_t id : c->leaf_ids)
        if (id >= G) return fail(c, SLIMRT_E_ARG, "scene BVH leaf id out of range");
    c->scene_height = height;
    if (c->cpu) return SLIMRT_OK;
    int rc = dev_upload(c, c->d_snodes, c->snodes.data(), c->snodes.size());
    return rc ? rc : dev_upload(c, c->d_leaf, c->leaf_ids.data(), c->leaf_ids.size());
}

int slimrt_upload_scene(slimrt_context *c, const slimrt_scene_desc *d) {
    if (!c || !d) return SLIMRT_E_ARG;
    if (d->geometry_count == 0 || !d->geometries || !d->scene_nodes || !d->scene_leaf_ids)
        return fail(c, SLIMRT_E_ARG, "a scene needs geometries and a scene BVH");
    c->loaded = false;
    // meshes: concatenate nodes and split triangle records
    c->meshes.clear(); c->mnodes.clear(); c->tri_hit.clear(); c->tri_attr.clear(); c->mnodes4.clear();
    c->mnodes8.clear(); c->leaf8.clear();
    c->mesh_height = 0;
    c->mesh_height4 = 0;
    c->mesh_stack4 = 0;
    c->mesh_nodes4_max = 0;
    c->mesh_stack8 = 0;
    c->mesh_nodes8_max = 0;
    c->mesh_leaf_max = 0;
    for (uint32_t k = 0; k < d->mesh_count; k++) {
        const slimrt_mesh &m = d->meshes[k];
        if (!m.nodes || !m.triangles || m.node_count == 0 || m.triangle_count == 0)
            return fail(c, SLIMRT_E_ARG, "mesh without nodes or triangles");
        MeshRec r;
        r.minx = m.aabb_min[0]; r.miny = m.aabb_min[1]; r.minz = m.aabb_min[2];
        r.maxx = m.aabb_max[0]; r.maxy = m.aabb_max[1]; r.maxz = m.aabb_max[2];
        r.node_base = (uint32_t)c->mnodes.size();
        r.tri_base = (uint32_t)c->tri_hit.size();
        r.node_count = m.node_count; r.tri_count = m.triangle_count;
        r.normals_count = m.normals_count; r.uvs_count = m.uvs_count;
        for (uint32_t i = 0; i < m.node_count; i++) {
            const slimrt_bvh_node &n = m.nodes[i];
            const bool bad = n.leaf_count ? (n.first_index + n.leaf_count > m.triangle_count)
                                          : (n.first_index <= i || n.first_index + 1 >= m.node_count);
            if (bad) return fail(c, SLIMRT_E_ARG, "mesh BVH node index out of range");
            c->mnodes.push_back(to_node(n));
        }
        r.node4_base = (uint32_t)c->mnodes4.size();
        std::vector<Node4> n4;
        uint32_t stack4 = 0;
        const uint32_t h4 = collapse_mesh_bvh4(c->mnodes.data() + r.node_base, n4, stack4);
        c->mesh_stack4 = std::max(c->mesh_stack4, stack4);
        c->mesh_nodes4_max = std::max(c->mesh_nodes4_max, (uint32_t)n4.size());
        c->mnodes4.insert(c->mnodes4.end(), n4.begin(), n4.end());
        r.pad = 0;
        c->mesh_height4 = std::max(c->mesh_height4, h4);
        for (uint32_t i = 0; i < m.triangle_count; i++) {
            const slimrt_triangle &t = m.triangles[i];
            const float *L = t.local_to_tangent;
            c->tri_hit.push_back(TriHit{L[0], L[1], L[3], L[4], L[6], L[7], t.position[0], t.position[1],
                                        t.position[2], t.normal[0], t.normal[1], t.normal[2]});
            c->tri_attr.push_back(TriAttr{t.n1[0], t.n1[1], t.n1[2], t.uv1[0], t.n2[0], t.n2[1], t.n2[2], t.uv1[1],
                                          t.n3[0], t.n3[1], t.n3[2], t.uv2[0], t.uv2[1], t.uv3[0], t.uv3[1],
                                          t.uv_coverage});
        }
        uint32_t leaf_max = 0;
        for (uint32_t i = 0; i < m.node_count; i++) leaf_max = std::max<uint32_t>(leaf_max, m.nodes[i].leaf_count);
        c->mesh_leaf_max = std::max(c->mesh_leaf_max, leaf_max);
        r.node8_base = (uint32_t)c->mnodes8.size();
        r.leaf8_base = (uint32_t)c->leaf8.size();
        if (leaf_max <= 254) {   // larger leaves cannot be described by the 8-wide nodes (check_width refuses them)
            std::vector<Node8> n8;
            std::vector<uint4> l8;
            const uint32_t stack8 = collapse_mesh_bvh8(c->mnodes.data() + r.node_base, c->tri_hit.data() + r.tri_base, n8, l8);
            c->mesh_stack8 = std::max(c->mesh_stack8, stack8);
            c->mesh_nodes8_max = std::max(c->mesh_nodes8_max, (uint32_t)n8.size());
            c->mnodes8.insert(c->mnodes8.end(), n8.begin(), n8.end());
            c->leaf8.insert(c->leaf8.end(), l8.begin(), l8.end());
        }
        c->meshes.push_back(r);
        c->mesh_height = std::max(c->mesh_height, m.height);
    }
    // textures: flatten the mips (texel quads, or bordered RGBA8 / RGBA16F texels) into one byte array, each mip
    // starting on a 16-byte boundary
    c->textures.clear(); c->mips.clear(); c->quads.clear();
    for (uint32_t k = 0; k < d->texture_count; k++) {
        const slimrt_texture &t = d->textures[k];
        if (!t.mips || t.mip_count == 0) return fail(c, SLIMRT_E_ARG, "texture without mips");
        const uint32_t fmt = (t.flags & kTexelsMask) >> kTexelsShift;
        if (fmt > kTexelsRGBA16F) return fail(c, SLIMRT_E_ARG, "unknown texel storage in texture flags");
        c->textures.push_back(TexRec{t.width, t.height, t.flags, t.mip_count, (uint32_t)c->mips.size(), 0, 0, 0});
        for (uint32_t m = 0; m < t.mip_count; m++) {
            const slimrt_texture_mip &mm = t.mips[m];
            if (!mm.texel_quads || mm.width == 0 || mm.height == 0) return fail(c, SLIMRT_E_ARG, "empty texture mip");
            const size_t bytes = fmt == kTexelQuads ? 12 * (size_t)(mm.width + 1) * (mm.height + 1)
                                                    : (fmt == kTexelsRGBA8 ? 4 : 8) * (size_t)(mm.width + 2) * (mm.height + 2);
            const size_t base = (c->quads.size() + 15) / 16;
            if (base > 0xFFFFFFFFull) return fail(c, SLIMRT_E_ARG, "textures beyond 64 GiB");
            c->quads.resize(16 * base, 0);
            c->mips.push_back(MipRec{mm.width, mm.height, (uint32_t)base, fmt});
            c->quads.insert(c->quads.end(), mm.texel_quads, mm.texel_quads + bytes);
        }
    }
    for (uint32_t i = 0; i < d->material_count; i++) {
        const slimrt_material &m = d->materials[i];
        for (uint32_t s = 0; s < std::min<uint32_t>(m.texture_count, 2); s++)
            if (m.texture_ids[s] >= d->texture_count) return fail(c, SLIMRT_E_ARG, "material texture id out of range");
    }
    for (uint32_t i = 0; i < d->geometry_count; i++) {
        const slimrt_geometry &g = d->geometries[i];
        if (g.material_id >= d->material_count) return fail(c, SLIMRT_E_ARG, "geometry material id out of range");
        if (g.type == GT_Mesh && g.id >= d->mesh_count) return fail(c, SLIMRT_E_ARG, "geometry mesh id out of range");
    }
    int rc;
    c->geos.clear();
    if ((rc = slimrt_upload_materials(c, d->materials, d->material_count))) return rc;
    if ((rc = slimrt_upload_geometries(c, d->geometries, d->geometry_count))) return rc;
    if ((rc = slimrt_upload_lights(c, d->lights, d->light_count))) return rc;
    if ((rc = slimrt_upload_scene_bvh(c, d->scene_nodes, d->scene_node_count, d->scene_leaf_ids, d->scene_height))) return rc;
    if (!c->cpu) {
        if ((rc = dev_upload(c, c->d_meshes, c->meshes.data(), c->meshes.size()))) return rc;
        if ((rc = dev_upload(c, c->d_mnodes, c->mnodes.data(), c->mnodes.size()))) return rc;
        if ((rc = dev_upload(c, c->d_mnodes4, c->mnodes4.data(), c->mnodes4.size()))) return rc;
        if ((rc = dev_upload(c, c->d_mnodes8, c->mnodes8.data(), c->mnodes8.size()))) return rc;
        if ((rc = dev_upload(c, c->d_leaf8, c->leaf8.data(), c->leaf8.size()))) return rc;
        if ((rc = dev_upload(c, c->d_tri_hit, c->tri_hit.data(), c->tri_hit.size()))) return rc;
        if ((rc = dev_upload(c, c->d_tri_attr, c->tri_attr.data(), c->tri_attr.size()))) return rc;
        if ((rc = dev_upload(c, c->d_tex, c->textures.data(), c->textures.size()))) return rc;
        if ((rc = dev_upload(c, c->d_mips, c->mips.data(), c->mips.size()))) return rc;
        if ((rc = dev_upload(c, c->d_quads, c->quads.data(), c->quads.size()))) return rc;
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    c->loaded = true;
    return SLIMRT_OK;
}

int slimrt_render_device(slimrt_context *c, const slimrt_render_params *p, void *d_pixels, void *d_depths) {
    if (!c) return SLIMRT_E_ARG;
    int rc = validate_params(c, p);
    if (rc) return rc;
    if (c->cpu) return fail(c, SLIMRT_E_STATE, "render_device needs a GPU context");
    if (!d_pixels || !d_depths) return fail(c, SLIMRT_E_ARG, "output buffers are NULL");
    c->timed = true;
    return launch(c, p, (float4 *)d_pixels, (float *)d_depths, false);
}

int slimrt_frame_flags(slimrt_context *c, uint32_t *flags) {
    if (!c || !flags) return SLIMRT_E_ARG;
    uint32_t st = 0;
    if (!c->cpu) {
        HIP_TRY(c, hipMemcpyAsync(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (st) {
            HIP_TRY(c, hipMemsetAsync(c->d_status, 0, sizeof st, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
    }
    *flags = st | c->host_status;
    c->host_status = 0;
    return SLIMRT_OK;
}

int slimrt_set_memory_limit(slimrt_context *c, uint64_t bytes) {
    if (!c) return SLIMRT_E_ARG;
    c->mem_limit = bytes;
    return SLIMRT_OK;
}

int slimrt_memory_usage(slimrt_context *c, uint64_t *used, uint64_t *limit) {
    if (!c) return SLIMRT_E_ARG;
    if (used) *used = c->mem_used;
    if (limit) *limit = c->mem_limit;
    return SLIMRT_OK;
}

int slimrt_render_device_noalloc(slimrt_context *c, const slimrt_render_params *p, void *d_pixels, void *d_depths) {
    if (!c) return SLIMRT_E_ARG;
    int rc = validate_params(c, p);
    if (rc) return rc;
    if (c->cpu) return fail(c, SLIMRT_E_STATE, "render_device_noalloc needs a GPU context");
    if (!d_pixels || !d_depths) return fail(c, SLIMRT_E_ARG, "output buffers are NULL");
    if (!c->noalloc_scratch) return fail(c, SLIMRT_E_STATE, "no scratch address for this device");
    if (p->flags & (SLIMRT_WAVEFRONT | SLIMRT_WAVEFRONT_COUNT))
        return fail(c, SLIMRT_E_ARG, "render_device_noalloc renders per-sample frames only: the wavefront forms "
                                     "need ray queues the library allocates");
    RenderParams rp;
    std::memcpy(&rp, p, sizeof rp);
    rp.ssaa = rp.ssaa ? 1u : 0u;
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;
    const uint32_t wtiles_x = (ws + kWaveTileX - 1) / kWaveTileX;
    const uint64_t wtiles64 = (uint64_t)wtiles_x * ((hs + kWaveTileY - 1) / kWaveTileY);
    if (wtiles64 > kNoAllocTiles)
        return fail(c, SLIMRT_E_ARG, "render_device_noalloc: the frame has " + std::to_string(wtiles64) +
                                     " tiles of 32 x 2 samples; at most 131072 (8,388,608 samples) fit its scratch");
    const uint32_t wtiles = (uint32_t)wtiles64;
    const uint32_t width = mesh_node_width(rp);
    const uint32_t scene_depth = c->scene_height + 1, mesh_depth = mesh_stack_depth(c, width);
    const size_t wlds = stack_lds_bytes(scene_depth, mesh_depth, width, 64);
    if (wlds > 160 * 1024) return fail(c, SLIMRT_E_ARG, "BVH too deep for the LDS traversal stacks");
    if (check_width(c, width)) return SLIMRT_E_ARG;
    const bool by_cost = !(rp.flags & SLIMRT_TILE_ORDER_FIXED);
    uint32_t *cost_tiles = c->noalloc_scratch, *cost = nullptr, *order = nullptr;
    uint32_t resident = 0;
    if (by_cost) {
        cost = c->noalloc_scratch + kNoAllocCost;
        order = c->noalloc_scratch + kNoAllocOrder;
        if ((rc = resident_waves(c, width, scene_lights(device_view(c)), wlds, resident))) return rc;
    }
    const SceneView v = device_view(c);
    c->timed = true;
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    // On the caller's stream, right before the frame: the order from the durations the last such frame on this
    // GPU recorded (the image order when they belong to another tile count); one stream, so a graph can hold it.
    if (by_cost) {
        launch_tile_order(c->stream, cost, order, wtiles, resident, cost_tiles, c->noalloc_scratch + kNoAllocOrderScratch);
        HIP_TRY(c, hipGetLastError());
    }
    const uint32_t split = by_cost ? std::min(wtiles / kSplitShare, kSplitMax) : 0u;
    const uint32_t quarter = std::min(quarter_tiles(wtiles, width), split), tp = kTopParts;
#define NOALLOC_LAUNCH(NW, LIGHTS)                                                                                \
    hipLaunchKernelGGL((render_kernel<false, NW, LIGHTS>), dim3(wtiles + split + (tp - 2) * quarter), dim3(64), wlds,      \
                       c->stream, v, rp, (float4 *)d_pixels, (float *)d_depths, scene_depth, mesh_depth, wtiles_x,     \
                       wtiles, c->d_status, nullptr, order, cost, split, quarter, tp, by_cost ? cost_tiles : nullptr, \
                       OrderWork{})
    const uint32_t lights = scene_lights(v);
    if (lights == kLightsAll) {
        if (width == 8) NOALLOC_LAUNCH(8, kLightsAll); else if (width == 4) NOALLOC_LAUNCH(4, kLightsAll); else NOALLOC_LAUNCH(2, kLightsAll);
    } else if (lights == kLightsQuads) {
        if (width == 8) NOALLOC_LAUNCH(8, kLightsQuads); else if (width == 4) NOALLOC_LAUNCH(4, kLightsQuads); else NOALLOC_LAUNCH(2, kLightsQuads);
    } else {
        if (width == 8) NOALLOC_LAUNCH(8, kLightsNone); else if (width == 4) NOALLOC_LAUNCH(4, kLightsNone); else NOALLOC_LAUNCH(2, kLightsNone);
    }
#undef NOALLOC_LAUNCH
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    return SLIMRT_OK;
}

int slimrt_synchronize(slimrt_context *c) {
    if (!c) return SLIMRT_E_ARG;
    if (c->cpu) return SLIMRT_OK;
    return check_status(c);
}

int slimrt_render(slimrt_context *c, const slimrt_render_params *p, slimrt_pixel *pixels, float *depths) {
    if (!c) return SLIMRT_E_ARG;
    int rc = validate_params(c, p);
    if (rc) return rc;
    if (!pixels || !depths) return fail(c, SLIMRT_E_ARG, "output buffers are NULL");
    if (c->cpu) return cpu_render(c, p, pixels, depths, nullptr);
    const size_t n = (size_t)p->width * p->height * (p->ssaa ? 4 : 1);
    if ((rc = ensure_outputs(c, n))) return rc;
    c->timed = true;
    if ((rc = launch(c, p, c->d_pixels.ptr, c->d_depths.ptr, false))) return rc;
    HIP_TRY(c, hipMemcpyAsync(pixels, c->d_pixels.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(depths, c->d_depths.ptr, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return check_status(c);
}

int slimrt_count_work(slimrt_context *c, const slimrt_render_params *p, uint64_t counters[8]) {
    if (!c || !counters) return SLIMRT_E_ARG;
    int rc = validate_params(c, p);
    if (rc) return rc;
    if (c->cpu) return cpu_render(c, p, nullptr, nullptr, counters);
    const size_t n = (size_t)p->width * p->height * (p->ssaa ? 4 : 1);
    if ((rc = ensure_outputs(c, n))) return rc;
    HIP_TRY(c, hipMemsetAsync(c->d_counters, 0, 8 * sizeof(unsigned long long), c->stream));
    c->timed = false;
    if ((rc = launch(c, p, c->d_pixels.ptr, c->d_depths.ptr, true))) return rc;
    HIP_TRY(c, hipMemcpyAsync(counters, c->d_counters, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    return check_status(c);
}

int slimrt_host_collapse_bvh4(const slimrt_bvh_node *nodes, uint32_t node_count, uint32_t triangle_count,
                              slimrt_bvh4_node *nodes_out, uint32_t node_cap, uint32_t *node_count_out,
                              uint32_t *height_out, uint32_t *stack_out) {
    if (!nodes || !node_count || !nodes_out || !node_count_out || !height_out || !stack_out) return SLIMRT_E_ARG;
    std::vector<Node> n2(node_count);
    for (uint32_t i = 0; i < node_count; i++) {   // the checks of the upload: the walk below stays in range
        const slimrt_bvh_node &n = nodes[i];
        const bool bad = n.leaf_count ? ((uint64_t)n.first_index + n.leaf_count > triangle_count)
                                      : (n.first_index <= i || (uint64_t)n.first_index + 1 >= node_count);
        if (bad) return SLIMRT_E_ARG;
        n2[i] = to_node(n);
    }
    std::vector<Node4> out;
    uint32_t stack4 = 0;
    const uint32_t h4 = collapse_mesh_bvh4(n2.data(), out, stack4);
    if (out.size() > node_cap) return SLIMRT_E_OVERFLOW;
    static_assert(sizeof(Node4) == sizeof(slimrt_bvh4_node), "4-wide node layout");
    memcpy(nodes_out, out.data(), out.size() * sizeof(Node4));
    *node_count_out = (uint32_t)out.size();
    *height_out = h4;
    *stack_out = stack4;
    return SLIMRT_OK;
}

int slimrt_host_collapse_bvh8(const slimrt_bvh_node *nodes, uint32_t node_count, const slimrt_triangle *triangles,
                              uint32_t triangle_count, slimrt_bvh8_node *nodes_out, uint32_t node_cap,
                              uint32_t *node_count_out, uint32_t *leaves_out, uint32_t leaf_cap,
                              uint32_t *leaf_count_out, uint32_t *stack_out) {
    if (!nodes || !node_count || !triangles || (!nodes_out && node_cap) || !node_count_out || (!leaves_out && leaf_cap) ||
        !leaf_count_out || !stack_out)
        return SLIMRT_E_ARG;
    std::vector<Node> n2(node_count);
    for (uint32_t i = 0; i < node_count; i++) {   // the checks of the upload: the walk below stays in range
        const slimrt_bvh_node &n = nodes[i];
        const bool bad = n.leaf_count ? ((uint64_t)n.first_index + n.leaf_count > triangle_count || n.leaf_count > 254)
                                      : (n.first_index <= i || (uint64_t)n.first_index + 1 >= node_count);
        if (bad) return SLIMRT_E_ARG;
        n2[i] = to_node(n);
    }
    std::vector<TriHit> tris(triangle_count);
    for (uint32_t i = 0; i < triangle_count; i++) {
        const slimrt_triangle &t = triangles[i];
        const float *L = t.local_to_tangent;
        tris[i] = TriHit{L[0], L[1], L[3], L[4], L[6], L[7], t.position[0], t.position[1], t.position[2],
                         t.normal[0], t.normal[1], t.normal[2]};
    }
    std::vector<Node8> out;
    std::vector<uint4> leaves;
    *stack_out = collapse_mesh_bvh8(n2.data(), tris.data(), out, leaves);
    *node_count_out = (uint32_t)out.size();
    *leaf_count_out = (uint32_t)leaves.size();
    if (out.size() > node_cap || leaves.size() > leaf_cap) return SLIMRT_E_OVERFLOW;
    static_assert(sizeof(Node8) == sizeof(slimrt_bvh8_node), "8-wide node layout");
    memcpy(nodes_out, out.data(), out.size() * sizeof(Node8));
    memcpy(leaves_out, leaves.data(), leaves.size() * sizeof(uint4));
    return SLIMRT_OK;
}

int slimrt_canvas_content(slimrt_context *c, const void *pixels, uint32_t width, uint32_t height, uint32_t ssaa,
                          void *content) {
    if (!c) return SLIMRT_E_ARG;
    if (!pixels || !content || !width || !height || width > 65535 || height > 65535)
        return fail(c, SLIMRT_E_ARG, "canvas_content: NULL buffer or dimensions out of range");
    const uint32_t n = width * height;
    if (c->cpu) {
        for (uint32_t i = 0; i < n; i++)
            ((uint32_t *)content)[i] = pixel_content((const float4 *)pixels + (size_t)i * (ssaa ? 4 : 1), ssaa != 0);
        return SLIMRT_OK;
    }
    hipLaunchKernelGGL(content_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream,
                       (const float4 *)pixels, n, ssaa ? 1u : 0u, (uint32_t *)content);
    HIP_TRY(c, hipGetLastError());
    return SLIMRT_OK;
}

int slimrt_pick(slimrt_context *c, const slimrt_render_params *p, int32_t x, int32_t y, slimrt_pick_result *out) {
    if (!c) return SLIMRT_E_ARG;
    int rc = validate_params(c, p);
    if (rc) return rc;
    if (!out) return fail(c, SLIMRT_E_ARG, "out is NULL");
    RenderParams rp;
    std::memcpy(&rp, p, sizeof rp);
    rp.ssaa = rp.ssaa ? 1u : 0u;
    int32_t res[7];
    if (c->cpu) {
        const SceneView v = host_view(c);
        std::vector<uint32_t> stack(c->scene_height + 1 + c->mesh_height + 1);
        Tracer tr;
        tr.scene_stack = Stack{stack.data(), 1};
        tr.mesh_stack = Stack{stack.data() + c->scene_height + 1, 1};
        tr.scene_depth = c->scene_height + 1;
        tr.mesh_depth = c->mesh_height + 1;
        tr.status = 0;
        tr.cnt = nullptr;
        float of[4];
        pick(v, rp, x, y, tr, res, of);
        std::memcpy(res + 3, of, sizeof of);
        if (tr.status & kStatusStack) return fail(c, SLIMRT_E_OVERFLOW, "traversal stack overflow");
    } else {
        if (!c->d_pick) HIP_TRY(c, ctx_malloc(c, &c->d_pick, 8 * sizeof(int32_t)));
        const uint32_t scene_depth = c->scene_height + 1, mesh_depth = mesh_stack_depth(c, 2);
        const size_t lds = stack_lds_bytes(scene_depth, mesh_depth, 2);
        if (lds > 160 * 1024) return fail(c, SLIMRT_E_ARG, "BVH too deep for the LDS traversal stacks");
        hipLaunchKernelGGL(pick_kernel, dim3(1), dim3(kBlock), lds, c->stream, device_view(c), rp, x, y, scene_depth,
                           mesh_depth, c->d_pick, c->d_status);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(res, c->d_pick, sizeof res, hipMemcpyDeviceToHost, c->stream));
        if ((rc = check_status(c))) return rc;   // synchronizes the stream
    }
    out->geometry = res[0];
    out->light = res[1];
    out->triangle = res[2];
    std::memcpy(&out->distance, res + 3, sizeof(float));
    std::memcpy(out->position, res + 4, 3 * sizeof(float));
    return SLIMRT_OK;
}

int slimrt_wavefront_counts(slimrt_context *c, uint32_t *counts, uint32_t n) {
    if (!c) return SLIMRT_E_ARG;
    if (!counts || n < kWfStats) return fail(c, SLIMRT_E_ARG, "counts needs 34 entries");
    if (c->cpu || !c->wf_stats) return fail(c, SLIMRT_E_STATE, "no frame rendered with SLIMRT_WAVEFRONT_COUNT");
    HIP_TRY(c, hipMemcpyAsync(counts, c->wf_stats, kWfStats * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return SLIMRT_OK;
}

int slimrt_tile_costs(slimrt_context *c, uint32_t *costs, uint32_t *order, uint32_t cap, uint32_t *tile_count,
                      uint32_t *resident) {
    if (!c || !tile_count) return SLIMRT_E_ARG;
    if (resident) *resident = c->cpu ? 0u : c->resident;
    const uint64_t g = c->tile_frame - 1u;   // the last cost-ordered frame
    const uint32_t last = (uint32_t)(g & 1u), last_cost = (uint32_t)(g % 3);
    const uint32_t n = (c->cpu || c->tile_frame == 0) ? 0u : c->cost_tiles[last_cost];
    *tile_count = n;
    if (n == 0) return SLIMRT_OK;
    if ((costs || order) && cap < n) return fail(c, SLIMRT_E_ARG, "cap is smaller than the tile count");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (costs) HIP_TRY(c, hipMemcpy(costs, c->tile_cost[last_cost], n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (order) {
        if (c->order_used[last]) HIP_TRY(c, hipMemcpy(order, c->tile_order[last], n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        else
            for (uint32_t i = 0; i < n; i++) order[i] = i;
    }
    return SLIMRT_OK;
}

int slimrt_last_kernel_ms(slimrt_context *c, float *ms) {
    if (!c || !ms) return SLIMRT_E_ARG;
    if (c->cpu || !c->timed) return fail(c, SLIMRT_E_STATE, "no timed GPU frame");
    HIP_TRY(c, hipEventSynchronize(c->ev1));
    HIP_TRY(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
    return SLIMRT_OK;
}

int slimrt_render_nrc(slimrt_context *c, const slimrt_render_params *p, slimrt_nrc *nrc,
                      const slimrt_nrc_render_params *np, void *d_pixels, void *d_depths) {
    if (!c) return SLIMRT_E_ARG;
    int rc = validate_params(c, p);
    if (rc) return rc;
    if (c->cpu) return fail(c, SLIMRT_E_STATE, "render_nrc needs a GPU context");
    if (!nrc || !np || !d_pixels || !d_depths) return fail(c, SLIMRT_E_ARG, "NULL cache, settings or output buffer");
    if (p->render_mode != SLIMRT_MODE_BEAUTY || p->output > SLIMRT_OUTPUT_RADIANCE)
        return fail(c, SLIMRT_E_ARG, "the cache renders beauty frames to the canvas or the radiance buffer");
    if (p->max_depth == 0 || p->max_depth > kNrcMaxDepth) return fail(c, SLIMRT_E_ARG, "max_depth must be 1..8 with the cache");
    if (np->gate_cells_log2 < 10 || np->gate_cells_log2 > 24 || !(np->gate_cell_size > 0.0f))
        return fail(c, SLIMRT_E_ARG, "gate table: cells_log2 10..24 and a positive cell size");
    const bool train = np->train_one_in != 0 && np->train_steps != 0;
    if (train && (np->train_steps > kMaxTrainSteps || np->train_batch == 0 || np->train_batch % 64 ||
                  np->train_batch > (1u << 22) || !(np->learning_rate > 0.0f)))
        return fail(c, SLIMRT_E_ARG, "training: 1..64 steps, batch a multiple of 64 up to 2^22, positive learning rate");
    RenderParams rp;
    std::memcpy(&rp, p, sizeof rp);
    rp.ssaa = rp.ssaa ? 1u : 0u;
    const uint32_t ws = rp.width << rp.ssaa, hs = rp.height << rp.ssaa;
    const uint32_t samples = ws * hs;
    // one-wave workgroups, a 32x2 strip of samples each (the render kernel's tiling)
    const uint32_t tiles_x = (ws + kWaveTileX - 1) / kWaveTileX, tiles_y = (hs + kWaveTileY - 1) / kWaveTileY;
    const uint32_t tiles = tiles_x * tiles_y;
    const uint32_t width = mesh_node_width(rp);
    const uint32_t scene_depth = c->scene_height + 1, mesh_depth = mesh_stack_depth(c, width);
    const size_t lds = stack_lds_bytes(scene_depth, mesh_depth, width, 64);
    if (lds > 160 * 1024) return fail(c, SLIMRT_E_ARG, "BVH too deep for the LDS traversal stacks");
    if (check_width(c, width)) return SLIMRT_E_ARG;
    // training waves are a hashed 1-in-N subset: twice the expected share (at least 64 waves) bounds the samples
    uint64_t t_cap = 0;
    if (np->train_one_in) {
        const uint64_t waves = ((uint64_t)tiles + np->train_one_in - 1) / np->train_one_in;
        t_cap = std::min<uint64_t>((uint64_t)tiles, 2 * waves + 64) * 64 * p->max_depth;
        t_cap = std::min<uint64_t>(t_cap, kLinkNone - 1);
    }
    bool cells_new;
    if ((rc = ensure_nrc(c, samples, (uint32_t)std::max<uint64_t>(t_cap, 64), np->gate_cells_log2, cells_new))) return rc;
    if ((rc = slimrt_nrc_set_stream(nrc, c->stream))) return fail(c, SLIMRT_E_STATE, "cache stream");

    NrcFrame nf;
    nf.q_raw = c->q_raw; nf.q_pixel = c->q_pixel; nf.q_prefix = c->q_prefix; nf.q_tp = c->q_tp; nf.q_cap = samples;
    nf.t_raw = c->t_raw; nf.t_target = c->t_target; nf.t_next = c->t_next; nf.t_link = c->t_link;
    nf.t_cap = (uint32_t)t_cap;
    c->nrc_frame_q_cap = nf.q_cap;
    c->nrc_frame_t_cap = nf.t_cap;
    nf.counters = c->nrc_counters;
    nf.cells = c->cells;
    nf.cell_mask = (1u << np->gate_cells_log2) - 1u;
    nf.cell_inv_size = 1.0f / np->gate_cell_size;
    nf.gate_threshold = np->gate_threshold;
    nf.gate_min_frames = np->gate_min_frames;
    nf.cache_vertex = np->cache_depth + 1;
    nf.train_one_in = np->train_one_in;
    nf.frame = np->frame;
    const uint32_t ncells = 1u << np->gate_cells_log2;

    c->timed = true;
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    if (cells_new || (np->flags & SLIMRT_NRC_RESET_GATE)) {
        hipLaunchKernelGGL(nrc_gate_clear_kernel, dim3(1024), dim3(256), 0, c->stream, c->cells, ncells * kGateFloats);
        HIP_TRY(c, hipGetLastError());
    }
    hipLaunchKernelGGL(nrc_reset_kernel, dim3(1), dim3(64), 0, c->stream, c->nrc_counters, c->nrc_losses, kMaxTrainSteps);
    HIP_TRY(c, hipGetLastError());
#define NRC_FRAME_LAUNCH(NW)                                                                                    \
    hipLaunchKernelGGL(nrc_frame_kernel<NW>, dim3(tiles), dim3(64), lds, c->stream, device_view(c), rp, nf,          \
                       (float4 *)d_pixels, (float *)d_depths, scene_depth, mesh_depth, tiles_x, tiles, c->d_status)
    if (width == 8) NRC_FRAME_LAUNCH(8u); else if (width == 4) NRC_FRAME_LAUNCH(4u); else NRC_FRAME_LAUNCH(2u);
#undef NRC_FRAME_LAUNCH
    HIP_TRY(c, hipGetLastError());
    if (np->cache_depth + 1 <= p->max_depth) {
        if (slimrt_nrc_infer_queue(nrc, c->q_raw, c->nrc_counters + NC_QUERIES, samples, c->q_out))
            return fail(c, SLIMRT_E_DEVICE, std::string("cache inference: ") + slimrt_nrc_last_error(nrc));
        const uint32_t blocks = std::min<uint32_t>((samples + 255) / 256, 4096);
        hipLaunchKernelGGL(nrc_composite_kernel, dim3(blocks), dim3(256), 0, c->stream, rp,
                           (const uint32_t *)(c->nrc_counters + NC_QUERIES), samples, (const uint32_t *)c->q_pixel,
                           (const float *)c->q_prefix, (const float *)c->q_tp, (const float *)c->q_out, (float4 *)d_pixels,
                           c->d_status);
        HIP_TRY(c, hipGetLastError());
    }
    c->nrc_last_steps = 0;
    if (train) {
        if (slimrt_nrc_train_queue(nrc, c->t_raw, c->t_target, c->nrc_counters + NC_SAMPLES, (uint32_t)t_cap,
                                   np->train_steps, np->train_batch, np->learning_rate, c->nrc_losses))
            return fail(c, SLIMRT_E_DEVICE, std::string("cache training: ") + slimrt_nrc_last_error(nrc));
        c->nrc_last_steps = np->train_steps;
    }
    if (np->train_one_in) {
        hipLaunchKernelGGL(nrc_gate_fold_kernel, dim3(std::min<uint32_t>((ncells + 255) / 256, 4096)), dim3(256), 0,
                           c->stream, c->cells, ncells);
        HIP_TRY(c, hipGetLastError());
    }
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    return SLIMRT_OK;
}

int slimrt_nrc_frame_samples(slimrt_context *c, float *raw, float *targets, uint32_t *pixel, uint32_t *vertex,
                             uint32_t cap, uint32_t *count) {
    if (!c || !count || (cap && (!raw || !targets || !pixel || !vertex))) return SLIMRT_E_ARG;
    if (c->cpu || !c->nrc_counters) return fail(c, SLIMRT_E_STATE, "no frame rendered with the cache");
    uint32_t cnt[NC_COUNT];
    HIP_TRY(c, hipMemcpyAsync(cnt, c->nrc_counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint32_t n = std::min(cnt[NC_SAMPLES], c->nrc_frame_t_cap);
    *count = n;
    const uint32_t m = std::min(n, cap);
    if (m) {
        std::vector<float> next((size_t)m * 3);
        HIP_TRY(c, hipMemcpyAsync(raw, c->t_raw, (size_t)m * kNrcRaw * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(targets, c->t_target, (size_t)m * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(pixel, c->t_link, (size_t)m * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(next.data(), c->t_next, next.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (uint32_t i = 0; i < m; i++) std::memcpy(&vertex[i], &next[3 * (size_t)i], 4);
    }
    return check_status(c);
}

int slimrt_nrc_frame_stats(slimrt_context *c, uint32_t stats[8]) {
    if (!c || !stats) return SLIMRT_E_ARG;
    if (c->cpu || !c->nrc_counters) return fail(c, SLIMRT_E_STATE, "no frame rendered with the cache");
    uint32_t cnt[NC_COUNT];
    float losses[kMaxTrainSteps];
    HIP_TRY(c, hipMemcpyAsync(cnt, c->nrc_counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(losses, c->nrc_losses, sizeof losses, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    stats[0] = std::min(cnt[NC_QUERIES], c->nrc_frame_q_cap);
    stats[1] = std::min(cnt[NC_SAMPLES], c->nrc_frame_t_cap);
    stats[2] = cnt[NC_FALLBACKS];
    stats[3] = cnt[NC_TRAIN_PIXELS];
    const float first = c->nrc_last_steps ? losses[0] : 0.0f, last = c->nrc_last_steps ? losses[c->nrc_last_steps - 1] : 0.0f;
    std::memcpy(&stats[4], &first, 4);
    std::memcpy(&stats[5], &last, 4);
    stats[6] = stats[7] = 0;
    return check_status(c);
}

}  // extern "C"
