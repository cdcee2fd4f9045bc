// Neural radiance cache on gfx950: fused BF16 MFMA inference, fused forward/backward with per-workgroup partial
// gradients, and a reduce + Adam kernel (include/slimrt_nrc.h).
//
// Matrix products use v_mfma_f32_16x16x16_bf16 with weights as the A operand and activations as the B operand
// (16 queries per tile).  With that orientation the accumulator layout of one layer - lane l holding rows
// 4*(l/16)+j of column l%16 - is exactly the B-operand layout of the next layer, so the forward pass chains
// layers in registers.  Weights are pre-swizzled into 512-byte A fragments (64 lanes x 4 bf16): "forward"
// fragments of W^T for the forward pass and "backward" fragments of W for the input gradients.
// The weight gradient sums over samples: dW^T = dZ X^T is computed with K = samples from LDS copies of the
// layer's inputs and output gradients stored as [feature][sample], so every operand read is 8 contiguous bytes.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <cstring>

#include "../../include/slimrt_nrc.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 96, NH = 64, NOUT = 3, NOUTP = 16, NBANDS = 12, NRAW = 14;
constexpr int W0 = 0, B0 = W0 + NIN * NH, W1 = B0 + NH, B1 = W1 + NH * NH, W2 = B1 + NH, B2 = W2 + NH * NH,
              W3 = B2 + NH, B3 = W3 + NH * NH, W4 = B3 + NH, B4 = W4 + NH * NOUTP, NPARAMS = B4 + NOUTP;
static_assert(NPARAMS == (int)SLIMRT_NRC_PARAMS, "parameter count");
constexpr int FWD_FRAGS = 24 + 3 * 16 + 4;   // layer 0: 4 out tiles x 6 in chunks; 1-3: 4x4; 4: 1x4
constexpr int BWD_FRAGS = 3 * 16 + 4;        // layers 1-3: 4 in tiles x 4 out chunks; 4: 4 in tiles x 1

__host__ __device__ inline int layer_w(int li) { return li == 0 ? W0 : li == 1 ? W1 : li == 2 ? W2 : li == 3 ? W3 : W4; }
__host__ __device__ inline int layer_b(int li) { return li == 0 ? B0 : li == 1 ? B1 : li == 2 ? B2 : li == 3 ? B3 : B4; }
__host__ __device__ inline int layer_fo(int li) { return li == 4 ? NOUTP : NH; }
__host__ __device__ inline int fwd_base(int li) { return li == 0 ? 0 : 24 + 16 * (li - 1); }

// Gradient partials are stored in "slot" order, as the MFMA accumulators hold them: weight tile t (numbered as
// the forward fragments: layer 0 tile mt*6+nt, layers 1-3 24+16(li-1)+mt*4+nt, layer 4 72+nt) occupies slots
// 256t .. 256t+255, lane l's four outputs j at 256t + 4l + j; the 272 biases follow in parameter order.
constexpr int kWeightSlots = 76 * 256;
static_assert(kWeightSlots + 4 * NH + NOUTP == NPARAMS, "every parameter has one slot");
__host__ __device__ inline int bias_slot(int li, int o) { return kWeightSlots + (li < 4 ? li * NH + o : 4 * NH + o); }
__host__ __device__ inline int slot_param(int s) {
    if (s >= kWeightSlots) {
        const int b = s - kWeightSlots;
        return b < 4 * NH ? layer_b(b / NH) + b % NH : B4 + (b - 4 * NH);
    }
    const int tile = s >> 8, lane = (s >> 2) & 63, j = s & 3;
    int li, mt, nt;
    if (tile < 24) { li = 0; mt = tile / 6; nt = tile % 6; }
    else if (tile < 72) { li = 1 + (tile - 24) / 16; mt = ((tile - 24) % 16) / 4; nt = (tile - 24) % 4; }
    else { li = 4; mt = 0; nt = tile - 72; }
    const int o = 16 * mt + 4 * (lane >> 4) + j, i = 16 * nt + (lane & 15);
    return layer_w(li) + i * layer_fo(li) + o;
}
__host__ __device__ inline int bwd_base(int li) { return 16 * (li - 1); }

__device__ __forceinline__ unsigned short to_bf16(float x) {   // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x4 as_bf16x4(uint2 v) { return __builtin_bit_cast(bf16x4, v); }

// Encoded input feature f of a query (frequency bands of the normalised position, then the raw attributes).
// Band b of coordinate p is sin and cos of 2^b * pi * p, evaluated as one turn times the fractional part of
// 2^(b-1) * p: the scaling by a power of two and the fractional part are exact in float32, so the hardware
// sine and cosine (which take turns) see a reduced argument in [0, 1].
__device__ __forceinline__ float feature(int f, const float *raw, const float *pn) {
    if (f < 72) {
        const int d = f / 24, r = f % 24, b = r >> 1;
        const float x = pn[d] * __builtin_bit_cast(float, (uint32_t)(126 + b) << 23);   // * 2^(b-1)
        const float turns = x - floorf(x);
        return (r & 1) ? __builtin_amdgcn_cosf(turns) : __builtin_amdgcn_sinf(turns);
    }
    if (f < 83) return raw[3 + (f - 72)];
    return 0.0f;
}

// Loads query q's raw record and builds this lane's six B-operand chunks of the layer-0 input:
// chunk kc holds features 16kc + 4*(lane/16) + j, j = 0..3, of query lane%16.
__device__ __forceinline__ void encode(const float *__restrict__ raw, uint32_t q, bool valid, float3 center,
                                       float3 inv_half, int lane, bf16x4 x[6]) {
    float r[NRAW];
    for (int i = 0; i < NRAW; i++) r[i] = valid ? raw[(size_t)q * NRAW + i] : 0.0f;
    float pn[3] = {(r[0] - center.x) * inv_half.x, (r[1] - center.y) * inv_half.y, (r[2] - center.z) * inv_half.z};
    for (int kc = 0; kc < 6; kc++)
        for (int j = 0; j < 4; j++) {
            const int f = 16 * kc + 4 * (lane >> 4) + j;
            x[kc][j] = (short)to_bf16(valid ? feature(f, r, pn) : 0.0f);
        }
}

// ---------------------------------------------------------------- weight packing
__global__ void pack_kernel(const float *__restrict__ p, uint2 *__restrict__ fwd, uint2 *__restrict__ bwd) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = t >> 6, l = t & 63;
    if (f >= FWD_FRAGS + BWD_FRAGS) return;
    unsigned short h[4];
    if (f < FWD_FRAGS) {   // A = W^T: lane l, j -> W[in = 16kc + 4(l/16) + j][out = 16mt + l%16]
        int li, mt, kc, kcn;
        if (f < 24) { li = 0; mt = f / 6; kc = f % 6; kcn = 6; }
        else if (f < 72) { li = 1 + (f - 24) / 16; int r = (f - 24) % 16; mt = r / 4; kc = r % 4; kcn = 4; }
        else { li = 4; mt = 0; kc = f - 72; kcn = 4; }
        (void)kcn;
        const int fo = layer_fo(li), wo = layer_w(li);
        for (int j = 0; j < 4; j++) h[j] = to_bf16(p[wo + (16 * kc + 4 * (l >> 4) + j) * fo + 16 * mt + (l & 15)]);
        fwd[f * 64 + l] = make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
    } else {               // A = W: lane l, j -> W[in = 16mt + l%16][out = 16kc + 4(l/16) + j]
        const int g = f - FWD_FRAGS;
        int li, mt, kc;
        if (g < 48) { li = 1 + g / 16; int r = g % 16; mt = r / 4; kc = r % 4; }
        else { li = 4; mt = g - 48; kc = 0; }
        const int fo = layer_fo(li), wo = layer_w(li);
        for (int j = 0; j < 4; j++) h[j] = to_bf16(p[wo + (16 * mt + (l & 15)) * fo + 16 * kc + 4 * (l >> 4) + j]);
        bwd[g * 64 + l] = make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
    }
}

// ---------------------------------------------------------------- inference
// 256 threads = 4 waves; each wave runs 16-query tiles through the network with the weights staged in LDS.
// With n_dev, the query count is min(*n_dev, n) (a queue filled on the device).
__global__ __launch_bounds__(256) void infer_kernel(const uint2 *__restrict__ frag_fwd, const float *__restrict__ p,
                                                    const float *__restrict__ raw, uint32_t n,
                                                    const uint32_t *__restrict__ n_dev, float3 center,
                                                    float3 inv_half, float *__restrict__ out) {
    if (n_dev) n = min(*n_dev, n);
    if (n == 0) return;
    __shared__ uint2 sw[FWD_FRAGS * 64];
    __shared__ float sb[4 * NH + NOUTP];
    for (int i = threadIdx.x; i < FWD_FRAGS * 64; i += 256) sw[i] = frag_fwd[i];
    for (int i = threadIdx.x; i < 4 * NH + NOUTP; i += 256) sb[i] = p[layer_b(i / NH) + (i % NH)];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t tiles = (n + 15) / 16;
    for (uint32_t tile = blockIdx.x * 4 + wave; tile < tiles; tile += gridDim.x * 4) {
        const uint32_t q = tile * 16 + (lane & 15);
        const bool valid = q < n;
        bf16x4 x[6], h[4];
        encode(raw, q, valid, center, inv_half, lane, x);
        for (int li = 0; li < 4; li++) {
            const int kcn = li == 0 ? 6 : 4;
            f32x4 acc[4];
            for (int mt = 0; mt < 4; mt++) {
                acc[mt] = f32x4{0, 0, 0, 0};
                for (int kc = 0; kc < kcn; kc++)
                    acc[mt] = mfma(as_bf16x4(sw[(fwd_base(li) + mt * kcn + kc) * 64 + lane]), li == 0 ? x[kc] : h[kc], acc[mt]);
            }
            for (int mt = 0; mt < 4; mt++)
                for (int j = 0; j < 4; j++) {
                    float z = acc[mt][j] + sb[li * NH + 16 * mt + 4 * (lane >> 4) + j];
                    h[mt][j] = (short)to_bf16(z > 0.0f ? z : 0.0f);
                }
        }
        f32x4 y = f32x4{0, 0, 0, 0};
        for (int kc = 0; kc < 4; kc++) y = mfma(as_bf16x4(sw[(72 + kc) * 64 + lane]), h[kc], y);
        if (lane < 16 && valid)
            for (int c = 0; c < NOUT; c++) out[(size_t)q * NOUT + c] = y[c] + sb[4 * NH + c];
    }
}

// ---------------------------------------------------------------- training: forward + backward, per-workgroup partials
// One workgroup of four waves per 64 samples (4 tiles of 16).  LDS: the weight fragments and biases, each
// layer's input X_l and two output-gradient buffers, all bf16 [feature][sample].
// One row per feature, 64 samples plus 8 halves of padding: rows 36 dwords apart spread a wave's 8-byte
// operand reads (16 rows x 4 sample groups) over all 64 LDS banks.
constexpr int kRow = 72;
struct TrainLds {
    uint2 wf[FWD_FRAGS * 64];          // forward and backward weight fragments, staged once per wave
    uint2 wb[BWD_FRAGS * 64];
    float bias[4 * NH + NOUTP];
    float bpart[2][4][NH];             // per-wave bias-gradient sums (double-buffered by layer parity)
    float b4part[4][NOUTP];
    float loss[4];
    unsigned short x0[NIN * kRow];
    unsigned short x[4][NH * kRow];   // inputs of layers 1..4 (outputs of hidden layers 0..3)
    unsigned short dz[2][NH * kRow];
};

__device__ __forceinline__ bf16x4 lds_bf16x4(const unsigned short *base) {   // 8 contiguous bytes
    return __builtin_bit_cast(bf16x4, *(const uint2 *)base);
}

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// Sum over the 16 lanes that share lane/16, as the butterfly v += v[lane ^ k] for k = 1, 2, 4, 8 would add:
// quad_perm [1,0,3,2] and [2,3,0,1] pair lanes within quads, then row_half_mirror and row_mirror bring in the
// other quad and the other half-row (whose lanes all hold the same partial sums by then).
__device__ __forceinline__ float sum16(float v) {
    v += dpp<0xB1>(v);
    v += dpp<0x4E>(v);
    v += dpp<0x141>(v);
    v += dpp<0x140>(v);
    return v;
}

// Batch slot j of a step maps to sample map(j) of the buffer.  Without a device count the batch is samples
// 0..n-1.  With one (count = min(*n_dev, cap)), slot offset + j of the frame's `total` slots takes sample
// (offset + j) * count / total when the buffer holds at least `total` samples (a frame's steps spread over the
// whole buffer), else sample (offset + j) mod count (every step full, samples repeated).  An empty buffer
// gives an empty batch: zero gradient.
struct SampleMap {
    uint32_t n_eff, count, offset, total;
    int mode;   // 0: identity, 1: stride, 2: wrap
    __device__ uint32_t operator()(uint32_t j) const {
        return mode == 1 ? (uint32_t)(((uint64_t)(offset + j) * count) / total)
                         : mode == 2 ? (offset + j) % count : j;
    }
};

__device__ __forceinline__ SampleMap sample_map(uint32_t n, const uint32_t *n_dev, uint32_t cap, uint32_t offset,
                                                uint32_t total) {
    SampleMap m{n, n, 0, n, 0};
    if (n_dev) {
        const uint32_t count = min(*n_dev, cap);
        m.count = count;
        m.offset = offset;
        m.total = total;
        m.mode = count >= total ? 1 : 2;
        m.n_eff = count ? n : 0u;
    }
    return m;
}

// Hidden layer LI of the forward pass for one 16-sample tile: in (B operands) -> out (next layer's B operands),
// with the bf16 activations also stored to LDS for the backward pass.  Every index is a compile-time constant,
// so the operand arrays stay in registers.
template <int LI>
__device__ __forceinline__ void train_fwd_layer(TrainLds &s, int lane, int sq, const bf16x4 *in, bf16x4 *out) {
    constexpr int KCN = LI == 0 ? 6 : 4;
    f32x4 acc[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        acc[mt] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int kc = 0; kc < KCN; kc++)
            acc[mt] = mfma(as_bf16x4(s.wf[(fwd_base(LI) + mt * KCN + kc) * 64 + lane]), in[kc], acc[mt]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = 16 * mt + 4 * (lane >> 4) + j;
            const float z = acc[mt][j] + s.bias[LI * NH + o];
            const unsigned short hb = to_bf16(z > 0.0f ? z : 0.0f);
            out[mt][j] = (short)hb;
            s.x[LI][o * kRow + sq] = hb;
        }
}

// Backward step of layer LI for the workgroup's 64 samples: its weight gradient (dW^T = dZ X^T, 16x16 tiles dealt
// round-robin to the four waves) and, for LI > 0, the output gradient of layer LI-1 (W dZ masked by that layer's
// ReLU; wave w takes samples 16w..16w+15) with its bias gradient (per-wave sums combined in wave order).
template <int LI>
__device__ __forceinline__ void train_bwd_layer(TrainLds &s, int wave, int lane, float *g) {
    constexpr int FO = LI == 4 ? NOUTP : NH, FI = LI == 0 ? NIN : NH, CUR = (4 - LI) & 1;
    constexpr int NT = FI / 16, TILES = (FO / 16) * NT;
    const unsigned short *X = LI == 0 ? s.x0 : s.x[LI == 0 ? 0 : LI - 1];
    const unsigned short *DZ = s.dz[CUR];
#pragma unroll
    for (int r = 0; r < TILES / 4; r++) {
        const int k = wave + 4 * r, mt = k / NT, nt = k % NT;
        f32x4 acc = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const int so = 16 * kk + 4 * (lane >> 4);
            acc = mfma(lds_bf16x4(DZ + (16 * mt + (lane & 15)) * kRow + so),
                       lds_bf16x4(X + (16 * nt + (lane & 15)) * kRow + so), acc);
        }
        constexpr int base = LI == 0 ? 0 : LI < 4 ? 24 + 16 * (LI - 1) : 72;
        *(f32x4 *)(g + (base + k) * 256 + lane * 4) = acc;
    }
    if constexpr (LI > 0) {
        constexpr int KCN = FO / 16;
        unsigned short *DZN = s.dz[CUR ^ 1];
        const int sq = 16 * wave + (lane & 15);
        bf16x4 b[KCN];
#pragma unroll
        for (int kc = 0; kc < KCN; kc++)
#pragma unroll
            for (int j = 0; j < 4; j++) b[kc][j] = (short)DZ[(16 * kc + 4 * (lane >> 4) + j) * kRow + sq];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            f32x4 acc = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int kc = 0; kc < KCN; kc++) {
                const int frag = LI == 4 ? 48 + mt : bwd_base(LI) + mt * 4 + kc;
                acc = mfma(as_bf16x4(s.wb[frag * 64 + lane]), b[kc], acc);
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int i = 16 * mt + 4 * (lane >> 4) + j;
                const float dzv = from_bf16(X[i * kRow + sq]) > 0.0f ? acc[j] : 0.0f;
                DZN[i * kRow + sq] = to_bf16(dzv);
                const float sum = sum16(dzv);
                if ((lane & 15) == 0) s.bpart[LI & 1][wave][i] = sum;
            }
        }
    }
    __syncthreads();
    if constexpr (LI > 0) {
        if (wave == 0) {
            const float (*bp)[NH] = s.bpart[LI & 1];
            g[bias_slot(LI - 1, lane)] = ((bp[0][lane] + bp[1][lane]) + bp[2][lane]) + bp[3][lane];
        }
    }
}

// One workgroup of four waves trains on 64 samples; wave w runs samples 16w..16w+15 through the forward pass.
__global__ __launch_bounds__(256) void train_kernel(const uint2 *__restrict__ frag_fwd, const uint2 *__restrict__ frag_bwd,
                                                    const float *__restrict__ p, const float *__restrict__ raw,
                                                    const float *__restrict__ targets, uint32_t n,
                                                    const uint32_t *__restrict__ n_dev, uint32_t cap, uint32_t offset,
                                                    uint32_t total, float3 center, float3 inv_half,
                                                    float *__restrict__ partial, float *__restrict__ partial_loss) {
    __shared__ TrainLds s;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t base = blockIdx.x * 64;
    const SampleMap map = sample_map(n, n_dev, cap, offset, total);
    const float inv_count = map.n_eff ? 1.0f / (3.0f * (float)map.n_eff) : 0.0f;
    for (int i = tid; i < FWD_FRAGS * 64; i += 256) s.wf[i] = frag_fwd[i];
    for (int i = tid; i < BWD_FRAGS * 64; i += 256) s.wb[i] = frag_bwd[i];
    for (int i = tid; i < 4 * NH + NOUTP; i += 256) s.bias[i] = p[layer_b(i / NH) + (i % NH)];
    __syncthreads();
    float *g = partial + (size_t)blockIdx.x * NPARAMS;

    // ---- forward: this wave's tile of 16 samples
    {
        const int sq = 16 * wave + (lane & 15);
        const bool valid = base + sq < map.n_eff;
        const uint32_t q = valid ? map(base + sq) : 0u;
        bf16x4 x[6], h0[4], h1[4];
        encode(raw, q, valid, center, inv_half, lane, x);
#pragma unroll
        for (int kc = 0; kc < 6; kc++)
#pragma unroll
            for (int j = 0; j < 4; j++) s.x0[(16 * kc + 4 * (lane >> 4) + j) * kRow + sq] = (unsigned short)x[kc][j];
        train_fwd_layer<0>(s, lane, sq, x, h0);
        train_fwd_layer<1>(s, lane, sq, h0, h1);
        train_fwd_layer<2>(s, lane, sq, h1, h0);
        train_fwd_layer<3>(s, lane, sq, h0, h1);
        f32x4 y = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int kc = 0; kc < 4; kc++) y = mfma(as_bf16x4(s.wf[(72 + kc) * 64 + lane]), h1[kc], y);
        // output gradient of the relative L2 loss; lanes 0..15 hold outputs j = 0..3 of sample lane
        f32x4 d = f32x4{0, 0, 0, 0};
        float loss = 0.0f;
        if (lane < 16 && valid) {
#pragma unroll
            for (int c = 0; c < NOUT; c++) {
                const float yc = y[c] + s.bias[4 * NH + c], tc = targets[(size_t)q * NOUT + c];
                const float den = yc * yc + 0.01f;
                const float e = yc - tc;
                loss += e * e / den;
                d[c] = 2.0f * e / den * inv_count;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = 4 * (lane >> 4) + j;
            s.dz[0][o * kRow + sq] = to_bf16(d[j]);   // rows 0..15 of buffer 0
            const float sum = sum16(d[j]);             // bias 4 gradient: fp32 sums of the output gradient
            if ((lane & 15) == 0) s.b4part[wave][o] = sum;
        }
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
        if (lane == 0) s.loss[wave] = loss;
    }
    __syncthreads();
    if (wave == 0 && lane < NOUTP)
        g[bias_slot(4, lane)] = ((s.b4part[0][lane] + s.b4part[1][lane]) + s.b4part[2][lane]) + s.b4part[3][lane];
    if (tid == 0) partial_loss[blockIdx.x] = (((s.loss[0] + s.loss[1]) + s.loss[2]) + s.loss[3]) * inv_count;

    // ---- backward, layer by layer
    train_bwd_layer<4>(s, wave, lane, g);
    train_bwd_layer<3>(s, wave, lane, g);
    train_bwd_layer<2>(s, wave, lane, g);
    train_bwd_layer<1>(s, wave, lane, g);
    train_bwd_layer<0>(s, wave, lane, g);
}

// Sum of the per-workgroup partials, then (optionally) one Adam step on the float32 master weights.  A workgroup of
// 4 waves owns 64 slots; wave w sums rows w, w+4, ... and the first wave combines the four sums (fixed order:
// the result does not depend on scheduling).
__global__ __launch_bounds__(256) void reduce_adam_kernel(const float *__restrict__ partial, uint32_t nblocks,
                                                          float *__restrict__ p, float *__restrict__ m,
                                                          float *__restrict__ v, float *__restrict__ grad_out,
                                                          int do_adam, float lr, float c1, float c2) {
    __shared__ float part[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    float acc = 0.0f;
    if (s < NPARAMS) {
        uint32_t b = wave;
        for (; b + 12 < nblocks; b += 16) {   // four independent loads in flight
            const float a0 = partial[(size_t)b * NPARAMS + s], a1 = partial[(size_t)(b + 4) * NPARAMS + s];
            const float a2 = partial[(size_t)(b + 8) * NPARAMS + s], a3 = partial[(size_t)(b + 12) * NPARAMS + s];
            acc += a0;
            acc += a1;
            acc += a2;
            acc += a3;
        }
        for (; b < nblocks; b += 4) acc += partial[(size_t)b * NPARAMS + s];
    }
    part[wave][lane] = acc;
    __syncthreads();
    if (wave != 0 || s >= NPARAMS) return;
    float gsum = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
    const int i = slot_param(s);
    const bool pad = (i >= W4 && i < B4 && (i - W4) % NOUTP >= NOUT) || (i >= B4 + NOUT);
    if (pad) gsum = 0.0f;
    if (grad_out) grad_out[i] = gsum;
    if (!do_adam || pad) return;
    const float b1 = 0.9f, b2 = 0.99f;
    const float mi = b1 * m[i] + (1.0f - b1) * gsum;
    const float vi = b2 * v[i] + (1.0f - b2) * gsum * gsum;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    p[i] = p[i] - lr * mh / (sqrtf(vh) + 1e-8f);
}

__global__ __launch_bounds__(64) void loss_sum_kernel(const float *__restrict__ partial_loss, uint32_t nblocks,
                                                      float *__restrict__ out) {
    float s = 0.0f;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 64) s += partial_loss[b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (threadIdx.x == 0) *out = s;
}

uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

struct slimrt_nrc {
    float *d_p = nullptr, *d_m = nullptr, *d_v = nullptr, *d_partial = nullptr, *d_ploss = nullptr, *d_grad = nullptr,
          *d_loss = nullptr;
    uint2 *d_fwd = nullptr, *d_bwd = nullptr;
    uint32_t partial_blocks = 0;
    hipStream_t stream = nullptr;
    int step = 0;
    float3 center, inv_half;
    std::string err;
};

namespace {
int nrc_fail(slimrt_nrc *c, int code, const std::string &m) {
    if (c) c->err = m;
    return code;
}
#define NRC_TRY(c, expr)                                                                             \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return nrc_fail(c, -2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int pack(slimrt_nrc *c) {
    const int threads = (FWD_FRAGS + BWD_FRAGS) * 64;
    hipLaunchKernelGGL(pack_kernel, dim3((threads + 255) / 256), dim3(256), 0, c->stream, c->d_p, c->d_fwd, c->d_bwd);
    NRC_TRY(c, hipGetLastError());
    return 0;
}

int ensure_partials(slimrt_nrc *c, uint32_t blocks) {
    if (blocks <= c->partial_blocks) return 0;
    if (c->d_partial) NRC_TRY(c, hipFree(c->d_partial));
    if (c->d_ploss) NRC_TRY(c, hipFree(c->d_ploss));
    c->d_partial = nullptr;
    c->d_ploss = nullptr;
    NRC_TRY(c, hipMalloc(&c->d_partial, (size_t)blocks * NPARAMS * sizeof(float)));
    NRC_TRY(c, hipMalloc(&c->d_ploss, (size_t)blocks * sizeof(float)));
    c->partial_blocks = blocks;
    return 0;
}

// One gradient evaluation (and optionally one Adam step) on a batch of n samples; see SampleMap for n_dev.
int train(slimrt_nrc *c, const float *raw, const float *targets, uint32_t n, bool adam, float lr, float *grad_host,
          float *loss_host, const uint32_t *n_dev = nullptr, uint32_t cap = 0, uint32_t offset = 0, uint32_t total = 0,
          float *loss_dev = nullptr) {
    if (!c || !raw || !targets || n == 0) return -1;
    const uint32_t blocks = (n + 63) / 64;
    int rc = ensure_partials(c, blocks);
    if (rc) return rc;
    hipLaunchKernelGGL(train_kernel, dim3(blocks), dim3(256), 0, c->stream, c->d_fwd, c->d_bwd, c->d_p, raw, targets, n,
                       n_dev, cap, offset, total, c->center, c->inv_half, c->d_partial, c->d_ploss);
    NRC_TRY(c, hipGetLastError());
    float c1 = 1.0f, c2 = 1.0f;
    if (adam) {
        c->step++;
        c1 = (float)(1.0 - pow(0.9, c->step));
        c2 = (float)(1.0 - pow(0.99, c->step));
    }
    hipLaunchKernelGGL(reduce_adam_kernel, dim3((NPARAMS + 63) / 64), dim3(256), 0, c->stream, c->d_partial, blocks,
                       c->d_p, c->d_m, c->d_v, grad_host ? c->d_grad : nullptr, adam ? 1 : 0, lr, c1, c2);
    NRC_TRY(c, hipGetLastError());
    if (adam && (rc = pack(c))) return rc;
    if (loss_host || loss_dev) {
        hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, c->stream, c->d_ploss, blocks, loss_dev ? loss_dev : c->d_loss);
        NRC_TRY(c, hipGetLastError());
    }
    if (loss_host) NRC_TRY(c, hipMemcpyAsync(loss_host, c->d_loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (grad_host) NRC_TRY(c, hipMemcpyAsync(grad_host, c->d_grad, NPARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (loss_host || grad_host) NRC_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}
}  // namespace

extern "C" {

int slimrt_nrc_create(slimrt_nrc **out, uint32_t seed, const float center[3], const float inv_half[3]) {
    if (!out || !center || !inv_half) return -1;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return -2;
    auto *c = new slimrt_nrc();
    c->center = make_float3(center[0], center[1], center[2]);
    c->inv_half = make_float3(inv_half[0], inv_half[1], inv_half[2]);
    // He-uniform initial weights from a SplitMix64 stream, zero biases, zero padded output columns
    std::vector<float> p(NPARAMS, 0.0f);
    uint64_t st = seed;
    for (int li = 0; li < 5; li++) {
        const int fi = li == 0 ? NIN : NH, fo = layer_fo(li), wo = layer_w(li);
        const float lim = sqrtf(6.0f / (float)fi);
        for (int i = 0; i < fi; i++)
            for (int o = 0; o < fo; o++) {
                const float u = (float)((splitmix(st) >> 40) * (1.0 / 16777216.0));
                p[wo + i * fo + o] = (li == 4 && o >= NOUT) ? 0.0f : (2.0f * u - 1.0f) * lim;
            }
    }
    if (hipMalloc(&c->d_p, NPARAMS * 4) != hipSuccess || hipMalloc(&c->d_m, NPARAMS * 4) != hipSuccess ||
        hipMalloc(&c->d_v, NPARAMS * 4) != hipSuccess || hipMalloc(&c->d_grad, NPARAMS * 4) != hipSuccess ||
        hipMalloc(&c->d_loss, 4) != hipSuccess || hipMalloc(&c->d_fwd, FWD_FRAGS * 64 * sizeof(uint2)) != hipSuccess ||
        hipMalloc(&c->d_bwd, BWD_FRAGS * 64 * sizeof(uint2)) != hipSuccess) {
        slimrt_nrc_destroy(c);
        return -3;
    }
    if (slimrt_nrc_set_params(c, p.data())) {
        slimrt_nrc_destroy(c);
        return -2;
    }
    *out = c;
    return 0;
}

int slimrt_nrc_destroy(slimrt_nrc *c) {
    if (!c) return -1;
    (void)hipStreamSynchronize(c->stream);
    for (void *ptr : {(void *)c->d_p, (void *)c->d_m, (void *)c->d_v, (void *)c->d_partial, (void *)c->d_ploss,
                      (void *)c->d_grad, (void *)c->d_loss, (void *)c->d_fwd, (void *)c->d_bwd})
        if (ptr) (void)hipFree(ptr);
    delete c;
    return 0;
}

int slimrt_nrc_set_stream(slimrt_nrc *c, void *stream) {
    if (!c) return -1;
    c->stream = (hipStream_t)stream;
    return 0;
}

int slimrt_nrc_get_params(slimrt_nrc *c, float *params) {
    if (!c || !params) return -1;
    NRC_TRY(c, hipMemcpyAsync(params, c->d_p, NPARAMS * 4, hipMemcpyDeviceToHost, c->stream));
    NRC_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

int slimrt_nrc_set_params(slimrt_nrc *c, const float *params) {
    if (!c || !params) return -1;
    NRC_TRY(c, hipMemcpyAsync(c->d_p, params, NPARAMS * 4, hipMemcpyHostToDevice, c->stream));
    NRC_TRY(c, hipMemsetAsync(c->d_m, 0, NPARAMS * 4, c->stream));
    NRC_TRY(c, hipMemsetAsync(c->d_v, 0, NPARAMS * 4, c->stream));
    c->step = 0;
    int rc = pack(c);
    if (rc) return rc;
    NRC_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

int slimrt_nrc_get_state(slimrt_nrc *c, float *params, float *m, float *v, uint32_t *step) {
    if (!c) return -1;
    if (params) NRC_TRY(c, hipMemcpyAsync(params, c->d_p, NPARAMS * 4, hipMemcpyDeviceToHost, c->stream));
    if (m) NRC_TRY(c, hipMemcpyAsync(m, c->d_m, NPARAMS * 4, hipMemcpyDeviceToHost, c->stream));
    if (v) NRC_TRY(c, hipMemcpyAsync(v, c->d_v, NPARAMS * 4, hipMemcpyDeviceToHost, c->stream));
    NRC_TRY(c, hipStreamSynchronize(c->stream));
    if (step) *step = (uint32_t)c->step;
    return 0;
}

int slimrt_nrc_set_state(slimrt_nrc *c, const float *params, const float *m, const float *v, uint32_t step) {
    if (!c || !params || !m || !v || step > 0x7FFFFFFFu) return -1;
    NRC_TRY(c, hipMemcpyAsync(c->d_p, params, NPARAMS * 4, hipMemcpyHostToDevice, c->stream));
    NRC_TRY(c, hipMemcpyAsync(c->d_m, m, NPARAMS * 4, hipMemcpyHostToDevice, c->stream));
    NRC_TRY(c, hipMemcpyAsync(c->d_v, v, NPARAMS * 4, hipMemcpyHostToDevice, c->stream));
    c->step = (int)step;
    int rc = pack(c);
    if (rc) return rc;
    NRC_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

int slimrt_nrc_infer(slimrt_nrc *c, const float *raw, uint32_t n, float *out) {
    if (!c || (!raw && n) || (!out && n)) return -1;
    if (n == 0) return 0;
    const uint32_t tiles = (n + 15) / 16;
    uint32_t blocks = (tiles + 3) / 4;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(infer_kernel, dim3(blocks), dim3(256), 0, c->stream, c->d_fwd, c->d_p, raw, n,
                       (const uint32_t *)nullptr, c->center, c->inv_half, out);
    NRC_TRY(c, hipGetLastError());
    return 0;
}

int slimrt_nrc_infer_queue(slimrt_nrc *c, const float *raw, const uint32_t *d_count, uint32_t cap, float *out) {
    if (!c || !raw || !d_count || !out || cap == 0) return nrc_fail(c, -1, "infer_queue: NULL buffer or zero capacity");
    const uint32_t tiles = (cap + 15) / 16;
    uint32_t blocks = (tiles + 3) / 4;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(infer_kernel, dim3(blocks), dim3(256), 0, c->stream, c->d_fwd, c->d_p, raw, cap, d_count,
                       c->center, c->inv_half, out);
    NRC_TRY(c, hipGetLastError());
    return 0;
}

int slimrt_nrc_train_queue(slimrt_nrc *c, const float *raw, const float *targets, const uint32_t *d_count, uint32_t cap,
                           uint32_t steps, uint32_t batch, float lr, float *d_losses) {
    if (!c || !raw || !targets || !d_count || cap == 0) return nrc_fail(c, -1, "train_queue: NULL buffer or zero capacity");
    if (steps == 0 || batch == 0 || batch % 64 || (uint64_t)steps * batch > 0xFFFFFFFFull)
        return nrc_fail(c, -1, "train_queue: batch must be a positive multiple of 64");
    for (uint32_t st = 0; st < steps; st++) {
        int rc = train(c, raw, targets, batch, true, lr, nullptr, nullptr, d_count, cap, st * batch, steps * batch,
                       d_losses ? d_losses + st : nullptr);
        if (rc) return rc;
    }
    return 0;
}

int slimrt_nrc_gradients(slimrt_nrc *c, const float *raw, const float *targets, uint32_t n, float *grad, float *loss) {
    float dummy;
    return train(c, raw, targets, n, false, 0.0f, grad, loss ? loss : &dummy);
}

int slimrt_nrc_train_step(slimrt_nrc *c, const float *raw, const float *targets, uint32_t n, float lr, float *loss) {
    return train(c, raw, targets, n, true, lr, nullptr, loss);
}

const char *slimrt_nrc_last_error(const slimrt_nrc *c) { return c ? c->err.c_str() : "null context"; }

}  // extern "C"
