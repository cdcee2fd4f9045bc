// Neural radiance cache on gfx950: fused BF16 MFMA inference, fused forward/backward with per-wave partial
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
constexpr float kPi = 3.14159265f;

__host__ __device__ inline int layer_w(int li) { return li == 0 ? W0 : li == 1 ? W1 : li == 2 ? W2 : li == 3 ? W3 : W4; }
__host__ __device__ inline int layer_b(int li) { return li == 0 ? B0 : li == 1 ? B1 : li == 2 ? B2 : li == 3 ? B3 : B4; }
__host__ __device__ inline int layer_fo(int li) { return li == 4 ? NOUTP : NH; }
__host__ __device__ inline int fwd_base(int li) { return li == 0 ? 0 : 24 + 16 * (li - 1); }
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
__device__ __forceinline__ float feature(int f, const float *raw, const float *pn) {
    if (f < 72) {
        const int d = f / 24, r = f % 24, b = r >> 1;
        const float arg = ((float)(1 << b) * kPi) * pn[d];
        return (r & 1) ? cosf(arg) : sinf(arg);
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

// ---------------------------------------------------------------- training: forward + backward, per-wave partials
// One wave per workgroup, 64 samples (4 tiles of 16).  LDS: each layer's input X_l and two output-gradient
// buffers, all bf16 [feature][sample].
struct TrainLds {
    unsigned short x0[NIN * 64];
    unsigned short x[4][NH * 64];   // inputs of layers 1..4 (outputs of hidden layers 0..3)
    unsigned short dz[2][NH * 64];
};

__device__ __forceinline__ bf16x4 lds_bf16x4(const unsigned short *base) {   // 8 contiguous bytes
    return __builtin_bit_cast(bf16x4, *(const uint2 *)base);
}

__device__ __forceinline__ float sum16(float v) {   // sum over the 16 lanes that share lane/16
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
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

__global__ __launch_bounds__(64) void train_kernel(const uint2 *__restrict__ frag_fwd, const uint2 *__restrict__ frag_bwd,
                                                   const float *__restrict__ p, const float *__restrict__ raw,
                                                   const float *__restrict__ targets, uint32_t n,
                                                   const uint32_t *__restrict__ n_dev, uint32_t cap, uint32_t offset,
                                                   uint32_t total, float3 center, float3 inv_half,
                                                   float *__restrict__ partial, float *__restrict__ partial_loss) {
    __shared__ TrainLds s;
    const int lane = threadIdx.x;
    const uint32_t base = blockIdx.x * 64;
    const SampleMap map = sample_map(n, n_dev, cap, offset, total);
    const float inv_count = map.n_eff ? 1.0f / (3.0f * (float)map.n_eff) : 0.0f;
    float *g = partial + (size_t)blockIdx.x * NPARAMS;
    float loss = 0.0f;
    bf16x4 dz4[4];                       // layer-4 output gradient per tile (B operand of dX4)

    // ---- forward, 4 tiles
    for (int t = 0; t < 4; t++) {
        const int sq = 16 * t + (lane & 15);         // sample within the wave
        const bool valid = base + sq < map.n_eff;
        const uint32_t q = valid ? map(base + sq) : 0u;
        bf16x4 x[6], h[4];
        encode(raw, q, valid, center, inv_half, lane, x);
        for (int kc = 0; kc < 6; kc++)
            for (int j = 0; j < 4; j++) s.x0[(16 * kc + 4 * (lane >> 4) + j) * 64 + sq] = (unsigned short)x[kc][j];
        for (int li = 0; li < 4; li++) {
            const int kcn = li == 0 ? 6 : 4;
            f32x4 acc[4];
            for (int mt = 0; mt < 4; mt++) {
                acc[mt] = f32x4{0, 0, 0, 0};
                for (int kc = 0; kc < kcn; kc++)
                    acc[mt] = mfma(as_bf16x4(frag_fwd[(fwd_base(li) + mt * kcn + kc) * 64 + lane]), li == 0 ? x[kc] : h[kc], acc[mt]);
            }
            for (int mt = 0; mt < 4; mt++)
                for (int j = 0; j < 4; j++) {
                    const int o = 16 * mt + 4 * (lane >> 4) + j;
                    float z = acc[mt][j] + p[layer_b(li) + o];
                    unsigned short hb = to_bf16(z > 0.0f ? z : 0.0f);
                    h[mt][j] = (short)hb;
                    s.x[li][o * 64 + sq] = hb;
                }
        }
        f32x4 y = f32x4{0, 0, 0, 0};
        for (int kc = 0; kc < 4; kc++) y = mfma(as_bf16x4(frag_fwd[(72 + kc) * 64 + lane]), h[kc], y);
        // output gradient of the relative L2 loss; lanes 0..15 hold outputs j = 0..3 of sample lane
        f32x4 d = f32x4{0, 0, 0, 0};
        if (lane < 16 && valid) {
            for (int c = 0; c < NOUT; c++) {
                const float yc = y[c] + p[B4 + c], tc = targets[(size_t)q * NOUT + c];
                const float den = yc * yc + 0.01f;
                const float e = yc - tc;
                loss += e * e / den;
                d[c] = 2.0f * e / den * inv_count;
            }
        }
        for (int j = 0; j < 4; j++) {
            const int o = 4 * (lane >> 4) + j;
            const unsigned short hb = to_bf16(d[j]);
            dz4[t][j] = (short)hb;
            s.dz[0][o * 64 + sq] = hb;                // rows 0..15 of buffer 0
        }
        // bias 4 gradient: fp32 sums of the output gradient
        for (int j = 0; j < 4; j++) {
            float v = sum16(d[j]);
            if ((lane & 15) == 0) {
                const int o = 4 * (lane >> 4) + j;
                if (t == 0) g[B4 + o] = v; else g[B4 + o] += v;
            }
        }
    }
    __syncthreads();

    // ---- backward, layer by layer.  cur = buffer holding dZ of the layer being processed.
    int cur = 0;
    for (int li = 4; li >= 0; li--) {
        const int fo = layer_fo(li), fi = li == 0 ? NIN : NH;
        const unsigned short *X = li == 0 ? s.x0 : s.x[li - 1];
        const unsigned short *DZ = s.dz[cur];
        // weight gradient dW^T[out][in] = sum over 64 samples of dZ[out][s] X[in][s]
        for (int mt = 0; mt < fo / 16; mt++)
            for (int nt = 0; nt < fi / 16; nt++) {
                f32x4 acc = f32x4{0, 0, 0, 0};
                for (int kk = 0; kk < 4; kk++) {
                    const int so = 16 * kk + 4 * (lane >> 4);
                    acc = mfma(lds_bf16x4(DZ + (16 * mt + (lane & 15)) * 64 + so),
                               lds_bf16x4(X + (16 * nt + (lane & 15)) * 64 + so), acc);
                }
                for (int j = 0; j < 4; j++) {
                    const int o = 16 * mt + 4 * (lane >> 4) + j, i = 16 * nt + (lane & 15);
                    g[layer_w(li) + i * fo + o] = acc[j];
                }
            }
        if (li == 0) break;
        // input gradient dX = W dZ per tile, masked by the ReLU of the previous layer -> dZ of layer li-1
        const int nxt = cur ^ 1;
        float bias_acc[4][4];
        for (int t = 0; t < 4; t++) {
            const int sq = 16 * t + (lane & 15);
            bf16x4 b[4];
            const int kcn = fo / 16;
            for (int kc = 0; kc < kcn; kc++) {
                if (li == 4) b[kc] = dz4[t];
                else
                    for (int j = 0; j < 4; j++) b[kc][j] = (short)DZ[(16 * kc + 4 * (lane >> 4) + j) * 64 + sq];
            }
            for (int mt = 0; mt < 4; mt++) {
                f32x4 acc = f32x4{0, 0, 0, 0};
                for (int kc = 0; kc < kcn; kc++) {
                    const int frag = li == 4 ? 48 + mt : bwd_base(li) + mt * 4 + kc;
                    acc = mfma(as_bf16x4(frag_bwd[frag * 64 + lane]), b[kc], acc);
                }
                for (int j = 0; j < 4; j++) {
                    const int i = 16 * mt + 4 * (lane >> 4) + j;
                    const float dzv = from_bf16(X[i * 64 + sq]) > 0.0f ? acc[j] : 0.0f;
                    s.dz[nxt][i * 64 + sq] = to_bf16(dzv);
                    const float sum = sum16(dzv);
                    bias_acc[mt][j] = t == 0 ? sum : bias_acc[mt][j] + sum;
                }
            }
        }
        if ((lane & 15) == 0)
            for (int mt = 0; mt < 4; mt++)
                for (int j = 0; j < 4; j++) g[layer_b(li - 1) + 16 * mt + 4 * (lane >> 4) + j] = bias_acc[mt][j];
        __syncthreads();
        cur = nxt;
    }
    // loss partial (sum over this wave's samples, unnormalised)
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if (lane == 0) partial_loss[blockIdx.x] = loss * inv_count;
}

// Sum of the per-wave partials, then (optionally) one Adam step on the float32 master weights.
__global__ void reduce_adam_kernel(const float *__restrict__ partial, uint32_t nblocks, float *__restrict__ p,
                                   float *__restrict__ m, float *__restrict__ v, float *__restrict__ grad_out,
                                   int do_adam, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAMS) return;
    float gsum = 0.0f;
    for (uint32_t b = 0; b < nblocks; b++) gsum += partial[(size_t)b * NPARAMS + i];
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

__global__ void loss_sum_kernel(const float *__restrict__ partial_loss, uint32_t nblocks, float *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.0f;
        for (uint32_t b = 0; b < nblocks; b++) s += partial_loss[b];
        *out = s;
    }
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
    hipLaunchKernelGGL(train_kernel, dim3(blocks), dim3(64), 0, c->stream, c->d_fwd, c->d_bwd, c->d_p, raw, targets, n,
                       n_dev, cap, offset, total, c->center, c->inv_half, c->d_partial, c->d_ploss);
    NRC_TRY(c, hipGetLastError());
    float c1 = 1.0f, c2 = 1.0f;
    if (adam) {
        c->step++;
        c1 = (float)(1.0 - pow(0.9, c->step));
        c2 = (float)(1.0 - pow(0.99, c->step));
    }
    hipLaunchKernelGGL(reduce_adam_kernel, dim3((NPARAMS + 255) / 256), dim3(256), 0, c->stream, c->d_partial, blocks,
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
