// The per-pixel integrator split around the neural radiance cache (device only).
//
// The reference integrator (renderPixelBeauty, src/slim/renderer/ray_tracer.h:19-108) accumulates, per loop
// iteration i, colour += throughput_i * current_i and then throughput_{i+1} = throughput_i * next_throughput_i.
// The radiance leaving the hit of iteration i toward the previous vertex is therefore
//     L_i = current_i + next_throughput_i * L_{i+1}          (L of the last iteration = its current)
// and the pixel's colour is sum_{j<i} throughput_j current_j + throughput_i L_i for any i.  The cache learns L at
// hit points; a cached pixel traces iterations 1..k-1 as the reference does, stops at the hit of iteration k
// (k = cache depth + 1, i.e. after that many bounces) and takes throughput_k * cache(hit) as the rest.
// Training pixels trace the full path and emit one sample (hit record, L_i) per shaded hit.  While a path runs,
// each recorded sample keeps (current, next throughput) of its own iteration, with any later iterations that
// could not be recorded folded in (current += next * current_j, next *= next_j); the pending pair of the latest
// sample lives in registers and earlier pairs in the sample buffers, chained by a link to the previous sample.
// At the end of the path the recursion runs backwards along the links.  A link word holds the previous sample
// (kLinkNone for none) and, in its top 3 bits, the sample's iteration; once the targets are written, a sample's
// link slot holds its canvas offset and the first float of its next slot its path vertex (1 = primary hit).
#pragma once

#include "slimrt_trace.h"

namespace slimrt {

constexpr uint32_t kNrcRaw = 14;
constexpr uint32_t kNrcMaxDepth = 8;   // iterations per path are bounded by max_depth
constexpr uint32_t kNoSample = ~0u;
constexpr uint32_t kLinkNone = 0x1FFFFFFFu;   // sample indices stay below this
constexpr uint32_t kGateFloats = 8;    // per cell: frame sum, frame sum of squares, frame count, mean, mean square, frames

// Frame counters (uint32): queries appended, training samples appended, gate fallbacks, training pixels.
enum : uint32_t { NC_QUERIES = 0, NC_SAMPLES = 1, NC_FALLBACKS = 2, NC_TRAIN_PIXELS = 3, NC_COUNT = 8 };

struct NrcFrame {
    float *q_raw;          // queries: hit records
    uint32_t *q_pixel;     // queries: canvas offset
    float *q_prefix;       // queries: colour accumulated before the cache vertex (rgb)
    float *q_tp;           // queries: throughput at the cache vertex (rgb)
    uint32_t q_cap;
    float *t_raw;          // training samples: hit records
    float *t_target;       // training samples: L at the hit (rgb); while the path runs, its current
    float *t_next;         // training samples: next throughput while the path runs
    uint32_t *t_link;      // training samples: the same path's previous sample
    uint32_t t_cap;
    uint32_t *counters;    // NC_COUNT values
    float *cells;          // confidence gate table
    uint32_t cell_mask;
    float cell_inv_size;
    float gate_threshold;  // <= 0: no gate
    uint32_t gate_min_frames;
    uint32_t cache_vertex; // iteration whose hit the cache answers (1 = primary hit)
    uint32_t train_one_in; // 0: no training waves
    uint32_t frame;
};

// Appends one slot per lane with pred set, with one atomic per wave; returns the slot (or ~0u without pred).
// Every lane that is active at the call takes part, so it may be called from divergent code.
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t *counter) {
    const uint64_t active = __ballot(1);
    const uint64_t mask = __ballot(pred);
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)active) - 1u;
    uint32_t base = 0;
    if (lane == leader && mask) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader, 64);
    return pred ? base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull)) : ~0u;
}

__device__ __forceinline__ uint32_t gate_cell(const NrcFrame &f, v3 p) {
    const int32_t ix = (int32_t)floorf(p.x * f.cell_inv_size), iy = (int32_t)floorf(p.y * f.cell_inv_size),
                  iz = (int32_t)floorf(p.z * f.cell_inv_size);
    uint32_t h = (uint32_t)ix * 73856093u ^ (uint32_t)iy * 19349663u ^ (uint32_t)iz * 83492791u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    return h & f.cell_mask;
}

// Confidence gate: the cell must have had training targets in enough frames, and their spread over the
// last ~16 frames (exponential window) must be small relative to their mean.
__device__ __forceinline__ bool gate_pass(const NrcFrame &f, v3 p) {
    if (f.gate_threshold <= 0.0f) return true;
    const float *c = f.cells + (size_t)gate_cell(f, p) * kGateFloats;
    if (c[5] < (float)f.gate_min_frames) return false;
    const float m1 = c[3], m2 = c[4];
    const float var = fmaxf(m2 - m1 * m1, 0.0f);
    return sqrtf(var) <= f.gate_threshold * (m1 + 0.01f);
}

__device__ __forceinline__ void write_record(float *dst, const Surface &sf) {
    const float r[kNrcRaw] = {sf.P.x, sf.P.y, sf.P.z, sf.N.x, sf.N.y, sf.N.z, sf.V.x, sf.V.y, sf.V.z,
                              sf.mat.roughness, sf.mat.metalness, sf.mat.ar * sf.albedo_from_map.r,
                              sf.mat.ag * sf.albedo_from_map.g, sf.mat.ab * sf.albedo_from_map.b};
    for (uint32_t i = 0; i < kNrcRaw; i++) dst[i] = r[i];
}

// The beauty integrator with a cache vertex (query) or full-path sample recording (training).  Returns true
// when the pixel is final (col holds its untone-mapped colour), false when a query was appended.
template <class TR>
__device__ __forceinline__ bool render_beauty_nrc(const SceneView &s, const RenderParams &rp, v3 direction, float scaling, TR &tr,
                                  const NrcFrame &nf, bool training, uint32_t pixel, color &col, float &depth) {
    Surface sf;
    Ray ray;
    Hit hit;
    hit.scaling_factor = scaling;
    color cur, next_tp = cs(0), tp = cs(1.0f);
    col = cs(0);
    depth = INFINITY;
    const Projection &p = rp.proj;
    ray_reset(ray, v3{p.camera_position[0], p.camera_position[1], p.camera_position[2]}, normalized(direction));
    uint32_t depth_left = rp.max_depth, iter = 0;
    uint32_t last = kNoSample;                 // latest recorded sample of this path
    color pend_cur = cs(0), pend_next = cs(0); // its pending (current, next throughput)
    while (depth_left && iter < kNrcMaxDepth) {
        cur = cs(0);
        const uint32_t it = iter++;
        bool recorded = false;
        sf.geo = scene_trace(s, ray, hit, false, INFINITY, tr);
        bool bounced = false;
        if (sf.geo >= 0) {
            hit.scaling_factor = keeps_ray_cones(rp) ? scaling : 0.0f;
            const Geo g = s.geos[sf.geo];
            const Mat m = s.mats[g.material_id];
            if ((mat_flags(m) & MAT_EMISSIVE) && g.type == GT_Quad && !hit.from_behind) {
                depth_left = 0;
                cur = color{m.er, m.eg, m.eb};
            } else {
                prepare_for_shading(sf, s, ray, hit, tr);
                if (depth_left == rp.max_depth) depth = depth_at(p, hit.position);
                if (!training && it + 1 == nf.cache_vertex) {
                    const bool want = gate_pass(nf, sf.P);
                    const uint32_t q = wave_append(want, nf.counters + NC_QUERIES);
                    const bool queued = want && q < nf.q_cap;
                    if (queued) {
                        write_record(nf.q_raw + (size_t)q * kNrcRaw, sf);
                        nf.q_pixel[q] = pixel;
                        nf.q_prefix[3 * q] = col.r; nf.q_prefix[3 * q + 1] = col.g; nf.q_prefix[3 * q + 2] = col.b;
                        nf.q_tp[3 * q] = tp.r; nf.q_tp[3 * q + 1] = tp.g; nf.q_tp[3 * q + 2] = tp.b;
                    }
                    (void)wave_append(!queued, nf.counters + NC_FALLBACKS);   // gate closed: full trace from here
                    if (queued) return false;
                }
                if (training) {
                    const uint32_t t = wave_append(true, nf.counters + NC_SAMPLES);
                    if (t < nf.t_cap) {
                        write_record(nf.t_raw + (size_t)t * kNrcRaw, sf);
                        if (last != kNoSample) {   // the previous sample's pair is complete
                            float *c = nf.t_target + 3 * (size_t)last, *n = nf.t_next + 3 * (size_t)last;
                            c[0] = pend_cur.r; c[1] = pend_cur.g; c[2] = pend_cur.b;
                            n[0] = pend_next.r; n[1] = pend_next.g; n[2] = pend_next.b;
                        }
                        nf.t_link[t] = (last == kNoSample ? kLinkNone : last) | (it << 29);
                        last = t;
                        recorded = true;
                    }
                }
                for (uint32_t i = 0; i < s.L; i++) shade_from_light(sf, s.lights[i], s, tr, cur);
                if (s.flags & 1u) shade_from_emissive_quads(sf, s, cur);
                if (rp.sky_irradiance >= 0 && rp.sky_radiance >= 0) {
                    sf.L = sf.N;
                    sf.NdotL = 1.0f;
                    color D = tex_sample_cube(s, (uint32_t)rp.sky_irradiance, sf.N.x, sf.N.y, sf.N.z);
                    color S = tex_sample_cube(s, (uint32_t)rp.sky_radiance, sf.R.x, sf.R.y, sf.R.z);
                    radiance_fraction(sf);
                    cur = cmul_add(D, sf.Fd, cmul_add(sf.Fs, S, cur));
                }
                if ((mat_flags(sf.mat) & (MAT_REFLECTIVE | MAT_REFRACTIVE)) && --depth_left) {
                    sf.F = schlick(clamp01(dot(sf.N, sf.R)), color{sf.mat.fr, sf.mat.fg, sf.mat.fb});
                    next_tp = sf.refracted ? one_minus(sf.F) : sf.F;
                    ray_reset(ray, hit.position, sf.RF);
                    bounced = true;
                } else depth_left = 0;
            }
        } else {
            depth_left = 0;
            if (rp.sky_color >= 0) cur = tex_sample_cube(s, (uint32_t)rp.sky_color, ray.direction.x, ray.direction.y, ray.direction.z);
        }
        for (uint32_t i = 0; i < s.L; i++) {
            const Light l = s.lights[i];
            float density;
            if (light_glow(l, ray, hit.distance, density)) cur = cscale_add(color{l.r, l.g, l.b}, glow_factor(density), cur);
        }
        if (training) {
            const color nt = bounced ? next_tp : cs(0);
            if (recorded) {
                pend_cur = cur;
                pend_next = nt;
            } else if (last != kNoSample) {
                pend_cur = cmul_add(pend_next, cur, pend_cur);
                pend_next = pend_next * nt;
            }
        }
        col = cmul_add(cur, tp, col);
        tp = tp * next_tp;
    }
    if (training && last != kNoSample) {   // L_i = current_i + next_i * L_{i+1} along the links; gate statistics
        color L = pend_cur;                  // the last sample's next throughput multiplies an empty rest
        for (uint32_t t = last;;) {
            float *dst = nf.t_target + 3 * (size_t)t;
            dst[0] = L.r; dst[1] = L.g; dst[2] = L.b;
            const float *r = nf.t_raw + (size_t)t * kNrcRaw;
            float *cell = nf.cells + (size_t)gate_cell(nf, v3{r[0], r[1], r[2]}) * kGateFloats;
            const float lum = (L.r + L.g + L.b) * (1.0f / 3.0f);
            atomicAdd(cell + 0, lum);
            atomicAdd(cell + 1, lum * lum);
            atomicAdd(cell + 2, 1.0f);
            const uint32_t w = nf.t_link[t];
            nf.t_link[t] = pixel;
            nf.t_next[3 * (size_t)t] = __uint_as_float((w >> 29) + 1u);
            t = w & kLinkNone;
            if (t == kLinkNone) break;
            const float *c = nf.t_target + 3 * (size_t)t, *n = nf.t_next + 3 * (size_t)t;
            L = color{c[0] + n[0] * L.r, c[1] + n[1] * L.g, c[2] + n[2] * L.b};
        }
    }
    return true;
}

}  // namespace slimrt
