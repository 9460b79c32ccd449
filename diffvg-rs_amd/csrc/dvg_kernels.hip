// gfx950 kernels of the render/backward path.
//
//   k_bin_count / k_scan / k_bin_write   tile lists (binning.rs:10-313, tiles.rs:12-38)
//   k_weights                            filter weight sums (forward.rs:1506-1565), by gather
//   k_forward                            per-sample scene evaluation + splat (forward.rs:1568-1735)
//   k_resolve                            splat resolve (forward.rs:295-334)
//   k_backward_interior                  colour adjoint per sample (backward/sampling.rs:76-224)
//   k_bnd_key / k_scan_sums / k_scan_chunks / k_bnd_scatter / k_bnd_eval
//                                        edge sampling (backward/boundary.rs:113-1148),
//                                        counting-sorted by tile before evaluation
//   k_bnd_eval_chk / k_bnd_redo          the same with uncertain inside tests
//                                        evaluated again in f64 (exact boundary mode)
//
// One workgroup of 256 lanes (4 wave64) owns one 16x16 tile; its samples are
// visited in 4x4-pixel blocks so that a wave covers a compact footprint and
// the tile's group list is wave-uniform (scalar loads).  Gradients of the
// tile's solid paints are staged in LDS and flushed once per tile; scalar
// gradients (background, filter radius) are wave-reduced before any global
// atomic.
#include "dvg_device.h"
#include "dvg_kernels.h"

namespace dvgk {

// Waves per SIMD the heavy kernels are compiled for.  The scene walks are
// VALU-issue bound with memory waits between groups, so more resident waves
// beat an unspilled register file: measured on the 1024x1024 / 1000-path
// scene, 7 waves (<= 72 VGPRs) is best for the forward and the colour adjoint
// (1.64 -> 1.38 ms from 4 to 7, with its LDS cut to 18 KiB a workgroup, 1.62
// at 8), 8 (<= 64 VGPRs) for the boundary pass's one walk of both query
// points (4.87, 3.90, 3.54, 3.53, 3.30, 3.21 ms from 3 to 8 waves).
#ifndef OCC_FORWARD_PF
#define OCC_FORWARD_PF 7
#endif
#ifndef OCC_FORWARD
#define OCC_FORWARD 7
#endif
#ifndef OCC_INTERIOR
#define OCC_INTERIOR 7
#endif
#ifndef OCC_BOUNDARY
#define OCC_BOUNDARY 8
#endif
// The redo of uncertain boundary samples runs beside the colour adjoint on a
// second stream; its own occupancy sets how much of each SIMD it takes
// (1024x1024 step: 4.95 ms at 8, 4.86 at 7 with the colour adjoint 1.24 ->
// 1.14 ms beside it, 5.14 at 6 and 5.11 at 4, where the redo outlasts it).
#ifndef OCC_REDO
#define OCC_REDO 7
#endif
// The prefiltered adjoint (closest-point adjoints, ~200 VGPRs unconstrained)
// is latency bound at 2 waves; 3 measured best on the 1024x1024 prefiltered
// scene (25.9 -> 20.4 ms; 4-6 spill enough to lose it again).
#ifndef OCC_PREFILTER
#define OCC_PREFILTER 3
#endif
// The batch's boundary evaluation (a persistent loop over its work queue):
// four 1024x1024 scenes step in 17.0 ms at 8 waves, 18.7 ms at 7 (with the
// backward's culled tile lists; 7 was the better before them).
#ifndef OCC_BATCH_BOUNDARY
#define OCC_BATCH_BOUNDARY 8
#endif

// ---------------------------------------------------------------------------
// Tile binning: group g belongs to tile t iff t lies in g's clamped tile range
// (binning.rs:30-110).  One wave per tile tests 64 groups at a time in
// ascending order (coalesced loads of the ranges); a hit's position in the
// tile's list is the count of hits on lower lanes, so the lists come out
// sorted by group id without the host sort of renderer.rs:364-367.
constexpr uint32_t BIN_TILES_PER_WG = BLOCK / 64;

__global__ __launch_bounds__(BLOCK) void k_bin_count(const int4 *__restrict__ rng, uint32_t G, uint32_t tiles_x,
                                                     uint32_t tiles_y, uint32_t *__restrict__ counts) {
    uint32_t t = blockIdx.x * BIN_TILES_PER_WG + threadIdx.x / 64u, lane = threadIdx.x & 63u;
    if (t >= tiles_x * tiles_y) return;  // (whole waves)
    int tx = (int)(t % tiles_x), ty = (int)(t / tiles_x);
    uint32_t c = 0;
    for (uint32_t g0 = 0; g0 < G; g0 += 64u) {
        uint32_t g = g0 + lane;
        bool hit = false;
        if (g < G) {
            int4 r = rng[g];
            hit = tx >= r.x && tx <= r.z && ty >= r.y && ty <= r.w;
        }
        c += (uint32_t)__popcll(__ballot(hit));
    }
    if (lane == 0) counts[t] = c;
}

__global__ __launch_bounds__(BLOCK) void k_bin_write(const int4 *__restrict__ rng, uint32_t G, uint32_t tiles_x,
                                                     uint32_t tiles_y, const uint32_t *__restrict__ offsets,
                                                     uint32_t *__restrict__ entries) {
    uint32_t t = blockIdx.x * BIN_TILES_PER_WG + threadIdx.x / 64u, lane = threadIdx.x & 63u;
    if (t >= tiles_x * tiles_y) return;  // (whole waves)
    int tx = (int)(t % tiles_x), ty = (int)(t / tiles_x);
    // the tile's slot [offsets[t], offsets[t+1]) bounds the writes (offsets
    // clamped to a buffer capacity cut lists short instead of overrunning)
    uint32_t o = offsets[t], e = offsets[t + 1];
    for (uint32_t g0 = 0; g0 < G && o < e; g0 += 64u) {
        uint32_t g = g0 + lane;
        bool hit = false;
        if (g < G) {
            int4 r = rng[g];
            hit = tx >= r.x && tx <= r.z && ty >= r.y && ty <= r.w;
        }
        uint64_t m = __ballot(hit);
        uint32_t pos = o + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (hit && pos < e) entries[pos] = g;
        o += (uint32_t)__popcll(m);
    }
}

// Exclusive scan of n counts into offsets[0..n] by one workgroup of NT
// threads (each sums a contiguous run, then a scan of the run totals).
template <int NT>
DVG_DEV void scan_unit(const uint32_t *__restrict__ counts, uint32_t n, uint32_t *__restrict__ offsets) {
    __shared__ uint32_t part[NT];
    uint32_t tid = threadIdx.x;
    uint32_t per = (n + NT - 1) / NT;
    uint32_t b = tid * per, e = min(n, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += counts[i];
    part[tid] = s;
    __syncthreads();
    for (uint32_t off = 1; off < NT; off <<= 1) {
        uint32_t v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = tid == 0 ? 0u : part[tid - 1];
    for (uint32_t i = b; i < e; ++i) {
        offsets[i] = run;
        run += counts[i];
    }
    if (tid == NT - 1) offsets[n] = part[NT - 1];
    __syncthreads();
}

__global__ __launch_bounds__(1024) void k_scan(const uint32_t *__restrict__ counts, uint32_t n,
                                               uint32_t *__restrict__ offsets) {
    scan_unit<1024>(counts, n, offsets);
}

// Exclusive scan of counts[0, n) into offsets[0, n] (offsets[n] = total)
// over scan_chunks(n) workgroups of 256: k_scan_sums adds up each chunk of
// SCAN_CHUNK counts, k_scan_chunks scans its chunk from the sum of the chunks
// before it (SCAN_PER_THREAD consecutive counts a thread).
DVG_DEV uint32_t block_sum_256(uint32_t v, uint32_t *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63u) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(256) void k_scan_sums(const uint32_t *__restrict__ counts, uint32_t n,
                                                   uint32_t *__restrict__ sums) {
    __shared__ uint32_t sh[4];
    uint32_t b = blockIdx.x * SCAN_CHUNK, e = min(n, b + SCAN_CHUNK);
    uint32_t s = 0;
    for (uint32_t i = b + threadIdx.x; i < e; i += 256u) s += counts[i];
    uint32_t t = block_sum_256(s, sh);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_scan_chunks(const uint32_t *__restrict__ counts, uint32_t n,
                                                     const uint32_t *__restrict__ sums,
                                                     uint32_t *__restrict__ offsets) {
    __shared__ uint32_t sh[4], wsum[4];
    uint32_t c = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += 256u) c += sums[j];
    uint32_t carry = block_sum_256(c, sh);
    uint32_t b = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD];
    uint32_t s = 0;
    for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
        v[k] = b + k < n ? counts[b + k] : 0u;
        s += v[k];
    }
    // exclusive prefix of s over the workgroup: within the wave, then the
    // totals of the waves before
    uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t inc = s;
    for (uint32_t o = 1; o < 64u; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63u) wsum[w] = inc;
    __syncthreads();
    uint32_t run = carry + inc - s;
    for (uint32_t k = 0; k < w; ++k) run += wsum[k];
    for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
        if (b + k < n) offsets[b + k] = run;
        run += v[k];
    }
    // the last thread of the last chunk holds the total
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255u) offsets[n] = run;
}

// ---------------------------------------------------------------------------
// Sample addressing inside a tile: pixel p of the tile is visited in square
// blocks of b = 1 << sblock_log2 pixels a side (b*b consecutive p share a
// block), sample s of pixel p at p*spp + s.
DVG_DEV void tile_sample(const DevScene &S, uint32_t tile, uint32_t s, uint32_t &x, uint32_t &y, uint32_t &sx,
                         uint32_t &sy) {
    uint32_t pl = s / S.spp, si = s - pl * S.spp;
    uint32_t L = S.sblock_log2, R = 4u - L;  // TILE = 16 = 1 << 4
    uint32_t blk = pl >> (2u * L), ip = pl & ((1u << (2u * L)) - 1u);
    uint32_t lx = ((blk & ((1u << R) - 1u)) << L) | (ip & ((1u << L) - 1u));
    uint32_t ly = ((blk >> R) << L) | (ip >> L);
    x = (tile % S.tiles_x) * TILE + lx;
    y = (tile / S.tiles_x) * TILE + ly;
    sx = si % S.nsx;
    sy = si / S.nsx;
}

// Sample position: forward.rs:1647-1650 multiplies by 1/n; the CPU backward
// (render.rs:108-109, filters.rs:45-46) divides by n.
template <bool CPU>
DVG_DEV void sample_point(const DevScene &S, uint32_t x, uint32_t y, uint32_t sx, uint32_t sy, float &px,
                          float &py) {
    float rx, ry;
    jitter_xy(x, y, sx, sy, S.W, S.nsx, S.nsy, S.seed, S.jitter != 0, rx, ry);
    if (CPU) {
        px = (float)x + ((float)sx + rx) / (float)S.nsx;
        py = (float)y + ((float)sy + ry) / (float)S.nsy;
    } else {
        float inv_sx = 1.0f / (float)S.nsx, inv_sy = 1.0f / (float)S.nsy;
        px = (float)x + ((float)sx + rx) * inv_sx;
        py = (float)y + ((float)sy + ry) * inv_sy;
    }
}

// ---- paints ------------------------------------------------------------------
// sampling.rs:11-139 (forward semantics: clamped radius and stop gap, t <= first)
DVG_DEV float4 paint_fwd(const DevScene &S, uint32_t kind, uint32_t gidx, const float *solid, float px, float py) {
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (kind == DVG_PAINT_SOLID) return make_float4(solid[0], solid[1], solid[2], solid[3]);
    if (kind != DVG_PAINT_LINEAR && kind != DVG_PAINT_RADIAL) return out;
    const float *gd = S.gradient_data + gidx * 8u;
    uint32_t gtype = (uint32_t)gd[0];
    float p0 = gd[1], p1 = gd[2], p2 = gd[3], p3 = gd[4];
    uint32_t so = (uint32_t)gd[5], sc = (uint32_t)gd[6];
    float t;
    if (gtype == 0) {
        float vx = p2 - p0, vy = p3 - p1;
        float denom = maxf_(vx * vx + vy * vy, 1.0e-3f);
        t = ((px - p0) * vx + (py - p1) * vy) / denom;
    } else {
        float dx = (px - p0) / maxf_(p2, 1.0e-3f);
        float dy = (py - p1) / maxf_(p3, 1.0e-3f);
        t = sqrtf(dx * dx + dy * dy);
    }
    if (sc == 0) return out;
    const float *offs = S.stop_offsets + so;
    const float *cols = S.stop_colors + so * 4u;
    if (t <= offs[0]) return make_float4(cols[0], cols[1], cols[2], cols[3]);
    if (t >= offs[sc - 1]) {
        const float *c = cols + (sc - 1) * 4u;
        return make_float4(c[0], c[1], c[2], c[3]);
    }
    for (uint32_t i = 0; i + 1 < sc; ++i) {
        float cu = offs[i], nx = offs[i + 1];
        if (t >= cu && t < nx) {
            float tt = (t - cu) / maxf_(nx - cu, 1.0e-5f);
            const float *c0 = cols + i * 4u, *c1 = c0 + 4;
            return make_float4(c0[0] * (1.f - tt) + c1[0] * tt, c0[1] * (1.f - tt) + c1[1] * tt,
                               c0[2] * (1.f - tt) + c1[2] * tt, c0[3] * (1.f - tt) + c1[3] * tt);
        }
    }
    return out;
}

// backward/paint.rs:8-74 (CPU semantics)
DVG_DEV float4 paint_cpu(const DevScene &S, uint32_t kind, uint32_t gidx, const float *solid, float px, float py) {
    if (kind == DVG_PAINT_SOLID) return make_float4(solid[0], solid[1], solid[2], solid[3]);
    float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    if (kind != DVG_PAINT_LINEAR && kind != DVG_PAINT_RADIAL) return zero;
    const float *gd = S.gradient_data + gidx * 8u;
    uint32_t gtype = (uint32_t)gd[0];
    float p0 = gd[1], p1 = gd[2], p2 = gd[3], p3 = gd[4];
    uint32_t so = (uint32_t)gd[5], sc = (uint32_t)gd[6];
    float t;
    if (gtype == 0) {
        float ex = p2 - p0, ey = p3 - p1;
        float denom = fmaxf(ex * ex + ey * ey, 1.0e-3f);
        t = ((px - p0) * ex + (py - p1) * ey) / denom;
    } else {
        float nx = (px - p0) / p2, ny = (py - p1) / p3;
        t = sqrtf(nx * nx + ny * ny);
    }
    if (sc == 0) return zero;
    const float *offs = S.stop_offsets + so;
    const float *cols = S.stop_colors + so * 4u;
    if (t < offs[0]) return make_float4(cols[0], cols[1], cols[2], cols[3]);
    for (uint32_t i = 0; i + 1 < sc; ++i) {
        float cu = offs[i], nx = offs[i + 1];
        if (t >= cu && t < nx) {
            float tt = (t - cu) / (nx - cu);
            const float *c0 = cols + i * 4u, *c1 = c0 + 4;
            return make_float4(c0[0] * (1.0f - tt) + c1[0] * tt, c0[1] * (1.0f - tt) + c1[1] * tt,
                               c0[2] * (1.0f - tt) + c1[2] * tt, c0[3] * (1.0f - tt) + c1[3] * tt);
        }
    }
    const float *c = cols + (sc - 1) * 4u;
    return make_float4(c[0], c[1], c[2], c[3]);
}

// ---- per-shape coverage --------------------------------------------------------
// `unsure` (backward inside tests of the boundary pass): raised where the
// reference CPU's f64 solve could decide differently.
struct Cov {
    float fill_min, winding, crossings, stroke_min, stroke_r, stroke_hit, unsure;
};

// Which inside tests a coverage walk runs (template argument M):
//   COV_FWD   the reference GPU forward's (forward.rs);
//   COV_CPU   the reference CPU backward's (distance/*.rs) in f32;
//   COV_CHK   COV_CPU raising Cov::unsure where f64 could decide otherwise;
//   COV_EXACT COV_CPU with the fill windings in f64 as the CPU computes them.
enum : int { COV_FWD = 0, COV_CPU = 1, COV_CHK = 2, COV_EXACT = 3 };

DVG_DEV float seg_radius(uint32_t kind, float t, float r0, float r1, float r2, float r3) {
    if (kind == 0) return r0 + t * (r1 - r0);
    float tt = 1.f - t;
    if (kind == 1) return tt * tt * r0 + 2.0f * tt * t * r1 + t * t * r2;
    return tt * tt * tt * r0 + 3.0f * tt * tt * t * r1 + 3.0f * tt * t * t * r2 + t * t * t * r3;
}

// The winding test of one segment whose ray box the point passed (shape_fill).
template <int M>
DVG_DEV void fill_segment(uint32_t k, float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                          float x3, float y3, const float *wc, uint32_t rule, float &lw, float &lc, float &unsure) {
    constexpr bool CPU = M != COV_FWD;
    {
        if (k == 0) {
            if (CPU) winding_line_cpu(px, py, x0, y0, x1, y1, rule, lw, lc);
            else winding_line(px, py, x0, y0, x1, y1, rule, lw, lc);
        } else if (M == COV_EXACT) {
            // the checked f32 test, and the f64 one where it is uncertain
            float tw = 0.f, tc = 0.f, tu = 0.f;
            if (k == 1) winding_quad_w<true>(px, py, x0, y0, x1, x2, wc, rule, tw, tc, &tu);
            else winding_cubic_bwd<true>(px, py, x0, y0, x1, x2, x3, wc, rule, tw, tc, tu);
            int n = (int)tw;
            bool odd = tc > 0.f;
            if (tu > 0.f) {
                float xs[4] = {x0, x1, x2, x3}, ys[4] = {y0, y1, y2, y3};
                n = (k == 1) ? winding_quad_f64(px, py, xs, ys) : winding_cubic_f64(px, py, xs, ys);
                odd = (n & 1) != 0;
            }
            if (rule == 1u) {
                if (odd) lc = 1.f - lc;
            } else {
                lw += (float)n;
            }
        } else if (k == 1) {
            // (the unchecked tests get no pointer into c: it would keep c in memory)
            if (M == COV_CHK) winding_quad_w<true>(px, py, x0, y0, x1, x2, wc, rule, lw, lc, &unsure);
            else winding_quad_w<false>(px, py, x0, y0, x1, x2, wc, rule, lw, lc);
        } else if (CPU) {
            float u = 0.f;
            winding_cubic_bwd<M == COV_CHK>(px, py, x0, y0, x1, x2, x3, wc, rule, lw, lc, M == COV_CHK ? unsure : u);
        } else {
            winding_cubic_w<false>(px, py, x0, y0, x1, x2, x3, wc, rule, lw, lc);
        }
    }
}

// accumulate_shape_fill, forward.rs:409-560 (distances only when prefiltering:
// the non-prefiltered blend reads the winding alone).  With CPU = true the
// winding follows the reference CPU's winding_number_segment
// (distance/winding.rs:118-192): the closed-interval line rule and the f64
// cubic solve on ill-conditioned segments.
//
// `cut` (prefiltered distances): distances at or beyond it give the same
// coverage as any larger one, so the running minimum starts there and
// segments farther away are not measured (the minimum comes out as `cut`
// when the shape is that far).
template <int M>
DVG_DEV void shape_fill(const DevScene &S, const float *sd, uint32_t rule, bool prefilter, float px, float py,
                        Cov &c, float cut = BIG) {
    constexpr bool CPU = M != COV_FWD;
    uint32_t kind = (uint32_t)sd[0];
    float q0 = sd[4], q1 = sd[5], q2 = sd[6], q3 = sd[7];
    if (kind != DVG_SHAPE_PATH) {
        float dist;
        if (kind == DVG_SHAPE_CIRCLE) {
            float dx = px - q0, dy = py - q1;
            dist = sqrtf(dx * dx + dy * dy) - q2;
        } else if (kind == DVG_SHAPE_ELLIPSE) {
            float dx = (px - q0) / maxf_(absf_(q2), 1.0e-3f);
            float dy = (py - q1) / maxf_(absf_(q3), 1.0e-3f);
            float len = sqrtf(dx * dx + dy * dy);
            dist = (len - 1.0f) * minf_(absf_(q2), absf_(q3));
        } else {
            dist = rect_sdf(px, py, q0, q1, q2, q3);
        }
        float ad = absf_(dist);
        if (ad < c.fill_min) c.fill_min = ad;
        if (dist <= 0.f) {
            if (rule == 1u) c.crossings = 1.f - c.crossings;
            else c.winding += 1.f;
        }
        return;
    }
    uint32_t off = (uint32_t)sd[12], cnt = (uint32_t)sd[13];
    bool approx = sd[14] > 0.5f;
    float lmin = cut, lw = 0.f, lc = 0.f;
    for (uint32_t s = 0; s < cnt; ++s) {
        const float *cv = S.curve_data + (off + s) * 13u;
        const float *wc = S.curve_wind + (size_t)(off + s) * WIND_STRIDE;
        uint32_t k = (uint32_t)cv[0];
        float x0 = cv[1], y0 = cv[2], x1 = cv[3], y1 = cv[4], x2 = cv[5], y2 = cv[6], x3 = cv[7], y3 = cv[8];
        // the curve's box from the segment's record (widened: see k_wind_coef)
        float bx1 = wc[14], by0 = wc[12], by1 = wc[13];
        if (prefilter) {
            // forward.rs:889-937: segments farther than the running minimum
            // cannot lower it (the curve's box, widened past the rounding of
            // the distances computed to it)
            float ddx = maxf_(maxf_(wc[15] - px, 0.f), px - bx1), ddy = maxf_(maxf_(by0 - py, 0.f), py - by1);
            if (sqrtf(ddx * ddx + ddy * ddy) <= lmin) {
                float d, t;
                if (k == 0) d = seg_dist(px, py, x0, y0, x1, y1);
                else if (k == 1) quad_closest<!CPU>(px, py, x0, y0, x1, y1, x2, y2, approx, d, t);
                else cubic_closest<!CPU>(px, py, x0, y0, x1, y1, x2, y2, x3, y3, approx, d, t);
                if (d < lmin) lmin = d;
            }
        }
        // forward.rs:938-1018: a +x ray only meets segments spanning py with
        // some point right of px.
        if (py < by0 || py > by1 || px > bx1) continue;
        fill_segment<M>(k, px, py, x0, y0, x1, y1, x2, y2, x3, y3, wc, rule, lw, lc, c.unsure);
    }
    if (lmin < c.fill_min) c.fill_min = lmin;
    if (rule == 1u) {
        if (lc > 0.f) c.crossings = 1.f - c.crossings;
    } else {
        c.winding += lw;
    }
}

// The CPU fill windings of one shape at the two query points of a boundary
// sample (shape space; a0 / a1: which points the group's root box admits).
// On a path the second point's winding is the first's when the path is a
// closed loop whose CPU winding is the curve's winding number (record slot
// 31 of its first segment, k_shape_topo) and no segment's curve box (record
// slots 12..15) meets the box spanned by the two points: the segment joining them then crosses
// no edge, so both lie in one component of the complement of the curve,
// where the winding number is constant.  Otherwise each point is tested.
template <int M>
DVG_DEV void shape_fill_pair(const DevScene &S, const float *sd, uint32_t rule, float px0, float py0, bool a0,
                             float px1, float py1, bool a1, Cov &c0, Cov &c1) {
    if ((uint32_t)sd[0] != DVG_SHAPE_PATH || !a0) {
        if (a0) shape_fill<M>(S, sd, rule, false, px0, py0, c0);
        if (a1) shape_fill<M>(S, sd, rule, false, px1, py1, c1);
        return;
    }
    uint32_t off = (uint32_t)sd[12], cnt = (uint32_t)sd[13];
    float ex0 = fminf(px0, px1), ex1 = fmaxf(px0, px1), ey0 = fminf(py0, py1), ey1 = fmaxf(py0, py1);
    bool same = a1 && cnt > 0u && S.curve_wind[(size_t)off * WIND_STRIDE + 31] > 0.f;
    float lw = 0.f, lc = 0.f;
    for (uint32_t s = 0; s < cnt; ++s) {
        const float *cv = S.curve_data + (off + s) * 13u;
        const float *wc = S.curve_wind + (size_t)(off + s) * WIND_STRIDE;
        float bx0 = wc[15], bx1 = wc[14], by0 = wc[12], by1 = wc[13];
        if (!(ex1 < bx0 || ex0 > bx1 || ey1 < by0 || ey0 > by1)) same = false;
        if (py0 < by0 || py0 > by1 || px0 > bx1) continue;
        fill_segment<M>((uint32_t)cv[0], px0, py0, cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7], cv[8], wc, rule, lw,
                        lc, c0.unsure);
    }
    if (rule == 1u) {
        if (lc > 0.f) c0.crossings = 1.f - c0.crossings;
    } else {
        c0.winding += lw;
    }
    if (same) {
        if (rule == 1u) {
            if (lc > 0.f) c1.crossings = 1.f - c1.crossings;
        } else {
            c1.winding += lw;
        }
    } else if (a1) {
        shape_fill<M>(S, sd, rule, false, px1, py1, c1);
    }
}

// accumulate_shape_stroke, forward.rs:565-709.  With CPU = true the ellipse
// uses the exact closest point of src/distance/shape.rs:117-164 and a shape
// needs stroke_width > 0, as the CPU query within_distance_bvh does
// (distance/api.rs:79-135).
// `cut`: as for shape_fill (a minimum left at `cut` keeps radius 0).
template <int M>
DVG_DEV void shape_stroke(const DevScene &S, const float *sd, bool prefilter, float px, float py, Cov &c,
                          float cut = BIG) {
    constexpr bool CPU = M != COV_FWD;
    uint32_t kind = (uint32_t)sd[0];
    float sw = sd[3];
    uint32_t off = (uint32_t)sd[12], cnt = (uint32_t)sd[13];
    if (CPU) {
        if (!(sw > 0.f)) return;
    } else {
        if (sw <= 0.f && cnt == 0) return;
    }
    float q0 = sd[4], q1 = sd[5], q2 = sd[6], q3 = sd[7];
    if (kind != DVG_SHAPE_PATH) {
        float dist;
        if (kind == DVG_SHAPE_CIRCLE) {
            float dx = px - q0, dy = py - q1;
            dist = absf_(sqrtf(dx * dx + dy * dy) - q2);
        } else if (kind == DVG_SHAPE_ELLIPSE) {
            if (CPU) {
                // ellipse_signed_distance (distance/shape.rs:204-212)
                float rx = fmaxf(absf_(q2), 1.0e-6f), ry = fmaxf(absf_(q3), 1.0e-6f);
                float ox, oy;
                ellipse_closest(q0, q1, rx, ry, px, py, ox, oy);
                dist = dist_pt(px, py, ox, oy);
            } else {
                float dx = (px - q0) / maxf_(absf_(q2), 1.0e-3f);
                float dy = (py - q1) / maxf_(absf_(q3), 1.0e-3f);
                float len = sqrtf(dx * dx + dy * dy);
                dist = absf_((len - 1.0f) * minf_(absf_(q2), absf_(q3)));
            }
        } else {
            dist = absf_(rect_sdf(px, py, q0, q1, q2, q3));
        }
        if (prefilter) {
            if (dist < c.stroke_min) { c.stroke_min = dist; c.stroke_r = sw; }
        } else if (dist < sw) {
            c.stroke_hit = 1.f;
        }
        return;
    }
    bool approx = sd[14] > 0.5f;
    float lmin = cut, lr = 0.f;
    for (uint32_t s = 0; s < cnt; ++s) {
        const float *cv = S.curve_data + (off + s) * 13u;
        uint32_t k = (uint32_t)cv[0];
        float x0 = cv[1], y0 = cv[2], x1 = cv[3], y1 = cv[4], x2 = cv[5], y2 = cv[6], x3 = cv[7], y3 = cv[8];
        float r0 = cv[9], r1 = cv[10], r2 = cv[11], r3 = cv[12];
        float pad = fmaxf(fmaxf(absf_(r0), absf_(r1)), fmaxf(absf_(r2), absf_(r3)));
        // the curve's box (record slots 12..15, widened past the rounding of
        // any distance computed to it), padded by the largest radius
        const float *wc = S.curve_wind + (size_t)(off + s) * WIND_STRIDE;
        float bx0 = wc[15] - pad, bx1 = wc[14] + pad, by0 = wc[12] - pad, by1 = wc[13] + pad;
        if (prefilter) {
            float ddx = maxf_(maxf_(bx0 + pad - px, 0.f), px - (bx1 - pad));
            float ddy = maxf_(maxf_(by0 + pad - py, 0.f), py - (by1 - pad));
            if (sqrtf(ddx * ddx + ddy * ddy) > lmin) continue;
        } else if (px < bx0 || px > bx1 || py < by0 || py > by1) {
            continue;  // forward.rs:1048-1050: outside the radius-padded bounds (the curve's)
        }
        float d, t;
        if (k == 0) seg_dist_t(px, py, x0, y0, x1, y1, d, t);
        else if (!CPU && prefilter) {
            if (k == 1) quad_closest<true>(px, py, x0, y0, x1, y1, x2, y2, approx, d, t);
            else cubic_closest<true>(px, py, x0, y0, x1, y1, x2, y2, x3, y3, approx, d, t);
        } else if (k == 1) {
            quad_closest(px, py, x0, y0, x1, y1, x2, y2, approx, d, t);
        } else {
            cubic_closest(px, py, x0, y0, x1, y1, x2, y2, x3, y3, approx, d, t);
        }
        float r = seg_radius(k, t, r0, r1, r2, r3);
        if (prefilter) {
            if (d < lmin) { lmin = d; lr = r; }
        } else if (d < r) {
            c.stroke_hit = 1.f;
            return;
        }
    }
    if (prefilter && lmin < c.stroke_min) {
        c.stroke_min = lmin;
        c.stroke_r = lr;
    }
}

// accumulate_group_shapes + accumulate_shape_in_group (forward.rs:1295-1503):
// every shape instance of group g whose padded bounds hold the group-local
// point contributes; (lx, ly) is the group-local point.
//
// With CPU = true the fill follows is_inside_bvh (distance/api.rs:259-309):
// the group's root box gates the test and then every shape's winding counts
// (an open path can wind outside its own box).
//
// fill_cut / stroke_cut: the prefiltered distance cut-offs of shape_fill /
// shape_stroke (BIG: none).
template <int M>
DVG_DEV void group_cov(const DevScene &S, uint32_t g, const float *gd, float lx, float ly, bool do_fill,
                       bool do_stroke, uint32_t rule, bool prefilter, Cov &c, float fill_cut = BIG,
                       float stroke_cut = BIG) {
    constexpr bool CPU = M != COV_FWD;
    uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
    if (CPU && do_fill) {
        const float *rb = S.group_cpu_bounds + g * 4u;
        bool has = rb[0] <= rb[2] && rb[1] <= rb[3];
        if (!has || (lx >= rb[0] && lx <= rb[2] && ly >= rb[1] && ly <= rb[3])) {
            for (uint32_t i = 0; i < cnt; ++i) {
                uint32_t si = (uint32_t)S.group_shapes[off + i];
                const float *xf = S.shape_xform + si * 6u;
                float spx = xf[0] * lx + xf[1] * ly + xf[2];
                float spy = xf[3] * lx + xf[4] * ly + xf[5];
                shape_fill<M>(S, S.shape_data + si * 15u, rule, false, spx, spy, c);
            }
        }
        do_fill = false;
    }
    for (uint32_t i = 0; i < cnt; ++i) {
        uint32_t si = (uint32_t)S.group_shapes[off + i];
        const float *b = S.shape_bounds + si * 4u;
        bool inb = b[0] <= b[2] && b[1] <= b[3];
        if (inb && (lx < b[0] || lx > b[2] || ly < b[1] || ly > b[3])) inb = false;
        if (!inb) continue;
        const float *xf = S.shape_xform + si * 6u;
        float spx = xf[0] * lx + xf[1] * ly + xf[2];
        float spy = xf[3] * lx + xf[4] * ly + xf[5];
        const float *sd = S.shape_data + si * 15u;
        if (do_fill) shape_fill<COV_FWD>(S, sd, rule, prefilter, spx, spy, c, fill_cut);
        if (do_stroke && !(c.stroke_hit > 0.f && !prefilter)) shape_stroke<M>(S, sd, prefilter, spx, spy, c, stroke_cut);
    }
}

DVG_DEV void group_local(const DevScene &S, uint32_t g, float px, float py, float &lx, float &ly) {
    const float *x = S.group_xform + g * 6u;
    lx = x[0] * px + x[1] * py + x[2];
    ly = x[3] * px + x[4] * py + x[5];
}

DVG_DEV void blend(float4 &o, float4 c, float a) {
    float inv = 1.f - a;
    o.x = c.x * a + o.x * inv;
    o.y = c.y * a + o.y * inv;
    o.z = c.z * a + o.z * inv;
    o.w = a + o.w * inv;
}

// eval_scene_tiled, forward.rs:1133-1290: premultiplied RGBA at (px, py).
// PF = S.prefilter, a template argument so that the analytic-coverage code
// (with its correctly rounded solves) stays out of the point-sampled kernel.
template <bool PF>
DVG_DEV float4 eval_forward(const DevScene &S, const uint32_t *list, uint32_t n, float px, float py, float4 bg) {
    float4 out = make_float4(bg.x * bg.w, bg.y * bg.w, bg.z * bg.w, bg.w);
    const bool prefilter = PF;
    for (uint32_t e = 0; e < n; ++e) {
        uint32_t g = list[e];
        const float *gd = S.group_data + g * 16u;
        uint32_t fk = (uint32_t)gd[2], sk = (uint32_t)gd[4], rule = (uint32_t)gd[7];
        if (fk == DVG_PAINT_NONE && sk == DVG_PAINT_NONE) continue;
        float lx, ly;
        group_local(S, g, px, py, lx, ly);
        Cov c = {BIG, 0.f, 0.f, BIG, 0.f, 0.f, 0.f};
        float inv_scale = S.group_inv_scale[g];
        float gscale = inv_scale > 0.f ? 1.f / inv_scale : 1.f;
        // prefiltered coverage saturates: smoothstep((d + 1) / 2) of the
        // scaled fill distance is 0 or 1 from |d| >= 1, the stroke's from
        // d >= r + 1 (r <= the group's largest radius); 2 / gscale keeps the
        // cut-offs clear of that in f32
        float fcut = BIG, scut = BIG;
        if (prefilter) {
            float unit = 2.0f / gscale;
            fcut = unit;
            scut = S.group_pf[2u * g] + unit;
        }
        group_cov<COV_FWD>(S, g, gd, lx, ly, fk != DVG_PAINT_NONE, sk != DVG_PAINT_NONE, rule, prefilter, c, fcut,
                           scut);
        // blend_group, forward.rs:339-404: stroke first, then fill
        if (sk != DVG_PAINT_NONE) {
            float cov = 0.f;
            if (prefilter) {
                float smin = c.stroke_min * gscale, sr = c.stroke_r * gscale;
                if (smin < BIG && sr > 0.f) cov = smoothstep_unit(smin + sr) - smoothstep_unit(smin - sr);
            } else if (c.stroke_hit > 0.f) {
                cov = 1.f;
            }
            if (cov > 0.f) {
                float4 col = paint_fwd(S, sk, (uint32_t)gd[5], gd + 12, px, py);
                blend(out, col, cov * col.w);
            }
        }
        if (fk != DVG_PAINT_NONE) {
            bool inside = (rule == 1u) ? c.crossings > 0.f : c.winding != 0.f;
            float cov = 0.f;
            if (prefilter) {
                float fmin = c.fill_min * gscale;
                if (fmin < BIG) cov = smoothstep_unit(inside ? fmin : -fmin);
            } else if (inside) {
                cov = 1.f;
            }
            if (cov > 0.f) {
                float4 col = paint_fwd(S, fk, (uint32_t)gd[3], gd + 8, px, py);
                blend(out, col, cov * col.w);
            }
        }
    }
    return out;
}

DVG_DEV float4 background_at(const DevScene &S, uint32_t x, uint32_t y) {
    if (S.bg_image) {
        const float *b = S.bg_image + ((size_t)y * S.W + x) * 4u;
        return make_float4(b[0], b[1], b[2], b[3]);
    }
    return make_float4(S.bg[0], S.bg[1], S.bg[2], S.bg[3]);
}

// ---------------------------------------------------------------------------
// Filter weight sum of every pixel, gathered from the samples of its
// neighbourhood (the same sum rasterize_weights_f32 scatters).
// Unit u covers the 16x16 pixel block u of the canvas (row-major blocks).
template <bool CPU>
DVG_DEV void weights_unit(const DevScene &S, uint32_t u, float *__restrict__ weight) {
    uint32_t gx = (S.W + 15u) / 16u;
    uint32_t x = (u % gx) * 16 + (threadIdx.x & 15), y = (u / gx) * 16 + (threadIdx.x >> 4);
    if (x >= S.W || y >= S.H) return;
    float xc = (float)x + 0.5f, yc = (float)y + 0.5f;
    float sum = 0.f;
    int ri = S.ri;
    for (int dy = -ri; dy <= ri; ++dy)
        for (int dx = -ri; dx <= ri; ++dx) {
            int sx_ = (int)x - dx, sy_ = (int)y - dy;
            if (sx_ < 0 || sy_ < 0 || sx_ >= (int)S.W || sy_ >= (int)S.H) continue;
            for (uint32_t sy = 0; sy < S.nsy; ++sy)
                for (uint32_t sx = 0; sx < S.nsx; ++sx) {
                    float px, py;
                    sample_point<CPU>(S, (uint32_t)sx_, (uint32_t)sy_, sx, sy, px, py);
                    float w = filter_weight(S.filter_type, xc - px, yc - py, S.filter_radius);
                    if (w > 0.f) sum += w;
                }
        }
    weight[(size_t)y * S.W + x] = sum;
}

template <bool CPU>
__global__ void k_weights(DevScene S, float *__restrict__ weight) {
    weights_unit<CPU>(S, blockIdx.x, weight);
}

// The splat of a tile's samples lands in the tile plus a halo of ri pixels.
// For ri <= SPLAT_HALO it is summed in LDS and flushed with one global atomic
// per touched pixel channel; wider filters splat straight to global memory.
constexpr int SPLAT_HALO = 4;
constexpr int SPLAT_EDGE = TILE + 2 * SPLAT_HALO;

template <bool PF>
DVG_DEV void forward_unit(const DevScene &S, uint32_t unit, const float *__restrict__ weight,
                          float *__restrict__ accum) {
    __shared__ float lacc[4][SPLAT_EDGE * SPLAT_EDGE];
    uint32_t upt = units_per_tile(S);
    uint32_t tile = unit / upt, s0 = (unit - tile * upt) * S.unit_samples;
    uint32_t lb = S.tile_offsets[tile], le = S.fwd_end[tile];
    const uint32_t *list = S.tile_entries + lb;
    uint32_t n = le - lb;
    uint32_t total = min(TILE * TILE * S.spp, s0 + S.unit_samples);
    int ri = S.ri;
    bool lds = ri <= SPLAT_HALO;
    int edge = TILE + 2 * ri;
    int x0 = (int)((tile % S.tiles_x) * TILE) - ri, y0 = (int)((tile / S.tiles_x) * TILE) - ri;
    if (lds) {
        for (int i = threadIdx.x; i < edge * edge; i += BLOCK)
            for (int c = 0; c < 4; ++c) lacc[c][i] = 0.f;
        __syncthreads();
    }
    for (uint32_t s = s0 + threadIdx.x; s < total; s += BLOCK) {
        uint32_t x, y, sx, sy;
        tile_sample(S, tile, s, x, y, sx, sy);
        if (x >= S.W || y >= S.H) continue;
        float px, py;
        sample_point<false>(S, x, y, sx, sy, px, py);
        float4 col = eval_forward<PF>(S, list, n, px, py, background_at(S, x, y));
        for (int dy = -ri; dy <= ri; ++dy)
            for (int dx = -ri; dx <= ri; ++dx) {
                int xx = (int)x + dx, yy = (int)y + dy;
                if (xx < 0 || yy < 0 || xx >= (int)S.W || yy >= (int)S.H) continue;
                size_t base = (size_t)yy * S.W + xx;
                float ws = weight[base];
                if (ws > 0.f) {
                    float w = filter_weight(S.filter_type, ((float)xx + 0.5f) - px, ((float)yy + 0.5f) - py,
                                            S.filter_radius);
                    if (w > 0.f) {
                        float wn = w / ws;
                        if (lds) {
                            int li = (yy - y0) * edge + (xx - x0);
                            atomicAdd(&lacc[0][li], col.x * wn);
                            atomicAdd(&lacc[1][li], col.y * wn);
                            atomicAdd(&lacc[2][li], col.z * wn);
                            atomicAdd(&lacc[3][li], col.w * wn);
                        } else {
                            float *a = accum + base * 4u;
                            unsafeAtomicAdd(a + 0, col.x * wn);
                            unsafeAtomicAdd(a + 1, col.y * wn);
                            unsafeAtomicAdd(a + 2, col.z * wn);
                            unsafeAtomicAdd(a + 3, col.w * wn);
                        }
                    }
                }
            }
    }
    if (lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < edge * edge; i += BLOCK) {
            int xx = x0 + i % edge, yy = y0 + i / edge;
            if (xx < 0 || yy < 0 || xx >= (int)S.W || yy >= (int)S.H) continue;
            float *a = accum + ((size_t)yy * S.W + xx) * 4u;
            for (int c = 0; c < 4; ++c) {
                float v = lacc[c][i];
                if (v != 0.f) unsafeAtomicAdd(a + c, v);
            }
        }
    }
}

template <bool PF>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(PF ? OCC_FORWARD_PF : OCC_FORWARD))) void k_forward(
    DevScene S, const float *__restrict__ weight, float *__restrict__ accum) {
    forward_unit<PF>(S, blockIdx.x, weight, accum);
}

// resolve_splat_f32 with the premultiplied background of renderer.rs:499-515.
// Unit u covers pixels [256 u, 256 u + 256).  With target and d_render set,
// also writes the L2 loss gradient d_render = 2 (out - target).
DVG_DEV void resolve_unit(const DevScene &S, uint32_t u, const float *__restrict__ weight,
                          const float *__restrict__ accum, float *__restrict__ out, const float *__restrict__ target,
                          float *__restrict__ d_render) {
    uint32_t i = u * BLOCK + threadIdx.x;
    if (i >= S.W * S.H) return;
    float a = S.bg[3];
    float r = S.bg[0] * a, g = S.bg[1] * a, b = S.bg[2] * a;
    if (weight[i] > 0.f) {
        float4 c = reinterpret_cast<const float4 *>(accum)[i];
        r = c.x; g = c.y; b = c.z; a = c.w;
    }
    if (a > 1.0e-8f) { r = r / a; g = g / a; b = b / a; }
    reinterpret_cast<float4 *>(out)[i] = make_float4(r, g, b, a);
    if (target && d_render) {
        float4 t = reinterpret_cast<const float4 *>(target)[i];
        reinterpret_cast<float4 *>(d_render)[i] =
            make_float4(2.0f * (r - t.x), 2.0f * (g - t.y), 2.0f * (b - t.z), 2.0f * (a - t.w));
    }
}

__global__ void k_resolve(DevScene S, const float *__restrict__ weight, const float *__restrict__ accum,
                          float *__restrict__ out) {
    resolve_unit(S, blockIdx.x, weight, accum, out, nullptr, nullptr);
}

// ---------------------------------------------------------------------------
// Backward (src/backward/*.rs semantics).

// compute_filter_weight (backward/filters.rs:159-188)
DVG_DEV float fw_cpu(uint32_t type, float radius, float dx, float dy) {
    if (radius <= 0.f) return 1.f;
    if (fabsf(dx) > radius || fabsf(dy) > radius) return 0.f;
    if (type == DVG_FILTER_BOX) { float d = 2.0f * radius; return 1.0f / (d * d); }
    if (type == DVG_FILTER_TENT) {
        float fx = radius - fabsf(dx), fy = radius - fabsf(dy);
        float norm = 1.0f / (radius * radius);
        return fx * fy * norm * norm;
    }
    if (type == DVG_FILTER_RADIAL_PARABOLIC) {
        float rx = 1.0f - (dx / radius) * (dx / radius), ry = 1.0f - (dy / radius) * (dy / radius);
        return (4.0f / 3.0f) * rx * (4.0f / 3.0f) * ry;
    }
    const float PI = 3.14159274101257324219f;  // core::f32::consts::PI
    float ndx = (dx / (2.0f * radius)) + 0.5f, ndy = (dy / (2.0f * radius)) + 0.5f;
    float fx = 0.5f * (1.0f - cosf(2.0f * PI * ndx)), fy = 0.5f * (1.0f - cosf(2.0f * PI * ndy));
    float norm = 1.0f / (radius * radius);
    return fx * fy * norm;
}

// d_compute_filter_weight (backward/filters.rs:190-239)
DVG_DEV float d_fw_cpu(uint32_t type, float radius, float dx, float dy, float d_ret) {
    if (type == DVG_FILTER_BOX) {
        float denom = 2.0f * radius;
        return denom != 0.f ? d_ret * (-2.0f) * denom / ((denom * denom) * denom) : 0.f;
    }
    if (type == DVG_FILTER_TENT) {
        float fx = radius - fabsf(dx), fy = radius - fabsf(dy);
        float norm = 1.0f / (radius * radius);
        float d_fx = d_ret * fy * norm, d_fy = d_ret * fx * norm, d_norm = d_ret * fx * fy;
        float r2 = radius * radius;
        return radius != 0.f ? d_fx + d_fy + (-4.0f) * d_norm / ((r2 * r2) * radius) : 0.f;
    }
    if (type == DVG_FILTER_RADIAL_PARABOLIC) {
        float r3 = radius * radius * radius;
        return r3 != 0.f ? -(2.0f * dx * dx + 2.0f * dy * dy) / r3 : 0.f;
    }
    const float PI = 3.14159274101257324219f;
    float ndx = (dx / (2.0f * radius)) + 0.5f, ndy = (dy / (2.0f * radius)) + 0.5f;
    float fx = 0.5f * (1.0f - cosf(2.0f * PI * ndx)), fy = 0.5f * (1.0f - cosf(2.0f * PI * ndy));
    float norm = 1.0f / (radius * radius);
    float d_fx = d_ret * fy * norm, d_fy = d_ret * fx * norm, d_norm = d_ret * fx * fy;
    float d_ndx = d_fx * 0.5f * sinf(2.0f * PI * ndx) * (2.0f * PI);
    float d_ndy = d_fy * 0.5f * sinf(2.0f * PI * ndy) * (2.0f * PI);
    float d2 = 2.0f * radius;
    return d_ndx * (-2.0f * dx / (d2 * d2)) + d_ndy * (-2.0f * dy / (d2 * d2)) +
           (-2.0f) * d_norm / ((radius * radius) * radius);
}

// gather_d_color (backward/filters.rs:70-103)
DVG_DEV float4 gather_d_color(const DevScene &S, const float *weight, const float *d_render, float ptx, float pty) {
    int x = (int)floorf(ptx), y = (int)floorf(pty);
    int ri = (int)ceilf(S.filter_radius);
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int dy = -ri; dy <= ri; ++dy)
        for (int dx = -ri; dx <= ri; ++dx) {
            int xx = x + dx, yy = y + dy;
            if (xx < 0 || yy < 0 || xx >= (int)S.W || yy >= (int)S.H) continue;
            float w = fw_cpu(S.filter_type, S.filter_radius, ((float)xx + 0.5f) - ptx, ((float)yy + 0.5f) - pty);
            size_t b = (size_t)yy * S.W + xx;
            float ws = weight[b];
            if (ws > 0.f) {
                float4 d = reinterpret_cast<const float4 *>(d_render)[b];
                o.x += (w / ws) * d.x;
                o.y += (w / ws) * d.y;
                o.z += (w / ws) * d.z;
                o.w += (w / ws) * d.w;
            }
        }
    return o;
}

// accumulate_filter_gradient (backward/filters.rs:106-156), returning d_radius.
DVG_DEV float filter_grad(const DevScene &S, const float *weight, const float *d_render, float4 col, float ptx,
                          float pty) {
    int x = (int)floorf(ptx), y = (int)floorf(pty);
    int ri = (int)ceilf(S.filter_radius);
    float acc = 0.f;
    for (int dy = -ri; dy <= ri; ++dy)
        for (int dx = -ri; dx <= ri; ++dx) {
            int xx = x + dx, yy = y + dy;
            if (xx < 0 || yy < 0 || xx >= (int)S.W || yy >= (int)S.H) continue;
            size_t b = (size_t)yy * S.W + xx;
            float ws = weight[b];
            if (ws <= 0.f) continue;
            float xc = (float)xx + 0.5f, yc = (float)yy + 0.5f;
            float w = fw_cpu(S.filter_type, S.filter_radius, xc - ptx, yc - pty);
            if (w <= 0.f) continue;
            float4 d = reinterpret_cast<const float4 *>(d_render)[b];
            float dot = d.x * col.x + d.y * col.y + d.z * col.z + d.w * col.w;
            float denom = ws * ws;
            float d_w = denom > 0.f ? (dot * ws - w * dot * (ws - w)) / denom : 0.f;
            if (d_w != 0.f) acc += d_fw_cpu(S.filter_type, S.filter_radius, xc - ptx, yc - pty, d_w);
        }
    return acc;
}

// Non-prefiltered fragment tests of one group (sampling.rs:95-118 with the
// CPU queries): stroke hit, then fill inside.
template <int M = COV_CPU>
DVG_DEV void group_frags(const DevScene &S, uint32_t g, float px, float py, bool &stroke, bool &fill, bool *uns = nullptr) {
    const float *gd = S.group_data + g * 16u;
    uint32_t fk = (uint32_t)gd[2], sk = (uint32_t)gd[4], rule = (uint32_t)gd[7];
    stroke = fill = false;
    if (fk == DVG_PAINT_NONE && sk == DVG_PAINT_NONE) return;
    float lx, ly;
    group_local(S, g, px, py, lx, ly);
    Cov c = {BIG, 0.f, 0.f, BIG, 0.f, 0.f, 0.f};
    group_cov<M>(S, g, gd, lx, ly, fk != DVG_PAINT_NONE, sk != DVG_PAINT_NONE, rule, false, c);
    stroke = sk != DVG_PAINT_NONE && c.stroke_hit > 0.f;
    fill = fk != DVG_PAINT_NONE && ((rule == 1u) ? c.crossings > 0.f : c.winding != 0.f);
    if (M == COV_CHK && c.unsure > 0.f) *uns = true;
}

// group_frags at the two query points of a boundary sample.
template <int M>
DVG_DEV void group_frags_pair(const DevScene &S, uint32_t g, float px0, float py0, float px1, float py1, bool &st0,
                              bool &fi0, bool &st1, bool &fi1, bool &uns) {
    const float *gd = S.group_data + g * 16u;
    uint32_t fk = (uint32_t)gd[2], sk = (uint32_t)gd[4], rule = (uint32_t)gd[7];
    st0 = fi0 = st1 = fi1 = false;
    if (fk == DVG_PAINT_NONE && sk == DVG_PAINT_NONE) return;
    float lx0, ly0, lx1, ly1;
    group_local(S, g, px0, py0, lx0, ly0);
    group_local(S, g, px1, py1, lx1, ly1);
    Cov c0 = {BIG, 0.f, 0.f, BIG, 0.f, 0.f, 0.f}, c1 = c0;
    uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
    if (fk != DVG_PAINT_NONE) {
        // the group's root box gates every shape's winding (group_cov)
        const float *rb = S.group_cpu_bounds + g * 4u;
        bool has = rb[0] <= rb[2] && rb[1] <= rb[3];
        bool a0 = !has || (lx0 >= rb[0] && lx0 <= rb[2] && ly0 >= rb[1] && ly0 <= rb[3]);
        bool a1 = !has || (lx1 >= rb[0] && lx1 <= rb[2] && ly1 >= rb[1] && ly1 <= rb[3]);
        if (a0 || a1) {
            for (uint32_t i = 0; i < cnt; ++i) {
                uint32_t si = (uint32_t)S.group_shapes[off + i];
                const float *xf = S.shape_xform + si * 6u;
                float sx0 = xf[0] * lx0 + xf[1] * ly0 + xf[2], sy0 = xf[3] * lx0 + xf[4] * ly0 + xf[5];
                float sx1 = xf[0] * lx1 + xf[1] * ly1 + xf[2], sy1 = xf[3] * lx1 + xf[4] * ly1 + xf[5];
                shape_fill_pair<M>(S, S.shape_data + si * 15u, rule, sx0, sy0, a0, sx1, sy1, a1, c0, c1);
            }
        }
    }
    if (sk != DVG_PAINT_NONE) {
        for (uint32_t i = 0; i < cnt; ++i) {
            uint32_t si = (uint32_t)S.group_shapes[off + i];
            const float *b = S.shape_bounds + si * 4u;
            bool ok = b[0] <= b[2] && b[1] <= b[3];
            bool in0 = ok && !(lx0 < b[0] || lx0 > b[2] || ly0 < b[1] || ly0 > b[3]) && !(c0.stroke_hit > 0.f);
            bool in1 = ok && !(lx1 < b[0] || lx1 > b[2] || ly1 < b[1] || ly1 > b[3]) && !(c1.stroke_hit > 0.f);
            if (!in0 && !in1) continue;
            const float *xf = S.shape_xform + si * 6u;
            const float *sd = S.shape_data + si * 15u;
            if (in0) shape_stroke<M>(S, sd, false, xf[0] * lx0 + xf[1] * ly0 + xf[2], xf[3] * lx0 + xf[4] * ly0 + xf[5], c0);
            if (in1) shape_stroke<M>(S, sd, false, xf[0] * lx1 + xf[1] * ly1 + xf[2], xf[3] * lx1 + xf[4] * ly1 + xf[5], c1);
        }
    }
    st0 = sk != DVG_PAINT_NONE && c0.stroke_hit > 0.f;
    st1 = sk != DVG_PAINT_NONE && c1.stroke_hit > 0.f;
    fi0 = fk != DVG_PAINT_NONE && ((rule == 1u) ? c0.crossings > 0.f : c0.winding != 0.f);
    fi1 = fk != DVG_PAINT_NONE && ((rule == 1u) ? c1.crossings > 0.f : c1.winding != 0.f);
    if (M == COV_CHK && (c0.unsure > 0.f || c1.unsure > 0.f)) uns = true;
}

DVG_DEV float4 frag_color(const DevScene &S, uint32_t g, bool stroke, float px, float py) {
    const float *gd = S.group_data + g * 16u;
    return stroke ? paint_cpu(S, (uint32_t)gd[4], (uint32_t)gd[5], gd + 12, px, py)
                  : paint_cpu(S, (uint32_t)gd[2], (uint32_t)gd[3], gd + 8, px, py);
}

// CPU blend of sampling.rs:155-160 on premultiplied colours.
DVG_DEV void blend_cpu(float4 &p, float4 c) {
    float a = c.w;
    p = make_float4(p.x * (1.0f - a) + c.x * a, p.y * (1.0f - a) + c.y * a, p.z * (1.0f - a) + c.z * a,
                    p.w * (1.0f - a) + a);
}

// One pass over the tile's groups: counts fragments, composites them all into
// `fin`, composites the first `start` into `pre`, and records the list
// positions of fragments [start, start + FRAG_WINDOW) as (pos << 1 | stroke)
// in ids[k * BLOCK] (the lane's column of an LDS table).
DVG_DEV uint32_t walk_window(const DevScene &S, const uint32_t *list, uint32_t n, float px, float py, uint32_t start,
                             uint32_t *ids, float4 &pre, float4 &fin) {
    uint32_t count = 0;
    for (uint32_t e = 0; e < n; ++e) {
        uint32_t g = list[e];
        bool st, fi;
        group_frags(S, g, px, py, st, fi);
        for (int k = 0; k < 2; ++k) {
            bool is_stroke = k == 0;
            if (!(is_stroke ? st : fi)) continue;
            float4 c = frag_color(S, g, is_stroke, px, py);
            blend_cpu(fin, c);
            if (count < start) blend_cpu(pre, c);
            else if (count < start + FRAG_WINDOW) ids[(count - start) * BLOCK] = (e << 1) | (is_stroke ? 1u : 0u);
            ++count;
        }
    }
    return count;
}

// Paint-gradient staging.  A tile's list entries e < LDS_GROUP_CAP get LDS
// slots for their fill and stroke paint adjoints (solid: 4 floats RGBA;
// gradient: 4 floats for p0..p3, then one offset and 4 colour floats per
// stop), allocated by a prefix sum over the entries; an entry whose slot does
// not fit in STAGE_FLOATS adds to global memory directly.  The slots are
// flushed with one global atomic per non-zero float at the end of the tile.
constexpr uint32_t STAGE_FLOATS = 2048;
constexpr uint32_t STAGE_NONE = 0xffffffffu;

template <bool LDS>
DVG_DEV void add_f(float *p, float v) {
    if (LDS) atomicAdd(p, v);
    else unsafeAtomicAdd(p, v);
}

DVG_DEV uint32_t paint_slot_size(const DevScene &S, uint32_t kind, uint32_t gidx) {
    if (kind == DVG_PAINT_SOLID) return 4u;
    if (kind == DVG_PAINT_LINEAR || kind == DVG_PAINT_RADIAL) return 4u + 5u * (uint32_t)S.gradient_data[gidx * 8u + 6u];
    return 0u;
}

// d_sample_paint for gradient paints (paint.rs:77-280).  dgp receives the
// p0..p3 adjoint (4 floats), doffs the stop offsets, dcols the stop colours.
template <bool LDS>
DVG_DEV void d_gradient_paint(const DevScene &S, uint32_t gidx, float px, float py, float4 d, float *dgp,
                              float *doffs, float *dcols, float &dtx, float &dty) {
    const float *gd = S.gradient_data + gidx * 8u;
    uint32_t gtype = (uint32_t)gd[0];
    float p0 = gd[1], p1 = gd[2], p2 = gd[3], p3 = gd[4];
    uint32_t so = (uint32_t)gd[5], sc = (uint32_t)gd[6];
    if (sc == 0) return;
    const float *offs = S.stop_offsets + so;
    const float *cols = S.stop_colors + so * 4u;
    float t, denom = 0.f, offx = 0.f, offy = 0.f, nx = 0.f, ny = 0.f;
    if (gtype == 0) {
        float ex = p2 - p0, ey = p3 - p1;
        denom = fmaxf(ex * ex + ey * ey, 1.0e-3f);
        t = ((px - p0) * ex + (py - p1) * ey) / denom;
    } else {
        offx = px - p0;
        offy = py - p1;
        nx = offx / p2;
        ny = offy / p3;
        t = sqrtf(nx * nx + ny * ny);
    }
    if (t < offs[0]) {
        add_f<LDS>(dcols + 0, d.x); add_f<LDS>(dcols + 1, d.y);
        add_f<LDS>(dcols + 2, d.z); add_f<LDS>(dcols + 3, d.w);
        return;
    }
    for (uint32_t i = 0; i + 1 < sc; ++i) {
        float cu = offs[i], nxo = offs[i + 1];
        if (t >= cu && t < nxo) {
            float tt = (t - cu) / (nxo - cu);
            const float *c0 = cols + i * 4u, *c1 = c0 + 4;
            float *d0 = dcols + i * 4u, *d1 = d0 + 4;
            add_f<LDS>(d0 + 0, d.x * (1.0f - tt)); add_f<LDS>(d0 + 1, d.y * (1.0f - tt));
            add_f<LDS>(d0 + 2, d.z * (1.0f - tt)); add_f<LDS>(d0 + 3, d.w * (1.0f - tt));
            add_f<LDS>(d1 + 0, d.x * tt); add_f<LDS>(d1 + 1, d.y * tt);
            add_f<LDS>(d1 + 2, d.z * tt); add_f<LDS>(d1 + 3, d.w * tt);
            float d_tt = d.x * (c1[0] - c0[0]) + d.y * (c1[1] - c0[1]) + d.z * (c1[2] - c0[2]) + d.w * (c1[3] - c0[3]);
            float dof = nxo - cu;
            if (fabsf(dof) > 0.f) {
                add_f<LDS>(doffs + i, d_tt * (tt - 1.0f) / dof);
                add_f<LDS>(doffs + i + 1, -d_tt * tt / dof);
            }
            float d_t = d_tt / dof;
            if (gtype == 0) {
                float dbx = d_t * (-(px - p0) - (p2 - p0)) / denom, dby = d_t * (-(py - p1) - (p3 - p1)) / denom;
                float dex = d_t * (px - p0) / denom, dey = d_t * (py - p1) / denom;
                float d_l = -d_t * t / denom;
                float l2 = (p2 - p0) * (p2 - p0) + (p3 - p1) * (p3 - p1);
                if (l2 > 1.0e-3f) {
                    add_f<LDS>(dgp + 0, dbx + 2.0f * d_l * (p0 - p2));
                    add_f<LDS>(dgp + 1, dby + 2.0f * d_l * (p1 - p3));
                    add_f<LDS>(dgp + 2, dex + 2.0f * d_l * (p2 - p0));
                    add_f<LDS>(dgp + 3, dey + 2.0f * d_l * (p3 - p1));
                } else {
                    add_f<LDS>(dgp + 0, dbx); add_f<LDS>(dgp + 1, dby);
                    add_f<LDS>(dgp + 2, dex); add_f<LDS>(dgp + 3, dey);
                }
                dtx += dbx + dex;
                dty += dby + dey;
            } else {
                // d_length (backward/math.rs:33-41)
                float l = sqrtf(nx * nx + ny * ny);
                float dnx = 0.f, dny = 0.f;
                if (l != 0.f) {
                    float dls = 0.5f * d_t / l;
                    dnx = nx * (2.0f * dls);
                    dny = ny * (2.0f * dls);
                }
                float dox = dnx / p2, doy = dny / p3;
                add_f<LDS>(dgp + 0, -dox);
                add_f<LDS>(dgp + 1, -doy);
                add_f<LDS>(dgp + 2, -dnx * offx / (p2 * p2));
                add_f<LDS>(dgp + 3, -dny * offy / (p3 * p3));
                dtx += -dox;
                dty += -doy;
            }
            return;
        }
    }
    float *dl = dcols + (sc - 1) * 4u;
    add_f<LDS>(dl + 0, d.x); add_f<LDS>(dl + 1, d.y);
    add_f<LDS>(dl + 2, d.z); add_f<LDS>(dl + 3, d.w);
}

// Paint adjoint of fragment (list entry e, stroke or fill) at (px, py).
DVG_DEV void paint_adjoint(const DevScene &S, const DevGrads &G, float *s_stage, const uint32_t *s_off,
                           const uint32_t *list, uint32_t e, bool is_stroke, float4 dv, float px, float py, float &dtx,
                           float &dty) {
    uint32_t g = list[e];
    const float *gd = S.group_data + g * 16u;
    uint32_t kind = (uint32_t)gd[is_stroke ? 4 : 2];
    uint32_t off = e < LDS_GROUP_CAP ? s_off[2u * e + (is_stroke ? 1u : 0u)] : STAGE_NONE;
    if (kind == DVG_PAINT_SOLID) {
        if (off != STAGE_NONE) {
            float *sp = s_stage + off;
            atomicAdd(sp + 0, dv.x); atomicAdd(sp + 1, dv.y);
            atomicAdd(sp + 2, dv.z); atomicAdd(sp + 3, dv.w);
        } else {
            float *dgp = G.d_group_data + g * 16u + (is_stroke ? 12u : 8u);
            unsafeAtomicAdd(dgp + 0, dv.x); unsafeAtomicAdd(dgp + 1, dv.y);
            unsafeAtomicAdd(dgp + 2, dv.z); unsafeAtomicAdd(dgp + 3, dv.w);
        }
    } else if (kind == DVG_PAINT_LINEAR || kind == DVG_PAINT_RADIAL) {
        uint32_t gidx = (uint32_t)gd[is_stroke ? 5 : 3];
        if (off != STAGE_NONE) {
            uint32_t sc = (uint32_t)S.gradient_data[gidx * 8u + 6u];
            float *b = s_stage + off;
            d_gradient_paint<true>(S, gidx, px, py, dv, b, b + 4, b + 4 + sc, dtx, dty);
        } else {
            uint32_t so = (uint32_t)S.gradient_data[gidx * 8u + 5u];
            d_gradient_paint<false>(S, gidx, px, py, dv, G.d_gradient_data + gidx * 8u + 1u, G.d_stop_offsets + so,
                                    G.d_stop_colors + so * 4u, dtx, dty);
        }
    }
}

// Slot allocation for a tile (all threads call it; wave 0 computes).
DVG_DEV void stage_build(const DevScene &S, const uint32_t *list, uint32_t n, uint32_t *s_off, float *s_stage,
                         uint32_t *s_total) {
    uint32_t m = n < LDS_GROUP_CAP ? n : LDS_GROUP_CAP;
    if (threadIdx.x < 64) {
        uint32_t lane = threadIdx.x;
        uint32_t carry = 0;
        for (uint32_t base = 0; base < m; base += 64) {
            uint32_t e = base + lane;
            uint32_t sz0 = 0, sz1 = 0;
            if (e < m) {
                const float *gd = S.group_data + list[e] * 16u;
                sz0 = paint_slot_size(S, (uint32_t)gd[2], (uint32_t)gd[3]);
                sz1 = paint_slot_size(S, (uint32_t)gd[4], (uint32_t)gd[5]);
            }
            uint32_t sz = sz0 + sz1, v = sz;
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t t = __shfl_up(v, o, 64);
                if ((int)lane >= o) v += t;
            }
            uint32_t start = carry + v - sz;
            if (e < m) {
                bool fits = start + sz <= STAGE_FLOATS;
                s_off[2u * e] = (fits && sz0) ? start : STAGE_NONE;
                s_off[2u * e + 1u] = (fits && sz1) ? start + sz0 : STAGE_NONE;
            }
            carry += __shfl(v, 63, 64);
        }
        if (lane == 0) *s_total = carry < STAGE_FLOATS ? carry : STAGE_FLOATS;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < *s_total; i += BLOCK) s_stage[i] = 0.f;
    __syncthreads();
}

// Flush of the staged paint adjoints (after a barrier).
DVG_DEV void stage_flush(const DevScene &S, const DevGrads &G, const uint32_t *list, uint32_t n,
                         const uint32_t *s_off, const float *s_stage) {
    uint32_t m = n < LDS_GROUP_CAP ? n : LDS_GROUP_CAP;
    for (uint32_t i = threadIdx.x; i < 2u * m; i += BLOCK) {
        uint32_t off = s_off[i];
        if (off == STAGE_NONE) continue;
        uint32_t e = i >> 1;
        bool is_stroke = (i & 1u) != 0;
        uint32_t g = list[e];
        const float *gd = S.group_data + g * 16u;
        uint32_t kind = (uint32_t)gd[is_stroke ? 4 : 2];
        const float *b = s_stage + off;
        if (kind == DVG_PAINT_SOLID) {
            float *dst = G.d_group_data + g * 16u + (is_stroke ? 12u : 8u);
            for (int c = 0; c < 4; ++c)
                if (b[c] != 0.f) unsafeAtomicAdd(dst + c, b[c]);
        } else {
            uint32_t gidx = (uint32_t)gd[is_stroke ? 5 : 3];
            uint32_t so = (uint32_t)S.gradient_data[gidx * 8u + 5u], sc = (uint32_t)S.gradient_data[gidx * 8u + 6u];
            for (int j = 0; j < 4; ++j)
                if (b[j] != 0.f) unsafeAtomicAdd(G.d_gradient_data + gidx * 8u + 1u + j, b[j]);
            for (uint32_t k = 0; k < sc; ++k)
                if (b[4 + k] != 0.f) unsafeAtomicAdd(G.d_stop_offsets + so + k, b[4 + k]);
            for (uint32_t k = 0; k < 4u * sc; ++k)
                if (b[4 + sc + k] != 0.f) unsafeAtomicAdd(G.d_stop_colors + so * 4u + k, b[4 + sc + k]);
        }
    }
}

DVG_DEV float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Reverse-mode step of one fragment (sampling.rs:177-218): colour adjoint of
// fragment (e, stroke) with colour c over prefix colour pc; the running
// adjoint (dr, dg, db, da) moves to the prefix.
DVG_DEV void frag_adjoint(const DevScene &S, const DevGrads &G, float *s_stage, const uint32_t *s_off,
                          const uint32_t *list, uint32_t e, bool is_stroke, float4 c, float4 pc, float px, float py,
                          float &dr, float &dg, float &db, float &da, float &dtx, float &dty) {
    float a = c.w;
    float d_prev_a = da * (1.0f - a);
    float d_alpha = da * (1.0f - pc.w) + (dr * (c.x - pc.x) + dg * (c.y - pc.y) + db * (c.z - pc.z));
    float4 dv = make_float4(dr * a, dg * a, db * a, d_alpha);
    paint_adjoint(S, G, s_stage, s_off, list, e, is_stroke, dv, px, py, dtx, dty);
    dr = dr * (1.0f - a); dg = dg * (1.0f - a); db = db * (1.0f - a); da = d_prev_a;
}

// Fragment flags of one sample as a bit set over the tile list: bit 2e is
// the stroke of list entry e, bit 2e + 1 its fill, so bit order is
// compositing order.  `mask` is the lane's column of an LDS table (stride
// BLOCK); MASK_WORDS words cover lists of up to 16 * MASK_WORDS groups.
constexpr int MASK_WORDS = FRAG_WINDOW;
constexpr uint32_t MASK_GROUPS = 16u * MASK_WORDS;
static_assert(MASK_GROUPS == LDS_MASK_GROUPS, "LDS bit set size");

// Fragments are taken back in blocks of ADJ_BLOCK (mask_adjoint).  With a
// checkpoint ring (global bit-set slots), the walk records for the last
// ADJ_RING blocks the composite before the block's first fragment and that
// fragment's bit, so the reverse pass starts those blocks without
// recompositing everything below them.
constexpr uint32_t ADJ_BLOCK = 32;
constexpr uint32_t ADJ_RING = 8;
static_assert(5 * ADJ_RING == 40, "ring words of interior_slot_words");

DVG_DEV void ring_put(float *ring, uint32_t blk, float4 c, uint32_t bit) {
    float *r = ring + (size_t)(blk % ADJ_RING) * 5u * BLOCK;
    r[0] = c.x; r[BLOCK] = c.y; r[2 * BLOCK] = c.z; r[3 * BLOCK] = c.w;
    r[4 * BLOCK] = __uint_as_float(bit);
}

DVG_DEV uint32_t walk_mask(const DevScene &S, const uint32_t *list, uint32_t n, float px, float py, uint32_t *mask,
                           float4 &fin, float *ring) {
    uint32_t nw = (2u * n + 31u) / 32u;
    for (uint32_t w = 0; w < nw; ++w) mask[w * BLOCK] = 0u;
    uint32_t count = 0;
    for (uint32_t e = 0; e < n; ++e) {
        uint32_t g = list[e];
        bool st, fi;
        group_frags(S, g, px, py, st, fi);
        if (!(st || fi)) continue;
        uint32_t bits = (st ? 1u : 0u) | (fi ? 2u : 0u);
        uint32_t bit = 2u * e;
        mask[(bit >> 5) * BLOCK] |= bits << (bit & 31u);
        if (st) {
            if (ring && count % ADJ_BLOCK == 0u) ring_put(ring, count / ADJ_BLOCK, fin, bit);
            blend_cpu(fin, frag_color(S, g, true, px, py));
            ++count;
        }
        if (fi) {
            if (ring && count % ADJ_BLOCK == 0u) ring_put(ring, count / ADJ_BLOCK, fin, bit + 1u);
            blend_cpu(fin, frag_color(S, g, false, px, py));
            ++count;
        }
    }
    return count;
}

// Colour adjoint of the nf fragments recorded in a lane's bit set (bit 2e:
// stroke of list entry e, bit 2e+1: its fill; fragment order = bit order),
// taken in reverse as sampling.rs:76-224 does.  The prefix composite before
// each fragment is recomposited from the bit set, not by a second walk over
// the geometry: blocks of ADJ_BLOCK fragments from the top, each block's
// prefix composited once from the background, then within the block from its
// start -- O(nf^2 / ADJ_BLOCK + nf ADJ_BLOCK) paint evaluations, the same
// blends in the same order as a full recomposition.
//
// The running adjoint shrinks by (1 - alpha) per fragment; once it is exactly
// zero every fragment below adds exactly zero (finite paints), so the walk
// down stops there -- under a deep stack of translucent fragments the f32
// adjoint reaches zero after a few hundred fragments.

DVG_DEV bool adjoint_zero(float dr, float dg, float db, float da) {
    return dr == 0.f && dg == 0.f && db == 0.f && da == 0.f;
}

DVG_DEV uint32_t next_bit(const uint32_t *mask, uint32_t nw, uint32_t &w, uint32_t &word) {
    while (word == 0u) {
        if (++w >= nw) return 0xffffffffu;
        word = mask[w * BLOCK];
    }
    uint32_t b = (uint32_t)__ffs(word) - 1u;
    word &= word - 1u;
    return w * 32u + b;
}

DVG_DEV void mask_adjoint(const DevScene &S, const DevGrads &G, float *s_stage, const uint32_t *s_off,
                          const uint32_t *list, uint32_t n, const uint32_t *mask, const float *ring, uint32_t nf,
                          float4 bg, float px, float py, float &dr, float &dg, float &db, float &da, float &dtx,
                          float &dty) {
    uint32_t nw = (2u * n + 31u) / 32u;
    if (nf == 0u || nw == 0u) return;
    uint32_t nblk = (nf + ADJ_BLOCK - 1u) / ADJ_BLOCK;
    for (int jb = (int)nblk - 1; jb >= 0; --jb) {
        if (adjoint_zero(dr, dg, db, da)) return;
        uint32_t first = (uint32_t)jb * ADJ_BLOCK, cnt = min(ADJ_BLOCK, nf - first);
        // prefix composite before fragment `first`; (w0, word0) = the scan
        // state at the block's first fragment
        float4 pre = bg;
        uint32_t w0 = 0u, word0 = mask[0];
        if (jb > 0 && ring && (uint32_t)jb + ADJ_RING >= nblk) {
            // recorded by the walk
            const float *r = ring + (size_t)((uint32_t)jb % ADJ_RING) * 5u * BLOCK;
            pre = make_float4(r[0], r[BLOCK], r[2 * BLOCK], r[3 * BLOCK]);
            uint32_t b0 = __float_as_uint(r[4 * BLOCK]);
            if (b0 >= 2u * n) return;
            w0 = b0 >> 5;
            word0 = mask[w0 * BLOCK] & (0xffffffffu << (b0 & 31u));
        } else {
            // (the bit set holds exactly nf bits: next_bit never runs out below)
            for (uint32_t k = 0; k < first; ++k) {
                uint32_t bit = next_bit(mask, nw, w0, word0);
                if (bit >= 2u * n) return;
                blend_cpu(pre, frag_color(S, list[bit >> 1], (bit & 1u) == 0u, px, py));
            }
        }
        for (int i = (int)cnt - 1; i >= 0; --i) {
            float4 pc = pre;
            uint32_t w = w0, word = word0;
            for (int k = 0; k < i; ++k) {
                uint32_t bit = next_bit(mask, nw, w, word);
                if (bit >= 2u * n) return;
                blend_cpu(pc, frag_color(S, list[bit >> 1], (bit & 1u) == 0u, px, py));
            }
            uint32_t bit = next_bit(mask, nw, w, word);
            if (bit >= 2u * n) return;
            uint32_t e = bit >> 1;
            bool is_stroke = (bit & 1u) == 0u;
            float4 c = frag_color(S, list[e], is_stroke, px, py);
            frag_adjoint(S, G, s_stage, s_off, list, e, is_stroke, c, pc, px, py, dr, dg, db, da, dtx, dty);
            if (adjoint_zero(dr, dg, db, da)) return;
        }
    }
}

// One unit of the colour adjoint.  `gmask` (or null) is this workgroup's
// global bit-set slot for lists longer than the LDS bit set holds: word w of
// lane l at gmask[w * BLOCK + l], for lists of up to `gcap` entries; longer
// lists fall back to fragment windows (one geometry walk per window).
DVG_DEV void interior_unit(const DevScene &S, uint32_t unit, const float *__restrict__ weight,
                           const float *__restrict__ d_render, const DevGrads &G, uint32_t *__restrict__ gmask,
                           uint32_t gcap) {
    __shared__ float s_stage[STAGE_FLOATS];
    __shared__ uint32_t s_off[2 * LDS_GROUP_CAP];
    __shared__ uint32_t s_total;
    __shared__ float s_red[4][6];
    // per lane: fragment bit set (lists <= MASK_GROUPS) or fragment window
    __shared__ uint32_t s_ids[FRAG_WINDOW * BLOCK];
    uint32_t *ids = s_ids + threadIdx.x;
    uint32_t upt = units_per_tile(S);
    uint32_t tile = unit / upt, s0 = (unit - tile * upt) * S.unit_samples;
    uint32_t lb = S.tile_offsets[tile], le = S.tile_end[tile];
    const uint32_t *list = S.bwd_entries + lb;
    uint32_t n = le - lb;
    bool use_mask = n <= MASK_GROUPS;
    bool use_gmask = !use_mask && gmask != nullptr && n <= gcap;
    uint32_t *gids = use_gmask ? gmask + threadIdx.x : nullptr;
    // the lane's checkpoint ring follows the slot's bit-set words
    float *gring = use_gmask ? reinterpret_cast<float *>(gmask + (size_t)((2u * gcap + 31u) / 32u) * BLOCK) + threadIdx.x
                             : nullptr;
    stage_build(S, list, n, s_off, s_stage, &s_total);
    float acc_bg[4] = {0.f, 0.f, 0.f, 0.f}, acc_fr = 0.f;
    uint32_t total = min(TILE * TILE * S.spp, s0 + S.unit_samples);
    float Wf = (float)S.W, Hf = (float)S.H;
    for (uint32_t s = s0 + threadIdx.x; s < total; s += BLOCK) {
        uint32_t x, y, sx, sy;
        tile_sample(S, tile, s, x, y, sx, sy);
        if (x >= S.W || y >= S.H) continue;
        float px0, py0;
        sample_point<true>(S, x, y, sx, sy, px0, py0);
        // render.rs:110-111 then sampling.rs:89: the point goes through the
        // normalized [0, 1] canvas and back.
        float px = (px0 / Wf) * Wf, py = (py0 / Hf) * Hf;
        size_t pix = (size_t)y * S.W + x;
        float4 bgs = background_at(S, x, y);
        float4 bg = make_float4(bgs.x * bgs.w, bgs.y * bgs.w, bgs.z * bgs.w, bgs.w);
        float4 dcol = gather_d_color(S, weight, d_render, px0, py0);
        float4 pre = bg, fin = bg;
        uint32_t nf = use_mask    ? walk_mask(S, list, n, px, py, ids, fin, nullptr)
                      : use_gmask ? walk_mask(S, list, n, px, py, gids, fin, gring)
                                  : walk_window(S, list, n, px, py, 0, ids, pre, fin);
        float4 col;
        float dtx = 0.f, dty = 0.f;
        if (nf == 0) {
            col = bg;
            acc_bg[0] += dcol.x; acc_bg[1] += dcol.y; acc_bg[2] += dcol.z; acc_bg[3] += dcol.w;
            if (S.bg_image && G.d_background_image) {
                float *db = G.d_background_image + pix * 4u;
                unsafeAtomicAdd(db + 0, dcol.x); unsafeAtomicAdd(db + 1, dcol.y);
                unsafeAtomicAdd(db + 2, dcol.z); unsafeAtomicAdd(db + 3, dcol.w);
            }
        } else {
            float fa = fin.w;
            float fr = fin.x, fg = fin.y, fb = fin.z;
            if (fa > 1.0e-6f) {
                float sc = 1.0f / fa;
                fr = fr * sc; fg = fg * sc; fb = fb * sc;
            }
            col = make_float4(fr, fg, fb, fa);
            float dr = dcol.x, dg = dcol.y, db = dcol.z, da = dcol.w;
            if (fa > 1.0e-6f) {
                float sc = 1.0f / fa;
                float dot = dr * fr + dg * fg + db * fb;
                dr = dr * sc; dg = dg * sc; db = db * sc;
                da -= dot / fa;
            }
            if (use_mask) {
                mask_adjoint(S, G, s_stage, s_off, list, n, ids, nullptr, nf, bg, px, py, dr, dg, db, da, dtx, dty);
            } else if (use_gmask) {
                mask_adjoint(S, G, s_stage, s_off, list, n, gids, gring, nf, bg, px, py, dr, dg, db, da, dtx, dty);
            } else {
                uint32_t nwin = (nf + FRAG_WINDOW - 1) / FRAG_WINDOW;
                for (int w = (int)nwin - 1; w >= 0; --w) {
                    if (adjoint_zero(dr, dg, db, da)) break;
                    uint32_t wstart = (uint32_t)w * FRAG_WINDOW;
                    float4 wpre = bg;
                    if (nwin > 1) {
                        float4 f2 = bg;
                        walk_window(S, list, n, px, py, wstart, ids, wpre, f2);
                    }
                    uint32_t m = min((uint32_t)FRAG_WINDOW, nf - wstart);
                    for (int i = (int)m - 1; i >= 0; --i) {
                        // prefix colour before fragment i of this window
                        float4 pc = wpre;
                        for (int j = 0; j < i; ++j) {
                            uint32_t id = ids[j * BLOCK];
                            blend_cpu(pc, frag_color(S, list[id >> 1], (id & 1u) != 0, px, py));
                        }
                        uint32_t id = ids[i * BLOCK];
                        uint32_t e = id >> 1;
                        bool is_stroke = (id & 1u) != 0;
                        float4 c = frag_color(S, list[e], is_stroke, px, py);
                        frag_adjoint(S, G, s_stage, s_off, list, e, is_stroke, c, pc, px, py, dr, dg, db, da, dtx, dty);
                    }
                }
            }
            acc_bg[0] += dr; acc_bg[1] += dg; acc_bg[2] += db; acc_bg[3] += da;
            if (S.bg_image && G.d_background_image) {
                float *dbp = G.d_background_image + pix * 4u;
                unsafeAtomicAdd(dbp + 0, dr); unsafeAtomicAdd(dbp + 1, dg);
                unsafeAtomicAdd(dbp + 2, db); unsafeAtomicAdd(dbp + 3, da);
            }
        }
        if (G.d_translation && (dtx != 0.f || dty != 0.f)) {
            unsafeAtomicAdd(G.d_translation + pix * 2u, dtx);
            unsafeAtomicAdd(G.d_translation + pix * 2u + 1, dty);
        }
        acc_fr += filter_grad(S, weight, d_render, col, px0, py0);
    }
    // scalar gradients: wave reduction, then one atomic per workgroup
    float v[5] = {acc_bg[0], acc_bg[1], acc_bg[2], acc_bg[3], acc_fr};
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (int k = 0; k < 5; ++k) {
        float r = wave_sum(v[k]);
        if (lane == 0) s_red[wave][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        float r = s_red[0][threadIdx.x] + s_red[1][threadIdx.x] + s_red[2][threadIdx.x] + s_red[3][threadIdx.x];
        if (threadIdx.x < 4) {
            if (!S.bg_image) unsafeAtomicAdd(G.d_background + threadIdx.x, r);
        } else {
            unsafeAtomicAdd(G.d_filter_radius, r);
        }
    }
    stage_flush(S, G, list, n, s_off, s_stage);
}

__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_INTERIOR))) void k_backward_interior(
    DevScene S, const float *__restrict__ weight, const float *__restrict__ d_render, DevGrads G) {
    interior_unit(S, blockIdx.x, weight, d_render, G, nullptr, 0u);
}

// Scenes with lists longer than the LDS bit set: a persistent grid takes the
// units from a queue; workgroup b owns bit-set slot b of `gmask`
// (nw words x BLOCK lanes a slot).
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_INTERIOR))) void k_backward_interior_q(
    DevScene S, const float *__restrict__ weight, const float *__restrict__ d_render, DevGrads G,
    uint32_t *__restrict__ counter, uint32_t n_units, uint32_t *__restrict__ gmask, uint32_t gcap) {
    __shared__ uint32_t s_unit;
    uint32_t *slot = gmask + (size_t)blockIdx.x * interior_slot_words(gcap);
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(counter, 1u);
        __syncthreads();
        uint32_t u = __builtin_amdgcn_readfirstlane(s_unit);
        __syncthreads();
        if (u >= n_units) break;
        interior_unit(S, u, weight, d_render, G, slot, gcap);
    }
}

// 3x3 helpers on affine matrices given by their first two rows.
struct M3 {
    float m[3][3];
};
DVG_DEV M3 m3_affine(const float *r) {
    M3 o;
    o.m[0][0] = r[0]; o.m[0][1] = r[1]; o.m[0][2] = r[2];
    o.m[1][0] = r[3]; o.m[1][1] = r[4]; o.m[1][2] = r[5];
    o.m[2][0] = 0.f; o.m[2][1] = 0.f; o.m[2][2] = 1.f;
    return o;
}
DVG_DEV M3 m3_mul(const M3 &a, const M3 &b) {  // math.rs:185-195
    M3 o;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return o;
}

// ---------------------------------------------------------------------------
// Prefiltered backward: the reference CPU sample_color_prefiltered
// (backward/sampling.rs:227-482) with its distance adjoint
// (backward/distance.rs:50-561).  Fragment coverage comes from the distance to
// the closest boundary point of each group (distance/api.rs:138-232).
//
// Closest-point queries skip shapes and path segments whose bounds lie
// farther than the best distance so far (or than the query radius): bounds
// contain the geometry, so the result is the one of the full scan.  A stroke
// fragment needs |d| < stroke width + 1 and a fill fragment |d| <= 1 or an
// inside point, so a group whose padded bounds miss the point contributes
// nothing and the tile lists (padded by one pixel for this mode) cull exactly.

constexpr int PF_WINDOW = 16;  // fragments recorded per lane before a re-walk

DVG_DEV float box_dist(float x0, float y0, float x1, float y1, float px, float py) {  // distance/bvh.rs:104-108
    float dx = maxf_(maxf_(x0 - px, 0.f), px - x1);
    float dy = maxf_(maxf_(y0 - py, 0.f), py - y1);
    return sqrtf(dx * dx + dy * dy);
}

// Smallest singular value of the 2x2 matrix [a b; c d].
DVG_DEV float min_scale2(float a, float b, float c, float d) {
    float q = a * a + b * b + c * c + d * d, det = a * d - b * c;
    float disc = sqrtf(fmaxf(q * q - 4.0f * det * det, 0.f));
    // a little under the exact value, so the pruning stays conservative
    return sqrtf(fmaxf(0.5f * (q - disc), 0.f)) * 0.999f;
}

struct PHit {
    float cx, cy, dist, t;  // canvas closest point, canvas distance, path parameter
    uint32_t prep, seg;     // prepared shape instance, segment index
    bool path;
};

// Closest point of prepared shape `prep` to the shape-local point (px, py),
// shape-local (closest_point_shape, closest.rs:364-456; closest_point_path_bvh,
// closest.rs:100-221).  `cut` bounds the useful distance (shape-local) for
// the segment pruning; the returned point is the true closest point whenever
// it lies within `cut`.
DVG_DEV bool closest_shape(const DevScene &S, uint32_t prep, float px, float py, float cut, float &ox, float &oy,
                           uint32_t &seg, float &tt, bool &is_path) {
    const float *sd = S.shape_data + prep * 15u;
    uint32_t kind = (uint32_t)sd[0];
    is_path = false;
    seg = 0;
    tt = 0.f;
    if (kind == DVG_SHAPE_CIRCLE) {  // shape.rs:107-114
        float cx = sd[4], cy = sd[5], r = sd[6];
        float dx = px - cx, dy = py - cy;
        float len = sqrtf(dx * dx + dy * dy);
        if (len <= 1.0e-8f) { ox = cx + r; oy = cy; }
        else { ox = cx + dx * (r / len); oy = cy + dy * (r / len); }
        return true;
    }
    if (kind == DVG_SHAPE_ELLIPSE) {
        ellipse_closest(sd[4], sd[5], sd[6], sd[7], px, py, ox, oy);
        return true;
    }
    if (kind == DVG_SHAPE_RECT) {  // shape.rs:167-185
        float x0 = sd[4], y0 = sd[5], x1 = sd[6], y1 = sd[7];
        bool inside = px >= x0 && px <= x1 && py >= y0 && py <= y1;
        if (!inside) {  // f32::clamp
            ox = px;
            if (ox < x0) ox = x0;
            if (ox > x1) ox = x1;
            oy = py;
            if (oy < y0) oy = y0;
            if (oy > y1) oy = y1;
            return true;
        }
        float dl = px - x0, dr = x1 - px, db = py - y0, dt = y1 - py;
        if (dl <= dr && dl <= db && dl <= dt) { ox = x0; oy = py; }
        else if (dr <= db && dr <= dt) { ox = x1; oy = py; }
        else if (db <= dt) { ox = px; oy = y0; }
        else { ox = px; oy = y1; }
        return true;
    }
    uint32_t off = (uint32_t)sd[12], cnt = (uint32_t)sd[13];
    bool approx = sd[14] > 0.5f;
    float best = BIG, bt = 0.f, bx = 0.f, by = 0.f;
    uint32_t bs = 0;
    bool any = false;
    for (uint32_t s = 0; s < cnt; ++s) {
        const float *cv = S.curve_data + (off + s) * 13u;
        uint32_t k = (uint32_t)cv[0];
        float x0 = cv[1], y0 = cv[2], x1 = cv[3], y1 = cv[4], x2 = cv[5], y2 = cv[6], x3 = cv[7], y3 = cv[8];
        const float *wc = S.curve_wind + (size_t)(off + s) * WIND_STRIDE;  // the curve's box
        float bx0 = wc[15], bx1 = wc[14], by0 = wc[12], by1 = wc[13];
        float lim = any ? best : cut;
        if (box_dist(bx0, by0, bx1, by1, px, py) > lim) continue;
        float d, t, qx, qy;
        if (k == 0) seg_closest_cpu(px, py, x0, y0, x1, y1, qx, qy, d, t);
        else if (k == 1) quad_closest_cpu(px, py, x0, y0, x1, y1, x2, y2, approx, qx, qy, d, t);
        else cubic_closest_cpu(px, py, x0, y0, x1, y1, x2, y2, x3, y3, approx, qx, qy, d, t);
        if (!any || d < best) {
            best = d;
            bt = t;
            bs = s;
            bx = qx;
            by = qy;
            any = true;
        }
    }
    if (!any) return false;
    ox = bx;
    oy = by;
    is_path = true;
    seg = bs;
    tt = bt;
    return true;
}

// compute_distance_bvh of group g at canvas point (px, py) within max_r:
// the closest point over the group's shapes, the last one on a tie.
DVG_DEV bool group_closest(const DevScene &S, uint32_t g, float px, float py, float max_r, PHit &h) {
    const float *gd = S.group_data + g * 16u;
    float lx, ly;
    group_local(S, g, px, py, lx, ly);
    const float *gs = S.group_shape_xform + g * 6u;
    uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
    // the reference's root node: group-local box distance against the query
    // radius (distance/api.rs:166-170)
    const float *rb = S.group_cpu_bounds + g * 4u;
    if (rb[0] <= rb[2] && rb[1] <= rb[3] && box_dist(rb[0], rb[1], rb[2], rb[3], lx, ly) > max_r) return false;
    // canvas distances are at least smin times group-local ones
    float smin = min_scale2(gs[0], gs[1], gs[3], gs[4]);
    float best = max_r;
    bool found = false;
    for (uint32_t i = 0; i < cnt; ++i) {
        uint32_t prep = (uint32_t)S.group_shapes[off + i];
        const float *b = S.shape_bounds + prep * 4u;
        // a shape whose bounds lie beyond the current best cannot win
        if (b[0] <= b[2] && b[1] <= b[3] && box_dist(b[0], b[1], b[2], b[3], lx, ly) * smin > best) continue;
        uint32_t si = S.prepared_to_scene[prep];
        const float *xf = S.shape_xform + prep * 6u;
        float spx = xf[0] * lx + xf[1] * ly + xf[2];
        float spy = xf[3] * lx + xf[4] * ly + xf[5];
        float ox, oy, t;
        uint32_t seg;
        bool is_path;
        const float *T = S.shape_transform + si * 6u;
        // shape-space distances beyond best / (smallest scale of the shape-to-
        // canvas map) cannot win: closest_shape need not resolve them
        float sc = smin * min_scale2(T[0], T[1], T[3], T[4]);
        float cut = (sc > 0.f && best < BIG) ? best / sc : BIG;
        if (!closest_shape(S, prep, spx, spy, cut, ox, oy, seg, t, is_path)) continue;
        float gx = T[0] * ox + T[1] * oy + T[2], gy = T[3] * ox + T[4] * oy + T[5];
        float cx = gs[0] * gx + gs[1] * gy + gs[2], cy = gs[3] * gx + gs[4] * gy + gs[5];
        float d = dist_pt(cx, cy, px, py);
        if (d <= best) {
            best = d;
            h.cx = cx; h.cy = cy; h.dist = d; h.t = t;
            h.prep = prep; h.seg = seg; h.path = is_path;
            found = true;
        }
    }
    return found;
}

// is_inside_bvh (api.rs:259-309) through the winding tests of the group
DVG_DEV bool group_inside(const DevScene &S, uint32_t g, float px, float py) {
    const float *gd = S.group_data + g * 16u;
    uint32_t rule = (uint32_t)gd[7];
    float lx, ly;
    group_local(S, g, px, py, lx, ly);
    Cov c = {BIG, 0.f, 0.f, BIG, 0.f, 0.f, 0.f};
    group_cov<COV_CPU>(S, g, gd, lx, ly, true, false, rule, false, c);
    return (rule == 1u) ? c.crossings > 0.f : c.winding != 0.f;
}

// Stroke fragment of group g: (signed) distance and coverage.
// Only a boundary nearer than the group's largest stroke width + 1 can give
// w > 0 (w = S(|d| + sw) - S(|d| - sw) of the nearest shape's sw, 0 once
// |d| - sw >= 1): the query is bounded there.
DVG_DEV bool pf_stroke(const DevScene &S, uint32_t g, float px, float py, PHit &h, float &w) {
    if (!group_closest(S, g, px, py, S.group_pf[2u * g + 1u] + 1.001f, h)) return false;
    float sw = S.shape_data[h.prep * 15u + 3u];
    float d = h.dist;
    w = smoothstep_unit(absf_(d) + sw) - smoothstep_unit(absf_(d) - sw);
    return w > 0.f;
}

// Fill fragment of group g; hit = a boundary point within one pixel.
DVG_DEV bool pf_fill(const DevScene &S, uint32_t g, float px, float py, PHit &h, bool &hit, float &d, float &w) {
    hit = group_closest(S, g, px, py, 1.0f, h);
    bool inside = group_inside(S, g, px, py);
    if (!hit && !inside) return false;
    d = hit ? h.dist : 1.0f;
    if (!inside) d = -d;
    w = smoothstep_unit(d);
    return w > 0.f;
}

// Colour (paint * coverage) of a recorded fragment: `covw` is the stroke
// coverage itself, or the signed fill distance.
DVG_DEV float4 pf_color(const DevScene &S, uint32_t g, bool stroke, float covw, float px, float py) {
    float4 c = frag_color(S, g, stroke, px, py);
    c.w *= stroke ? covw : smoothstep_unit(covw);
    return c;
}

// One pass over the tile's groups: composites every prefiltered fragment into
// `fin`, the first `start` into `pre`, and records fragments
// [start, start + PF_WINDOW) as (list pos << 2 | within << 1 | stroke) and a
// coverage word (the stroke coverage w, or the signed fill distance d) in the
// lane's LDS columns.
DVG_DEV uint32_t pf_walk(const DevScene &S, const uint32_t *list, uint32_t n, float px, float py, uint32_t start,
                         uint32_t *ids, float *cov, float4 &pre, float4 &fin) {
    uint32_t count = 0;
    for (uint32_t e = 0; e < n; ++e) {
        uint32_t g = list[e];
        const float *gd = S.group_data + g * 16u;
        uint32_t fk = (uint32_t)gd[2], sk = (uint32_t)gd[4];
        for (int k = 0; k < 2; ++k) {
            bool is_stroke = k == 0;
            if ((is_stroke ? sk : fk) == DVG_PAINT_NONE) continue;
            PHit h;
            float w, covw;
            bool within = true;
            if (is_stroke) {
                if (!pf_stroke(S, g, px, py, h, w)) continue;
                covw = w;
            } else {
                float d;
                if (!pf_fill(S, g, px, py, h, within, d, w)) continue;
                covw = d;
            }
            float4 c = frag_color(S, g, is_stroke, px, py);
            c.w *= w;
            blend_cpu(fin, c);
            if (count < start) blend_cpu(pre, c);
            else if (count < start + PF_WINDOW) {
                uint32_t slot = (count - start) * BLOCK;
                ids[slot] = (e << 2) | (within ? 2u : 0u) | (is_stroke ? 1u : 0u);
                cov[slot] = covw;
            }
            ++count;
        }
    }
    return count;
}

// 3x3 helpers for the transform adjoints (row-major, full matrices).
DVG_DEV void m3_zero(M3 &m) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m.m[i][j] = 0.f;
}
DVG_DEV M3 m3_transpose(const M3 &a) {
    M3 o;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o.m[i][j] = a.m[j][i];
    return o;
}

// d_xform_pt (backward/math.rs:101-120)
DVG_DEV void d_xform_pt(const M3 &m, float ptx, float pty, float dox, float doy, M3 &dm, float &dpx, float &dpy) {
    float t0 = m.m[0][0] * ptx + m.m[0][1] * pty + m.m[0][2];
    float t1 = m.m[1][0] * ptx + m.m[1][1] * pty + m.m[1][2];
    float t2 = m.m[2][0] * ptx + m.m[2][1] * pty + m.m[2][2];
    float ox = t0 / t2, oy = t1 / t2;
    float d0 = dox / t2, d1 = doy / t2;
    float d2 = -(dox * ox + doy * oy) / t2;
    dm.m[0][0] += d0 * ptx; dm.m[0][1] += d0 * pty; dm.m[0][2] += d0;
    dm.m[1][0] += d1 * ptx; dm.m[1][1] += d1 * pty; dm.m[1][2] += d1;
    dm.m[2][0] += d2 * ptx; dm.m[2][1] += d2 * pty; dm.m[2][2] += d2;
    dpx += d0 * m.m[0][0] + d1 * m.m[1][0] + d2 * m.m[2][0];
    dpy += d0 * m.m[0][1] + d1 * m.m[1][1] + d2 * m.m[2][1];
}

// d_length (backward/math.rs:33-41)
DVG_DEV void d_length2(float vx, float vy, float d_l, float &ox, float &oy) {
    float l = sqrtf(vx * vx + vy * vy);
    ox = oy = 0.f;
    if (l == 0.f) return;
    float d_l_sq = 0.5f * d_l / l;
    ox = vx * (2.0f * d_l_sq);
    oy = vy * (2.0f * d_l_sq);
}

// d_closest_point_path (backward/distance.rs:188-404): point gradients of the
// closest point on segment `base` of scene path si at parameter t.
DVG_DEV void d_closest_path(const DevScene &S, const DevGrads &G, uint32_t si, uint32_t base, float t, float px,
                            float py, float dcx, float dcy, float &dpx, float &dpy) {
    uint32_t n = S.sp_cnt[si];
    if (n == 0) return;
    bool closed = S.s_closed[si] != 0;
    const uint32_t *ctrl = S.path_ctrl + S.sc_off[si];
    uint32_t nseg = S.sc_cnt[si];
    uint32_t pid = 0;
    uint32_t lim = base < nseg ? base : nseg;
    for (uint32_t i = 0; i < lim; ++i) {
        uint32_t c = ctrl[i];
        pid += c == 0u ? 1u : (c == 1u ? 2u : (c == 2u ? 3u : 0u));
    }
    const float *pts = S.path_points + (size_t)S.sp_off[si] * 2u;
    float *dpt = G.d_shape_points + (size_t)S.sp_off[si] * 2u;
    auto gx = [&](uint32_t i) { return closed ? pts[(i % n) * 2u] : pts[(i < n - 1 ? i : n - 1) * 2u]; };
    auto gy = [&](uint32_t i) { return closed ? pts[(i % n) * 2u + 1u] : pts[(i < n - 1 ? i : n - 1) * 2u + 1u]; };
    auto add = [&](uint32_t i, float vx, float vy) {
        uint32_t index = closed ? i % n : i;
        if (index < n) {
            unsafeAtomicAdd(dpt + index * 2u, vx);
            unsafeAtomicAdd(dpt + index * 2u + 1u, vy);
        }
    };
    uint32_t nc = base < nseg ? ctrl[base] : 0u;
    if (nc == 0u) {
        uint32_t i0 = pid, i1 = (pid + 1u) % n;
        float p0x = gx(i0), p0y = gy(i0), p1x = gx(i1), p1y = gy(i1);
        float ex = p1x - p0x, ey = p1y - p0y;
        float tl = ((px - p0x) * ex + (py - p0y) * ey) / (ex * ex + ey * ey);
        if (tl < 0.f) add(i0, dcx, dcy);
        else if (tl > 1.f) add(i1, dcx, dcy);
        else {
            add(i0, dcx * (1.0f - tl), dcy * (1.0f - tl));
            add(i1, dcx * tl, dcy * tl);
        }
    } else if (nc == 1u) {
        uint32_t i0 = pid, i1 = pid + 1u, i2 = (pid + 2u) % n;
        float p0x = gx(i0), p0y = gy(i0), p1x = gx(i1), p1y = gy(i1), p2x = gx(i2), p2y = gy(i2);
        float d0x = 0.f, d0y = 0.f, d1x = 0.f, d1y = 0.f, d2x = 0.f, d2y = 0.f;
        if (t == 0.f) { d0x += dcx; d0y += dcy; }
        else if (t == 1.f) { d2x += dcx; d2y += dcy; }
        else {
            float ax = (p0x - p1x * 2.0f) + p2x, ay = (p0y - p1y * 2.0f) + p2y;
            float bx = p1x - p0x, by = p1y - p0y;
            float qx = p0x - px, qy = p0y - py;
            float A = ax * ax + ay * ay;
            float B = 3.0f * (ax * bx + ay * by);
            float C = 2.0f * (bx * bx + by * by) + (ax * qx + ay * qy);
            float u = 1.0f - t;
            float d_tt = 2.0f * u * (dcx * p0x + dcy * p0y) + 2.0f * t * (dcx * p1x + dcy * p1y);
            float d_t = -d_tt + 2.0f * u * (dcx * p1x + dcy * p1y) + 2.0f * t * (dcx * p2x + dcy * p2y);
            d0x += dcx * (u * u); d0y += dcy * (u * u);
            d1x += dcx * (2.0f * u * t); d1y += dcy * (2.0f * u * t);
            d2x += dcx * (t * t); d2y += dcy * (t * t);
            float pd = 3.0f * A * t * t + 2.0f * B * t + C;
            if (absf_(pd) > 1.0e-6f) {
                float dA = -(d_t / pd) * t * t * t, dB = -(d_t / pd) * t * t, dC = -(d_t / pd) * t, dD = -(d_t / pd);
                d0x += (((ax * (2.0f * dA) + (bx - ax) * (3.0f * dB)) + (bx * -4.0f) * dC) + (qx + ax) * dC) + (bx - qx) * (2.0f * dD);
                d0y += (((ay * (2.0f * dA) + (by - ay) * (3.0f * dB)) + (by * -4.0f) * dC) + (qy + ay) * dC) + (by - qy) * (2.0f * dD);
                d1x += (((ax * (-4.0f * dA) + (ax + bx * -2.0f) * (3.0f * dB)) + (bx * 4.0f) * dC) + qx * (-2.0f * dC)) + qx * dD;
                d1y += (((ay * (-4.0f * dA) + (ay + by * -2.0f) * (3.0f * dB)) + (by * 4.0f) * dC) + qy * (-2.0f * dC)) + qy * dD;
                d2x += (ax * (2.0f * dA) + bx * (3.0f * dB)) + qx * dC;
                d2y += (ay * (2.0f * dA) + by * (3.0f * dB)) + qy * dC;
                dpx += ax * (-dC) + bx * dD;
                dpy += ay * (-dC) + by * dD;
            }
        }
        add(i0, d0x, d0y);
        add(i1, d1x, d1y);
        add(i2, d2x, d2y);
    } else if (nc == 2u) {
        uint32_t i0 = pid, i1 = pid + 1u, i2 = pid + 2u, i3 = (pid + 3u) % n;
        float p0x = gx(i0), p0y = gy(i0), p1x = gx(i1), p1y = gy(i1);
        float p2x = gx(i2), p2y = gy(i2), p3x = gx(i3), p3y = gy(i3);
        float d0x = 0.f, d0y = 0.f, d1x = 0.f, d1y = 0.f, d2x = 0.f, d2y = 0.f, d3x = 0.f, d3y = 0.f;
        if (t == 0.f) { d0x += dcx; d0y += dcy; }
        else if (t == 1.f) { d3x += dcx; d3y += dcy; }
        else {
            float ax = ((p0x * -1.0f + p1x * 3.0f) + p2x * -3.0f) + p3x, ay = ((p0y * -1.0f + p1y * 3.0f) + p2y * -3.0f) + p3y;
            float bx = (p0x * 3.0f + p1x * -6.0f) + p2x * 3.0f, by = (p0y * 3.0f + p1y * -6.0f) + p2y * 3.0f;
            float cx = p0x * -3.0f + p1x * 3.0f, cy = p0y * -3.0f + p1y * 3.0f;
            float A = 3.0f * (ax * ax + ay * ay);
            if (absf_(A) < 1.0e-10f) return;
            float qx = p0x - px, qy = p0y - py;
            float B = 5.0f * (ax * bx + ay * by);
            float C = 4.0f * (ax * cx + ay * cy) + 2.0f * (bx * bx + by * by);
            float D = 3.0f * ((bx * cx + by * cy) + (ax * qx + ay * qy));
            float E = (cx * cx + cy * cy) + 2.0f * (qx * bx + qy * by);
            float F = qx * cx + qy * cy;
            B = B / A; C = C / A; D = D / A; E = E / A; F = F / A;
            float u = 1.0f - t;
            float dp0 = dcx * p0x + dcy * p0y, dp1 = dcx * p1x + dcy * p1y, dp2 = dcx * p2x + dcy * p2y,
                  dp3 = dcx * p3x + dcy * p3y;
            float d_tt = 3.0f * u * u * dp0 + 6.0f * u * t * dp1 + 3.0f * t * t * dp2;
            float d_t = -d_tt + 3.0f * u * u * dp1 + 6.0f * u * t * dp2 + 3.0f * t * t * dp3;
            d0x += dcx * (u * u * u); d0y += dcy * (u * u * u);
            d1x += dcx * (3.0f * u * u * t); d1y += dcy * (3.0f * u * u * t);
            d2x += dcx * (3.0f * u * t * t); d2y += dcy * (3.0f * u * t * t);
            d3x += dcx * (t * t * t); d3y += dcy * (t * t * t);
            float pd = 5.0f * t * t * t * t + 4.0f * B * t * t * t + 3.0f * C * t * t + 2.0f * D * t + E;
            if (absf_(pd) > 1.0e-10f) {
                float dB = -(d_t / pd) * t * t * t * t, dC = -(d_t / pd) * t * t * t, dD = -(d_t / pd) * t * t;
                float dE = -(d_t / pd) * t, dF = -(d_t / pd);
                float dA = -dB * B / A - dC * C / A - dD * D / A - dE * E / A - dF * F / A;
                dB /= A; dC /= A; dD /= A; dE /= A; dF /= A;
                d0x += ax * (3.0f * -1.0f * 2.0f * dA); d0y += ay * (3.0f * -1.0f * 2.0f * dA);
                d1x += ax * (3.0f * 3.0f * 2.0f * dA); d1y += ay * (3.0f * 3.0f * 2.0f * dA);
                d2x += ax * (3.0f * -3.0f * 2.0f * dA); d2y += ay * (3.0f * -3.0f * 2.0f * dA);
                d3x += ax * (3.0f * 1.0f * 2.0f * dA); d3y += ay * (3.0f * 1.0f * 2.0f * dA);
                d0x += (bx * -1.0f + ax * 3.0f) * (5.0f * dB); d0y += (by * -1.0f + ay * 3.0f) * (5.0f * dB);
                d1x += (bx * 3.0f + ax * -6.0f) * (5.0f * dB); d1y += (by * 3.0f + ay * -6.0f) * (5.0f * dB);
                d2x += (bx * -3.0f + ax * 3.0f) * (5.0f * dB); d2y += (by * -3.0f + ay * 3.0f) * (5.0f * dB);
                d3x += bx * (5.0f * dB); d3y += by * (5.0f * dB);
                d0x += (cx * -1.0f + ax * -3.0f) * (4.0f * dC) + bx * (3.0f * 2.0f * dC);
                d0y += (cy * -1.0f + ay * -3.0f) * (4.0f * dC) + by * (3.0f * 2.0f * dC);
                d1x += (cx * 3.0f + ax * 3.0f) * (4.0f * dC) + bx * (-6.0f * 2.0f * dC);
                d1y += (cy * 3.0f + ay * 3.0f) * (4.0f * dC) + by * (-6.0f * 2.0f * dC);
                d2x += cx * (-3.0f * dC * 4.0f) + bx * (3.0f * 2.0f * dC);
                d2y += cy * (-3.0f * dC * 4.0f) + by * (3.0f * 2.0f * dC);
                d3x += cx * (4.0f * dC); d3y += cy * (4.0f * dC);
                d0x += (cx * 3.0f + bx * -3.0f) * (3.0f * dD) + (ax - qx) * (3.0f * dD);
                d0y += (cy * 3.0f + by * -3.0f) * (3.0f * dD) + (ay - qy) * (3.0f * dD);
                d1x += (cx * -6.0f + bx * 3.0f) * (3.0f * dD) + qx * (3.0f * 3.0f * dD);
                d1y += (cy * -6.0f + by * 3.0f) * (3.0f * dD) + qy * (3.0f * 3.0f * dD);
                d2x += cx * (3.0f * 3.0f * dD) + qx * (-3.0f * 3.0f * dD);
                d2y += cy * (3.0f * 3.0f * dD) + qy * (-3.0f * 3.0f * dD);
                dpx += ax * (-1.0f * 3.0f * dD); dpy += ay * (-1.0f * 3.0f * dD);
                d0x += cx * (-3.0f * 2.0f * dE) + (bx + qx * 3.0f) * (2.0f * dE);
                d0y += cy * (-3.0f * 2.0f * dE) + (by + qy * 3.0f) * (2.0f * dE);
                d1x += cx * (3.0f * 2.0f * dE) + qx * (-6.0f * 2.0f * dE);
                d1y += cy * (3.0f * 2.0f * dE) + qy * (-6.0f * 2.0f * dE);
                d2x += qx * (3.0f * 2.0f * dE); d2y += qy * (3.0f * 2.0f * dE);
                dpx += bx * (-1.0f * 2.0f * dE); dpy += by * (-1.0f * 2.0f * dE);
                d0x += cx * dF + qx * (-3.0f * dF); d0y += cy * dF + qy * (-3.0f * dF);
                d1x += qx * (3.0f * dF); d1y += qy * (3.0f * dF);
                dpx += cx * (-1.0f * dF); dpy += cy * (-1.0f * dF);
            }
        }
        add(i0, d0x, d0y);
        add(i1, d1x, d1y);
        add(i2, d2x, d2y);
        add(i3, d3x, d3y);
    }
}

// d_closest_point_ellipse (backward/distance.rs:489-561), including its
// omission of the direct d(closest)/d(center) term.
DVG_DEV void d_closest_ellipse(float cx, float cy, float rx_, float ry_, float px, float py, float dcx, float dcy,
                               float *dp, float &dpx, float &dpy) {
    float rx = absf_(rx_), ry = absf_(ry_);
    const float eps = 1.0e-6f;
    float lx = px - cx, ly = py - cy;
    float dcen_x = 0.f, dcen_y = 0.f, drx = 0.f, dry = 0.f;
    if (rx < eps && ry < eps) {
        dcen_x += dcx; dcen_y += dcy;
    } else if (rx < eps) {
        float hit = (ly >= -ry && ly <= ry) ? 1.0f : 0.0f;
        dcen_x += dcx; dcen_y += dcy;
        dpy += dcy * hit;
        if (ly > ry) dry += dcy;
        else if (ly < -ry) dry -= dcy;
    } else if (ry < eps) {
        float hit = (lx >= -rx && lx <= rx) ? 1.0f : 0.0f;
        dcen_x += dcx; dcen_y += dcy;
        dpx += dcx * hit;
        if (lx > rx) drx += dcx;
        else if (lx < -rx) drx -= dcx;
    } else {
        float sx = lx < 0.f ? -1.0f : 1.0f, sy = ly < 0.f ? -1.0f : 1.0f;
        float x = absf_(lx), y = absf_(ly);
        float t = atan2f(y * rx, x * ry);
        float g_t = 0.f;
        for (int i = 0; i < 20; ++i) {
            float s = sinf(t), c = cosf(t);
            float g = rx * x * s - ry * y * c + (ry * ry - rx * rx) * s * c;
            g_t = rx * x * c + ry * y * s + (ry * ry - rx * rx) * (c * c - s * s);
            if (absf_(g_t) < 1.0e-12f) break;
            float next = t - g / g_t;
            next = next < 0.f ? 0.f : (next > 1.57079632679489661923f ? 1.57079632679489661923f : next);
            if (absf_(next - t) < 1.0e-6f) { t = next; break; }
            t = next;
        }
        float s = sinf(t), c = cosf(t);
        float d_t = 0.f;
        drx += dcx * sx * c;
        d_t += dcx * sx * (-rx * s);
        dry += dcy * sy * s;
        d_t += dcy * sy * (ry * c);
        float g_a = x * s - 2.0f * rx * s * c;
        float g_b = -y * c + 2.0f * ry * s * c;
        float g_x = rx * s, g_y = -ry * c;
        if (absf_(g_t) > 1.0e-12f) {
            float inv = -d_t / g_t;
            drx += inv * g_a;
            dry += inv * g_b;
            float d_x = inv * g_x, d_y = inv * g_y;
            dpx += d_x * sx;
            dpy += d_y * sy;
            dcen_x -= d_x * sx;
            dcen_y -= d_y * sy;
        }
    }
    unsafeAtomicAdd(dp + 0, dcen_x);
    unsafeAtomicAdd(dp + 1, dcen_y);
    unsafeAtomicAdd(dp + 2, drx);
    unsafeAtomicAdd(dp + 3, dry);
}

// d_closest_point_rect (backward/distance.rs:407-486)
DVG_DEV void d_closest_rect(float x0, float y0, float x1, float y1, float px, float py, float dcx, float dcy,
                            float *dp, float &dpx, float &dpy) {
    // corners lt (x0,y0), rt (x1,y0), lb (x0,y1), rb (x1,y1); edges
    // 0 lt-lb, 1 lt-rt, 2 rt-rb, 3 lb-rb
    float ex[4][4] = {{x0, y0, x0, y1}, {x0, y0, x1, y0}, {x1, y0, x1, y1}, {x0, y1, x1, y1}};
    auto seg_d = [&](int k) {
        float ax = ex[k][0], ay = ex[k][1], bx = ex[k][2], by = ex[k][3];
        float vx = bx - ax, vy = by - ay;
        float t = ((px - ax) * vx + (py - ay) * vy) / (vx * vx + vy * vy);
        if (t < 0.f) return dist_pt(ax, ay, px, py);
        if (t > 1.f) return dist_pt(bx, by, px, py);
        return dist_pt(ax + vx * t, ay + vy * t, px, py);
    };
    int id = 0;
    float md = seg_d(0);
    for (int k = 1; k < 4; ++k) {
        float dk = seg_d(k);
        if (dk < md) { md = dk; id = k; }
    }
    float ax = ex[id][0], ay = ex[id][1], bx = ex[id][2], by = ex[id][3];
    float d0x = 0.f, d0y = 0.f, d1x = 0.f, d1y = 0.f;
    float vx = bx - ax, vy = by - ay;
    float t = ((px - ax) * vx + (py - ay) * vy) / (vx * vx + vy * vy);
    if (t < 0.f) { d0x += dcx; d0y += dcy; }
    else if (t > 1.f) { d1x += dcx; d1y += dcy; }
    else {
        d0x += dcx * (1.0f - t); d0y += dcy * (1.0f - t);
        d1x += dcx * t; d1y += dcy * t;
        float d_t = dcx * vx + dcy * vy;
        float den = vx * vx + vy * vy;
        float d_num = d_t / den, d_den = d_t * (-t) / den;
        dpx += vx * d_num; dpy += vy * d_num;
        d1x += (px - ax) * d_num; d1y += (py - ay) * d_num;
        d0x += (((ax - bx) + ax) - px) * d_num; d0y += (((ay - by) + ay) - py) * d_num;
        d1x += vx * (2.0f * d_den); d1y += vy * (2.0f * d_den);
        d0x += (ax - bx) * (2.0f * d_den); d0y += (ay - by) * (2.0f * d_den);
    }
    // corner gradients -> min / max (distance.rs:479-485)
    float dmin_x = 0.f, dmin_y = 0.f, dmax_x = 0.f, dmax_y = 0.f;
    auto corner = [&](int which, float gxv, float gyv) {  // 0 lt, 1 rt, 2 lb, 3 rb
        if (which == 0) { dmin_x += gxv; dmin_y += gyv; }
        else if (which == 1) { dmax_x += gxv; dmin_y += gyv; }
        else if (which == 2) { dmin_x += gxv; dmax_y += gyv; }
        else { dmax_x += gxv; dmax_y += gyv; }
    };
    const int ends[4][2] = {{0, 2}, {0, 1}, {1, 3}, {2, 3}};
    corner(ends[id][0], d0x, d0y);
    corner(ends[id][1], d1x, d1y);
    unsafeAtomicAdd(dp + 0, dmin_x);
    unsafeAtomicAdd(dp + 1, dmin_y);
    unsafeAtomicAdd(dp + 2, dmax_x);
    unsafeAtomicAdd(dp + 3, dmax_y);
}

// d_compute_distance (backward/distance.rs:50-126) for the closest point of a
// fragment; adds -d(pt) into (dtx, dty).
DVG_DEV void d_distance_query(const DevScene &S, const DevGrads &G, uint32_t g, const PHit &h, float px, float py,
                              float d_dist, float &dtx, float &dty) {
    float ddx = px - h.cx, ddy = py - h.cy;
    if (ddx * ddx + ddy * ddy < 1.0e-10f) return;
    uint32_t prep = h.prep, si = S.prepared_to_scene[prep];
    M3 C2S = m3_affine(S.group_xform + g * 6u), S2C = m3_affine(S.group_shape_xform + g * 6u);
    M3 T = m3_affine(S.shape_transform + si * 6u), Ti = m3_affine(S.shape_xform + prep * 6u);
    float lpgx = C2S.m[0][0] * px + C2S.m[0][1] * py + C2S.m[0][2];
    float lpgy = C2S.m[1][0] * px + C2S.m[1][1] * py + C2S.m[1][2];
    float lcgx = C2S.m[0][0] * h.cx + C2S.m[0][1] * h.cy + C2S.m[0][2];
    float lcgy = C2S.m[1][0] * h.cx + C2S.m[1][1] * h.cy + C2S.m[1][2];
    float lpsx = Ti.m[0][0] * lpgx + Ti.m[0][1] * lpgy + Ti.m[0][2];
    float lpsy = Ti.m[1][0] * lpgx + Ti.m[1][1] * lpgy + Ti.m[1][2];
    float lcsx = Ti.m[0][0] * lcgx + Ti.m[0][1] * lcgy + Ti.m[0][2];
    float lcsy = Ti.m[1][0] * lcgx + Ti.m[1][1] * lcgy + Ti.m[1][2];
    float vx, vy;
    d_length2(ddx, ddy, d_dist, vx, vy);  // d_distance(closest, pt)
    float dptx = vx, dpty = vy, dclx = -vx, dcly = -vy;
    M3 d_s2c, d_st, d_si, d_c2s;
    m3_zero(d_s2c); m3_zero(d_st); m3_zero(d_si); m3_zero(d_c2s);
    float dlcgx = 0.f, dlcgy = 0.f;
    d_xform_pt(S2C, lcgx, lcgy, dclx, dcly, d_s2c, dlcgx, dlcgy);
    float dlcsx = 0.f, dlcsy = 0.f;
    d_xform_pt(T, lcsx, lcsy, dlcgx, dlcgy, d_st, dlcsx, dlcsy);
    float dlpsx = 0.f, dlpsy = 0.f;
    const float *sd = S.shape_data + prep * 15u;
    uint32_t kind = (uint32_t)sd[0];
    float *dp = G.d_shape_params + si * 8u;
    if (kind == DVG_SHAPE_CIRCLE) {
        float cx = sd[4], cy = sd[5], r = sd[6];
        float vx2 = lpsx - cx, vy2 = lpsy - cy;
        float l = sqrtf(vx2 * vx2 + vy2 * vy2);
        float nx = 0.f, ny = 0.f;
        if (l != 0.f) { nx = vx2 / l; ny = vy2 / l; }
        float dcenx = dlcsx, dceny = dlcsy;
        float drad = dlcsx * nx + dlcsy * ny;
        float dnx = dlcsx * r, dny = dlcsy * r;
        float dvx = 0.f, dvy = 0.f;
        if (l != 0.f) {  // d_normalize (backward/math.rs:54-64)
            float d_l = -(dnx * nx + dny * ny) / l;
            float lx2, ly2;
            d_length2(vx2, vy2, d_l, lx2, ly2);
            dvx = dnx / l + lx2;
            dvy = dny / l + ly2;
        }
        dcenx -= dvx; dceny -= dvy;
        dlpsx += dvx; dlpsy += dvy;
        unsafeAtomicAdd(dp + 0, dcenx);
        unsafeAtomicAdd(dp + 1, dceny);
        unsafeAtomicAdd(dp + 2, drad);
    } else if (kind == DVG_SHAPE_ELLIPSE) {
        d_closest_ellipse(sd[4], sd[5], sd[6], sd[7], lpsx, lpsy, dlcsx, dlcsy, dp, dlpsx, dlpsy);
    } else if (kind == DVG_SHAPE_RECT) {
        d_closest_rect(sd[4], sd[5], sd[6], sd[7], lpsx, lpsy, dlcsx, dlcsy, dp, dlpsx, dlpsy);
    } else if (h.path) {
        d_closest_path(S, G, si, h.seg, h.t, lpsx, lpsy, dlcsx, dlcsy, dlpsx, dlpsy);
    }
    float dlpgx = 0.f, dlpgy = 0.f;
    d_xform_pt(Ti, lpgx, lpgy, dlpsx, dlpsy, d_si, dlpgx, dlpgy);
    d_xform_pt(C2S, px, py, dlpgx, dlpgy, d_c2s, dptx, dpty);
    // matrix-inverse corrections -M^T dM M^T (distance.rs:116-124)
    M3 tc = m3_transpose(C2S), ts = m3_transpose(Ti);
    M3 cg = m3_mul(m3_mul(tc, d_c2s), tc), cs = m3_mul(m3_mul(ts, d_si), ts);
    float *dgt = G.d_group_transform + g * 9u;
    float *dst = G.d_shape_transform + si * 9u;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            unsafeAtomicAdd(dgt + i * 3 + j, d_s2c.m[i][j] + (-cg.m[i][j]));
            unsafeAtomicAdd(dst + i * 3 + j, (-cs.m[i][j]) + d_st.m[i][j]);
        }
    dtx -= dptx;
    dty -= dpty;
}

DVG_DEV void prefiltered_unit(const DevScene &S, uint32_t unit, const float *__restrict__ weight,
                              const float *__restrict__ d_render, const DevGrads &G) {
    __shared__ float s_stage[STAGE_FLOATS];
    __shared__ uint32_t s_off[2 * LDS_GROUP_CAP];
    __shared__ uint32_t s_total;
    __shared__ float s_red[4][6];
    __shared__ uint32_t s_ids[PF_WINDOW * BLOCK];
    __shared__ float s_cov[PF_WINDOW * BLOCK];
    uint32_t *ids = s_ids + threadIdx.x;
    float *cov = s_cov + threadIdx.x;
    uint32_t upt = units_per_tile(S);
    uint32_t tile = unit / upt, s0 = (unit - tile * upt) * S.unit_samples;
    uint32_t lb = S.tile_offsets[tile], le = S.tile_end[tile];
    const uint32_t *list = S.bwd_entries + lb;
    uint32_t n = le - lb;
    stage_build(S, list, n, s_off, s_stage, &s_total);
    float acc_bg[4] = {0.f, 0.f, 0.f, 0.f}, acc_fr = 0.f;
    uint32_t total = min(TILE * TILE * S.spp, s0 + S.unit_samples);
    float Wf = (float)S.W, Hf = (float)S.H;
    for (uint32_t s = s0 + threadIdx.x; s < total; s += BLOCK) {
        uint32_t x, y, sx, sy;
        tile_sample(S, tile, s, x, y, sx, sy);
        if (x >= S.W || y >= S.H) continue;
        // render.rs:96-111: sample centres (no jitter when prefiltering)
        float px0 = (float)x + ((float)sx + 0.5f) / (float)S.nsx;
        float py0 = (float)y + ((float)sy + 0.5f) / (float)S.nsy;
        float px = (px0 / Wf) * Wf, py = (py0 / Hf) * Hf;
        size_t pix = (size_t)y * S.W + x;
        float4 bgs = background_at(S, x, y);
        float4 bg = make_float4(bgs.x * bgs.w, bgs.y * bgs.w, bgs.z * bgs.w, bgs.w);
        float4 dcol = gather_d_color(S, weight, d_render, px0, py0);
        float4 pre = bg, fin = bg;
        uint32_t nf = pf_walk(S, list, n, px, py, 0, ids, cov, pre, fin);
        float4 col;
        float dtx = 0.f, dty = 0.f;
        if (nf == 0) {
            col = bg;
            acc_bg[0] += dcol.x; acc_bg[1] += dcol.y; acc_bg[2] += dcol.z; acc_bg[3] += dcol.w;
            if (S.bg_image && G.d_background_image) {
                float *db = G.d_background_image + pix * 4u;
                unsafeAtomicAdd(db + 0, dcol.x); unsafeAtomicAdd(db + 1, dcol.y);
                unsafeAtomicAdd(db + 2, dcol.z); unsafeAtomicAdd(db + 3, dcol.w);
            }
        } else {
            float fa = fin.w;
            float fr = fin.x, fg = fin.y, fb = fin.z;
            if (fa > 1.0e-6f) {
                float sc = 1.0f / fa;
                fr = fr * sc; fg = fg * sc; fb = fb * sc;
            }
            col = make_float4(fr, fg, fb, fa);
            float dr = dcol.x, dg = dcol.y, db = dcol.z, da = dcol.w;
            if (fa > 1.0e-6f) {
                float sc = 1.0f / fa;
                float dot = dr * fr + dg * fg + db * fb;
                dr = dr * sc; dg = dg * sc; db = db * sc;
                da -= dot / fa;
            }
            uint32_t nwin = (nf + PF_WINDOW - 1) / PF_WINDOW;
            for (int w = (int)nwin - 1; w >= 0; --w) {
                uint32_t wstart = (uint32_t)w * PF_WINDOW;
                float4 wpre = bg;
                if (nwin > 1) {
                    float4 f2 = bg;
                    pf_walk(S, list, n, px, py, wstart, ids, cov, wpre, f2);
                }
                uint32_t m = min((uint32_t)PF_WINDOW, nf - wstart);
                for (int i = (int)m - 1; i >= 0; --i) {
                    float4 pc = wpre;
                    for (int j = 0; j < i; ++j) {
                        uint32_t id = ids[j * BLOCK];
                        blend_cpu(pc, pf_color(S, list[id >> 2], (id & 1u) != 0, cov[j * BLOCK], px, py));
                    }
                    uint32_t id = ids[i * BLOCK];
                    uint32_t e = id >> 2;
                    bool is_stroke = (id & 1u) != 0, within = (id & 2u) != 0;
                    uint32_t g = list[e];
                    float4 c = pf_color(S, g, is_stroke, cov[i * BLOCK], px, py);
                    float a = c.w;
                    float d_prev_a = da * (1.0f - a);
                    float d_alpha = da * (1.0f - pc.w) + (dr * (c.x - pc.x) + dg * (c.y - pc.y) + db * (c.z - pc.z));
                    // the fragment's geometry again (one closest-point query)
                    PHit h;
                    float wcov, d;
                    bool ok;
                    if (is_stroke) {
                        ok = group_closest(S, g, px, py, BIG, h);
                        d = h.dist;
                        float sw = S.shape_data[h.prep * 15u + 3u];
                        float ap = absf_(d) + sw, am = absf_(d) - sw;
                        wcov = smoothstep_unit(ap) - smoothstep_unit(am);
                        if (ok && wcov != 0.f) {
                            float d_w = wcov > 0.f ? (a / wcov) * d_alpha : 0.f;
                            float4 dv = make_float4(dr * a, dg * a, db * a, d_alpha * wcov);
                            paint_adjoint(S, G, s_stage, s_off, list, e, true, dv, px, py, dtx, dty);
                            float d_ap = d_smoothstep_unit(ap, d_w);
                            float d_am = -d_smoothstep_unit(am, d_w);
                            float d_d = d_ap + d_am;
                            if (d < 0.f) d_d = -d_d;
                            unsafeAtomicAdd(G.d_shape_stroke_width + S.prepared_to_scene[h.prep], d_ap - d_am);
                            if (absf_(d_d) > 1.0e-10f) d_distance_query(S, G, g, h, px, py, d_d, dtx, dty);
                        }
                    } else {
                        d = cov[i * BLOCK];
                        wcov = smoothstep_unit(d);
                        if (wcov != 0.f) {
                            float d_w = wcov > 0.f ? (a / wcov) * d_alpha : 0.f;
                            float4 dv = make_float4(dr * a, dg * a, db * a, d_alpha * wcov);
                            paint_adjoint(S, G, s_stage, s_off, list, e, false, dv, px, py, dtx, dty);
                            float d_d = d_smoothstep_unit(d, d_w);
                            if (d < 0.f) d_d = -d_d;
                            if (absf_(d_d) > 1.0e-10f && within && group_closest(S, g, px, py, 1.0f, h))
                                d_distance_query(S, G, g, h, px, py, d_d, dtx, dty);
                        }
                    }
                    dr = dr * (1.0f - a); dg = dg * (1.0f - a); db = db * (1.0f - a); da = d_prev_a;
                }
            }
            acc_bg[0] += dr; acc_bg[1] += dg; acc_bg[2] += db; acc_bg[3] += da;
            if (S.bg_image && G.d_background_image) {
                float *dbp = G.d_background_image + pix * 4u;
                unsafeAtomicAdd(dbp + 0, dr); unsafeAtomicAdd(dbp + 1, dg);
                unsafeAtomicAdd(dbp + 2, db); unsafeAtomicAdd(dbp + 3, da);
            }
        }
        if (G.d_translation && (dtx != 0.f || dty != 0.f)) {
            unsafeAtomicAdd(G.d_translation + pix * 2u, dtx);
            unsafeAtomicAdd(G.d_translation + pix * 2u + 1, dty);
        }
        acc_fr += filter_grad(S, weight, d_render, col, px0, py0);
    }
    float v[5] = {acc_bg[0], acc_bg[1], acc_bg[2], acc_bg[3], acc_fr};
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (int k = 0; k < 5; ++k) {
        float r = wave_sum(v[k]);
        if (lane == 0) s_red[wave][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        float r = s_red[0][threadIdx.x] + s_red[1][threadIdx.x] + s_red[2][threadIdx.x] + s_red[3][threadIdx.x];
        if (threadIdx.x < 4) {
            if (!S.bg_image) unsafeAtomicAdd(G.d_background + threadIdx.x, r);
        } else {
            unsafeAtomicAdd(G.d_filter_radius, r);
        }
    }
    stage_flush(S, G, list, n, s_off, s_stage);
}

__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_PREFILTER))) void k_backward_prefiltered(
    DevScene S, const float *__restrict__ weight,
                                                                const float *__restrict__ d_render, DevGrads G) {
    prefiltered_unit(S, blockIdx.x, weight, d_render, G);
}

// ---------------------------------------------------------------------------
// Boundary sampling (backward/boundary.rs:113-1148).

// sample_cdf (boundary.rs:454-469): first entry with u <= cdf[i], by binary
// search on the non-decreasing table.
DVG_DEV uint32_t cdf_search(const float *cdf, uint32_t n, float u, float &t_out) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (u > cdf[mid]) lo = mid + 1;
        else hi = mid;
    }
    uint32_t idx = lo >= n ? n - 1 : lo;
    float prev = idx == 0 ? 0.f : cdf[idx - 1];
    float denom = fmaxf(cdf[idx] - prev, 1.0e-6f);
    float t = (u - prev) / denom;
    t_out = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
    return idx;
}

// The shape CDF's block ends in LDS, for a two-level search: block
// j = [j B, min((j + 1) B, n)) of the table ends at piv[j], so the first block
// whose end is >= u holds the table's first entry >= u, and a search inside
// that block finds the same index as the search over the whole table.
constexpr uint32_t CDF_PIVOTS = 1024;
struct CdfPivots {
    float piv[CDF_PIVOTS];
    uint32_t nblk, bsz;
};

// All threads of the workgroup call it (it synchronizes).
DVG_DEV void stage_cdf_pivots(const DevScene &S, CdfPivots *cp) {
    uint32_t n = S.n_cdf;
    uint32_t bsz = n > CDF_PIVOTS ? (n + CDF_PIVOTS - 1) / CDF_PIVOTS : 1u;
    uint32_t nblk = n > 0 ? (n + bsz - 1) / bsz : 0u;
    for (uint32_t j = threadIdx.x; j < nblk; j += BLOCK) {
        uint32_t end = min((j + 1u) * bsz, n) - 1u;
        cp->piv[j] = S.shape_cdf[end];
    }
    if (threadIdx.x == 0) { cp->nblk = nblk; cp->bsz = bsz; }
    __syncthreads();
}

// cdf_search over the shape CDF through its staged block ends.
DVG_DEV uint32_t shape_cdf_search(const DevScene &S, const CdfPivots *cp, float u) {
    uint32_t nblk = cp->nblk, bsz = cp->bsz, n = S.n_cdf;
    uint32_t lo = 0, hi = nblk;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (u > cp->piv[mid]) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= nblk) return n - 1;  // no entry >= u: the last one, as cdf_search
    uint32_t b0 = lo * bsz, b1 = min(b0 + bsz, n);
    lo = b0;
    hi = b1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (u > S.shape_cdf[mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

DVG_DEV float2 normalize2(float x, float y) {
    float len = sqrtf(x * x + y * y);
    if (len == 0.f) return make_float2(0.f, 0.f);
    return make_float2(x / len, y / len);
}

struct BData {
    uint32_t base_point_id, point_id;
    float t;
    bool is_stroke;
};

DVG_DEV bool path_pt(const DevScene &S, uint32_t si, uint32_t idx, float default_r, float &x, float &y, float &r) {
    uint32_t np = S.sp_cnt[si];
    if (np == 0) return false;
    bool closed = S.s_closed[si] != 0;
    uint32_t i;
    if (closed) i = idx % np;
    else {
        if (idx >= np) return false;
        i = idx;
    }
    const float *p = S.path_points + (size_t)(S.sp_off[si] + i) * 2u;
    x = p[0];
    y = p[1];
    uint32_t nt = S.st_cnt[si];
    r = (nt > 0 && i < nt) ? S.path_thickness[S.st_off[si] + i] * 1.0f : default_r;
    return true;
}

// sample_boundary_path (boundary.rs:586-843)
DVG_DEV bool sample_boundary_path(const DevScene &S, uint32_t si, float path_length, float t, float dir, float sw,
                                  float &ox, float &oy, float &nx, float &ny, float &pdf, BData &bd) {
    const float PI = 3.14159274101257324219f;
    uint32_t np = S.sp_cnt[si], nb = S.sc_cnt[si];
    if (np == 0 || nb == 0) return false;
    bool closed = S.s_closed[si] != 0;
    uint32_t nt = S.st_cnt[si];
    const float *th = S.path_thickness + S.st_off[si];
    const float *pts = S.path_points + (size_t)S.sp_off[si] * 2u;
    const uint32_t *ctrl = S.path_ctrl + S.sc_off[si];
    if (dir != 0.f && !closed) {
        float cap_length = 0.f;
        if (nt > 0) cap_length = PI * (th[0] + th[nt - 1]);
        else cap_length = 2.0f * PI * sw;
        float denom = cap_length + path_length;
        if (denom > 0.f) {
            float cap_prob = cap_length / denom;
            if (t < cap_prob) {
                t /= cap_prob;
                pdf *= cap_prob;
                float r0 = sw, r1 = sw;
                if (nt > 0) { r0 = th[0]; r1 = th[nt - 1]; }
                float angle = 2.0f * PI * t;
                if (dir < 0.f) {
                    float offx = r0 * cosf(angle), offy = r0 * sinf(angle);
                    float2 nn = normalize2(offx, offy);
                    nx = nn.x; ny = nn.y;
                    pdf /= 2.0f * PI * r0;
                    bd.base_point_id = 0; bd.point_id = 0; bd.t = 0.f;
                    ox = pts[0] + offx; oy = pts[1] + offy;
                    return true;
                }
                float offx = r1 * cosf(angle), offy = r1 * sinf(angle);
                float2 nn = normalize2(offx, offy);
                nx = nn.x; ny = nn.y;
                pdf /= 2.0f * PI * r1;
                bd.base_point_id = nb - 1;
                uint32_t sub = 2u + ctrl[nb - 1];
                bd.point_id = np >= sub ? np - sub : 0u;
                bd.t = 1.f;
                ox = pts[(np - 1) * 2] + offx; oy = pts[(np - 1) * 2 + 1] + offy;
                return true;
            }
            t = (t - cap_prob) / (1.0f - cap_prob);
            pdf *= 1.0f - cap_prob;
        }
    }
    uint32_t ccnt = S.path_cdf_cnt[si];
    if (ccnt == 0) return false;
    float lt;
    uint32_t sid = cdf_search(S.path_cdf + S.path_cdf_off[si], ccnt, t, lt);
    uint32_t pid = S.path_point_ids[S.path_point_off[si] + sid];
    bd.base_point_id = sid;
    bd.point_id = pid;
    bd.t = lt;
    if (lt < -1.0e-3f || lt > 1.0f + 1.0e-3f) return false;
    uint32_t nc = sid < nb ? ctrl[sid] : 0u;
    float pmf = S.path_pmf[S.path_cdf_off[si] + sid];
    auto next_index = [&](uint32_t idx) -> uint32_t { return closed ? idx % np : (idx < np - 1 ? idx : np - 1); };
    float x0, y0, r0, x1, y1, r1, x2, y2, r2, x3, y3, r3;
    float ex, ey, tx, ty, rr;
    float tt = 1.0f - lt;
    if (nc == 0) {
        if (!path_pt(S, si, pid, sw, x0, y0, r0) || !path_pt(S, si, next_index(pid + 1), sw, x1, y1, r1)) return false;
        tx = x1 - x0; ty = y1 - y0;
        ex = x0 + tx * lt; ey = y0 + ty * lt;
        rr = r0 + lt * (r1 - r0);
    } else if (nc == 1) {
        if (!path_pt(S, si, pid, sw, x0, y0, r0) || !path_pt(S, si, pid + 1, sw, x1, y1, r1) ||
            !path_pt(S, si, next_index(pid + 2), sw, x2, y2, r2)) return false;
        ex = ((x0 * (tt * tt)) + (x1 * (2.0f * tt * lt))) + (x2 * (lt * lt));
        ey = ((y0 * (tt * tt)) + (y1 * (2.0f * tt * lt))) + (y2 * (lt * lt));
        tx = ((x1 - x0) * (2.0f * tt)) + ((x2 - x1) * (2.0f * lt));
        ty = ((y1 - y0) * (2.0f * tt)) + ((y2 - y1) * (2.0f * lt));
        rr = ((r0 * (tt * tt)) + (r1 * (2.0f * tt * lt))) + (r2 * (lt * lt));
    } else if (nc == 2) {
        if (!path_pt(S, si, pid, sw, x0, y0, r0) || !path_pt(S, si, pid + 1, sw, x1, y1, r1) ||
            !path_pt(S, si, pid + 2, sw, x2, y2, r2) || !path_pt(S, si, next_index(pid + 3), sw, x3, y3, r3))
            return false;
        ex = (((x0 * (tt * tt * tt)) + (x1 * (3.0f * tt * tt * lt))) + (x2 * (3.0f * tt * lt * lt))) + (x3 * (lt * lt * lt));
        ey = (((y0 * (tt * tt * tt)) + (y1 * (3.0f * tt * tt * lt))) + (y2 * (3.0f * tt * lt * lt))) + (y3 * (lt * lt * lt));
        tx = (((x1 - x0) * (3.0f * tt * tt)) + ((x2 - x1) * (6.0f * tt * lt))) + ((x3 - x2) * (3.0f * lt * lt));
        ty = (((y1 - y0) * (3.0f * tt * tt)) + ((y2 - y1) * (6.0f * tt * lt))) + ((y3 - y2) * (3.0f * lt * lt));
        rr = (((r0 * (tt * tt * tt)) + (r1 * (3.0f * tt * tt * lt))) + (r2 * (3.0f * tt * lt * lt))) + (r3 * (lt * lt * lt));
    } else {
        return false;
    }
    float tl = sqrtf(tx * tx + ty * ty);
    if (tl == 0.f) return false;
    nx = -ty / tl;
    ny = tx / tl;
    pdf *= pmf / tl;
    ox = ex;
    oy = ey;
    if (dir != 0.f) {
        ox += nx * (dir * rr);
        oy += ny * (dir * rr);
        if (dir < 0.f) { nx = nx * -1.0f; ny = ny * -1.0f; }
    }
    return true;
}

// within_distance_edge / is_inside_edge (sampling.rs:485-549) on the sampled
// shape alone, in a group of several shapes.
template <int M>
DVG_DEV bool edge_shape_hit(const DevScene &S, uint32_t g, uint32_t eq_prep, float px, float py, bool &uns) {
    const float *gd = S.group_data + g * 16u;
    uint32_t fk = (uint32_t)gd[2], sk = (uint32_t)gd[4], rule = (uint32_t)gd[7];
    float lx, ly;
    group_local(S, g, px, py, lx, ly);
    const float *xf = S.shape_xform + eq_prep * 6u;
    float spx = xf[0] * lx + xf[1] * ly + xf[2];
    float spy = xf[3] * lx + xf[4] * ly + xf[5];
    const float *sd = S.shape_data + eq_prep * 15u;
    Cov c = {BIG, 0.f, 0.f, BIG, 0.f, 0.f, 0.f};
    if (sk != DVG_PAINT_NONE) shape_stroke<M>(S, sd, false, spx, spy, c);
    if (fk != DVG_PAINT_NONE) shape_fill<M>(S, sd, rule, false, spx, spy, c);
    if (M == COV_CHK && c.unsure > 0.f) uns = true;
    bool et = c.stroke_hit > 0.f;
    if (fk != DVG_PAINT_NONE && ((rule == 1u) ? c.crossings > 0.f : c.winding != 0.f)) et = true;
    return et;
}

// One query point's fragment of group g in the EdgeQuery compositing loop
// (sampling.rs:145-160): blend, count, and track the hit for an initial
// false (h0) and an initial true (h1).
DVG_DEV void edge_blend(const DevScene &S, uint32_t g, bool is_stroke, bool eq, float px, float py, float4 &acc,
                        uint32_t &nfr, bool &h0, bool &h1) {
    float4 c = frag_color(S, g, is_stroke, px, py);
    if (c.w >= 1.0f) { h0 = false; h1 = false; }
    if (eq) { h0 = true; h1 = true; }
    blend_cpu(acc, c);
    ++nfr;
}

// (with no fragment the composite is still the background it started from)
DVG_DEV float4 edge_finish(float4 acc, uint32_t nfr) {
    if (nfr == 0) return acc;
    float fa = acc.w;
    if (fa > 1.0e-6f) {
        float sc = 1.0f / fa;
        return make_float4(acc.x * sc, acc.y * sc, acc.z * sc, fa);
    }
    return acc;
}

// Colours of the two query points of a boundary sample with the EdgeQuery hit
// tracking of sampling.rs:76-224 (no adjoint), in one walk of a tile list:
// both points' fragment tests per group (group_frags_pair), two composites.
template <int M>
DVG_DEV void edge_walk_pair(const DevScene &S, uint32_t tile, float px0, float py0, float px1, float py1, float4 bg,
                            uint32_t eq_group, uint32_t eq_prep, float4 &col0, float4 &col1, bool &hit0, bool &hit1,
                            bool &uns) {
    uint32_t lb = S.tile_offsets[tile], le = S.tile_end[tile];
    float4 acc0 = bg, acc1 = bg;
    uint32_t nfr0 = 0, nfr1 = 0;
    bool et0 = false, et1 = false;  // hit set while collecting fragments
    bool h00 = false, h01 = true, h10 = false, h11 = true;
    for (uint32_t e = lb; e < le; ++e) {
        uint32_t g = S.bwd_entries[e];
        bool st0, fi0, st1, fi1;
        group_frags_pair<M>(S, g, px0, py0, px1, py1, st0, fi0, st1, fi1, uns);
        bool eq = g == eq_group;
        if (eq) {
            // for a group of one shape the sampled shape's own tests are the
            // group's tests above (bounds are conservative)
            if ((uint32_t)S.group_data[g * 16u + 1] == 1u) {
                et0 = et0 || st0 || fi0;
                et1 = et1 || st1 || fi1;
            } else {
                if (edge_shape_hit<M>(S, g, eq_prep, px0, py0, uns)) et0 = true;
                if (edge_shape_hit<M>(S, g, eq_prep, px1, py1, uns)) et1 = true;
            }
        }
        if (st0) edge_blend(S, g, true, eq, px0, py0, acc0, nfr0, h00, h01);
        if (st1) edge_blend(S, g, true, eq, px1, py1, acc1, nfr1, h10, h11);
        if (fi0) edge_blend(S, g, false, eq, px0, py0, acc0, nfr0, h00, h01);
        if (fi1) edge_blend(S, g, false, eq, px1, py1, acc1, nfr1, h10, h11);
    }
    hit0 = et0 ? h01 : h00;
    hit1 = et1 ? h11 : h10;
    col0 = edge_finish(acc0, nfr0);
    col1 = edge_finish(acc1, nfr1);
}

DVG_DEV uint32_t tile_of(const DevScene &S, float px, float py) {
    int tx = (int)floorf(px / (float)TILE), ty = (int)floorf(py / (float)TILE);
    tx = tx < 0 ? 0 : (tx >= (int)S.tiles_x ? (int)S.tiles_x - 1 : tx);
    ty = ty < 0 ? 0 : (ty >= (int)S.tiles_y ? (int)S.tiles_y - 1 : ty);
    return (uint32_t)ty * S.tiles_x + (uint32_t)tx;
}

// The colours of a boundary sample's query points (normalized (qx0, qy0) and
// (qx1, qy1)) through the tile list of the sampled pixel (pxi, pyi): binning
// pads every group's box by tile_pad pixels, more than either point's
// distance from the sampled point, so that list holds every group either
// point can be inside.  Samples are sorted by pixel block, so the lanes of a
// wave almost always share the tile, and the walk then runs with it in a
// scalar register (list, group and curve data through scalar loads).  A lane
// whose points leave the padded tile (a sampled point up to a pixel left of
// or above the canvas, which truncates into pixel 0) walks each point through
// its own tile, one per pass.
template <int M>
DVG_DEV void edge_colors(const DevScene &S, uint32_t pxi, uint32_t pyi, float qx0, float qy0, float qx1, float qy1,
                         float4 bg, uint32_t eq_group, uint32_t eq_prep, float4 &c0, float4 &c1, bool &h0, bool &h1,
                         bool &uns) {
    float Wf = (float)S.W, Hf = (float)S.H;
    float px0 = qx0 * Wf, py0 = qy0 * Hf, px1 = qx1 * Wf, py1 = qy1 * Hf;
    uint32_t tx = pxi / (uint32_t)TILE, ty = pyi / (uint32_t)TILE;
    float pad = tile_pad(S.W, S.H) - 0.01f;  // less a margin for rounding
    float rx0 = (float)(tx * TILE) - pad, rx1 = (float)((tx + 1) * TILE) + pad;
    float ry0 = (float)(ty * TILE) - pad, ry1 = (float)((ty + 1) * TILE) + pad;
    bool split = !(fminf(px0, px1) >= rx0 && fmaxf(px0, px1) <= rx1 && fminf(py0, py1) >= ry0 &&
                   fmaxf(py0, py1) <= ry1);
    uint32_t ptile = ty * S.tiles_x + tx;
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && __ballot(split) == 0) break;
        if (pass == 0 || split) {
            float ax = px0, ay = py0, bx = px1, by = py1;
            uint32_t t = ptile;
            if (split) {
                if (pass == 0) { bx = px0; by = py0; }
                else { ax = px1; ay = py1; }
                t = tile_of(S, ax, ay);
            }
            float4 ca, cb;
            bool ha, hb;
            uint32_t t0 = __builtin_amdgcn_readfirstlane(t);
            if (__ballot(t != t0) == 0)
                edge_walk_pair<M>(S, t0, ax, ay, bx, by, bg, eq_group, eq_prep, ca, cb, ha, hb, uns);
            else
                edge_walk_pair<M>(S, t, ax, ay, bx, by, bg, eq_group, eq_prep, ca, cb, ha, hb, uns);
            if (pass == 0) { c0 = ca; h0 = ha; }
            if (pass == 1 || !split) { c1 = cb; h1 = hb; }
        }
    }
}

// One boundary sample of boundary_sampling (boundary.rs:138-197): shape by
// length CDF, point and normal on it, canvas position and pixel.
struct BPoint {
    float t, pdf, lx, ly, ncx, ncy, qx, qy;
    uint32_t si, gi, prep, kind, pxi, pyi;
    BData bd;
};

DVG_DEV bool boundary_point(const DevScene &S, const CdfPivots *cp, size_t idx, BPoint &P) {
    const float PI = 3.14159274101257324219f;
    Pcg32 rng((uint64_t)idx, (uint64_t)S.seed);
    float u = rng.next_f32();
    float t = rng.next_f32();
    float unused;
    uint32_t sid = cp ? shape_cdf_search(S, cp, u) : cdf_search(S.shape_cdf, S.n_cdf, u, unused);
    uint32_t si = S.shape_ids[sid], gi = S.group_ids[sid];
    float spmf = S.shape_pmf[sid];
    if (spmf <= 0.f) return false;
    uint32_t prep = sid;  // prepared instances follow the CDF order
    float slen = S.shape_lengths[si];
    const float *gd = S.group_data + gi * 16u;
    bool has_fill = (uint32_t)gd[2] != DVG_PAINT_NONE, has_stroke = (uint32_t)gd[4] != DVG_PAINT_NONE;
    // sample_boundary_point (boundary.rs:472-583)
    if (!has_fill && !has_stroke) return false;
    float pdf = 1.0f, lt = t;
    bool stroke = false;
    if (has_fill && has_stroke) {
        if (lt < 0.5f) { lt *= 2.0f; pdf = 0.5f; }
        else { stroke = true; lt = 2.0f * (lt - 0.5f); pdf = 0.5f; }
    } else if (has_stroke) {
        stroke = true;
    }
    float dir = 0.f;
    if (stroke) {
        if (lt < 0.5f) { dir = -1.0f; lt *= 2.0f; pdf *= 0.5f; }
        else { dir = 1.0f; lt = 2.0f * (lt - 0.5f); pdf *= 0.5f; }
    }
    BData bd = {0u, 0u, 0.f, stroke};
    const float *sd = S.shape_data + prep * 15u;
    uint32_t kind = (uint32_t)sd[0];
    float sw = sd[3];
    float lx = 0.f, ly = 0.f, nx = 0.f, ny = 0.f;
    if (kind == DVG_SHAPE_CIRCLE) {
        float r = sd[6];
        if (r <= 0.f) return false;
        float angle = 2.0f * PI * lt;
        float offx = r * cosf(angle), offy = r * sinf(angle);
        float2 nn = normalize2(offx, offy);
        nx = nn.x; ny = nn.y;
        pdf /= 2.0f * PI * r;
        lx = sd[4] + offx; ly = sd[5] + offy;
        if (dir != 0.f) {
            lx += nx * (dir * sw); ly += ny * (dir * sw);
            if (dir < 0.f) { nx = -nx; ny = -ny; }
        }
    } else if (kind == DVG_SHAPE_ELLIPSE) {
        float rx = sd[6], ry = sd[7];
        if (rx <= 0.f || ry <= 0.f) return false;
        float angle = 2.0f * PI * lt;
        float s_ = sinf(angle), c_ = cosf(angle);
        float offx = rx * c_, offy = ry * s_;
        float dxdt = -rx * s_ * 2.0f * PI, dydt = ry * c_ * 2.0f * PI;
        float2 nn = normalize2(dydt, -dxdt);
        nx = nn.x; ny = nn.y;
        pdf /= sqrtf(dxdt * dxdt + dydt * dydt);
        lx = sd[4] + offx; ly = sd[5] + offy;
        if (dir != 0.f) {
            lx += nx * (dir * sw); ly += ny * (dir * sw);
            if (dir < 0.f) { nx = -nx; ny = -ny; }
        }
    } else if (kind == DVG_SHAPE_RECT) {
        float mnx = sd[4], mny = sd[5], mxx = sd[6], mxy = sd[7];
        float w = mxx - mnx, h = mxy - mny;
        if (w == 0.f && h == 0.f) return false;
        pdf /= 2.0f * (w + h);
        if (lt <= w / (w + h)) {
            lt *= (w + h) / w;
            if (lt < 0.5f) { nx = 0.f; ny = -1.f; lx = mnx + 2.0f * lt * (mxx - mnx); ly = mny + 0.f; }
            else { nx = 0.f; ny = 1.f; lx = mnx + 2.0f * (lt - 0.5f) * (mxx - mnx); ly = mxy + 0.f; }
        } else {
            lt = (lt - w / (w + h)) * (w + h) / h;
            if (lt < 0.5f) { nx = -1.f; ny = 0.f; lx = mnx + 0.f; ly = mny + 2.0f * lt * (mxy - mny); }
            else { nx = 1.f; ny = 0.f; lx = mxx + 0.f; ly = mny + 2.0f * (lt - 0.5f) * (mxy - mny); }
        }
        if (dir != 0.f) {
            lx += nx * (dir * sw); ly += ny * (dir * sw);
            if (dir < 0.f) { nx = nx * -1.0f; ny = ny * -1.0f; }
        }
    } else {
        if (S.path_cdf_cnt[si] == 0 || !(slen > 0.f)) return false;
        if (!sample_boundary_path(S, si, slen, lt, dir, sw, lx, ly, nx, ny, pdf, bd)) return false;
    }
    if (!isfinite(pdf) || pdf <= 0.f) return false;
    pdf = spmf * pdf;
    // boundary.rs:168-174: local -> canvas, normal by the inverse transpose
    M3 Gs = m3_affine(S.group_shape_xform + gi * 6u);
    M3 T = m3_affine(S.shape_transform + si * 6u);
    M3 s2c = m3_mul(Gs, T);
    float bx = s2c.m[0][0] * lx + s2c.m[0][1] * ly + s2c.m[0][2];
    float by = s2c.m[1][0] * lx + s2c.m[1][1] * ly + s2c.m[1][2];
    M3 Si = m3_affine(S.shape_xform + prep * 6u);
    M3 Gi = m3_affine(S.group_xform + gi * 6u);
    M3 ci = m3_mul(Si, Gi);
    float2 nc = normalize2(ci.m[0][0] * nx + ci.m[1][0] * ny, ci.m[0][1] * nx + ci.m[1][1] * ny);
    float Wf = (float)S.W, Hf = (float)S.H;
    float qx = bx / Wf, qy = by / Hf;
    // render_edge_sample (boundary.rs:845-867): the pixel under the point
    int pxi = (int)(qx * Wf), pyi = (int)(qy * Hf);
    if (pxi < 0 || pxi >= (int)S.W || pyi < 0 || pyi >= (int)S.H) return false;
    P.t = t; P.pdf = pdf; P.lx = lx; P.ly = ly; P.ncx = nc.x; P.ncy = nc.y; P.qx = qx; P.qy = qy;
    P.si = si; P.gi = gi; P.prep = prep; P.kind = kind; P.pxi = (uint32_t)pxi; P.pyi = (uint32_t)pyi;
    P.bd = bd;
    return true;
}

DVG_DEV uint32_t bucket_of(const DevScene &S, uint32_t x, uint32_t y) {
    uint32_t tile = (y / TILE) * S.tiles_x + x / TILE;
    uint32_t L = S.bnd_log2, per_row = (uint32_t)TILE >> L;
    uint32_t blk = ((y % TILE) >> L) * per_row + ((x % TILE) >> L);
    return tile * per_row * per_row + blk;
}

// Pass 1: bucket key of every boundary sample (nb when it contributes
// nothing), the histogram of the contributing ones, and each sample's rank
// in its bucket (the histogram atomic's return value).
DVG_DEV void bnd_key_unit(const DevScene &S, uint32_t u, uint32_t *__restrict__ keys, uint32_t *__restrict__ ranks,
                          uint32_t *__restrict__ counts, uint32_t nb) {
    __shared__ CdfPivots s_cdf;
    stage_cdf_pivots(S, &s_cdf);
    size_t idx = (size_t)u * BLOCK + threadIdx.x;
    size_t total = (size_t)S.W * S.H * S.spp;
    if (idx >= total) return;
    BPoint P;
    uint32_t key = nb;
    if (S.n_cdf > 0 && boundary_point(S, &s_cdf, idx, P)) key = bucket_of(S, P.pxi, P.pyi);
    keys[idx] = key;
    // samples without a contribution are not counted: a shared counter for
    // them would serialise every such atomic on one address
    if (key < nb) ranks[idx] = atomicAdd(counts + key, 1u);
}

__global__ __launch_bounds__(BLOCK) void k_bnd_key(DevScene S, uint32_t *__restrict__ keys,
                                                   uint32_t *__restrict__ ranks, uint32_t *__restrict__ counts,
                                                   uint32_t nb) {
    bnd_key_unit(S, blockIdx.x, keys, ranks, counts, nb);
}

// Pass 2: counting-sort scatter of sample indices by bucket (bucket offset
// plus the rank from pass 1: no atomics).
DVG_DEV void bnd_scatter_unit(const DevScene &S, uint32_t u, const uint32_t *__restrict__ keys,
                              const uint32_t *__restrict__ ranks, const uint32_t *__restrict__ offsets,
                              uint32_t *__restrict__ order, uint32_t nb) {
    size_t idx = (size_t)u * BLOCK + threadIdx.x;
    size_t total = (size_t)S.W * S.H * S.spp;
    if (idx >= total) return;
    uint32_t key = keys[idx];
    if (key >= nb) return;
    order[offsets[key] + ranks[idx]] = (uint32_t)idx;
}

__global__ __launch_bounds__(BLOCK) void k_bnd_scatter(DevScene S, const uint32_t *__restrict__ keys,
                                                       const uint32_t *__restrict__ ranks,
                                                       const uint32_t *__restrict__ offsets,
                                                       uint32_t *__restrict__ order, uint32_t nb) {
    bnd_scatter_unit(S, blockIdx.x, keys, ranks, offsets, order, nb);
}

DVG_DEV float sel3(uint32_t i, float a, float b, float c) { return i == 0u ? a : (i == 1u ? b : c); }
DVG_DEV float sel4(uint32_t i, float a, float b, float c, float d) {
    return i < 2u ? (i == 0u ? a : b) : (i == 2u ? c : d);
}
DVG_DEV float sel8(uint32_t i, const float *r) {
    return i < 4u ? sel4(i, r[0], r[1], r[2], r[3]) : sel4(i - 4u, r[4], r[5], r[6], r[7]);
}

// Sum over the lanes of the wave whose `mine` is set; every lane returns it.
DVG_DEV float masked_wave_sum(bool mine, float v) { return wave_sum(mine ? v : 0.f); }

// Pass 3: boundary samples in tile order.  Consecutive lanes sit on the same
// few shapes of one tile, so the tile's group list is shared by the wave and
// the shape gradients are reduced per shape instance before the atomics.
// Uncertain boundary samples (a query point within rounding distance of an
// edge), per unit of the bucket-ordered pass: up to FLAG_SLOTS sample indices
// a unit and their count (beyond it the f32 result stands).  The redo pass
// gives each unit FLAG_SLOTS adjacent lanes, so a wave holds two neighbouring
// units of the bucket order and its walks stay close to uniform.
constexpr uint32_t FLAG_SLOTS = FLAG_UNIT_SLOTS;
struct FlagSlots {
    uint32_t *count, *slots;
};

// LIST = false, M = COV_CHK: the samples of unit u in bucket order; those
// with uncertain tests are left to the redo (slots of unit u).  LIST = true:
// the samples in the slots of units u * FLAG_UNITS_PER_WG + k (k <
// FLAG_UNITS_PER_WG; FLAG_SLOTS lanes a unit), with the tests of M.
constexpr uint32_t FLAG_UNITS_PER_WG = BLOCK / FLAG_SLOTS;
template <int M, bool LIST>
DVG_DEV void bnd_eval_unit(const DevScene &S, uint32_t u, const float *__restrict__ weight,
                           const float *__restrict__ d_render, const DevGrads &G, const uint32_t *__restrict__ offsets,
                           const uint32_t *__restrict__ order, uint32_t nb, FlagSlots fs) {
    const float PI = 3.14159274101257324219f;
    // (no staged CDF here: its LDS would cost a workgroup per CU)
    __shared__ BPoint s_pt[BLOCK];
    __shared__ uint32_t s_nflag;
    uint32_t n_valid = offsets[nb];
    uint32_t pos = u * BLOCK + threadIdx.x;
    uint32_t lane = threadIdx.x & 63u;
    size_t slot_base = 0;
    if (LIST) {
        // this lane's unit and slot; units past the samples have no slots
        // (no barrier follows in this mode: lanes without a sample idle)
        uint32_t lu = u * FLAG_UNITS_PER_WG + threadIdx.x / FLAG_SLOTS;
        uint32_t cnt = (lu * BLOCK < n_valid) ? min(fs.count[lu], FLAG_SLOTS) : 0u;
        slot_base = (size_t)lu * FLAG_SLOTS;
        pos = threadIdx.x % FLAG_SLOTS;
        n_valid = cnt;
    } else {
        if (u * BLOCK >= n_valid) return;  // whole workgroup past the end
        if (M == COV_CHK) {
            if (threadIdx.x == 0) s_nflag = 0;
            __syncthreads();
        }
    }
    BPoint P;
    bool valid = false;
    float contrib = 0.f, nnx = 0.f, nny = 0.f;
    size_t pix = 0;
    if (pos < n_valid) {
        size_t idx = LIST ? fs.slots[slot_base + pos] : order[pos];
        valid = boundary_point(S, nullptr, idx, P);
        if (valid) {
            float4 bgs = background_at(S, P.pxi, P.pyi);
            float4 bg = make_float4(bgs.x * bgs.w, bgs.y * bgs.w, bgs.z * bgs.w, bgs.w);
            bool ih, oh;
            float qix = P.qx - P.ncx * 1.0e-4f, qiy = P.qy - P.ncy * 1.0e-4f;
            float qox = P.qx + P.ncx * 1.0e-4f, qoy = P.qy + P.ncy * 1.0e-4f;
            uint32_t gi = P.gi, prep = P.prep, pxi = P.pxi, pyi = P.pyi;
            // the point waits in LDS, not in registers, during the scene walk
            // (the empty asm is a compiler barrier: P is reloaded)
            s_pt[threadIdx.x] = P;
            asm volatile("" ::: "memory");
            bool uns = false;
            float4 cin, cout;
            edge_colors<M>(S, pxi, pyi, qix, qiy, qox, qoy, bg, gi, prep, cin, cout, ih, oh, uns);
            asm volatile("" ::: "memory");
            P = s_pt[threadIdx.x];
            pix = (size_t)P.pyi * S.W + P.pxi;
            valid = ih || oh;
            if (M == COV_CHK && !LIST && uns) {
                uint32_t slot = atomicAdd(&s_nflag, 1u);
                if (slot < FLAG_SLOTS) {
                    // (the sample index read again: not held through the walk)
                    fs.slots[(size_t)u * FLAG_SLOTS + slot] = order[pos];
                    valid = false;
                }
            }
            nnx = P.ncx;
            nny = P.ncy;
            if (valid && !ih) {
                nnx = nnx * -1.0f; nny = nny * -1.0f;
                float4 tmp = cin; cin = cout; cout = tmp;
            }
            if (valid) {
                float Wf = (float)S.W, Hf = (float)S.H;
                float4 dc = gather_d_color(S, weight, d_render, P.qx * Wf, P.qy * Hf);
                float norm = Wf * Hf;
                if (norm > 0.f) { dc.x /= norm; dc.y /= norm; dc.z /= norm; dc.w /= norm; }
                contrib = ((cin.x - cout.x) * dc.x + (cin.y - cout.y) * dc.y + (cin.z - cout.z) * dc.z +
                           (cin.w - cout.w) * dc.w) / P.pdf;
                valid = isfinite(contrib);
            }
        }
    }
    // ---- accumulate_boundary_gradient (boundary.rs:970-1148) ----
    // per instance: d(shape_to_canvas) (9), stroke width (1), shape params (4)
    float v[14];
    for (int k = 0; k < 14; ++k) v[k] = 0.f;
    uint32_t key = 0xffffffffu;
    if (valid) {
        key = P.prep;
        M3 Gs = m3_affine(S.group_shape_xform + P.gi * 6u);
        M3 T = m3_affine(S.shape_transform + P.si * 6u);
        M3 s2c = m3_mul(Gs, T);
        float t0 = s2c.m[0][0] * P.lx + s2c.m[0][1] * P.ly + s2c.m[0][2];
        float t1 = s2c.m[1][0] * P.lx + s2c.m[1][1] * P.ly + s2c.m[1][2];
        float t2 = s2c.m[2][0] * P.lx + s2c.m[2][1] * P.ly + s2c.m[2][2];
        float dox = nnx * contrib, doy = nny * contrib;
        float ox = t0 / t2, oy = t1 / t2;
        float dt0 = dox / t2, dt1 = doy / t2, dt2 = -(dox * ox + doy * oy) / t2;
        v[0] = dt0 * P.lx; v[1] = dt0 * P.ly; v[2] = dt0;
        v[3] = dt1 * P.lx; v[4] = dt1 * P.ly; v[5] = dt1;
        v[6] = dt2 * P.lx; v[7] = dt2 * P.ly; v[8] = dt2;
        if (P.bd.is_stroke) {
            uint32_t nt = S.st_cnt[P.si];
            if (P.kind == DVG_SHAPE_PATH && nt > 0) {
                float *dth = G.d_shape_thickness + S.st_off[P.si];
                uint32_t np = S.sp_cnt[P.si];
                bool closed = S.s_closed[P.si] != 0;
                uint32_t nc_ = P.bd.base_point_id < S.sc_cnt[P.si] ? S.path_ctrl[S.sc_off[P.si] + P.bd.base_point_id] : 0u;
                uint32_t pid = P.bd.point_id;
                float tv = P.bd.t, tt = 1.0f - tv;
                auto add = [&](uint32_t i, float x) { if (i < nt) unsafeAtomicAdd(dth + i, x); };
                if (nc_ == 0) {
                    add(pid, (1.0f - tv) * contrib);
                    add(closed ? (pid + 1) % np : pid + 1, tv * contrib);
                } else if (nc_ == 1) {
                    add(pid, tt * tt * contrib);
                    add(pid + 1, 2.0f * tt * tv * contrib);
                    add(closed ? (pid + 2) % np : pid + 2, tv * tv * contrib);
                } else if (nc_ == 2) {
                    add(pid, tt * tt * tt * contrib);
                    add(pid + 1, 3.0f * tt * tt * tv * contrib);
                    add(pid + 2, 3.0f * tt * tv * tv * contrib);
                    add(closed ? (pid + 3) % np : pid + 3, tv * tv * tv * contrib);
                }
            } else {
                v[9] = contrib;
            }
        }
        if (P.kind == DVG_SHAPE_CIRCLE) {
            v[10] = nnx * contrib; v[11] = nny * contrib; v[12] = contrib;
        } else if (P.kind == DVG_SHAPE_ELLIPSE) {
            v[10] = nnx * contrib; v[11] = nny * contrib;
            float angle = 2.0f * PI * P.t;
            v[12] = cosf(angle) * nnx * contrib;
            v[13] = sinf(angle) * nny * contrib;
        } else if (P.kind == DVG_SHAPE_RECT) {
            if (nnx == -1.f && nny == 0.f) v[10] = -contrib;
            else if (nnx == 1.f && nny == 0.f) v[12] = contrib;
            else if (nnx == 0.f && nny == -1.f) v[11] = -contrib;
            else if (nnx == 0.f && nny == 1.f) v[13] = contrib;
        }
        if (G.d_translation) {
            unsafeAtomicAdd(G.d_translation + pix * 2u, nnx * contrib);
            unsafeAtomicAdd(G.d_translation + pix * 2u + 1, nny * contrib);
        }
    }
    // reduce per instance across the wave; the leader applies
    // d_G += dM T^T and d_T += G^T dM (the products are linear in dM)
    uint64_t pending = __ballot(valid);
    while (pending) {
        int leader = __ffsll((unsigned long long)pending) - 1;
        uint32_t lkey = __shfl(key, leader, 64);
        bool mine = valid && key == lkey;
        float r[14];
        for (int k = 0; k < 14; ++k) r[k] = masked_wave_sum(mine, v[k]);
        // the 23 outputs go out from 23 lanes in one instruction each
        uint32_t lgi = __shfl(P.gi, leader, 64), lsi = __shfl(P.si, leader, 64);
        if (lane < 18u) {
            uint32_t e = lane < 9u ? lane : lane - 9u, i = e / 3u, j = e - 3u * i;
            float val;
            float *dst;
            if (lane < 9u) {  // d_G[i][j] = sum_k dM[i][k] T[j][k]
                const float *tr = S.shape_transform + lsi * 6u + j * 3u;
                float t0 = j < 2u ? tr[0] : 0.f, t1 = j < 2u ? tr[1] : 0.f, t2 = j < 2u ? tr[2] : 1.f;
                float m0 = sel3(i, r[0], r[3], r[6]), m1 = sel3(i, r[1], r[4], r[7]), m2 = sel3(i, r[2], r[5], r[8]);
                val = m0 * t0 + m1 * t1 + m2 * t2;
                dst = G.d_group_transform + lgi * 9u + e;
            } else {  // d_T[i][j] = sum_k G[k][i] dM[k][j]
                const float *gr = S.group_shape_xform + lgi * 6u;
                float g0 = gr[i], g1 = gr[3u + i], g2 = i == 2u ? 1.f : 0.f;
                float m0 = sel3(j, r[0], r[1], r[2]), m1 = sel3(j, r[3], r[4], r[5]), m2 = sel3(j, r[6], r[7], r[8]);
                val = g0 * m0 + g1 * m1 + g2 * m2;
                dst = G.d_shape_transform + lsi * 9u + e;
            }
            unsafeAtomicAdd(dst, val);
        } else if (lane == 18u) {
            if (r[9] != 0.f) unsafeAtomicAdd(G.d_shape_stroke_width + lsi, r[9]);
        } else if (lane < 23u) {
            uint32_t k = lane - 19u;
            float val = sel4(k, r[10], r[11], r[12], r[13]);
            if (val != 0.f) unsafeAtomicAdd(G.d_shape_params + lsi * 8u + k, val);
        }
        pending &= ~__ballot(mine);
    }
    // path points, reduced per segment (same shape and first point)
    uint32_t pkey = 0xffffffffu;
    float pv[8];
    for (int k = 0; k < 8; ++k) pv[k] = 0.f;
    uint32_t nc_ = 0;
    bool pvalid = valid && P.kind == DVG_SHAPE_PATH;
    if (pvalid) {
        nc_ = P.bd.base_point_id < S.sc_cnt[P.si] ? S.path_ctrl[S.sc_off[P.si] + P.bd.base_point_id] : 0u;
        pvalid = nc_ <= 2;
        float tv = P.bd.t, tt = 1.0f - tv;
        float w[4] = {0.f, 0.f, 0.f, 0.f};
        if (nc_ == 0) { w[0] = 1.0f - tv; w[1] = tv; }
        else if (nc_ == 1) { w[0] = tt * tt; w[1] = 2.0f * tt * tv; w[2] = tv * tv; }
        else { w[0] = tt * tt * tt; w[1] = 3.0f * tt * tt * tv; w[2] = 3.0f * tt * tv * tv; w[3] = tv * tv * tv; }
        for (int k = 0; k < 4; ++k) {
            pv[2 * k] = nnx * (w[k] * contrib);
            pv[2 * k + 1] = nny * (w[k] * contrib);
        }
        pkey = S.sp_off[P.si] + P.bd.point_id;
    }
    pending = __ballot(pvalid);
    while (pending) {
        int leader = __ffsll((unsigned long long)pending) - 1;
        uint32_t lkey = __shfl(pkey, leader, 64);
        bool mine = pvalid && pkey == lkey;
        float r[8];
        for (int k = 0; k < 8; ++k) r[k] = masked_wave_sum(mine, pv[k]);
        uint32_t lsi = __shfl(P.si, leader, 64), lpid = __shfl(P.bd.point_id, leader, 64);
        uint32_t lnc = __shfl(nc_, leader, 64);
        uint32_t k = lane >> 1;
        if (lane < 8u && k <= lnc + 1u) {
            uint32_t np = S.sp_cnt[lsi];
            bool closed = S.s_closed[lsi] != 0;
            uint32_t i = lpid + k;
            uint32_t index = closed ? i % np : i;
            float val = sel8(lane, r);
            if (index < np) unsafeAtomicAdd(G.d_shape_points + ((size_t)S.sp_off[lsi] + index) * 2u + (lane & 1u), val);
        }
        pending &= ~__ballot(mine);
    }
    if (M == COV_CHK && !LIST) {
        __syncthreads();
        if (threadIdx.x == 0) fs.count[u] = min(s_nflag, FLAG_SLOTS);
    }
}

__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_BOUNDARY))) void k_bnd_eval(
    DevScene S, const float *__restrict__ weight, const float *__restrict__ d_render, DevGrads G,
    const uint32_t *__restrict__ offsets, const uint32_t *__restrict__ order, uint32_t nb) {
    bnd_eval_unit<COV_CPU, false>(S, blockIdx.x, weight, d_render, G, offsets, order, nb, FlagSlots{nullptr, nullptr});
}

// The exact boundary mode's pass: uncertain samples go to the redo's slots.
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_BOUNDARY))) void k_bnd_eval_chk(
    DevScene S, const float *__restrict__ weight, const float *__restrict__ d_render, DevGrads G,
    const uint32_t *__restrict__ offsets, const uint32_t *__restrict__ order, uint32_t nb, FlagSlots fs) {
    bnd_eval_unit<COV_CHK, false>(S, blockIdx.x, weight, d_render, G, offsets, order, nb, fs);
}

// The samples k_bnd_eval left to the redo, unit by unit, with the f64
// windings where the f32 ones are uncertain.
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_REDO))) void k_bnd_redo(DevScene S, const float *__restrict__ weight,
                                                    const float *__restrict__ d_render, DevGrads G,
                                                    const uint32_t *__restrict__ offsets, uint32_t nb, FlagSlots fs) {
    bnd_eval_unit<COV_EXACT, true>(S, blockIdx.x, weight, d_render, G, offsets, nullptr, nb, fs);
}

// Zero up to 14 float arrays in one launch (used instead of memset nodes so
// that the calls stay capturable in a HIP graph).
struct ZeroList {
    float *p[14];
    uint32_t n[14];
};
__global__ void k_zero(ZeroList z) {
    uint32_t k = blockIdx.y;
    float *p = z.p[k];
    if (!p) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < z.n[k]; i += gridDim.x * blockDim.x) p[i] = 0.f;
}

// ---------------------------------------------------------------------------
// Device-side scene update: the packing steps whose inputs change in an
// optimisation loop (control points, colours) restated as kernels, so an
// Adam step never leaves the device.  The arena buffers written here are the
// scene's own (DevScene holds them as const for the render kernels).
// prepare_shape / bounds: prepare.rs:459-561, 207-237; binning.rs:10-313;
// compute_shape_lengths / build_shape_cdf / build_path_cdfs:
// backward/boundary.rs:200-451.

__global__ void k_copy_f32(float *__restrict__ dst, const float *__restrict__ src, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Control points of every curve from the path points (path.rs:100-173).
// ---- self-test of the winding solve --------------------------------------
// The cubic winding test with libm's acos / cos / pow throughout (the
// formula of curves.rs:170-214 and curves.rs:897-946 as written), the
// reference the per-segment record and its fast transcendentals are checked
// against.
DVG_DEV void winding_cubic_libm(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                                float x3, float y3, float &w) {
    float a = -y0 + 3.0f * y1 - 3.0f * y2 + y3;
    float b = 3.0f * y0 - 6.0f * y1 + 3.0f * y2;
    float c = -3.0f * y0 + 3.0f * y1;
    float d = y0 - py;
    float r[3];
    int n;
    const float eps = 1.0e-6f;
    if (absf_(a) < eps) {
        n = solve_quadratic(b, c, d, r[0], r[1]);
    } else {
        float bb = b / a, cc = c / a, dd = d / a;
        const float third = 1.0f / 3.0f;
        float q = (bb * bb - 3.0f * cc) / 9.0f;
        float rr = (2.0f * bb * bb * bb - 9.0f * bb * cc + 27.0f * dd) / 54.0f;
        float r2 = rr * rr, q3 = q * q * q;
        if (r2 < q3) {
            float theta = acosf(rr / sqrtf(q3));
            float sq = sqrtf(q);
            const float two_pi = 2.0f * 3.14159265f;
            float off = bb * third;
            r[0] = -2.0f * sq * cosf(theta * third) - off;
            r[1] = -2.0f * sq * cosf((theta + two_pi) * third) - off;
            r[2] = -2.0f * sq * cosf((theta - two_pi) * third) - off;
            n = 3;
        } else {
            float ar = (rr > 0.f) ? -cbrt_ref(rr + sqrtf(r2 - q3)) : cbrt_ref(-rr + sqrtf(r2 - q3));
            float br = (absf_(ar) > eps) ? q / ar : 0.f;
            r[0] = (ar + br) - bb * third;
            n = 1;
        }
    }
    for (int i = 0; i < n; ++i) {
        float t = r[i];
        if (t >= 0.f && t <= 1.f) {
            float tt = 1.f - t;
            float x = (tt * tt * tt) * x0 + (3.0f * tt * tt * t) * x1 + (3.0f * tt * t * t) * x2 + (t * t * t) * x3;
            if (x > px) {
                float dy = 3.0f * a * t * t + 2.0f * b * t + c;
                w += (dy > 0.f) ? 1.f : -1.f;
            }
        }
    }
}

// Random cubic segments (control points within 64 px of a centre in a
// 128 px square) and points at a random height inside each segment's
// y-range: the cases the walk sends to the solver.  Counts the pairs whose
// winding differs between the record path and the libm formula.
__global__ void k_selftest_winding(uint32_t n, uint32_t seed, uint32_t *__restrict__ mismatches) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Pcg32 rng((uint64_t)i, (uint64_t)seed);
    float cx = 128.0f * rng.next_f32(), cy = 128.0f * rng.next_f32();
    float x[4], y[4];
    for (int k = 0; k < 4; ++k) {
        x[k] = cx + 128.0f * rng.next_f32() - 64.0f;
        y[k] = cy + 128.0f * rng.next_f32() - 64.0f;
    }
    float ymin = fminf(fminf(y[0], y[1]), fminf(y[2], y[3])), ymax = fmaxf(fmaxf(y[0], y[1]), fmaxf(y[2], y[3]));
    float px = cx + 128.0f * rng.next_f32() - 64.0f;
    float py = ymin + (ymax - ymin) * rng.next_f32();
    float rec[WIND_STRIDE];
    wind_coef_cubic(y[0], y[1], y[2], y[3], rec, x);
    float w0 = 0.f, w1 = 0.f, cr = 0.f;
    winding_cubic_libm(px, py, x[0], y[0], x[1], y[1], x[2], y[2], x[3], y[3], w0);
    winding_cubic_w<true>(px, py, x[0], y[0], x[1], x[2], x[3], rec, 0u, w1, cr);
    if (w0 != w1) atomicAdd(mismatches, 1u);
}

// ---- self-test of the backward's inside test ------------------------------
// Random segment / point pairs in four families (pair i is family i % 4):
// 0 random control points in a 128 px square, py in the segment's y-range;
// 1 near-quadratic cubics (|a| <= 0.25 against control points 128 px apart),
// the ill-conditioned case; 2 py within 0.05 px of a y-extremum of the curve,
// where two roots merge; 3 points within 0.01 px of the curve.  counts[f]
// receives the pairs of family f whose winding differs between
// winding_cubic_bwd and the f64 restatement although winding_cubic_bwd did
// not flag them as uncertain, counts[4 + f] the pairs it flagged, counts[8 + f]
// the pairs where the unchecked test (the default boundary pass's) differs.
__global__ void k_selftest_inside(uint32_t n, uint32_t seed, uint32_t *__restrict__ counts) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t fam = i & 3u;
    Pcg32 rng((uint64_t)i, (uint64_t)seed);
    float cx = 128.0f * rng.next_f32(), cy = 128.0f * rng.next_f32();
    float x[4], y[4];
    for (int k = 0; k < 4; ++k) {
        x[k] = cx + 128.0f * rng.next_f32() - 64.0f;
        y[k] = cy + 128.0f * rng.next_f32() - 64.0f;
    }
    if (fam == 1u) y[3] = y[0] - 3.0f * y[1] + 3.0f * y[2] + 0.5f * (rng.next_f32() - 0.5f);
    float ymin = fminf(fminf(y[0], y[1]), fminf(y[2], y[3])), ymax = fmaxf(fmaxf(y[0], y[1]), fmaxf(y[2], y[3]));
    float px = cx + 128.0f * rng.next_f32() - 64.0f;
    float py = ymin + (ymax - ymin) * rng.next_f32();
    float rec[WIND_STRIDE];
    wind_coef_cubic(y[0], y[1], y[2], y[3], rec, x);
    if (fam == 2u) {
        // a y-extremum of the curve: record slot 16 or 17 when inside (0, 1)
        float tc = rec[16] < 1.f ? rec[16] : (rec[17] < 1.f ? rec[17] : 0.5f);
        float tt = 1.f - tc;
        float yc = (tt * tt * tt) * y[0] + (3.0f * tt * tt * tc) * y[1] + (3.0f * tt * tc * tc) * y[2] +
                   (tc * tc * tc) * y[3];
        py = yc + 0.1f * (rng.next_f32() - 0.5f);
    } else if (fam == 3u) {
        float t = rng.next_f32(), tt = 1.f - t;
        float bx = (tt * tt * tt) * x[0] + (3.0f * tt * tt * t) * x[1] + (3.0f * tt * t * t) * x[2] + (t * t * t) * x[3];
        float by = (tt * tt * tt) * y[0] + (3.0f * tt * tt * t) * y[1] + (3.0f * tt * t * t) * y[2] + (t * t * t) * y[3];
        px = bx + 0.02f * (rng.next_f32() - 0.5f);
        py = by + 0.02f * (rng.next_f32() - 0.5f);
    }
    if (py < ymin || py > ymax) return;  // the walks cull those (no root either way)
    float w = 0.f, cr = 0.f, uns = 0.f;
    winding_cubic_bwd<true>(px, py, x[0], y[0], x[1], x[2], x[3], rec, 0u, w, cr, uns);
    int w64 = winding_cubic_f64(px, py, x, y);
    if (uns > 0.f) atomicAdd(counts + 4 + fam, 1u);
    else if ((int)w != w64) atomicAdd(counts + fam, 1u);
    float wd = 0.f, crd = 0.f, ud = 0.f;
    winding_cubic_bwd<false>(px, py, x[0], y[0], x[1], x[2], x[3], rec, 0u, wd, crd, ud);
    if ((int)wd != w64) atomicAdd(counts + 8 + fam, 1u);
}

// ---- signed distance (Renderer::render_sdf / eval_positions) ------------------
//
// The reference's query (compute_signed_distance, sdf.rs:9-117, over
// compute_distance_group, distance.rs:298-513) walks every group from the last
// to the first; in a group with a BVH it skips a node whose group-local box
// lies farther than the best canvas distance so far and measures every shape
// of a visited leaf.  Two walks give its result here:
//
//   sdf_walk       that walk itself, over the scene's group BVHs, for any
//                  scene (a node test in group-local units against a canvas
//                  distance is exact only when the group does not shrink;
//                  the walk keeps the reference's test, so it keeps the
//                  reference's result when a group does shrink);
//   k_sdf_tiles    for scenes whose groups do not shrink, where the walk's
//                  result is the smallest distance over all shapes (the
//                  highest group on a tie): per 16x16 tile, the shapes within
//                  a radius U of the tile (U = the distance at the tile
//                  centre + the half diagonal); a sample's minimum over them
//                  is the scene minimum whenever it is below U (every other
//                  shape is farther than U), else the sample takes sdf_walk.
//
// Both measure a shape as the reference does (same transforms, same solver)
// and skip only work that cannot change the result: shapes whose canvas box
// lies beyond the best distance, and path segments ruled out by lower bounds
// (control-point boxes, then SDF_PIECES stadiums around sub-curves).

DVG_DEV float box4_dist(kfloat *b, float px, float py) {  // INF for an empty box
    float x0 = b[0], y0 = b[1], x1 = b[2], y1 = b[3];
    if (!(x0 <= x1 && y0 <= y1)) return __builtin_inff();
    return box_dist(x0, y0, x1, y1, px, py);
}

// True when every stadium of curve cv is farther than `cut` from (x, y): the
// curve is then farther than cut.
DVG_DEV bool pieces_beyond(const SdfView &V, uint32_t cv, float x, float y, float cut) {
    kfloat *pc = V.piece + (size_t)cv * (SDF_PIECES * 4u);
    kfloat *ph = V.piece_h + (size_t)cv * (SDF_PIECES * 2u);
    for (uint32_t i = 0; i < SDF_PIECES; ++i) {
        float wx = x - pc[4u * i], wy = y - pc[4u * i + 1u], vx = pc[4u * i + 2u], vy = pc[4u * i + 3u];
        float t = clamp01((wx * vx + wy * vy) * ph[2u * i + 1u]);
        float dx = wx - t * vx, dy = wy - t * vy;
        float r = cut + ph[2u * i];
        if (dx * dx + dy * dy <= r * r) return false;
    }
    return true;
}

// closest_point_shape, distance.rs:235-296: the closest boundary point of
// prepared shape si to the shape-space point (px, py); false when a path has
// no segment.  A path segment is skipped when a lower bound of its distance
// exceeds the best segment distance so far (it cannot be the first nearest
// segment) or `gcut` (the shape is then farther than gcut whichever segment is
// nearest; gcut = INF when no such bound is known); the result is the full
// scan's whenever the nearest point lies within gcut.
DVG_DEV bool closest_point_shape_cut(const DevScene &S, const SdfView &V, uint32_t si, float px, float py,
                                     float gcut, float &ox, float &oy) {
    kfloat *sd = S.shape_data + si * 15u;
    uint32_t kind = (uint32_t)sd[0];
    float q0 = sd[4], q1 = sd[5], q2 = sd[6], q3 = sd[7];
    if (kind == DVG_SHAPE_CIRCLE) {  // closest_point_circle, distance.rs:33-49
        float r = absf_(q2);
        float dx = px - q0, dy = py - q1;
        float len = sqrtf(dx * dx + dy * dy);
        if (len <= 1.0e-8f) {
            ox = q0 + r;
            oy = q1;
        } else {
            float inv = r / len;
            ox = q0 + dx * inv;
            oy = q1 + dy * inv;
        }
        return true;
    }
    if (kind == DVG_SHAPE_ELLIPSE) {  // closest_point_ellipse, distance.rs:83-139
        ellipse_closest(q0, q1, q2, q3, px, py, ox, oy);
        return true;
    }
    if (kind == DVG_SHAPE_RECT) {  // closest_point_rect, distance.rs:52-80
        bool inside = px >= q0 && px <= q2 && py >= q1 && py <= q3;
        if (!inside) {
            ox = fminf(fmaxf(px, q0), q2);
            oy = fminf(fmaxf(py, q1), q3);
        } else {
            float dl = px - q0, dr = q2 - px, db = py - q1, dt = q3 - py;
            if (dl <= dr && dl <= db && dl <= dt) { ox = q0; oy = py; }
            else if (dr <= db && dr <= dt) { ox = q2; oy = py; }
            else if (db <= dt) { ox = px; oy = q1; }
            else { ox = px; oy = q3; }
        }
        return true;
    }
    // closest_point_path, distance.rs:142-231: the first segment at the
    // smallest distance, its point from the parameter.  The approximate
    // cubic distance is measured to a chord polygon, not to the curve, so no
    // lower bound of the curve's distance applies to it: such paths scan
    // every segment.
    uint32_t off = (uint32_t)sd[12], cnt = (uint32_t)sd[13];
    bool approx = sd[14] > 0.5f;
    float eps = 1.0e-6f * (absf_(px) + absf_(py)) + 1.0e-6f;  // rounding of the bounds
    float best = BIG;
    for (uint32_t s = 0; s < cnt; ++s) {
        uint32_t cvi = off + s;
        kfloat *cv = S.curve_data + cvi * 13u;
        uint32_t k = (uint32_t)cv[0];
        if (!approx) {
            float cut = fminf(best, gcut) * 1.00001f + eps;
            if (box4_dist(V.cvbox + cvi * 4u, px, py) > cut) continue;
            if (k != 0u && pieces_beyond(V, cvi, px, py, cut)) continue;
        }
        float x0 = cv[1], y0 = cv[2], x1 = cv[3], y1 = cv[4], x2 = cv[5], y2 = cv[6], x3 = cv[7], y3 = cv[8];
        float d, t;
        if (k == 0) seg_dist_t(px, py, x0, y0, x1, y1, d, t);
        else if (k == 1) quad_closest<true>(px, py, x0, y0, x1, y1, x2, y2, approx, d, t);
        else cubic_closest<true>(px, py, x0, y0, x1, y1, x2, y2, x3, y3, approx, d, t);
        if (d < best) {
            best = d;
            float tt = 1.0f - t;
            if (k == 0) {
                ox = x0 + t * (x1 - x0);
                oy = y0 + t * (y1 - y0);
            } else if (k == 1) {
                ox = tt * tt * x0 + 2.0f * tt * t * x1 + t * t * x2;
                oy = tt * tt * y0 + 2.0f * tt * t * y1 + t * t * y2;
            } else {
                ox = tt * tt * tt * x0 + 3.0f * tt * tt * t * x1 + 3.0f * tt * t * t * x2 + t * t * t * x3;
                oy = tt * tt * tt * y0 + 3.0f * tt * tt * t * y1 + 3.0f * tt * t * t * y2 + t * t * t * y3;
            }
        }
    }
    return best < BIG;
}

// The canvas distance from (px, py) to prepared shape si of group g as
// compute_distance_group measures it (distance.rs:346-413): group-local point
// (lx, ly), shape-space closest point, back through the shape and group
// transforms.  INF when the shape has no boundary, or when it provably lies
// farther than `best` (a similarity map of scale s turns best into the
// shape-space bound best / s).
DVG_DEV float sdf_shape_dist(const DevScene &S, const SdfView &V, uint32_t si, uint32_t g, float px, float py,
                             float lx, float ly, float best) {
    kfloat *xf = S.shape_xform + si * 6u;
    float spx = xf[0] * lx + xf[1] * ly + xf[2];
    float spy = xf[3] * lx + xf[4] * ly + xf[5];
    float s = V.sim[si];
    float gcut = s > 0.f ? (best * 1.00001f + 1.0e-3f) / s : __builtin_inff();
    float cx = 0.f, cy = 0.f;
    if (!closest_point_shape_cut(S, V, si, spx, spy, gcut, cx, cy)) return __builtin_inff();
    kfloat *T = S.shape_transform + S.prepared_to_scene[si] * 6u;
    float gx = T[0] * cx + T[1] * cy + T[2], gy = T[3] * cx + T[4] * cy + T[5];
    kfloat *gs = S.group_shape_xform + g * 6u;
    float ccx = gs[0] * gx + gs[1] * gy + gs[2], ccy = gs[3] * gx + gs[4] * gy + gs[5];
    float dx = ccx - px, dy = ccy - py;
    return sqrtf(dx * dx + dy * dy);
}

struct SdfBest {
    float dist;
    uint32_t group;
    bool found;
};

// Measures shape si of group g unless its canvas box lies beyond the best
// distance; the nearer shape wins, the higher group on a tie (the reference
// visits groups from the last, keeping the first strict improvement).
DVG_DEV void sdf_visit(const DevScene &S, const SdfView &V, uint32_t si, uint32_t g, float px, float py, float lx,
                       float ly, SdfBest &b) {
    if (box4_dist(V.cbox + si * 4u, px, py) > b.dist) return;
    float d = sdf_shape_dist(S, V, si, g, px, py, lx, ly, b.dist);
    if (d < b.dist || (d == b.dist && g > b.group)) {
        b.dist = d;
        b.group = g;
        b.found = true;
    }
}

// The reference's walk (compute_signed_distance's group loop with
// compute_distance_group's BVH traversal).  The BVHs were checked on the
// host: left and skip > node, leaves within the index range.  BVH = false
// scans every group's shapes instead: the same result when no group shrinks
// (no node the reference skips could hold a nearer shape).
template <bool BVH>
DVG_DEV SdfBest sdf_walk(const DevScene &S, const SdfView &V, float px, float py) {
    SdfBest b = {BIG, 0u, false};
    for (uint32_t gg = S.num_groups; gg > 0; --gg) {
        uint32_t g = gg - 1u;
        float lx, ly;
        group_local(S, g, px, py, lx, ly);
        kuint *meta = S.gbvh_meta + g * 4u;
        if (BVH && meta[1] > 0u && meta[3] > 0u) {
            uint32_t noff = meta[0], ioff = meta[2], node = 0;
            while (node != SDF_NONE) {
                kfloat *nb = S.gbvh_bounds + (size_t)(noff + node) * 4u;
                kuint *nd = S.gbvh_nodes + (size_t)(noff + node) * 4u;
                if (box_dist(nb[0], nb[1], nb[2], nb[3], lx, ly) > b.dist) {  // bounds_distance, distance.rs:9-15
                    node = nd[1];
                    continue;
                }
                uint32_t cnt = nd[3];
                if (cnt == 0u) {
                    node = nd[0];
                    continue;
                }
                for (uint32_t i = 0; i < cnt; ++i)
                    sdf_visit(S, V, S.gbvh_indices[ioff + nd[2] + i], g, px, py, lx, ly, b);
                node = nd[1];
            }
        } else {
            kfloat *gd = S.group_data + g * 16u;
            uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
            for (uint32_t i = 0; i < cnt; ++i) sdf_visit(S, V, (uint32_t)S.group_shapes[off + i], g, px, py, lx, ly, b);
        }
    }
    return b;
}

// Signed value of a walk's result: negative inside the fill of the nearest
// shape's group (fill_inside_group, backward.rs:1529-1606, which is the
// forward's fill test); 0 when no shape has a boundary.
DVG_DEV float sdf_signed(const DevScene &S, const SdfBest &b, float px, float py) {
    if (!b.found) return 0.f;
    kfloat *gd = S.group_data + b.group * 16u;
    uint32_t fk = (uint32_t)gd[2], rule = (uint32_t)gd[7];
    bool inside = false;
    if (fk != DVG_PAINT_NONE) {
        float lx, ly;
        group_local(S, b.group, px, py, lx, ly);
        Cov c = {BIG, 0.f, 0.f, BIG, 0.f, 0.f, 0.f};
        group_cov<COV_FWD>(S, b.group, gd, lx, ly, true, false, rule, false, c);
        inside = rule == 1u ? c.crossings > 0.f : c.winding != 0.f;
    }
    return inside ? -b.dist : b.dist;
}

// ---- work arrays, derived from the scene at every call -----------------------
// The group of every prepared shape (group_data's shape lists).
__global__ void k_sdf_prep_groups(DevScene S, uint32_t *__restrict__ pgroup) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= S.num_groups) return;
    kfloat *gd = S.group_data + g * 16u;
    uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
    for (uint32_t i = 0; i < cnt; ++i) pgroup[(uint32_t)S.group_shapes[off + i]] = g;
}

DVG_DEV float2 lerp2(float2 a, float2 b, float t) { return make_float2(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t); }

// Canvas boxes and similarity scales of the prepared shapes (after
// k_sdf_prep_groups), control-point boxes and stadiums of the curves.
__global__ void k_sdf_prep(DevScene S, SdfWork Wk, uint32_t ncurves) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S.num_prepared) {
        kfloat *sd = S.shape_data + i * 15u;
        uint32_t kind = (uint32_t)sd[0];
        float x0 = __builtin_inff(), y0 = __builtin_inff(), x1 = -__builtin_inff(), y1 = -__builtin_inff();
        if (kind == DVG_SHAPE_CIRCLE || kind == DVG_SHAPE_ELLIPSE) {
            float rx = absf_(sd[6]), ry = kind == DVG_SHAPE_CIRCLE ? rx : absf_(sd[7]);
            x0 = sd[4] - rx; x1 = sd[4] + rx; y0 = sd[5] - ry; y1 = sd[5] + ry;
        } else if (kind == DVG_SHAPE_RECT) {
            x0 = fminf(sd[4], sd[6]); x1 = fmaxf(sd[4], sd[6]); y0 = fminf(sd[5], sd[7]); y1 = fmaxf(sd[5], sd[7]);
        } else {
            uint32_t off = (uint32_t)sd[12], cnt = (uint32_t)sd[13];
            for (uint32_t s = 0; s < cnt; ++s) {
                kfloat *cv = S.curve_data + (off + s) * 13u;
                for (int j = 0; j < 4; ++j) {
                    x0 = fminf(x0, cv[1 + 2 * j]); x1 = fmaxf(x1, cv[1 + 2 * j]);
                    y0 = fminf(y0, cv[2 + 2 * j]); y1 = fmaxf(y1, cv[2 + 2 * j]);
                }
            }
        }
        float4 box = make_float4(1.f, 1.f, 0.f, 0.f);  // empty
        float sim = 0.f;
        uint32_t g = Wk.pgroup[i];
        if (g >= S.num_groups) g = 0;  // every prepared shape is in a group; stay in bounds regardless
        kfloat *T = S.shape_transform + S.prepared_to_scene[i] * 6u;
        kfloat *G = S.group_shape_xform + g * 6u;
        if (x0 <= x1 && y0 <= y1) {
            float cx0 = __builtin_inff(), cy0 = __builtin_inff(), cx1 = -__builtin_inff(), cy1 = -__builtin_inff();
            for (int c = 0; c < 4; ++c) {
                float ux = (c & 1) ? x1 : x0, uy = (c & 2) ? y1 : y0;
                float gx = T[0] * ux + T[1] * uy + T[2], gy = T[3] * ux + T[4] * uy + T[5];
                float vx = G[0] * gx + G[1] * gy + G[2], vy = G[3] * gx + G[4] * gy + G[5];
                cx0 = fminf(cx0, vx); cx1 = fmaxf(cx1, vx); cy0 = fminf(cy0, vy); cy1 = fmaxf(cy1, vy);
            }
            float mag = fmaxf(fmaxf(absf_(cx0), absf_(cx1)), fmaxf(absf_(cy0), absf_(cy1)));
            float pad = 1.0e-5f * mag + 1.0e-4f;
            box = make_float4(cx0 - pad, cy0 - pad, cx1 + pad, cy1 + pad);
        }
        // shape-to-canvas linear map A = G T: a similarity when its columns
        // are orthogonal and of one length
        float a = G[0] * T[0] + G[1] * T[3], b = G[0] * T[1] + G[1] * T[4];
        float c = G[3] * T[0] + G[4] * T[3], d = G[3] * T[1] + G[4] * T[4];
        float n = absf_(a) + absf_(b) + absf_(c) + absf_(d), tol = 1.0e-6f * n;
        bool rot = absf_(a - d) <= tol && absf_(b + c) <= tol, refl = absf_(a + d) <= tol && absf_(b - c) <= tol;
        bool approx = kind == DVG_SHAPE_PATH && sd[14] > 0.5f;
        if ((rot || refl) && !approx) sim = sqrtf(a * a + c * c);
        reinterpret_cast<float4 *>(Wk.cbox)[i] = box;
        Wk.sim[i] = sim;
    }
    if (i < ncurves) {
        kfloat *cv = S.curve_data + (size_t)i * 13u;
        uint32_t k = (uint32_t)cv[0];
        float2 p[4];
        for (int j = 0; j < 4; ++j) p[j] = make_float2(cv[1 + 2 * j], cv[2 + 2 * j]);
        float4 bx = make_float4(p[0].x, p[0].y, p[0].x, p[0].y);
        for (int j = 1; j < 4; ++j)
            bx = make_float4(fminf(bx.x, p[j].x), fminf(bx.y, p[j].y), fmaxf(bx.z, p[j].x), fmaxf(bx.w, p[j].y));
        reinterpret_cast<float4 *>(Wk.cvbox)[i] = bx;
        float4 *pc = reinterpret_cast<float4 *>(Wk.piece) + (size_t)i * SDF_PIECES;
        float2 *ph = reinterpret_cast<float2 *>(Wk.piece_h) + (size_t)i * SDF_PIECES;
        for (uint32_t q = 0; q < SDF_PIECES; ++q) {
            if (k == 0u) {  // lines are measured directly
                pc[q] = make_float4(0.f, 0.f, 0.f, 0.f);
                ph[q] = make_float2(0.f, 0.f);
                continue;
            }
            // control points of the sub-curve on [ta, tb]: blossoms
            // (ta, .., ta, tb, .., tb), by de Casteljau steps
            float ta = (float)q / (float)SDF_PIECES, tb = (float)(q + 1u) / (float)SDF_PIECES;
            int deg = k == 1u ? 2 : 3;
            float2 Q[4];
            for (int m = 0; m <= deg; ++m) {
                float2 w[4] = {p[0], p[1], p[2], p[3]};
                for (int lev = 0; lev < deg; ++lev) {
                    float tt = lev < deg - m ? ta : tb;
                    for (int j = 0; j < deg - lev; ++j) w[j] = lerp2(w[j], w[j + 1], tt);
                }
                Q[m] = w[0];
            }
            // stadium: the chord line, the control points' extent along it
            // and their largest distance from it
            float2 u = make_float2(Q[deg].x - Q[0].x, Q[deg].y - Q[0].y);
            float len = sqrtf(u.x * u.x + u.y * u.y);
            float mag = 0.f;
            for (int m = 0; m <= deg; ++m) mag = fmaxf(mag, fmaxf(absf_(Q[m].x), absf_(Q[m].y)));
            float pad = 1.0e-5f * mag + 1.0e-5f;
            if (len > 1.0e-6f * (mag + 1.0f)) {
                u.x /= len; u.y /= len;
                float smin = 0.f, smax = 0.f, h = 0.f;
                for (int m = 1; m <= deg; ++m) {
                    float wx = Q[m].x - Q[0].x, wy = Q[m].y - Q[0].y;
                    float sp = wx * u.x + wy * u.y;
                    smin = fminf(smin, sp); smax = fmaxf(smax, sp);
                    h = fmaxf(h, absf_(wx * u.y - wy * u.x));
                }
                float ex = Q[0].x + smin * u.x, ey = Q[0].y + smin * u.y;
                float vx = (smax - smin) * u.x, vy = (smax - smin) * u.y;
                float vv = vx * vx + vy * vy;
                pc[q] = make_float4(ex, ey, vx, vy);
                ph[q] = make_float2(h + pad, vv > 0.f ? 1.0f / vv : 0.f);
            } else {  // a point-like piece: a disc around Q0
                float h = 0.f;
                for (int m = 1; m <= deg; ++m)
                    h = fmaxf(h, sqrtf((Q[m].x - Q[0].x) * (Q[m].x - Q[0].x) + (Q[m].y - Q[0].y) * (Q[m].y - Q[0].y)));
                pc[q] = make_float4(Q[0].x, Q[0].y, 0.f, 0.f);
                ph[q] = make_float2(h + pad, 0.f);
            }
        }
    }
}

// One wave per tile: the nearest shape to the tile centre (lanes over the
// shapes, then a wave minimum), the candidate radius U, and the shapes whose
// canvas box lies within U of the tile, in prepared order (a hit's position is
// the count of hits on lower lanes).  Lists longer than SDF_LCAP are cut; the
// count is kept, and such tiles take the walk.
constexpr uint32_t SDF_LIST_TILES_PER_WG = BLOCK / 64;
__global__ __launch_bounds__(BLOCK) void k_sdf_lists(DevScene S, SdfView V, SdfWork Wk) {
    uint32_t t = blockIdx.x * SDF_LIST_TILES_PER_WG + threadIdx.x / 64u, lane = threadIdx.x & 63u;
    uint32_t ntiles = S.tiles_x * S.tiles_y;
    if (t >= ntiles) return;  // the whole wave
    float x0 = (float)((t % S.tiles_x) * (uint32_t)TILE), y0 = (float)((t / S.tiles_x) * (uint32_t)TILE);
    float cx = x0 + 0.5f * (float)TILE, cy = y0 + 0.5f * (float)TILE;
    float lb = BIG;
    uint32_t la = SDF_NONE;
    for (uint32_t si = lane; si < S.num_prepared; si += 64u) {
        if (box4_dist(V.cbox + si * 4u, cx, cy) > lb) continue;
        uint32_t g = V.pgroup[si];
        float lx, ly;
        group_local(S, g, cx, cy, lx, ly);
        float d = sdf_shape_dist(S, V, si, g, cx, cy, lx, ly, lb);
        if (d < lb) {
            lb = d;
            la = si;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        float ob = __shfl_xor(lb, o);
        uint32_t oa = __shfl_xor(la, o);
        if (ob < lb || (ob == lb && oa < la)) {
            lb = ob;
            la = oa;
        }
    }
    // a sample of the tile lies within the half diagonal of the centre
    float U = lb < BIG ? lb * 1.0001f + 0.5f * 1.41422f * (float)TILE + 0.1f : __builtin_inff();
    float x1 = x0 + (float)TILE, y1 = y0 + (float)TILE;
    uint32_t count = 0;
    uint32_t *list = Wk.tile_list + (size_t)t * SDF_LCAP;
    for (uint32_t base = 0; base < S.num_prepared; base += 64u) {
        uint32_t si = base + lane;
        bool hit = false;
        if (si < S.num_prepared) {
            kfloat *b = V.cbox + si * 4u;
            if (b[0] <= b[2] && b[1] <= b[3]) {
                float dx = fmaxf(fmaxf(b[0] - x1, 0.f), x0 - b[2]);
                float dy = fmaxf(fmaxf(b[1] - y1, 0.f), y0 - b[3]);
                hit = sqrtf(dx * dx + dy * dy) <= U;
            }
        }
        uint64_t m = __ballot(hit);
        if (hit) {
            uint32_t pos = count + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (pos < SDF_LCAP) list[pos] = si;
        }
        count += (uint32_t)__popcll(m);
    }
    if (lane == 0u) {
        Wk.tile_u[t] = U;
        Wk.tile_seed[t] = la;
        Wk.tile_count[t] = count;
    }
}

// render_sdf_kernel, sdf.rs:119-200, for scenes whose groups do not shrink:
// one workgroup a tile, one pixel a lane, the mean signed distance of its
// samples (the forward's sample positions).  A sample measures the tile's
// candidates, the one nearest the centre first; a minimum below the tile's U
// is the scene minimum (every shape off the list lies farther than U), else
// (or for a cut list) the sample takes the walk.
__global__ __launch_bounds__(256) void k_sdf_tiles(DevScene S, SdfView V, float *__restrict__ out) {
    uint32_t t = blockIdx.x;
    uint32_t x = (t % S.tiles_x) * (uint32_t)TILE + threadIdx.x % (uint32_t)TILE;
    uint32_t y = (t / S.tiles_x) * (uint32_t)TILE + threadIdx.x / (uint32_t)TILE;
    if (x >= S.W || y >= S.H) return;
    uint32_t n = V.tile_count[t], seed = V.tile_seed[t];
    float U = V.tile_u[t];
    bool listed = n <= SDF_LCAP;
    kuint *list = V.tile_list + (size_t)t * SDF_LCAP;
    float acc = 0.f;
    for (uint32_t sy = 0; sy < S.nsy; ++sy)
        for (uint32_t sx = 0; sx < S.nsx; ++sx) {
            float px, py;
            sample_point<false>(S, x, y, sx, sy, px, py);
            SdfBest b = {BIG, 0u, false};
            if (listed) {
                if (seed != SDF_NONE) {
                    uint32_t g = V.pgroup[seed];
                    float lx, ly;
                    group_local(S, g, px, py, lx, ly);
                    sdf_visit(S, V, seed, g, px, py, lx, ly, b);
                }
                for (uint32_t k = 0; k < n; ++k) {
                    uint32_t si = list[k];
                    if (si == seed || box4_dist(V.cbox + si * 4u, px, py) > b.dist) continue;
                    uint32_t g = V.pgroup[si];
                    float lx, ly;
                    group_local(S, g, px, py, lx, ly);
                    sdf_visit(S, V, si, g, px, py, lx, ly, b);
                }
            }
            if (!listed || !(b.found && b.dist < U)) b = sdf_walk<false>(S, V, px, py);
            acc += sdf_signed(S, b, px, py);
        }
    out[(size_t)y * S.W + x] = acc * (1.0f / (float)(S.nsx * S.nsy));
}

// render_sdf_kernel by the walk (scenes with a shrinking group): one pixel a
// lane.
__global__ __launch_bounds__(256) void k_sdf_walk(DevScene S, SdfView V, float *__restrict__ out) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= S.W * S.H) return;
    uint32_t x = i % S.W, y = i / S.W;
    float acc = 0.f;
    for (uint32_t sy = 0; sy < S.nsy; ++sy)
        for (uint32_t sx = 0; sx < S.nsx; ++sx) {
            float px, py;
            sample_point<false>(S, x, y, sx, sy, px, py);
            acc += sdf_signed(S, sdf_walk<true>(S, V, px, py), px, py);
        }
    out[i] = acc * (1.0f / (float)(S.nsx * S.nsy));
}

// eval_positions_kernel, sdf.rs:207-262: the signed distance at n canvas
// points (x, y pairs), by the walk.
__global__ __launch_bounds__(256) void k_sdf_points(DevScene S, SdfView V, const float *__restrict__ pos, uint32_t n,
                                                    float *__restrict__ out) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float px = pos[2u * i], py = pos[2u * i + 1u];
    out[i] = S.sdf_tiled ? sdf_signed(S, sdf_walk<false>(S, V, px, py), px, py)
                         : sdf_signed(S, sdf_walk<true>(S, V, px, py), px, py);
}

// sample_distance (backward/sampling.rs:14-68) for the adjoint d_dist of the
// signed distance at (px, py): the nearest boundary over the groups from last
// to first (the CPU distance query), d_compute_distance of it with the sign of
// the group's CPU fill test; -d(pt) into translation pixel pix (< 0: none).
DVG_DEV void sdf_point_grad(const DevScene &S, const DevGrads &G, float px, float py, float d_dist, long pix) {
    float best = __builtin_inff();
    uint32_t best_g = 0;
    bool found = false;
    PHit bh = {};
    for (uint32_t gg = S.num_groups; gg > 0; --gg) {
        PHit h;
        if (group_closest(S, gg - 1u, px, py, best, h) && h.dist < best) {
            best = h.dist;
            best_g = gg - 1u;
            bh = h;
            found = true;
        }
    }
    if (!found) return;
    const float *gd = S.group_data + best_g * 16u;
    bool inside = (uint32_t)gd[2] != DVG_PAINT_NONE && group_inside(S, best_g, px, py);
    float dtx = 0.f, dty = 0.f;
    d_distance_query(S, G, best_g, bh, px, py, inside ? -d_dist : d_dist, dtx, dty);
    if (G.d_translation && pix >= 0) {
        unsafeAtomicAdd(G.d_translation + (size_t)pix * 2u, dtx);
        unsafeAtomicAdd(G.d_translation + (size_t)pix * 2u + 1u, dty);
    }
}

// render_backward_positions (backward/render.rs:213-280), one point a lane;
// translation into the pixel holding the point (coordinates truncated toward
// zero, as `as i32`).
__global__ __launch_bounds__(256) void k_sdf_points_backward(DevScene S, const float *__restrict__ pos, uint32_t n,
                                                             const float *__restrict__ d_sdf, DevGrads G) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float d_dist = d_sdf[i];
    if (d_dist == 0.f) return;
    float px = pos[2u * i], py = pos[2u * i + 1u];
    float tx = truncf(px), ty = truncf(py);
    long pix = (tx >= 0.f && ty >= 0.f && tx < (float)S.W && ty < (float)S.H) ? (long)ty * (long)S.W + (long)tx : -1;
    sdf_point_grad(S, G, px, py, d_dist, pix);
}

// The d_sdf image term of render_backward (backward/render.rs:171-186), one
// sample a lane: sample_distance at the CPU sample position with the pixel's
// d_sdf; translation into that pixel.
__global__ __launch_bounds__(256) void k_sdf_image_backward(DevScene S, const float *__restrict__ d_sdf, DevGrads G) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= S.W * S.H * S.spp) return;
    uint32_t sx = i % S.nsx, sy = (i / S.nsx) % S.nsy, p = i / S.spp;
    float d_dist = d_sdf[p];
    if (d_dist == 0.f) return;
    float px, py;
    sample_point<true>(S, p % S.W, p / S.W, sx, sy, px, py);
    sdf_point_grad(S, G, px, py, d_dist, (long)p);
}

// Winding terms of every segment (the record described with WIND_STRIDE).
__global__ void k_wind_coef(const float *__restrict__ curve_data, uint32_t ncurves, float *__restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncurves) return;
    const float *cv = curve_data + (size_t)i * 13u;
    float *o = out + (size_t)i * WIND_STRIDE;
    uint32_t k = (uint32_t)cv[0];
    if (k == 2u) {
        const float xs[4] = {cv[1], cv[3], cv[5], cv[7]};
        wind_coef_cubic(cv[2], cv[4], cv[6], cv[8], o, xs);
    } else if (k == 1u) {
        wind_coef_quad(cv[2], cv[4], cv[6], o);
    } else {
        for (int j = 0; j < WIND_STRIDE; ++j) o[j] = 0.f;
    }
    float x0 = cv[1], y0 = cv[2], x1 = cv[3], y1 = cv[4], x2 = cv[5], y2 = cv[6], x3 = cv[7], y3 = cv[8];
    // the ray test's box (slots 12..15: y min, y max, x max, x min): the
    // curve's own box over t in [0, 1], widened, within the box of the four
    // stored points.  Widening: the f32 evaluation of a crossing's x and the
    // f32 coefficients of the CPU's polynomial (1e-6 of the coordinates, plus
    // 1e-5 px); on cubics also the f32 solve of the forward, whose roots are
    // off by up to ~1e-5 (2 sqrt|q| + |b / 3a| + 1) in t near the ends, and
    // whose three-root branch can wrongly open within ~2e-7 |a| (2 |b/a|^3 +
    // 9 |b c| / a^2 + 27 |d / a|) of a y-extremum (ten times both).  Segments
    // with an ill-conditioned solve (record bit 4) keep the stored points'
    // box in y.  A point beyond the box then meets no crossing in any of the
    // solves, so the box test skips exactly segments that add nothing.
    float cx0 = fminf(fminf(x0, x1), fminf(x2, x3)), cx1 = fmaxf(fmaxf(x0, x1), fmaxf(x2, x3));
    float cy0 = fminf(fminf(y0, y1), fminf(y2, y3)), cy1 = fmaxf(fmaxf(y0, y1), fmaxf(y2, y3));
    int np = k == 2u ? 4 : (k == 1u ? 3 : 2);
    double xs[4] = {x0, x1, x2, x3}, ys[4] = {y0, y1, y2, y3};
    double xl, xh, yl, yh, sx = 0.0, sy = 0.0;
    bezier_range_f64(xs, np, xl, xh);
    bezier_range_f64(ys, np, yl, yh);
    for (int j = 0; j < np; ++j) {
        sx += fabs(xs[j]);
        sy += fabs(ys[j]);
    }
    double mx = 1.0e-6 * sx + 1.0e-5, my = 1.0e-6 * sy + 1.0e-5;
    bool keep_y = false;
    if (k == 2u) {
        double a = o[0], b = o[1], c = o[2];
        if (((uint32_t)o[11] & 16u) || !(fabs(a) > 0.0)) {
            keep_y = true;
        } else {
            double ba = fabs(b / a);
            double et = 1.0e-4 * (2.0 * sqrt(fabs((double)o[3])) + ba / 3.0 + 1.0);
            my += et * 3.0 * (fabs(a) + fabs(b) + fabs(c)) +
                  2.0e-6 * fabs(a) * (2.0 * ba * ba * ba + 9.0 * fabs(b * c) / (a * a) + 27.0 * (sy + fabs(y0)) / fabs(a));
        }
    } else if (k == 1u) {
        my += 1.0e-4 * 4.0 * sy;  // the forward's f32 quadratic, near its extremum
    }
    o[12] = keep_y ? cy0 : fmaxf(cy0, (float)(yl - my));
    o[13] = keep_y ? cy1 : fminf(cy1, (float)(yh + my));
    o[14] = fminf(cx1, (float)(xh + mx));
    o[15] = fmaxf(cx0, (float)(xl - mx));
}

// Slot 31 of a path's first segment record (after k_wind_coef): 1 when the
// reference CPU's winding of the path (distance/winding.rs:118-192) is the
// winding number of its curve, which is constant off the curve: the segments
// form one closed loop (each begins where the previous ends, the last ends
// where the first begins, coordinates finite) and no cubic takes the CPU's
// quadratic branch with a near-zero quadratic term (distance/utils.rs:44-89:
// |a| < 1e-6 drops the cubic term, and with b near zero the linear root is
// returned twice, or the dropped term's root can fall inside [0, 1]).
__global__ void k_shape_topo(const float *__restrict__ shape_data, uint32_t nprep, const float *__restrict__ curve_data,
                             uint32_t ncurves, float *__restrict__ curve_wind) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprep) return;
    const float *sd = shape_data + (size_t)p * 15u;
    if ((uint32_t)sd[0] != DVG_SHAPE_PATH) return;
    uint32_t off = (uint32_t)sd[12], cnt = (uint32_t)sd[13];
    if (cnt == 0u || off >= ncurves || cnt > ncurves - off) return;
    bool ok = true;
    float sx = curve_data[(size_t)off * 13u + 1], sy = curve_data[(size_t)off * 13u + 2], ex = sx, ey = sy;
    for (uint32_t s = 0; s < cnt; ++s) {
        const float *cv = curve_data + (size_t)(off + s) * 13u;
        uint32_t k = (uint32_t)cv[0];
        if (cv[1] != ex || cv[2] != ey) ok = false;
        for (int j = 1; j < 9; ++j)
            if (!isfinite(cv[j])) ok = false;
        if (k == 2u) {
            float a = -cv[2] + 3.0f * cv[4] - 3.0f * cv[6] + cv[8];
            float b = 3.0f * cv[2] - 6.0f * cv[4] + 3.0f * cv[6];
            if (!(fabsf(a) >= 1.0e-6f) && !(fabsf(b) >= 1.0e-4f)) ok = false;
        }
        ex = k == 0u ? cv[3] : (k == 1u ? cv[5] : cv[7]);
        ey = k == 0u ? cv[4] : (k == 1u ? cv[6] : cv[8]);
    }
    if (ex != sx || ey != sy) ok = false;
    curve_wind[(size_t)off * WIND_STRIDE + 31] = ok ? 1.f : 0.f;
}

// ---- tile-list culling ----------------------------------------------------------
// A tile list holds every group whose padded box meets the tile.  A group
// whose curves all miss the tile's padded region and whose fill is outside
// there adds nothing to the backward's inside tests anywhere in the region:
// the winding number of a closed loop is constant off its curve, and those
// tests find the winding number (f64, or f32 where f64 would decide the
// same) -- the f64 winding at the region's centre.  The backward's lists
// (bwd_entries, from tile_offsets[t] to tile_end[t], for the colour adjoint,
// the boundary pass and the prefiltered adjoint) go without such entries.
// The forward's f32 solve counts crossings of ill-conditioned segments (record
// bit 4) and quadratic segments wrongly over whole bands of rows far from
// their curve, and the images follow it there; on lines and well-conditioned
// cubics it can go wrong only within ~1e-4 px of a segment end's y.  So the
// forward's lists (the full lists compacted in place, to fwd_end[t]) drop a
// filled group only when every segment of it is a line or a
// well-conditioned cubic.  A stroke adds nothing where its curve lies farther than the
// segment's radius (record radii) from every point of the region: the
// distance to any point of the curve is at least the least one, so no test
// finds a hit there (prefiltered coverage saturates one pixel beyond: the
// region of a prefiltered scene is widened by that band).  Kept in both:
// every group with a shape other than a path, or with a fill and a path
// that is not a closed loop (record slot 31 of its first segment).

// Whether a Bezier segment (kind k: 0 line, 1 quadratic, 2 cubic) may meet
// the box: false when the control box of each of its 32 equal-t pieces
// misses it (a piece on [t0, t1] has control points P(t0), P(t0) + h P'(t0)
// / 3, P(t1) - h P'(t1) / 3, P(t1), h = t1 - t0, exact for degree <= 3).
DVG_DEV bool bezier_meets_box(uint32_t k, const float *xs, const float *ys, float bx0, float by0, float bx1,
                              float by1) {
    // power basis P(t) = p0 + c1 t + c2 t^2 + c3 t^3
    float cx1, cx2 = 0.f, cx3 = 0.f, cy1, cy2 = 0.f, cy3 = 0.f;
    if (k == 0u) {
        cx1 = xs[1] - xs[0]; cy1 = ys[1] - ys[0];
    } else if (k == 1u) {
        cx1 = 2.0f * (xs[1] - xs[0]); cx2 = xs[0] - 2.0f * xs[1] + xs[2];
        cy1 = 2.0f * (ys[1] - ys[0]); cy2 = ys[0] - 2.0f * ys[1] + ys[2];
    } else {
        cx1 = 3.0f * (xs[1] - xs[0]); cx2 = 3.0f * (xs[0] - 2.0f * xs[1] + xs[2]);
        cx3 = -xs[0] + 3.0f * (xs[1] - xs[2]) + xs[3];
        cy1 = 3.0f * (ys[1] - ys[0]); cy2 = 3.0f * (ys[0] - 2.0f * ys[1] + ys[2]);
        cy3 = -ys[0] + 3.0f * (ys[1] - ys[2]) + ys[3];
    }
    constexpr int N = 32;
    const float h = 1.0f / (float)N, h3 = h / 3.0f;
    float x0 = xs[0], y0 = ys[0], dx0 = cx1, dy0 = cy1;
    for (int i = 1; i <= N; ++i) {
        float t = (float)i * h;
        float x1 = xs[0] + ((cx3 * t + cx2) * t + cx1) * t, y1 = ys[0] + ((cy3 * t + cy2) * t + cy1) * t;
        float dx1 = (3.0f * cx3 * t + 2.0f * cx2) * t + cx1, dy1 = (3.0f * cy3 * t + 2.0f * cy2) * t + cy1;
        float qx1 = x0 + h3 * dx0, qx2 = x1 - h3 * dx1, qy1 = y0 + h3 * dy0, qy2 = y1 - h3 * dy1;
        float mnx = fminf(fminf(x0, x1), fminf(qx1, qx2)), mxx = fmaxf(fmaxf(x0, x1), fmaxf(qx1, qx2));
        float mny = fminf(fminf(y0, y1), fminf(qy1, qy2)), mxy = fmaxf(fmaxf(y0, y1), fmaxf(qy1, qy2));
        if (!(mnx > bx1 || mxx < bx0 || mny > by1 || mxy < by0)) return true;
        x0 = x1; y0 = y1; dx0 = dx1; dy0 = dy1;
    }
    return false;
}

// 0 keep; 1 drop from the backward's list; 3 drop from the forward's too.
DVG_DEV uint32_t tile_entry_drops(const DevScene &S, uint32_t t, uint32_t g) {
    const float *gd = S.group_data + g * 16u;
    uint32_t fk = (uint32_t)gd[2], sk = (uint32_t)gd[4], rule = (uint32_t)gd[7];
    bool fill = fk != DVG_PAINT_NONE, stroke = sk != DVG_PAINT_NONE;
    if (!fill && !stroke) return 3u;  // no paint: every walk skips it
    // the region: the tile padded as binning pads group boxes, plus a pixel
    // (and the prefiltered coverage band); on the canvas border 4 px + 2e-4
    // of the canvas further out (points clamped into border tiles from
    // beyond the canvas)
    uint32_t tx = t % S.tiles_x, ty = t / S.tiles_x;
    float pad = tile_pad(S.W, S.H) + 1.0f + (S.prefilter ? 1.5f : 0.f);
    float far = pad + 4.0f + 2.0e-4f * (float)(S.W > S.H ? S.W : S.H);
    float rx0 = (float)(tx * TILE) - (tx == 0u ? far : pad);
    float rx1 = (float)((tx + 1u) * TILE) + (tx + 1u == S.tiles_x ? far : pad);
    float ry0 = (float)(ty * TILE) - (ty == 0u ? far : pad);
    float ry1 = (float)((ty + 1u) * TILE) + (ty + 1u == S.tiles_y ? far : pad);
    float qx[4] = {rx0, rx1, rx0, rx1}, qy[4] = {ry0, ry0, ry1, ry1};
    float lx[4], ly[4], lcx, lcy;
    for (int c = 0; c < 4; ++c) group_local(S, g, qx[c], qy[c], lx[c], ly[c]);
    group_local(S, g, 0.5f * (rx0 + rx1), 0.5f * (ry0 + ry1), lcx, lcy);
    uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
    float wsum = 0.f, cr = 0.f;
    bool fwd = true;  // no fill winding the forward's f32 solve could get wrong far away
    for (uint32_t i = 0; i < cnt; ++i) {
        uint32_t si = (uint32_t)S.group_shapes[off + i];
        const float *sd = S.shape_data + si * 15u;
        if ((uint32_t)sd[0] != DVG_SHAPE_PATH) return 0u;
        uint32_t co = (uint32_t)sd[12], cc = (uint32_t)sd[13];
        if (cc == 0u) continue;  // no segment: no winding, no stroke
        if (fill && !(S.curve_wind[(size_t)co * WIND_STRIDE + 31] > 0.f)) return 0u;
        // the region's box in shape space, widened for rounding
        const float *xf = S.shape_xform + si * 6u;
        float sx0 = __builtin_inff(), sy0 = __builtin_inff(), sx1 = -__builtin_inff(), sy1 = -__builtin_inff();
        for (int c = 0; c < 4; ++c) {
            float x = xf[0] * lx[c] + xf[1] * ly[c] + xf[2], y = xf[3] * lx[c] + xf[4] * ly[c] + xf[5];
            sx0 = fminf(sx0, x); sx1 = fmaxf(sx1, x); sy0 = fminf(sy0, y); sy1 = fmaxf(sy1, y);
        }
        float m = 1.0e-4f * fmaxf(fmaxf(fabsf(sx0), fabsf(sx1)), fmaxf(fabsf(sy0), fabsf(sy1))) + 1.0e-3f;
        sx0 -= m; sy0 -= m; sx1 += m; sy1 += m;
        if (!(sx0 <= sx1 && sy0 <= sy1)) return 0u;  // non-finite transform
        float scx = xf[0] * lcx + xf[1] * lcy + xf[2], scy = xf[3] * lcx + xf[4] * lcy + xf[5];
        float lw = 0.f, lc = 0.f;
        for (uint32_t s = 0; s < cc; ++s) {
            const float *cv = S.curve_data + (size_t)(co + s) * 13u;
            const float *wc = S.curve_wind + (size_t)(co + s) * WIND_STRIDE;
            uint32_t k = (uint32_t)cv[0];
            if (fill && (k == 1u || (k == 2u && ((uint32_t)wc[11] & 16u)))) fwd = false;
            float xs[4] = {cv[1], cv[3], cv[5], cv[7]}, ys[4] = {cv[2], cv[4], cv[6], cv[8]};
            // a stroke reaches as far as the segment's largest radius
            float rr = 0.f;
            if (stroke) {
                rr = fmaxf(fmaxf(fabsf(cv[9]), fabsf(cv[10])), fmaxf(fabsf(cv[11]), fabsf(cv[12])));
                if (!(rr < 1.0e30f)) return 0u;
                rr += 1.0e-4f * rr + 1.0e-3f;
            }
            float bx0 = sx0 - rr, by0 = sy0 - rr, bx1 = sx1 + rr, by1 = sy1 + rr;
            bool near = !(wc[15] > bx1 || wc[14] < bx0 || wc[12] > by1 || wc[13] < by0);
            if (near && bezier_meets_box(k, xs, ys, bx0, by0, bx1, by1)) return 0u;
            if (!fill) continue;
            if (k == 0u) {
                winding_line_cpu(scx, scy, xs[0], ys[0], xs[1], ys[1], rule, lw, lc);
            } else {
                int n = (k == 1u) ? winding_quad_f64(scx, scy, xs, ys) : winding_cubic_f64(scx, scy, xs, ys);
                if (rule == 1u) {
                    if (n & 1) lc = 1.f - lc;
                } else {
                    lw += (float)n;
                }
            }
        }
        if (rule == 1u) {
            if (lc > 0.f) cr = 1.f - cr;
        } else {
            wsum += lw;
        }
    }
    if (fill && ((rule == 1u) ? cr > 0.f : wsum != 0.f)) return 0u;
    return fwd ? 3u : 1u;
}

// Pass 1, one lane per entry of the full lists: out[e] = the drop bits of
// entry e (tile_entry_drops).  Pass 2, one wave per tile, 64 entries at a
// time: an entry kept in a list goes to the count of entries kept in it
// before it (the lists keep their group order; a write never passes a flag
// or an entry not yet read).
__global__ __launch_bounds__(BLOCK) void k_tile_keep(DevScene S, const uint32_t *__restrict__ entries,
                                                     uint32_t *__restrict__ out, uint32_t cull) {
    uint32_t ntiles = S.tiles_x * S.tiles_y;
    uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= S.tile_offsets[ntiles]) return;
    // the tile holding entry e: the last t with tile_offsets[t] <= e
    uint32_t lo = 0, hi = ntiles;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (S.tile_offsets[mid] <= e) lo = mid;
        else hi = mid;
    }
    out[e] = cull ? tile_entry_drops(S, lo, entries[e]) : 0u;
}

__global__ __launch_bounds__(BLOCK) void k_tile_compact(DevScene S, uint32_t *__restrict__ entries,
                                                        uint32_t *__restrict__ out, uint32_t *__restrict__ tile_end,
                                                        uint32_t *__restrict__ fwd_end) {
    uint32_t t = blockIdx.x * BIN_TILES_PER_WG + threadIdx.x / 64u, lane = threadIdx.x & 63u;
    if (t >= S.tiles_x * S.tiles_y) return;  // (whole waves)
    uint32_t lb = S.tile_offsets[t], le = S.tile_offsets[t + 1];
    uint32_t wb = lb, wf = lb;
    for (uint32_t e0 = lb; e0 < le; e0 += 64u) {
        uint32_t e = e0 + lane;
        uint32_t drop = e < le ? out[e] : 3u;
        uint32_t g = e < le ? entries[e] : 0u;
        bool kb = !(drop & 1u), kf = !(drop & 2u);
        uint64_t mb = __ballot(kb), mf = __ballot(kf);
        uint64_t below = (1ull << lane) - 1ull;
        if (kb) out[wb + (uint32_t)__popcll(mb & below)] = g;
        if (kf) entries[wf + (uint32_t)__popcll(mf & below)] = g;
        wb += (uint32_t)__popcll(mb);
        wf += (uint32_t)__popcll(mf);
    }
    if (lane == 0u) {
        tile_end[t] = wb;
        fwd_end[t] = wf;
    }
}

void launch_tile_cull(const DevScene &S, uint32_t *entries, uint32_t *out, uint32_t *tile_end, uint32_t *fwd_end,
                      uint32_t cap, bool cull, hipStream_t stream) {
    uint32_t ntiles = S.tiles_x * S.tiles_y;
    if (ntiles == 0) return;
    if (cap > 0)
        hipLaunchKernelGGL(k_tile_keep, dim3((cap + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, S, entries, out,
                           cull ? 1u : 0u);
    hipLaunchKernelGGL(k_tile_compact, dim3((ntiles + BIN_TILES_PER_WG - 1) / BIN_TILES_PER_WG), dim3(BLOCK), 0,
                       stream, S, entries, out, tile_end, fwd_end);
}

__global__ void k_upd_curves(float *__restrict__ curve_data, const uint32_t *__restrict__ curve_pt,
                             const float *__restrict__ pts, uint32_t ncurves) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncurves) return;
    float *cv = curve_data + (size_t)c * 13u;
    for (int j = 0; j < 4; ++j) {
        uint32_t p = curve_pt[4u * c + j];
        cv[1 + 2 * j] = pts[2u * p];
        cv[2 + 2 * j] = pts[2u * p + 1u];
    }
}

// Per prepared instance: path box (shape_data params), padded group-local
// bounds (the packer's), and the reference CPU shape_bounds box
// (distance/shape.rs:10-91) into cpu_inst_bounds.
__global__ void k_upd_instances(DevScene S, float *__restrict__ shape_data, float *__restrict__ shape_bounds,
                                const float *__restrict__ inst_pad, float *__restrict__ cpu_b) {
    uint32_t prep = blockIdx.x * blockDim.x + threadIdx.x;
    if (prep >= S.num_prepared) return;
    float *sd = shape_data + prep * 15u;
    uint32_t si = S.prepared_to_scene[prep], kind = (uint32_t)sd[0];
    const float *T = S.shape_transform + si * 6u;
    float scale = fmaxf(fmaxf(sqrtf(T[0] * T[0] + T[3] * T[3]), sqrtf(T[1] * T[1] + T[4] * T[4])), 0.f);
    float r = fabsf(sd[3]) * scale;
    if (kind == DVG_SHAPE_PATH && S.st_cnt[si] > 0) {
        const float *th = S.path_thickness + S.st_off[si];
        float mt = 0.f;
        for (uint32_t k = 0; k < S.st_cnt[si]; ++k) mt = fmaxf(mt, fabsf(th[k]));
        r = fmaxf(r, mt * scale);
    }
    if (r > 0.f && (uint32_t)sd[9] == DVG_JOIN_MITER) r *= fmaxf(sd[11], 1.0f);
    float cpad = fmaxf(r, 0.f);
    float *cb = cpu_b + prep * 4u;
    if (kind == DVG_SHAPE_PATH) {
        uint32_t np = S.sp_cnt[si];
        if (np == 0) {
            cb[0] = 1.f; cb[1] = 1.f; cb[2] = 0.f; cb[3] = 0.f;
            return;
        }
        const float *pts = S.path_points + (size_t)S.sp_off[si] * 2u;
        float bx0 = INFINITY, by0 = INFINITY, bx1 = -INFINITY, by1 = -INFINITY;
        float tx0 = INFINITY, ty0 = INFINITY, tx1 = -INFINITY, ty1 = -INFINITY;
        for (uint32_t k = 0; k < np; ++k) {
            float x = pts[2u * k], y = pts[2u * k + 1u];
            bx0 = fminf(bx0, x); by0 = fminf(by0, y); bx1 = fmaxf(bx1, x); by1 = fmaxf(by1, y);
            float gx = T[0] * x + T[1] * y + T[2], gy = T[3] * x + T[4] * y + T[5];
            tx0 = fminf(tx0, gx); ty0 = fminf(ty0, gy); tx1 = fmaxf(tx1, gx); ty1 = fmaxf(ty1, gy);
        }
        sd[4] = bx0; sd[5] = by0; sd[6] = bx1; sd[7] = by1;
        float cx[4] = {bx0, bx1, bx1, bx0}, cy[4] = {by0, by0, by1, by1};
        float mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
        for (int k = 0; k < 4; ++k) {
            float gx = T[0] * cx[k] + T[1] * cy[k] + T[2], gy = T[3] * cx[k] + T[4] * cy[k] + T[5];
            mnx = fminf(mnx, gx); mny = fminf(mny, gy); mxx = fmaxf(mxx, gx); mxy = fmaxf(mxy, gy);
        }
        float pad = inst_pad[prep];
        float *b = shape_bounds + prep * 4u;
        b[0] = mnx - pad; b[1] = mny - pad; b[2] = mxx + pad; b[3] = mxy + pad;
        cb[0] = tx0 - cpad; cb[1] = ty0 - cpad; cb[2] = tx1 + cpad; cb[3] = ty1 + cpad;
        return;
    }
    float mnx, mny, mxx, mxy;
    if (kind == DVG_SHAPE_CIRCLE || kind == DVG_SHAPE_ELLIPSE) {
        float rx = fabsf(sd[6]), ry = kind == DVG_SHAPE_CIRCLE ? fabsf(sd[6]) : fabsf(sd[7]);
        float cx = T[0] * sd[4] + T[1] * sd[5] + T[2], cy = T[3] * sd[4] + T[4] * sd[5] + T[5];
        float ex = sqrtf((T[0] * rx) * (T[0] * rx) + (T[1] * ry) * (T[1] * ry));
        float ey = sqrtf((T[3] * rx) * (T[3] * rx) + (T[4] * ry) * (T[4] * ry));
        mnx = cx - ex; mny = cy - ey; mxx = cx + ex; mxy = cy + ey;
    } else {
        float xs[4] = {sd[4], sd[6], sd[6], sd[4]}, ys[4] = {sd[5], sd[5], sd[7], sd[7]};
        mnx = mny = INFINITY;
        mxx = mxy = -INFINITY;
        for (int k = 0; k < 4; ++k) {
            float gx = T[0] * xs[k] + T[1] * ys[k] + T[2], gy = T[3] * xs[k] + T[4] * ys[k] + T[5];
            mnx = fminf(mnx, gx); mny = fminf(mny, gy); mxx = fmaxf(mxx, gx); mxy = fmaxf(mxy, gy);
        }
    }
    cb[0] = mnx - cpad; cb[1] = mny - cpad; cb[2] = mxx + cpad; cb[3] = mxy + cpad;
}

// Per group: tile range from the union of its bounds, the CPU root box, and
// the colours.
__global__ void k_upd_groups(DevScene S, const float *__restrict__ shape_bounds, const float *__restrict__ cpu_b,
                             float *__restrict__ group_data, const float *__restrict__ colors, int4 *__restrict__ rng,
                             float *__restrict__ cpu_root) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= S.num_groups) return;
    float *gd = group_data + g * 16u;
    if (colors)
        for (int k = 0; k < 8; ++k) gd[8 + k] = colors[g * 8u + k];
    uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
    float u[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY}, c[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
    bool any = false, cany = false;
    for (uint32_t i = 0; i < cnt; ++i) {
        uint32_t prep = (uint32_t)S.group_shapes[off + i];
        const float *b = shape_bounds + prep * 4u;
        if (b[0] <= b[2] && b[1] <= b[3]) {
            u[0] = fminf(u[0], b[0]); u[1] = fminf(u[1], b[1]); u[2] = fmaxf(u[2], b[2]); u[3] = fmaxf(u[3], b[3]);
            any = true;
        }
        const float *q = cpu_b + prep * 4u;
        if (q[0] <= q[2] && q[1] <= q[3]) {
            c[0] = fminf(c[0], q[0]); c[1] = fminf(c[1], q[1]); c[2] = fmaxf(c[2], q[2]); c[3] = fmaxf(c[3], q[3]);
            cany = true;
        }
    }
    float *cr = cpu_root + g * 4u;
    if (cany) { cr[0] = c[0]; cr[1] = c[1]; cr[2] = c[2]; cr[3] = c[3]; }
    else { cr[0] = 1.f; cr[1] = 1.f; cr[2] = 0.f; cr[3] = 0.f; }
    int4 r = make_int4(1, 1, 0, 0);
    if (any && S.tiles_x > 0 && S.tiles_y > 0) {
        const float *m = S.group_shape_xform + g * 6u;
        float xs[4] = {m[0] * u[0] + m[1] * u[1] + m[2], m[0] * u[2] + m[1] * u[1] + m[2],
                       m[0] * u[2] + m[1] * u[3] + m[2], m[0] * u[0] + m[1] * u[3] + m[2]};
        float ys[4] = {m[3] * u[0] + m[4] * u[1] + m[5], m[3] * u[2] + m[4] * u[1] + m[5],
                       m[3] * u[2] + m[4] * u[3] + m[5], m[3] * u[0] + m[4] * u[3] + m[5]};
        float mnx = fminf(fminf(xs[0], xs[1]), fminf(xs[2], xs[3])), mxx = fmaxf(fmaxf(xs[0], xs[1]), fmaxf(xs[2], xs[3]));
        float mny = fminf(fminf(ys[0], ys[1]), fminf(ys[2], ys[3])), mxy = fmaxf(fmaxf(ys[0], ys[1]), fmaxf(ys[2], ys[3]));
        const float pad = tile_pad(S.W, S.H);  // as the host's binning
        mnx -= pad; mny -= pad; mxx += pad; mxy += pad;
        if (!(mxx < 0.f || mxy < 0.f || mnx >= (float)S.W || mny >= (float)S.H)) {
            int tx0 = (int)floorf(mnx / (float)TILE), tx1 = (int)floorf(mxx / (float)TILE);
            int ty0 = (int)floorf(mny / (float)TILE), ty1 = (int)floorf(mxy / (float)TILE);
            int hx = (int)S.tiles_x - 1, hy = (int)S.tiles_y - 1;
            tx0 = tx0 < 0 ? 0 : (tx0 > hx ? hx : tx0); tx1 = tx1 < 0 ? 0 : (tx1 > hx ? hx : tx1);
            ty0 = ty0 < 0 ? 0 : (ty0 > hy ? hy : ty0); ty1 = ty1 < 0 ? 0 : (ty1 > hy ? hy : ty1);
            if (tx0 <= tx1 && ty0 <= ty1) r = make_int4(tx0, ty0, tx1, ty1);
        }
    }
    rng[g] = r;
}

// Offsets past the tile-list capacity are clamped (every list stays inside
// the buffer) and reported.
__global__ void k_clamp_offsets(uint32_t *__restrict__ offsets, uint32_t n, uint32_t cap, uint32_t *flag) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (offsets[i] > cap) {
        offsets[i] = cap;
        *flag = 1u;
    }
}

// polyline_length (backward/boundary.rs:230-309) of packed curve c.
DVG_DEV float curve_polyline_length(const float *cv) {
    uint32_t k = (uint32_t)cv[0];
    float x0 = cv[1], y0 = cv[2], x1 = cv[3], y1 = cv[4], x2 = cv[5], y2 = cv[6], x3 = cv[7], y3 = cv[8];
    if (k == 0) return sqrtf((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
    if (k == 1) {
        float t = 0.5f, tt = 1.0f - t;
        float vx = ((x0 * (tt * tt)) + (x1 * (2.0f * tt * t))) + (x2 * (t * t));
        float vy = ((y0 * (tt * tt)) + (y1 * (2.0f * tt * t))) + (y2 * (t * t));
        return sqrtf((vx - x0) * (vx - x0) + (vy - y0) * (vy - y0)) +
               sqrtf((x2 - vx) * (x2 - vx) + (y2 - vy) * (y2 - vy));
    }
    float t1 = 1.0f / 3.0f, t2 = 2.0f / 3.0f;
    float u1 = 1.0f - t1, u2 = 1.0f - t2;
    float ax = (((x0 * (u1 * u1 * u1)) + (x1 * (3.0f * u1 * u1 * t1))) + (x2 * (3.0f * u1 * t1 * t1))) + (x3 * (t1 * t1 * t1));
    float ay = (((y0 * (u1 * u1 * u1)) + (y1 * (3.0f * u1 * u1 * t1))) + (y2 * (3.0f * u1 * t1 * t1))) + (y3 * (t1 * t1 * t1));
    float bx = (((x0 * (u2 * u2 * u2)) + (x1 * (3.0f * u2 * u2 * t2))) + (x2 * (3.0f * u2 * t2 * t2))) + (x3 * (t2 * t2 * t2));
    float by = (((y0 * (u2 * u2 * u2)) + (y1 * (3.0f * u2 * u2 * t2))) + (y2 * (3.0f * u2 * t2 * t2))) + (y3 * (t2 * t2 * t2));
    return (sqrtf((ax - x0) * (ax - x0) + (ay - y0) * (ay - y0)) + sqrtf((bx - ax) * (bx - ax) + (by - ay) * (by - ay))) +
           sqrtf((x3 - bx) * (x3 - bx) + (y3 - by) * (y3 - by));
}

// Per scene path: total length and its segment CDF / PMF (written in place;
// the table layout is fixed at packing).
__global__ void k_upd_lengths(DevScene S, float *__restrict__ shape_lengths, float *__restrict__ path_cdf,
                              float *__restrict__ path_pmf) {
    uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= S.num_shapes) return;
    uint32_t prep = S.scene_to_prepared[si];
    const float *sd = S.shape_data + prep * 15u;
    if ((uint32_t)sd[0] != DVG_SHAPE_PATH) return;
    uint32_t off = (uint32_t)sd[12], cnt = (uint32_t)sd[13];
    float len = 0.f;
    for (uint32_t k = 0; k < cnt; ++k) len += curve_polyline_length(S.curve_data + (size_t)(off + k) * 13u);
    shape_lengths[si] = len;
    uint32_t n = S.path_cdf_cnt[si];
    if (n == 0) return;
    float *cdf = path_cdf + S.path_cdf_off[si], *pmf = path_pmf + S.path_cdf_off[si];
    bool ok = len > 0.f && isfinite(len);
    float inv = ok ? 1.0f / len : 0.f, prev = 0.f;
    for (uint32_t k = 0; k < n; ++k) {
        float sl = k < cnt ? curve_polyline_length(S.curve_data + (size_t)(off + k) * 13u) : 0.f;
        float sn = sl * inv;
        pmf[k] = sn;
        prev = sn + prev;
        cdf[k] = prev;
    }
}

// Shape CDF over prepared instances (build_shape_cdf, backward/boundary.rs:
// 312-347): the lengths are gathered into LDS by every lane, a chunk at a
// time, and summed in order by one lane (the host's order, so the tables
// match a host packing bit for bit); then every lane normalises.
constexpr uint32_t CDF_CHUNK = 8192, CDF_THREADS = 1024;

__global__ __launch_bounds__(CDF_THREADS) void k_upd_shape_cdf(DevScene S, float *__restrict__ cdf,
                                                               float *__restrict__ pmf) {
    __shared__ float s_len[CDF_CHUNK];
    __shared__ float s_total;
    float accum = 0.f;  // (lane 0)
    for (uint32_t base = 0; base < S.n_cdf; base += CDF_CHUNK) {
        uint32_t n = min(CDF_CHUNK, S.n_cdf - base);
        for (uint32_t i = threadIdx.x; i < n; i += CDF_THREADS) {
            float len = S.shape_lengths[S.shape_ids[base + i]];
            s_len[i] = len;
            pmf[base + i] = len;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (uint32_t i = 0; i < n; ++i) {
                accum += s_len[i];
                cdf[base + i] = accum;
            }
        __syncthreads();  // (s_len is refilled; lane 0's cdf stores land before the last pass)
    }
    if (threadIdx.x == 0) s_total = accum;
    __syncthreads();
    float total = s_total;
    bool ok = total > 0.f && isfinite(total);
    for (uint32_t i = threadIdx.x; i < S.n_cdf; i += CDF_THREADS) {
        cdf[i] = ok ? cdf[i] / total : 0.f;
        pmf[i] = ok ? pmf[i] / total : 0.f;
    }
}

// ---------------------------------------------------------------------------
// Batches: persistent kernels over a device work queue.  Every workgroup
// takes the next item with one atomic, so uneven scenes and tiles balance
// themselves; the loop ends when the queue is drained (every workgroup
// reaches the exit test on each iteration).

#define DVG_QUEUE_LOOP(...)                                            \
    __shared__ uint32_t s_item;                                        \
    for (;;) {                                                         \
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);         \
        __syncthreads();                                               \
        uint32_t it = __builtin_amdgcn_readfirstlane(s_item);          \
        __syncthreads();                                               \
        if (it >= n_items) break;                                      \
        const WorkItem w = items[it];                                  \
        const BatchScene &B = scenes[w.scene];                         \
        __VA_ARGS__;                                                   \
    }

#define DVG_QUEUE_ARGS \
    const BatchScene *__restrict__ scenes, const WorkItem *__restrict__ items, uint32_t n_items, uint32_t *counter

__global__ __launch_bounds__(BLOCK) void k_batch_zero(DVG_QUEUE_ARGS) {
    DVG_QUEUE_LOOP({
        uint32_t k = w.unit >> 20, chunk = w.unit & 0xfffffu;
        float *p = B.zp[k];
        uint32_t n = B.zn[k];
        uint32_t b = chunk * ZERO_CHUNK, e = min(n, b + ZERO_CHUNK);
        for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) p[i] = 0.f;
    })
}

template <bool CPU>
__global__ __launch_bounds__(BLOCK) void k_batch_weights(DVG_QUEUE_ARGS) {
    DVG_QUEUE_LOOP(weights_unit<CPU>(B.S, w.unit, B.weight))
}

template <bool PF>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(PF ? OCC_FORWARD_PF : OCC_FORWARD))) void k_batch_forward(
    DVG_QUEUE_ARGS) {
    DVG_QUEUE_LOOP(forward_unit<PF>(B.S, w.unit, B.weight, B.accum))
}

__global__ __launch_bounds__(BLOCK) void k_batch_resolve(DVG_QUEUE_ARGS, int loss) {
    DVG_QUEUE_LOOP(resolve_unit(B.S, w.unit, B.weight, B.accum, B.img, loss ? B.target : nullptr,
                                loss ? B.d_render : nullptr))
}

__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_INTERIOR))) void k_batch_interior(
    DVG_QUEUE_ARGS, uint32_t *__restrict__ gmask, uint32_t gcap) {
    uint32_t *slot = gmask ? gmask + (size_t)blockIdx.x * interior_slot_words(gcap) : nullptr;
    DVG_QUEUE_LOOP(interior_unit(B.S, w.unit, B.weight, B.d_render, B.G, slot, gcap))
}

__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_PREFILTER))) void k_batch_prefiltered(
    DVG_QUEUE_ARGS) {
    DVG_QUEUE_LOOP(prefiltered_unit(B.S, w.unit, B.weight, B.d_render, B.G))
}

__global__ __launch_bounds__(BLOCK) void k_batch_bnd_key(DVG_QUEUE_ARGS) {
    DVG_QUEUE_LOOP(bnd_key_unit(B.S, w.unit, B.keys, B.ranks, B.counts, B.nb))
}

__global__ __launch_bounds__(BLOCK) void k_batch_scan(DVG_QUEUE_ARGS) {
    DVG_QUEUE_LOOP(scan_unit<BLOCK>(B.counts, B.nb + 1u, B.offsets))
}

__global__ __launch_bounds__(BLOCK) void k_batch_bnd_scatter(DVG_QUEUE_ARGS) {
    DVG_QUEUE_LOOP(bnd_scatter_unit(B.S, w.unit, B.keys, B.ranks, B.offsets, B.order, B.nb))
}

__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_BATCH_BOUNDARY))) void k_batch_bnd_eval(
    DVG_QUEUE_ARGS) {
    DVG_QUEUE_LOOP(bnd_eval_unit<COV_CPU, false>(B.S, w.unit, B.weight, B.d_render, B.G, B.offsets, B.order, B.nb,
                                                   FlagSlots{nullptr, nullptr}))
}


__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Launchers
static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

void launch_zero(float *const *ptrs, const uint32_t *lens, int count, hipStream_t stream) {
    ZeroList z;
    uint32_t mx = 0;
    for (int k = 0; k < 14; ++k) {
        z.p[k] = k < count ? ptrs[k] : nullptr;
        z.n[k] = k < count ? lens[k] : 0u;
        if (z.p[k] && z.n[k] > mx) mx = z.n[k];
    }
    if (mx == 0) return;
    uint32_t bx = cdiv(mx, 256);
    if (bx > 1024) bx = 1024;
    hipLaunchKernelGGL(k_zero, dim3(bx, 14), dim3(256), 0, stream, z);
}

void launch_bin_count(const float *group_aabb, uint32_t G, uint32_t tiles_x, uint32_t tiles_y, uint32_t *counts,
                      hipStream_t stream) {
    uint32_t nt = tiles_x * tiles_y;
    hipLaunchKernelGGL(k_bin_count, dim3(cdiv(nt, BIN_TILES_PER_WG)), dim3(BLOCK), 0, stream,
                       reinterpret_cast<const int4 *>(group_aabb), G, tiles_x, tiles_y, counts);
}
void launch_scan(const uint32_t *counts, uint32_t n, uint32_t *offsets, hipStream_t stream) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, stream, counts, n, offsets);
}
void launch_bin_write(const float *group_aabb, uint32_t G, uint32_t tiles_x, uint32_t tiles_y,
                      const uint32_t *offsets, uint32_t *entries, hipStream_t stream) {
    uint32_t nt = tiles_x * tiles_y;
    hipLaunchKernelGGL(k_bin_write, dim3(cdiv(nt, BIN_TILES_PER_WG)), dim3(BLOCK), 0, stream,
                       reinterpret_cast<const int4 *>(group_aabb), G, tiles_x, tiles_y, offsets, entries);
}
void launch_weights(const DevScene &S, float *weight, bool cpu, hipStream_t stream) {
    uint32_t nb = cdiv(S.W, 16) * cdiv(S.H, 16);
    if (cpu) hipLaunchKernelGGL(k_weights<true>, dim3(nb), dim3(256), 0, stream, S, weight);
    else hipLaunchKernelGGL(k_weights<false>, dim3(nb), dim3(256), 0, stream, S, weight);
}
void launch_forward(const DevScene &S, const float *weight, float *accum, hipStream_t stream) {
    dim3 grid(S.tiles_x * S.tiles_y * units_per_tile(S));
    if (S.prefilter) hipLaunchKernelGGL(k_forward<true>, grid, dim3(BLOCK), 0, stream, S, weight, accum);
    else hipLaunchKernelGGL(k_forward<false>, grid, dim3(BLOCK), 0, stream, S, weight, accum);
}
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
size_t sdf_work_bytes(uint32_t np, uint32_t nc, uint32_t nt) {
    size_t n = std::max(np, 1u), c = std::max(nc, 1u), t = std::max(nt, 1u);
    return align256(n * 16) + align256(n * 4) + align256(c * 16) + align256(c * SDF_PIECES * 16) +
           align256(c * SDF_PIECES * 8) + align256(t * 4) + align256(n * 4) + align256(t * 4) + align256(t * 4) +
           align256(t * SDF_LCAP * 4);
}
SdfWork sdf_work_carve(void *buf, uint32_t np, uint32_t nc, uint32_t nt) {
    size_t n = std::max(np, 1u), c = std::max(nc, 1u), t = std::max(nt, 1u);
    uint8_t *p = (uint8_t *)buf;
    auto take = [&](size_t bytes) {
        void *r = p;
        p += align256(bytes);
        return r;
    };
    SdfWork w;
    w.cbox = (float *)take(n * 16);
    w.sim = (float *)take(n * 4);
    w.cvbox = (float *)take(c * 16);
    w.piece = (float *)take(c * SDF_PIECES * 16);
    w.piece_h = (float *)take(c * SDF_PIECES * 8);
    w.tile_u = (float *)take(t * 4);
    w.pgroup = (uint32_t *)take(n * 4);
    w.tile_seed = (uint32_t *)take(t * 4);
    w.tile_count = (uint32_t *)take(t * 4);
    w.tile_list = (uint32_t *)take(t * SDF_LCAP * 4);
    return w;
}
static SdfView sdf_view(const SdfWork &w) {
    SdfView v;
    v.cbox = (kfloat *)w.cbox;
    v.sim = (kfloat *)w.sim;
    v.cvbox = (kfloat *)w.cvbox;
    v.piece = (kfloat *)w.piece;
    v.piece_h = (kfloat *)w.piece_h;
    v.tile_u = (kfloat *)w.tile_u;
    v.pgroup = (kuint *)w.pgroup;
    v.tile_seed = (kuint *)w.tile_seed;
    v.tile_count = (kuint *)w.tile_count;
    v.tile_list = (kuint *)w.tile_list;
    return v;
}
static void launch_sdf_prep(const DevScene &S, uint32_t ncurves, const SdfWork &w, hipStream_t stream) {
    if (S.num_groups)
        hipLaunchKernelGGL(k_sdf_prep_groups, dim3(cdiv(S.num_groups, 256)), dim3(256), 0, stream, S, w.pgroup);
    uint32_t n = std::max(S.num_prepared, ncurves);
    if (n) hipLaunchKernelGGL(k_sdf_prep, dim3(cdiv(n, 256)), dim3(256), 0, stream, S, w, ncurves);
}
void launch_sdf(const DevScene &S, uint32_t ncurves, void *work, float *out, hipStream_t stream) {
    uint64_t npix = (uint64_t)S.W * S.H;
    if (npix == 0) return;
    uint32_t ntiles = S.tiles_x * S.tiles_y;
    SdfWork w = sdf_work_carve(work, S.num_prepared, ncurves, ntiles);
    SdfView v = sdf_view(w);
    launch_sdf_prep(S, ncurves, w, stream);
    if (S.sdf_tiled) {
        hipLaunchKernelGGL(k_sdf_lists, dim3(cdiv(ntiles, SDF_LIST_TILES_PER_WG)), dim3(BLOCK), 0, stream, S, v, w);
        hipLaunchKernelGGL(k_sdf_tiles, dim3(ntiles), dim3(256), 0, stream, S, v, out);
    } else {
        hipLaunchKernelGGL(k_sdf_walk, dim3(cdiv(npix, 256)), dim3(256), 0, stream, S, v, out);
    }
}

void launch_sdf_points(const DevScene &S, uint32_t ncurves, void *work, const float *pos, uint32_t n, float *out,
                       hipStream_t stream) {
    if (n == 0) return;
    SdfWork w = sdf_work_carve(work, S.num_prepared, ncurves, S.tiles_x * S.tiles_y);
    launch_sdf_prep(S, ncurves, w, stream);
    hipLaunchKernelGGL(k_sdf_points, dim3(cdiv(n, 256)), dim3(256), 0, stream, S, sdf_view(w), pos, n, out);
}

void launch_sdf_image_backward(const DevScene &S, const float *d_sdf, const DevGrads &G, hipStream_t stream) {
    uint64_t ns = (uint64_t)S.W * S.H * S.spp;
    if (ns == 0) return;
    hipLaunchKernelGGL(k_sdf_image_backward, dim3(cdiv(ns, 256)), dim3(256), 0, stream, S, d_sdf, G);
}

void launch_sdf_points_backward(const DevScene &S, const float *pos, uint32_t n, const float *d_sdf,
                                const DevGrads &G, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_sdf_points_backward, dim3(cdiv(n, 256)), dim3(256), 0, stream, S, pos, n, d_sdf, G);
}

void launch_resolve(const DevScene &S, const float *weight, const float *accum, float *out, hipStream_t stream) {
    hipLaunchKernelGGL(k_resolve, dim3(cdiv((uint64_t)S.W * S.H, 256)), dim3(256), 0, stream, S, weight, accum, out);
}
void launch_backward_interior(const DevScene &S, const float *weight, const float *d_render, const DevGrads &G,
                              const InteriorWork &iw, hipStream_t stream) {
    uint32_t n_units = S.tiles_x * S.tiles_y * units_per_tile(S);
    if (!iw.gmask || iw.slots == 0) {
        hipLaunchKernelGGL(k_backward_interior, dim3(n_units), dim3(BLOCK), 0, stream, S, weight, d_render, G);
        return;
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream, iw.counter, 1u, 0u);
    hipLaunchKernelGGL(k_backward_interior_q, dim3(std::min(n_units, iw.slots)), dim3(BLOCK), 0, stream, S, weight,
                       d_render, G, iw.counter, n_units, iw.gmask, iw.gcap);
}
void launch_backward_prefiltered(const DevScene &S, const float *weight, const float *d_render, const DevGrads &G,
                                 hipStream_t stream) {
    hipLaunchKernelGGL(k_backward_prefiltered, dim3(S.tiles_x * S.tiles_y * units_per_tile(S)), dim3(BLOCK), 0,
                       stream, S, weight, d_render, G);
}
void launch_boundary_sort(const DevScene &S, const BoundaryWork &bw, hipStream_t stream) {
    uint64_t total = (uint64_t)S.W * S.H * S.spp;
    uint32_t nbk = boundary_buckets(S);
    uint32_t grid = cdiv(total, BLOCK);
    // counts[nbk + 1] start at zero
    float *zp[1] = {reinterpret_cast<float *>(bw.counts)};
    uint32_t zn[1] = {nbk + 1};
    launch_zero(zp, zn, 1, stream);
    hipLaunchKernelGGL(k_bnd_key, dim3(grid), dim3(BLOCK), 0, stream, S, bw.keys, bw.ranks, bw.counts, nbk);
    uint32_t nc = scan_chunks(nbk + 1);
    hipLaunchKernelGGL(k_scan_sums, dim3(nc), dim3(256), 0, stream, bw.counts, nbk + 1, bw.chunk_sums);
    hipLaunchKernelGGL(k_scan_chunks, dim3(nc), dim3(256), 0, stream, bw.counts, nbk + 1, bw.chunk_sums, bw.offsets);
    hipLaunchKernelGGL(k_bnd_scatter, dim3(grid), dim3(BLOCK), 0, stream, S, bw.keys, bw.ranks, bw.offsets, bw.order,
                       nbk);
}

void launch_boundary_eval(const DevScene &S, const float *weight, const float *d_render, const DevGrads &G,
                          const BoundaryWork &bw, hipStream_t stream) {
    uint64_t total = (uint64_t)S.W * S.H * S.spp;
    FlagSlots fs = {bw.flag_count, bw.flag_slots};
    if (fs.count)
        hipLaunchKernelGGL(k_bnd_eval_chk, dim3(cdiv(total, BLOCK)), dim3(BLOCK), 0, stream, S, weight, d_render, G,
                           bw.offsets, bw.order, boundary_buckets(S), fs);
    else
        hipLaunchKernelGGL(k_bnd_eval, dim3(cdiv(total, BLOCK)), dim3(BLOCK), 0, stream, S, weight, d_render, G,
                           bw.offsets, bw.order, boundary_buckets(S));
}

void launch_boundary_redo(const DevScene &S, const float *weight, const float *d_render, const DevGrads &G,
                          const BoundaryWork &bw, hipStream_t stream) {
    uint64_t total = (uint64_t)S.W * S.H * S.spp;
    FlagSlots fs = {bw.flag_count, bw.flag_slots};
    hipLaunchKernelGGL(k_bnd_redo, dim3(cdiv(cdiv(total, BLOCK), FLAG_UNITS_PER_WG)), dim3(BLOCK), 0, stream, S,
                       weight, d_render, G, bw.offsets, boundary_buckets(S), fs);
}

void launch_batch_phase(int phase, const BatchScene *scenes, const WorkItem *items, uint32_t n_items,
                        uint32_t *counter, uint32_t grid, bool loss, uint32_t *gmask, uint32_t gcap,
                        hipStream_t stream) {
    if (n_items == 0 || grid == 0) return;
    dim3 g(grid), b(BLOCK);
    switch (phase) {
        case PH_ZERO: hipLaunchKernelGGL(k_batch_zero, g, b, 0, stream, scenes, items, n_items, counter); break;
        case PH_WEIGHTS_FWD:
            hipLaunchKernelGGL(k_batch_weights<false>, g, b, 0, stream, scenes, items, n_items, counter);
            break;
        case PH_FORWARD:
            hipLaunchKernelGGL(k_batch_forward<false>, g, b, 0, stream, scenes, items, n_items, counter);
            break;
        case PH_FORWARD_PF:
            hipLaunchKernelGGL(k_batch_forward<true>, g, b, 0, stream, scenes, items, n_items, counter);
            break;
        case PH_RESOLVE:
            hipLaunchKernelGGL(k_batch_resolve, g, b, 0, stream, scenes, items, n_items, counter, loss ? 1 : 0);
            break;
        case PH_WEIGHTS_BWD:
            hipLaunchKernelGGL(k_batch_weights<true>, g, b, 0, stream, scenes, items, n_items, counter);
            break;
        case PH_INTERIOR:
            hipLaunchKernelGGL(k_batch_interior, g, b, 0, stream, scenes, items, n_items, counter, gmask, gcap);
            break;
        case PH_PREFILTERED:
            hipLaunchKernelGGL(k_batch_prefiltered, g, b, 0, stream, scenes, items, n_items, counter);
            break;
        case PH_BND_KEY: hipLaunchKernelGGL(k_batch_bnd_key, g, b, 0, stream, scenes, items, n_items, counter); break;
        case PH_SCAN: hipLaunchKernelGGL(k_batch_scan, g, b, 0, stream, scenes, items, n_items, counter); break;
        case PH_BND_SCATTER:
            hipLaunchKernelGGL(k_batch_bnd_scatter, g, b, 0, stream, scenes, items, n_items, counter);
            break;
        case PH_BND_EVAL: hipLaunchKernelGGL(k_batch_bnd_eval, g, b, 0, stream, scenes, items, n_items, counter); break;
        default: break;
    }
}

void launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(cdiv(n, 256)), dim3(256), 0, stream, p, n, v);
}

void launch_selftest_winding(uint32_t n, uint32_t seed, uint32_t *mismatches, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream, mismatches, 1u, 0u);
    hipLaunchKernelGGL(k_selftest_winding, dim3(cdiv(n, 256)), dim3(256), 0, stream, n, seed, mismatches);
}

void launch_selftest_inside(uint32_t n, uint32_t seed, uint32_t *counts, hipStream_t stream) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream, counts, 12u, 0u);
    if (n == 0) return;
    hipLaunchKernelGGL(k_selftest_inside, dim3(cdiv(n, 256)), dim3(256), 0, stream, n, seed, counts);
}

void launch_wind_coef(const float *curve_data, uint32_t ncurves, float *curve_wind, const float *shape_data,
                      uint32_t nprep, hipStream_t stream) {
    if (ncurves == 0) return;
    hipLaunchKernelGGL(k_wind_coef, dim3(cdiv(ncurves, 256)), dim3(256), 0, stream, curve_data, ncurves, curve_wind);
    if (nprep)
        hipLaunchKernelGGL(k_shape_topo, dim3(cdiv(nprep, 256)), dim3(256), 0, stream, shape_data, nprep, curve_data,
                           ncurves, curve_wind);
}

// Group BVH boxes refitted to the updated shape bounds, one lane a group: a
// node's box is the union of the shapes of the leaves in its subtree (nodes
// [k, skip) in preorder).  The tree keeps the split of the scene it was built
// for.
__global__ void k_upd_group_bvh(DevScene S, float *__restrict__ bounds) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= S.num_groups) return;
    kuint *m = S.gbvh_meta + g * 4u;
    uint32_t noff = m[0], ncnt = m[1], ioff = m[2];
    if (ncnt == 0u || m[3] == 0u) return;
    for (uint32_t k = 0; k < ncnt; ++k) {
        uint32_t skip = S.gbvh_nodes[(size_t)(noff + k) * 4u + 1u];
        uint32_t end = skip == SDF_NONE ? ncnt : skip;
        float x0 = __builtin_inff(), y0 = __builtin_inff(), x1 = -__builtin_inff(), y1 = -__builtin_inff();
        for (uint32_t j = k; j < end; ++j) {
            kuint *nd = S.gbvh_nodes + (size_t)(noff + j) * 4u;
            for (uint32_t i = 0; i < nd[3]; ++i) {
                kfloat *b = S.shape_bounds + S.gbvh_indices[ioff + nd[2] + i] * 4u;
                if (!(b[0] <= b[2] && b[1] <= b[3])) continue;
                x0 = fminf(x0, b[0]); y0 = fminf(y0, b[1]); x1 = fmaxf(x1, b[2]); y1 = fmaxf(y1, b[3]);
            }
        }
        float *o = bounds + (size_t)(noff + k) * 4u;
        o[0] = x0; o[1] = y0; o[2] = x1; o[3] = y1;
    }
}

void launch_scene_update(const DevScene &S, const UpdatePlan &P, const float *path_points, const float *group_colors,
                         hipStream_t stream) {
    // the scene's own arena buffers, written in place
    float *pts = (float *)(S.path_points);
    float *curve_data = (float *)(S.curve_data);
    float *shape_data = (float *)(S.shape_data);
    float *shape_bounds = (float *)(S.shape_bounds);
    float *group_data = (float *)(S.group_data);
    float *cpu_root = (float *)(S.group_cpu_bounds);
    uint32_t *offsets = (uint32_t *)(S.tile_offsets);
    uint32_t *entries = (uint32_t *)(S.tile_entries);
    uint32_t ntiles = S.tiles_x * S.tiles_y;
    if (path_points && P.n_points)
        hipLaunchKernelGGL(k_copy_f32, dim3(cdiv(2ull * P.n_points, 256)), dim3(256), 0, stream, pts, path_points,
                           2u * P.n_points);
    if (P.ncurves)
        hipLaunchKernelGGL(k_upd_curves, dim3(cdiv(P.ncurves, 256)), dim3(256), 0, stream, curve_data, P.curve_pt, pts,
                           P.ncurves);
    launch_wind_coef(curve_data, P.ncurves, (float *)(S.curve_wind), shape_data, S.num_prepared, stream);
    if (S.num_prepared)
        hipLaunchKernelGGL(k_upd_instances, dim3(cdiv(S.num_prepared, 256)), dim3(256), 0, stream, S, shape_data,
                           shape_bounds, P.inst_pad, P.cpu_inst_bounds);
    if (S.num_groups)
        hipLaunchKernelGGL(k_upd_groups, dim3(cdiv(S.num_groups, 256)), dim3(256), 0, stream, S, shape_bounds,
                           P.cpu_inst_bounds, group_data, group_colors, reinterpret_cast<int4 *>(P.ranges), cpu_root);
    if (S.num_groups)
        hipLaunchKernelGGL(k_upd_group_bvh, dim3(cdiv(S.num_groups, 256)), dim3(256), 0, stream, S,
                           (float *)(S.gbvh_bounds));
    if (ntiles > 0) {
        launch_bin_count(reinterpret_cast<const float *>(P.ranges), S.num_groups, S.tiles_x, S.tiles_y, P.tile_counts,
                         stream);
        launch_scan(P.tile_counts, ntiles, offsets, stream);
        hipLaunchKernelGGL(k_clamp_offsets, dim3(cdiv(ntiles + 1, 256)), dim3(256), 0, stream, offsets, ntiles,
                           P.tile_cap, P.overflow);
        launch_bin_write(reinterpret_cast<const float *>(P.ranges), S.num_groups, S.tiles_x, S.tiles_y, offsets,
                         entries, stream);
        launch_tile_cull(S, entries, (uint32_t *)(S.bwd_entries), (uint32_t *)(S.tile_end), (uint32_t *)(S.fwd_end),
                         P.tile_cap, true, stream);
    }
    if (S.n_cdf > 0 && !S.prefilter) {
        hipLaunchKernelGGL(k_upd_lengths, dim3(cdiv(S.num_shapes, 256)), dim3(256), 0, stream, S,
                           (float *)(S.shape_lengths), (float *)(S.path_cdf),
                           (float *)(S.path_pmf));
        hipLaunchKernelGGL(k_upd_shape_cdf, dim3(1), dim3(CDF_THREADS), 0, stream, S, (float *)(S.shape_cdf),
                           (float *)(S.shape_pmf));
    }
}

}  // namespace dvgk
