// C ABI of libdvg (include/dvg.h): context, device-resident scenes, and the
// synchronous scene / packed entry points built on them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dvg.h"
#include "dvg_internal.h"
#include "dvg_kernels.h"
#include "dvg_pack.h"

#define DVG_VERSION_STRING "dvg-mi355x 0.1 (gfx950)"

namespace {

// Bump allocator over one device allocation.
struct Arena {
    std::vector<std::pair<const void *, size_t>> srcs;
    std::vector<size_t> offs;
    size_t total = 0;
    size_t add(const void *src, size_t bytes) {
        size_t o = (total + 255) & ~(size_t)255;
        srcs.push_back({src, bytes});
        offs.push_back(o);
        total = o + std::max(bytes, (size_t)4);
        return offs.size() - 1;
    }
};

int ensure(dvg_ctx *ctx, float *&buf, size_t &cap, size_t n) {
    if (cap >= n) return DVG_OK;
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    HIP_TRY(ctx, hipMalloc(&buf, std::max(n, (size_t)1) * sizeof(float)));
    cap = n;
    return DVG_OK;
}

}  // namespace

extern "C" {

const char *dvg_version(void) { return DVG_VERSION_STRING; }

int32_t dvg_selftest_winding(dvg_ctx *ctx, uint32_t n, uint32_t seed, uint32_t *mismatches) {
    if (!ctx || !mismatches) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    uint32_t *d = nullptr;
    HIP_TRY(ctx, hipMalloc(&d, sizeof(uint32_t)));
    dvgk::launch_selftest_winding(n, seed, d, ctx->stream);
    hipError_t e = hipMemcpyAsync(mismatches, d, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_err(ctx, e, "winding self-test");
    if (n == 0) *mismatches = 0;
    return DVG_OK;
}

int32_t dvg_ctx_set_exact_boundary(dvg_ctx *ctx, int enable) {
    if (!ctx) return DVG_INVALID_SCENE;
    if (enable && !ctx->side) {
        // the redo's stream (default priority) and its fork / join events
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        if (!ctx->ev_fork) HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
        if (!ctx->ev_join) HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
        HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    }
    ctx->exact_boundary = enable != 0;
    return DVG_OK;
}

int32_t dvg_selftest_inside(dvg_ctx *ctx, uint32_t n, uint32_t seed, uint32_t *counts) {
    if (!ctx || !counts) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    uint32_t *d = nullptr;
    HIP_TRY(ctx, hipMalloc(&d, 12 * sizeof(uint32_t)));
    dvgk::launch_selftest_inside(n, seed, d, ctx->stream);
    hipError_t e = hipMemcpyAsync(counts, d, 12 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_err(ctx, e, "inside-test self-test");
    return DVG_OK;
}

dvg_ctx *dvg_ctx_create(int device_ordinal) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device_ordinal < 0 || device_ordinal >= n) return nullptr;
    if (hipSetDevice(device_ordinal) != hipSuccess) return nullptr;
    dvg_ctx *ctx = new dvg_ctx();
    ctx->device = device_ordinal;
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return nullptr;
    }
    ctx->stream = ctx->own_stream;
    // the no-allocation entry point's scratch address on this device; a
    // failure is reported by that entry point
    (void)dvgn::noalloc_setup(device_ordinal);
    // the checked boundary pass is the default (its redo stream and events)
    if (dvg_ctx_set_exact_boundary(ctx, 1) != DVG_OK) {
        dvg_ctx_destroy(ctx);
        return nullptr;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device_ordinal) == hipSuccess && ncu > 0)
        ctx->n_cu = (uint32_t)ncu;
    else
        ctx->n_cu = 256;
    return ctx;
}

void dvg_ctx_destroy(dvg_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->weight) (void)hipFree(ctx->weight);
    if (ctx->accum) (void)hipFree(ctx->accum);
    if (ctx->bwork) (void)hipFree(ctx->bwork);
    if (ctx->imask) (void)hipFree(ctx->imask);
    if (ctx->sdfw) (void)hipFree(ctx->sdfw);
    for (const TimedLaunch &tl : ctx->pending) ctx->event_pool.push_back({tl.start, tl.stop});
    for (auto &e : ctx->event_pool) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamDestroy(ctx->side);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

const char *dvg_last_error(const dvg_ctx *ctx) { return ctx ? ctx->last_error : "no context"; }

int32_t dvg_ctx_set_stream(dvg_ctx *ctx, void *s) {
    if (!ctx) return DVG_DEVICE_ERROR;
    ctx->stream = (hipStream_t)s;
    return DVG_OK;
}

int32_t dvg_ctx_set_timing(dvg_ctx *ctx, int enable) {
    if (!ctx) return DVG_DEVICE_ERROR;
    ctx->timing = enable != 0;
    return DVG_OK;
}

int32_t dvg_kernel_times(dvg_ctx *ctx, double *total_ms, uint32_t *calls, int n, int reset) {
    if (!ctx) return DVG_DEVICE_ERROR;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (const TimedLaunch &tl : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, tl.start, tl.stop) == hipSuccess) {
            ctx->kernel_ms[tl.kernel] += ms;
            ctx->kernel_calls[tl.kernel] += 1;
        }
        ctx->event_pool.push_back({tl.start, tl.stop});
    }
    ctx->pending.clear();
    for (int k = 0; k < n && k < DVG_NUM_KERNELS; ++k) {
        if (total_ms) total_ms[k] = ctx->kernel_ms[k];
        if (calls) calls[k] = ctx->kernel_calls[k];
    }
    if (reset)
        for (int k = 0; k < DVG_NUM_KERNELS; ++k) {
            ctx->kernel_ms[k] = 0.0;
            ctx->kernel_calls[k] = 0;
        }
    return DVG_OK;
}

int32_t dvg_sync(dvg_ctx *ctx) {
    if (!ctx) return DVG_DEVICE_ERROR;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DVG_OK;
}

int32_t dvg_prepare(const dvg_scene_desc *scene, const dvg_render_options *options, dvg_prepared **out,
                    const char **error_message) {
    const char *err = "";
    dvg_prepared *p = new dvg_prepared();
    int st = dvgh::prepare(scene, options, *p, &err);
    if (error_message) *error_message = err;
    if (st != DVG_OK) {
        delete p;
        *out = nullptr;
        return st;
    }
    *out = p;
    return DVG_OK;
}

void dvg_prepared_free(dvg_prepared *p) { delete p; }

void dvg_prepared_pack(const dvg_prepared *p, dvg_scene_pack *pack, dvg_backward_pack *bpack,
                       dvg_render_params *params) {
    dvgh::pack_views(*p, pack, bpack, params);
}

// Root box of the reference CPU group BVH for every group, in group-local
// space: the union of shape_bounds (distance/shape.rs:10-91) over the group's
// shapes -- geometry bounds under the shape transform, inflated by the
// largest stroke radius.  min > max for a group with none.  Without the
// backward tables (forward-only scenes) every box is "none".
static std::vector<float> group_cpu_bounds(const dvg_scene_pack *pk, const dvg_backward_pack *bp) {
    uint32_t G = pk->num_groups;
    std::vector<float> out((size_t)std::max(G, 1u) * 4);
    for (uint32_t g = 0; g < std::max(G, 1u); ++g) {
        out[4 * g] = 1.f; out[4 * g + 1] = 1.f; out[4 * g + 2] = 0.f; out[4 * g + 3] = 0.f;
    }
    if (!bp || !bp->prepared_to_scene || !bp->shape_transform) return out;
    for (uint32_t g = 0; g < G; ++g) {
        const float *gd = pk->group_data + 16 * (size_t)g;
        uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
        float u[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
        bool any = false;
        for (uint32_t i = 0; i < cnt; ++i) {
            uint32_t prep = (uint32_t)pk->group_shapes[off + i];
            uint32_t si = bp->prepared_to_scene[prep];
            const float *sd = pk->shape_data + 15 * (size_t)prep;
            const float *T = bp->shape_transform + 6 * (size_t)si;
            auto tx = [&](float x, float y) { return T[0] * x + T[1] * y + T[2]; };
            auto ty = [&](float x, float y) { return T[3] * x + T[4] * y + T[5]; };
            uint32_t kind = (uint32_t)sd[0];
            float mnx, mny, mxx, mxy;
            if (kind == DVG_SHAPE_CIRCLE || kind == DVG_SHAPE_ELLIPSE) {
                float rx = std::fabs(sd[6]), ry = kind == DVG_SHAPE_CIRCLE ? std::fabs(sd[6]) : std::fabs(sd[7]);
                float cx = tx(sd[4], sd[5]), cy = ty(sd[4], sd[5]);
                float ex = std::sqrt((T[0] * rx) * (T[0] * rx) + (T[1] * ry) * (T[1] * ry));
                float ey = std::sqrt((T[3] * rx) * (T[3] * rx) + (T[4] * ry) * (T[4] * ry));
                mnx = cx - ex; mny = cy - ey; mxx = cx + ex; mxy = cy + ey;
            } else if (kind == DVG_SHAPE_RECT) {
                float xs[4] = {sd[4], sd[6], sd[6], sd[4]}, ys[4] = {sd[5], sd[5], sd[7], sd[7]};
                mnx = mny = INFINITY;
                mxx = mxy = -INFINITY;
                for (int k = 0; k < 4; ++k) {
                    float a = tx(xs[k], ys[k]), b = ty(xs[k], ys[k]);
                    mnx = std::fmin(mnx, a); mny = std::fmin(mny, b); mxx = std::fmax(mxx, a); mxy = std::fmax(mxy, b);
                }
            } else {
                uint32_t np = bp->shape_path_point_counts[si];
                if (np == 0) continue;
                const float *pts = bp->path_points + 2 * (size_t)bp->shape_path_offsets[si];
                mnx = mny = INFINITY;
                mxx = mxy = -INFINITY;
                for (uint32_t k = 0; k < np; ++k) {
                    float a = tx(pts[2 * k], pts[2 * k + 1]), b = ty(pts[2 * k], pts[2 * k + 1]);
                    mnx = std::fmin(mnx, a); mny = std::fmin(mny, b); mxx = std::fmax(mxx, a); mxy = std::fmax(mxy, b);
                }
            }
            // max_stroke_radius (distance/shape.rs:62-79)
            float scale = std::fmax(std::fmax(std::sqrt(T[0] * T[0] + T[3] * T[3]), std::sqrt(T[1] * T[1] + T[4] * T[4])), 0.f);
            float r = std::fabs(sd[3]) * scale;
            if (kind == DVG_SHAPE_PATH && bp->shape_path_thickness_counts[si] > 0) {
                const float *th = bp->path_thickness + bp->shape_path_thickness_offsets[si];
                float mt = 0.f;
                for (uint32_t k = 0; k < bp->shape_path_thickness_counts[si]; ++k) mt = std::fmax(mt, std::fabs(th[k]));
                r = std::fmax(r, mt * scale);
            }
            if (r > 0.f && (uint32_t)sd[9] == DVG_JOIN_MITER) r *= std::fmax(sd[11], 1.0f);
            float pad = std::fmax(r, 0.f);
            u[0] = std::fmin(u[0], mnx - pad); u[1] = std::fmin(u[1], mny - pad);
            u[2] = std::fmax(u[2], mxx + pad); u[3] = std::fmax(u[3], mxy + pad);
            any = true;
        }
        if (any) std::memcpy(&out[4 * (size_t)g], u, sizeof(u));
    }
    return out;
}

// Plan of device-side geometry updates: for every packed curve the global
// path-point indices of its 4 control points (the walk of path.rs:100-173),
// and for every prepared instance the padding of its bounds -- the packer's
// stroke padding with the radius taken as max(|stroke width|, max |thickness|)
// (an upper bound of the flattened radius, so culling stays conservative)
// plus the prefiltering band.
static bool make_update_plan(const dvg_scene_pack *pk, const dvg_backward_pack *bp, const dvg_render_params *prm,
                             std::vector<uint32_t> &curve_pt, std::vector<float> &inst_pad) {
    uint32_t SP = pk->num_prepared_shapes, G = pk->num_groups;
    if (!bp->prepared_to_scene || !bp->shape_transform) return false;
    size_t ncurves = pk->curve_data_len / 13;
    curve_pt.assign(ncurves * 4, 0u);
    inst_pad.assign(std::max(SP, 1u), 0.f);
    for (uint32_t g = 0; g < G; ++g) {
        const float *gd = pk->group_data + 16 * (size_t)g;
        uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
        float aa = prm->use_prefiltering ? pk->group_inv_scale[g] : 0.f;
        for (uint32_t i = 0; i < cnt; ++i) {
            uint32_t prep = (uint32_t)pk->group_shapes[off + i];
            const float *sd = pk->shape_data + 15 * (size_t)prep;
            uint32_t si = bp->prepared_to_scene[prep];
            const float *T = bp->shape_transform + 6 * (size_t)si;
            float r = std::fabs(sd[3]);
            if ((uint32_t)sd[0] == DVG_SHAPE_PATH) {
                const float *th = bp->path_thickness + bp->shape_path_thickness_offsets[si];
                for (uint32_t k = 0; k < bp->shape_path_thickness_counts[si]; ++k) r = std::fmax(r, std::fabs(th[k]));
            }
            float scale = std::fmax(std::fmax(std::sqrt(T[0] * T[0] + T[3] * T[3]), std::sqrt(T[1] * T[1] + T[4] * T[4])), 0.f);
            float sp = r * scale;
            if ((uint32_t)sd[9] == DVG_JOIN_MITER) sp *= std::fmax(sd[11], 1.0f);
            inst_pad[prep] = std::fmax(aa + sp, 0.f);
            if ((uint32_t)sd[0] != DVG_SHAPE_PATH) continue;
            uint32_t coff = (uint32_t)sd[12], ccnt = (uint32_t)sd[13];
            uint32_t np = bp->shape_path_point_counts[si], nc = bp->shape_path_ctrl_counts[si];
            bool closed = bp->shape_path_is_closed[si] != 0;
            const uint32_t *ctrl = bp->path_num_control_points + bp->shape_path_ctrl_offsets[si];
            uint32_t base = bp->shape_path_offsets[si];
            uint32_t pid = 0, made = 0;
            for (uint32_t k = 0; k < nc && made < ccnt && np > 0; ++k) {
                uint32_t c = ctrl[k];
                if (c > 2) break;
                uint32_t idx[4];
                bool ok = true;
                for (uint32_t j = 0; j <= c + 1; ++j) {
                    uint32_t q = pid + j;
                    if (closed) q %= np;
                    else if (q >= np) { ok = false; break; }
                    idx[j] = q;
                }
                if (!ok) break;
                for (uint32_t j = c + 2; j < 4; ++j) idx[j] = idx[c + 1];
                size_t cv = (size_t)coff + made;
                if (cv >= ncurves) return false;
                for (int j = 0; j < 4; ++j) curve_pt[cv * 4 + j] = base + idx[j];
                ++made;
                pid += c + 1;
            }
            if (made != ccnt) return false;
        }
    }
    return true;
}

int32_t dvg_device_scene_create(dvg_ctx *ctx, const dvg_scene_pack *pk, const dvg_backward_pack *bp,
                                const dvg_render_params *prm, dvg_device_scene **out) {
    if (!ctx || !pk || !prm || !out) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    *out = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const uint32_t G = pk->num_groups, SP = pk->num_prepared_shapes;
    const uint32_t W = prm->width, H = prm->height;
    uint32_t nsx = std::max(prm->samples_x, 1u), nsy = std::max(prm->samples_y, 1u);
    if ((uint64_t)W * H * nsx * nsy > 0xffffffffull)
        return set_err(ctx, DVG_INVALID_SCENE, "too many samples for 1d launch");
    uint32_t tiles_x = (W + dvgk::TILE - 1) / dvgk::TILE, tiles_y = (H + dvgk::TILE - 1) / dvgk::TILE;
    uint32_t ntiles = tiles_x * tiles_y;
    const size_t ncurves = pk->curve_data ? pk->curve_data_len / 13 : 0;

    // Tile range of every group (binning.rs:30-110): its group-local bounds
    // mapped to canvas by shape_to_canvas, floored to 16-px tiles, clamped.
    std::vector<int32_t> ranges((size_t)std::max(G, 1u) * 4, -1);
    uint64_t n_entries = 0;
    for (uint32_t g = 0; g < G; ++g) {
        const float *b = pk->group_bounds + 4 * (size_t)g;
        int32_t *r = &ranges[4 * (size_t)g];
        r[0] = r[1] = 1;
        r[2] = r[3] = 0;  // empty
        if (!(b[0] <= b[2] && b[1] <= b[3]) || tiles_x == 0 || tiles_y == 0) continue;
        const float *m = pk->group_shape_xform + 6 * (size_t)g;
        float xs[4] = {m[0] * b[0] + m[1] * b[1] + m[2], m[0] * b[2] + m[1] * b[1] + m[2],
                       m[0] * b[2] + m[1] * b[3] + m[2], m[0] * b[0] + m[1] * b[3] + m[2]};
        float ys[4] = {m[3] * b[0] + m[4] * b[1] + m[5], m[3] * b[2] + m[4] * b[1] + m[5],
                       m[3] * b[2] + m[4] * b[3] + m[5], m[3] * b[0] + m[4] * b[3] + m[5]};
        float mnx = std::min(std::min(xs[0], xs[1]), std::min(xs[2], xs[3]));
        float mxx = std::max(std::max(xs[0], xs[1]), std::max(xs[2], xs[3]));
        float mny = std::min(std::min(ys[0], ys[1]), std::min(ys[2], ys[3]));
        float mxy = std::max(std::max(ys[0], ys[1]), std::max(ys[2], ys[3]));
        // pad in canvas space, whatever the scaling (reaches of the
        // prefiltered adjoint and of the boundary pass's query points)
        const float pad = dvgk::tile_pad(W, H);
        mnx -= pad; mny -= pad; mxx += pad; mxy += pad;
        if (mxx < 0.f || mxy < 0.f || mnx >= (float)W || mny >= (float)H) continue;
        auto clampi = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
        int tx0 = clampi((int)std::floor(mnx / (float)dvgk::TILE), (int)tiles_x - 1);
        int tx1 = clampi((int)std::floor(mxx / (float)dvgk::TILE), (int)tiles_x - 1);
        int ty0 = clampi((int)std::floor(mny / (float)dvgk::TILE), (int)tiles_y - 1);
        int ty1 = clampi((int)std::floor(mxy / (float)dvgk::TILE), (int)tiles_y - 1);
        if (tx0 > tx1 || ty0 > ty1) continue;
        r[0] = tx0; r[1] = ty0; r[2] = tx1; r[3] = ty1;
        n_entries += (uint64_t)(tx1 - tx0 + 1) * (uint64_t)(ty1 - ty0 + 1);
    }
    // longest tile list: the group rectangles summed on a 2-D difference grid
    uint32_t list_max = 0;
    if (ntiles > 0) {
        std::vector<int32_t> diff((size_t)(tiles_x + 1) * (tiles_y + 1), 0);
        for (uint32_t g = 0; g < G; ++g) {
            const int32_t *r = &ranges[4 * (size_t)g];
            if (r[0] > r[2] || r[1] > r[3]) continue;
            diff[(size_t)r[1] * (tiles_x + 1) + r[0]] += 1;
            diff[(size_t)r[1] * (tiles_x + 1) + r[2] + 1] -= 1;
            diff[(size_t)(r[3] + 1) * (tiles_x + 1) + r[0]] -= 1;
            diff[(size_t)(r[3] + 1) * (tiles_x + 1) + r[2] + 1] += 1;
        }
        for (uint32_t y = 0; y <= tiles_y; ++y)
            for (uint32_t x = 0; x <= tiles_x; ++x) {
                size_t i = (size_t)y * (tiles_x + 1) + x;
                if (x > 0) diff[i] += diff[i - 1];
                if (y > 0) diff[i] += diff[i - (tiles_x + 1)];
                if (x > 0 && y > 0) diff[i] -= diff[i - (tiles_x + 1) - 1];
                if (x < tiles_x && y < tiles_y) list_max = std::max(list_max, (uint32_t)diff[i]);
            }
    }

    // scene_to_prepared: first instance of each scene shape
    uint32_t S = bp ? bp->num_shapes : 0;
    std::vector<uint32_t> s2p(std::max(S, 1u), 0);
    if (bp && bp->prepared_to_scene)
        for (uint32_t i = SP; i-- > 0;) s2p[bp->prepared_to_scene[i]] = i;
    std::vector<float> cpu_bounds = group_cpu_bounds(pk, bp);
    // group BVHs (renderer/bvh.rs), checked so that a skip-pointer walk
    // always moves forward and stays in its arrays; a null meta is no BVH
    std::vector<uint32_t> gbvh_meta((size_t)std::max(G, 1u) * 4, 0);
    if (pk->group_bvh_meta) std::copy(pk->group_bvh_meta, pk->group_bvh_meta + 4 * (size_t)G, gbvh_meta.begin());
    for (uint32_t g = 0; g < G; ++g) {
        const uint32_t *m = &gbvh_meta[4 * (size_t)g];
        if (m[1] == 0 || m[3] == 0) continue;
        uint64_t nodes_end = (uint64_t)m[0] + m[1], idx_end = (uint64_t)m[2] + m[3];
        if (!pk->group_bvh_nodes || !pk->group_bvh_bounds || !pk->group_bvh_indices ||
            nodes_end * 4 > pk->group_bvh_nodes_len || nodes_end * 4 > pk->group_bvh_bounds_len ||
            idx_end > pk->group_bvh_indices_len)
            return set_err(ctx, DVG_INVALID_SCENE, "group BVH outside its arrays");
        for (uint32_t k = 0; k < m[1]; ++k) {
            const uint32_t *nd = pk->group_bvh_nodes + 4 * ((size_t)m[0] + k);
            bool skip_ok = nd[1] == 0xffffffffu || (nd[1] > k && nd[1] < m[1]);
            bool node_ok = nd[3] == 0 ? (nd[0] > k && nd[0] < m[1]) : ((uint64_t)nd[2] + nd[3] <= m[3]);
            if (!skip_ok || !node_ok) return set_err(ctx, DVG_INVALID_SCENE, "invalid group BVH node");
        }
        for (uint32_t k = 0; k < m[3]; ++k)
            if (pk->group_bvh_indices[(size_t)m[2] + k] >= SP)
                return set_err(ctx, DVG_INVALID_SCENE, "group BVH index out of range");
    }
    // prefiltered cut-offs: largest stroke radius and stroke width per group
    std::vector<float> group_pf((size_t)std::max(G, 1u) * 2, 0.f);
    for (uint32_t g = 0; g < G; ++g) {
        const float *gd = pk->group_data + 16 * (size_t)g;
        uint32_t off = (uint32_t)gd[0], cnt = (uint32_t)gd[1];
        float rmax = 0.f, swmax = 0.f;
        for (uint32_t i = 0; i < cnt && (uint64_t)off + i < pk->group_shapes_len; ++i) {
            uint32_t prep = (uint32_t)pk->group_shapes[off + i];
            if (prep >= SP) continue;
            const float *sd = pk->shape_data + 15 * (size_t)prep;
            float sw = sd[3];
            swmax = std::max(swmax, sw);
            rmax = std::max(rmax, std::fabs(sw));
            uint32_t coff = (uint32_t)sd[12], ccnt = (uint32_t)sd[13];
            for (uint32_t k = 0; k < ccnt && (uint64_t)coff + k < ncurves; ++k)
                for (int j = 9; j < 13; ++j) rmax = std::max(rmax, std::fabs(pk->curve_data[13 * ((size_t)coff + k) + j]));
        }
        group_pf[2 * (size_t)g] = rmax;
        group_pf[2 * (size_t)g + 1] = swmax;
    }
    // tiled signed distance: no group's shape-to-canvas map shrinks
    uint32_t sdf_tiled = 1;
    for (uint32_t g = 0; g < G; ++g) {
        const float *m = pk->group_shape_xform + 6 * (size_t)g;
        double a = m[0], b = m[1], c = m[3], d = m[4];
        double q = a * a + b * b + c * c + d * d, det = a * d - b * c;
        double smin = std::sqrt(std::max(0.5 * (q - std::sqrt(std::max(q * q - 4.0 * det * det, 0.0))), 0.0));
        if (!(smin >= 1.0 - 1.0e-6)) sdf_tiled = 0;
    }

    Arena A;
    auto addv = [&](const void *p, size_t bytes) { return A.add(p, p ? bytes : 0); };
    size_t i_shape = addv(pk->shape_data, SP * 15 * 4), i_curve = addv(pk->curve_data, pk->curve_data_len * 4),
           i_sb = addv(pk->shape_bounds, SP * 16), i_gd = addv(pk->group_data, (size_t)G * 64),
           i_gx = addv(pk->group_xform, (size_t)G * 24), i_gsx = addv(pk->group_shape_xform, (size_t)G * 24),
           i_gis = addv(pk->group_inv_scale, (size_t)G * 4), i_gs = addv(pk->group_shapes, pk->group_shapes_len * 4),
           i_sx = addv(pk->shape_xform, SP * 24), i_grd = addv(pk->gradient_data, pk->gradient_data_len * 4),
           i_so = addv(pk->stop_offsets, pk->stop_offsets_len * 4),
           i_sc = addv(pk->stop_colors, pk->stop_offsets_len * 16), i_rng = addv(ranges.data(), ranges.size() * 4),
           i_toff = addv(nullptr, 0), i_tent = addv(nullptr, 0), i_tcnt = addv(nullptr, 0), i_tend = addv(nullptr, 0),
           i_tbwd = addv(nullptr, 0), i_fend = addv(nullptr, 0);
    // tile work buffers (no host source)
    A.srcs[i_toff].second = ((size_t)ntiles + 1) * 4;
    // room for the lists to grow when a device-side update moves shapes
    uint64_t tile_cap = std::max<uint64_t>(2 * n_entries + 4ull * ntiles, 1);
    A.srcs[i_tent].second = tile_cap * 4;
    A.srcs[i_tcnt].second = (size_t)std::max(ntiles, 1u) * 4;
    A.srcs[i_tend].second = (size_t)std::max(ntiles, 1u) * 4;
    A.srcs[i_tbwd].second = tile_cap * 4;
    A.srcs[i_fend].second = (size_t)std::max(ntiles, 1u) * 4;
    // recompute offsets after resizing the sourceless entries
    {
        size_t t = 0;
        for (size_t k = 0; k < A.offs.size(); ++k) {
            size_t o = (t + 255) & ~(size_t)255;
            A.offs[k] = o;
            t = o + std::max(A.srcs[k].second, (size_t)4);
        }
        A.total = t;
    }
    const float *bg_src = prm->background_image;
    size_t i_bg = addv(bg_src, bg_src ? (size_t)W * H * 16 : 0);
    size_t i_p2s = 0, i_st = 0, i_pp = 0, i_pc = 0, i_pt = 0, i_spo = 0, i_spc = 0, i_sco = 0, i_scc = 0, i_sto = 0,
           i_stc = 0, i_scl = 0, i_sl = 0, i_cdf = 0, i_pmf = 0, i_sid = 0, i_gid = 0, i_pcdf = 0, i_ppmf = 0,
           i_ppid = 0, i_pco = 0, i_pcc = 0, i_ppo = 0, i_s2p = 0;
    uint32_t n_cdf = 0;
    if (bp) {
        i_p2s = addv(bp->prepared_to_scene, SP * 4);
        i_st = addv(bp->shape_transform, (size_t)S * 24);
        i_pp = addv(bp->path_points, bp->path_points_len * 4);
        i_pc = addv(bp->path_num_control_points, bp->path_num_control_points_len * 4);
        i_pt = addv(bp->path_thickness, bp->path_thickness_len * 4);
        i_spo = addv(bp->shape_path_offsets, (size_t)S * 4);
        i_spc = addv(bp->shape_path_point_counts, (size_t)S * 4);
        i_sco = addv(bp->shape_path_ctrl_offsets, (size_t)S * 4);
        i_scc = addv(bp->shape_path_ctrl_counts, (size_t)S * 4);
        i_sto = addv(bp->shape_path_thickness_offsets, (size_t)S * 4);
        i_stc = addv(bp->shape_path_thickness_counts, (size_t)S * 4);
        i_scl = addv(bp->shape_path_is_closed, (size_t)S * 4);
        n_cdf = (uint32_t)bp->shape_cdf_len;
        if (n_cdf > 0 && n_cdf != SP)
            return set_err(ctx, DVG_INVALID_SCENE, "boundary tables do not match the prepared shapes");
        i_sl = addv(bp->shape_lengths, n_cdf ? (size_t)S * 4 : 0);
        i_cdf = addv(bp->shape_cdf, (size_t)n_cdf * 4);
        i_pmf = addv(bp->shape_pmf, (size_t)n_cdf * 4);
        i_sid = addv(bp->shape_ids, (size_t)n_cdf * 4);
        i_gid = addv(bp->group_ids, (size_t)n_cdf * 4);
        i_pcdf = addv(bp->path_cdf, bp->path_cdf_len * 4);
        i_ppmf = addv(bp->path_pmf, bp->path_cdf_len * 4);
        i_ppid = addv(bp->path_point_ids, bp->path_point_ids_len * 4);
        i_pco = addv(bp->path_cdf_offsets, n_cdf ? (size_t)S * 4 : 0);
        i_pcc = addv(bp->path_cdf_counts, n_cdf ? (size_t)S * 4 : 0);
        i_ppo = addv(bp->path_point_offsets, n_cdf ? (size_t)S * 4 : 0);
        i_s2p = addv(s2p.data(), (size_t)std::max(S, 1u) * 4);
    }
    size_t i_gcb = addv(cpu_bounds.data(), cpu_bounds.size() * 4);
    size_t i_gpf = addv(group_pf.data(), group_pf.size() * 4);
    size_t i_gbm = addv(gbvh_meta.data(), gbvh_meta.size() * 4);
    size_t i_gbb = addv(pk->group_bvh_bounds, pk->group_bvh_bounds ? pk->group_bvh_bounds_len * 4 : 0);
    size_t i_gbn = addv(pk->group_bvh_nodes, pk->group_bvh_nodes ? pk->group_bvh_nodes_len * 4 : 0);
    size_t i_gbi = addv(pk->group_bvh_indices, pk->group_bvh_indices ? pk->group_bvh_indices_len * 4 : 0);
    // plan of device-side updates: control-point indices of every curve and
    // the padding of every instance's bounds
    std::vector<uint32_t> curve_pt;
    std::vector<float> inst_pad;
    bool plan_ok = bp && make_update_plan(pk, bp, prm, curve_pt, inst_pad);
    size_t i_cpt = addv(plan_ok ? curve_pt.data() : nullptr, plan_ok ? curve_pt.size() * 4 : 0);
    size_t i_pad = addv(plan_ok ? inst_pad.data() : nullptr, plan_ok ? inst_pad.size() * 4 : 0);
    size_t i_cib = addv(nullptr, 0), i_flag = addv(nullptr, 0), i_wind = addv(nullptr, 0);
    A.srcs[i_cib].second = (size_t)std::max(SP, 1u) * 16;
    A.srcs[i_flag].second = 4;
    A.srcs[i_wind].second = std::max<size_t>(ncurves, 1) * dvgk::WIND_STRIDE * 4;
    {  // offsets again: sizes of sourceless entries changed
        size_t t = 0;
        for (size_t k = 0; k < A.offs.size(); ++k) {
            size_t o = (t + 255) & ~(size_t)255;
            A.offs[k] = o;
            t = o + std::max(A.srcs[k].second, (size_t)4);
        }
        A.total = t;
    }

    dvg_device_scene *ds = new dvg_device_scene();
    hipError_t e = hipMalloc(&ds->arena, A.total);
    if (e != hipSuccess) {
        delete ds;
        return hip_err(ctx, e, "hipMalloc(scene arena)");
    }
    ds->arena_bytes = A.total;
    // stage everything in one pinned-free host buffer, then one copy
    std::vector<uint8_t> host(A.total, 0);
    for (size_t k = 0; k < A.srcs.size(); ++k)
        if (A.srcs[k].first && A.srcs[k].second)
            std::memcpy(host.data() + A.offs[k], A.srcs[k].first, A.srcs[k].second);
    e = hipMemcpyAsync(ds->arena, host.data(), A.total, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(ds->arena);
        delete ds;
        return hip_err(ctx, e, "scene upload");
    }
    uint8_t *base = (uint8_t *)ds->arena;
    auto at = [&](size_t i) { return (void *)(base + A.offs[i]); };
    dvgk::DevScene &D = ds->S;
    std::memset(&D, 0, sizeof(D));
    D.W = W; D.H = H; D.nsx = nsx; D.nsy = nsy; D.spp = nsx * nsy; D.seed = prm->seed;
    D.jitter = prm->use_prefiltering ? 0u : prm->jitter;
    D.prefilter = prm->use_prefiltering;
    D.filter_type = prm->filter_type;
    D.sblock_log2 = 2;  // 4x4 pixel sample blocks
    D.bnd_log2 = 1;     // 2x2 pixel boundary buckets (1024x1024 step 4.63 -> 4.59 ms against 4x4; 256x256 / 64 paths 6.14 -> 6.01)
    // tile work in units of one sample per lane: many more workgroups than
    // compute units even for small canvases, and no tile on the critical path
    D.unit_samples = std::min((uint32_t)dvgk::BLOCK, (uint32_t)dvgk::TILE * dvgk::TILE * D.spp);
    D.filter_radius = prm->filter_radius;
    D.ri = (int)std::fmax(std::ceil(prm->filter_radius), 0.0f);
    for (int c = 0; c < 4; ++c) D.bg[c] = prm->background[c];
    D.bg_image = bg_src ? (dvgk::kfloat *)at(i_bg) : nullptr;
    D.num_groups = G;
    D.num_prepared = SP;
    D.shape_data = (dvgk::kfloat *)at(i_shape);
    D.curve_data = (dvgk::kfloat *)at(i_curve);
    D.curve_wind = (dvgk::kfloat *)at(i_wind);
    D.shape_bounds = (dvgk::kfloat *)at(i_sb);
    D.group_data = (dvgk::kfloat *)at(i_gd);
    D.group_xform = (dvgk::kfloat *)at(i_gx);
    D.group_shape_xform = (dvgk::kfloat *)at(i_gsx);
    D.group_inv_scale = (dvgk::kfloat *)at(i_gis);
    D.group_shapes = (dvgk::kfloat *)at(i_gs);
    D.shape_xform = (dvgk::kfloat *)at(i_sx);
    D.gradient_data = (dvgk::kfloat *)at(i_grd);
    D.stop_offsets = (dvgk::kfloat *)at(i_so);
    D.stop_colors = (dvgk::kfloat *)at(i_sc);
    D.tiles_x = tiles_x;
    D.tiles_y = tiles_y;
    D.tile_offsets = (dvgk::kuint *)at(i_toff);
    D.tile_entries = (dvgk::kuint *)at(i_tent);
    D.tile_end = (dvgk::kuint *)at(i_tend);
    D.bwd_entries = (dvgk::kuint *)at(i_tbwd);
    D.fwd_end = (dvgk::kuint *)at(i_fend);
    if (bp) {
        D.num_shapes = S;
        D.prepared_to_scene = (dvgk::kuint *)at(i_p2s);
        D.shape_transform = (dvgk::kfloat *)at(i_st);
        D.path_points = (dvgk::kfloat *)at(i_pp);
        D.path_ctrl = (dvgk::kuint *)at(i_pc);
        D.path_thickness = (dvgk::kfloat *)at(i_pt);
        D.sp_off = (dvgk::kuint *)at(i_spo);
        D.sp_cnt = (dvgk::kuint *)at(i_spc);
        D.sc_off = (dvgk::kuint *)at(i_sco);
        D.sc_cnt = (dvgk::kuint *)at(i_scc);
        D.st_off = (dvgk::kuint *)at(i_sto);
        D.st_cnt = (dvgk::kuint *)at(i_stc);
        D.s_closed = (dvgk::kuint *)at(i_scl);
        D.n_cdf = n_cdf;
        D.shape_lengths = (dvgk::kfloat *)at(i_sl);
        D.shape_cdf = (dvgk::kfloat *)at(i_cdf);
        D.shape_pmf = (dvgk::kfloat *)at(i_pmf);
        D.shape_ids = (dvgk::kuint *)at(i_sid);
        D.group_ids = (dvgk::kuint *)at(i_gid);
        D.path_cdf = (dvgk::kfloat *)at(i_pcdf);
        D.path_pmf = (dvgk::kfloat *)at(i_ppmf);
        D.path_point_ids = (dvgk::kuint *)at(i_ppid);
        D.path_cdf_off = (dvgk::kuint *)at(i_pco);
        D.path_cdf_cnt = (dvgk::kuint *)at(i_pcc);
        D.path_point_off = (dvgk::kuint *)at(i_ppo);
        D.scene_to_prepared = (dvgk::kuint *)at(i_s2p);
    }
    D.group_cpu_bounds = (dvgk::kfloat *)at(i_gcb);
    D.group_pf = (dvgk::kfloat *)at(i_gpf);
    D.gbvh_meta = (dvgk::kuint *)at(i_gbm);
    D.gbvh_bounds = (dvgk::kfloat *)at(i_gbb);
    D.gbvh_nodes = (dvgk::kuint *)at(i_gbn);
    D.gbvh_indices = (dvgk::kuint *)at(i_gbi);
    D.sdf_tiled = sdf_tiled;
    if (plan_ok) {
        dvgk::UpdatePlan &P = ds->plan;
        P.ncurves = (uint32_t)(curve_pt.size() / 4);
        P.n_points = (uint32_t)(bp->path_points_len / 2);
        P.tile_cap = (uint32_t)tile_cap;
        P.curve_pt = (const uint32_t *)at(i_cpt);
        P.inst_pad = (const float *)at(i_pad);
        P.cpu_inst_bounds = (float *)at(i_cib);
        P.ranges = (int32_t *)at(i_rng);
        P.tile_counts = (uint32_t *)at(i_tcnt);
        P.overflow = (uint32_t *)at(i_flag);
    }
    ds->num_groups = G;
    ds->num_shapes = S;
    ds->list_max = list_max;
    ds->ncurves = (uint32_t)ncurves;
    ds->gcap = prm->use_prefiltering ? 0u : dvgk::interior_gcap(list_max);
    {
        size_t npix = (size_t)W * H;
        uint64_t lens[14] = {8ull * S,
                             bp ? bp->path_points_len : 0,
                             bp ? bp->path_thickness_len : 0,
                             S,
                             9ull * S,
                             9ull * G,
                             16ull * G,
                             pk->gradient_data_len,
                             pk->stop_offsets_len,
                             4 * pk->stop_offsets_len,
                             1,
                             4,
                             bg_src ? npix * 4 : 0,
                             npix * 2};
        for (int k = 0; k < 14; ++k) ds->grad_len[k] = (uint32_t)lens[k];
    }
    dvgk::launch_wind_coef(D.curve_data, (uint32_t)ncurves, (float *)at(i_wind), D.shape_data, D.num_prepared,
                           ctx->stream);
    // tile lists: count, scan, write (sorted by group id by construction)
    if (ntiles > 0) {
        dvgk::launch_bin_count((const float *)at(i_rng), G, tiles_x, tiles_y, (uint32_t *)at(i_tcnt), ctx->stream);
        dvgk::launch_scan((const uint32_t *)at(i_tcnt), ntiles, (uint32_t *)at(i_toff), ctx->stream);
        dvgk::launch_bin_write((const float *)at(i_rng), G, tiles_x, tiles_y, (const uint32_t *)at(i_toff),
                               (uint32_t *)at(i_tent), ctx->stream);
        dvgk::launch_tile_cull(D, (uint32_t *)at(i_tent), (uint32_t *)at(i_tbwd), (uint32_t *)at(i_tend),
                               (uint32_t *)at(i_fend), (uint32_t)tile_cap, true, ctx->stream);
        e = hipGetLastError();
        if (e != hipSuccess) {
            (void)hipFree(ds->arena);
            delete ds;
            return hip_err(ctx, e, "tile binning launch");
        }
    }
    *out = ds;
    return DVG_OK;
}

void dvg_device_scene_destroy(dvg_device_scene *ds) {
    if (!ds) return;
    if (ds->arena) (void)hipFree(ds->arena);
    delete ds;
}

// Renderer::render_sdf / eval_positions (renderer.rs:586-848) on a device
// scene; the signed distance needs its backward tables.
int32_t dvg_device_render_sdf(dvg_ctx *ctx, dvg_device_scene *ds, float *out) {
    if (!ctx || !ds || !out) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if (!ds->S.prepared_to_scene || !ds->S.shape_transform)
        return set_err(ctx, DVG_INVALID_SCENE, "device scene created without a backward pack");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const dvgk::DevScene &S = ds->S;
    size_t bytes = dvgk::sdf_work_bytes(S.num_prepared, ds->ncurves, S.tiles_x * S.tiles_y);
    int st = ensure(ctx, ctx->sdfw, ctx->sdfw_cap, (bytes + 3) / 4);
    if (st) return st;
    dvgk::launch_sdf(S, ds->ncurves, ctx->sdfw, out, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    return DVG_OK;
}

int32_t dvg_device_eval_positions(dvg_ctx *ctx, dvg_device_scene *ds, const float *positions, uint64_t count,
                                  float *out) {
    if (!ctx || !ds || (count && (!positions || !out))) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if (count > 0xffffffffull) return set_err(ctx, DVG_INVALID_SCENE, "too many positions for 1d launch");
    if (!ds->S.prepared_to_scene || !ds->S.shape_transform)
        return set_err(ctx, DVG_INVALID_SCENE, "device scene created without a backward pack");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const dvgk::DevScene &S = ds->S;
    size_t bytes = dvgk::sdf_work_bytes(S.num_prepared, ds->ncurves, S.tiles_x * S.tiles_y);
    int st = ensure(ctx, ctx->sdfw, ctx->sdfw_cap, (bytes + 3) / 4);
    if (st) return st;
    dvgk::launch_sdf_points(S, ds->ncurves, ctx->sdfw, positions, (uint32_t)count, out, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    return DVG_OK;
}

int32_t dvg_device_render(dvg_ctx *ctx, dvg_device_scene *ds, float *out) {
    if (!ctx || !ds || !out) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    const dvgk::DevScene &S = ds->S;
    size_t npix = (size_t)S.W * S.H;
    if (npix == 0) return DVG_OK;
    int st = ensure(ctx, ctx->weight, ctx->weight_cap, npix);
    if (st) return st;
    st = ensure(ctx, ctx->accum, ctx->accum_cap, npix * 4);
    if (st) return st;
    {
        float *p = ctx->accum;
        uint32_t n = (uint32_t)(npix * 4);
        KernelTimer kt(ctx, DVG_KERNEL_ZERO);
        dvgk::launch_zero(&p, &n, 1, ctx->stream);
    }
    {
        KernelTimer kt(ctx, DVG_KERNEL_WEIGHTS);
        dvgk::launch_weights(S, ctx->weight, false, ctx->stream);
    }
    {
        KernelTimer kt(ctx, DVG_KERNEL_FORWARD);
        dvgk::launch_forward(S, ctx->weight, ctx->accum, ctx->stream);
    }
    {
        KernelTimer kt(ctx, DVG_KERNEL_RESOLVE);
        dvgk::launch_resolve(S, ctx->weight, ctx->accum, out, ctx->stream);
    }
    HIP_TRY(ctx, hipGetLastError());
    return DVG_OK;
}

static void grads_to_dev(const dvg_grads *g, dvgk::DevGrads &G) {
    G.d_shape_params = g->d_shape_params;
    G.d_shape_points = g->d_shape_points;
    G.d_shape_thickness = g->d_shape_thickness;
    G.d_shape_stroke_width = g->d_shape_stroke_width;
    G.d_shape_transform = g->d_shape_transform;
    G.d_group_transform = g->d_group_transform;
    G.d_group_data = g->d_group_data;
    G.d_gradient_data = g->d_gradient_data;
    G.d_stop_offsets = g->d_stop_offsets;
    G.d_stop_colors = g->d_stop_colors;
    G.d_filter_radius = g->d_filter_radius;
    G.d_background = g->d_background;
    G.d_background_image = g->d_background_image;
    G.d_translation = g->d_translation;
}

// Renderer::eval_positions_backward on a device scene: zeroes the device
// gradient arrays, then accumulates the points' gradients.
int32_t dvg_device_eval_positions_backward(dvg_ctx *ctx, dvg_device_scene *ds, const float *positions,
                                           uint64_t count, const float *d_sdf, const dvg_grads *grads) {
    if (!ctx || !ds || !grads || (count && (!positions || !d_sdf)))
        return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if (count > 0xffffffffull) return set_err(ctx, DVG_INVALID_SCENE, "too many positions for 1d launch");
    if (!ds->S.prepared_to_scene || !ds->S.shape_transform)
        return set_err(ctx, DVG_INVALID_SCENE, "device scene created without a backward pack");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    dvgk::DevGrads G;
    grads_to_dev(grads, G);
    dvgk::launch_zero(&G.d_shape_params, ds->grad_len, 14, ctx->stream);
    if (ds->S.num_groups > 0) dvgk::launch_sdf_points_backward(ds->S, positions, (uint32_t)count, d_sdf, G, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    return DVG_OK;
}

// The d_sdf image term of render_backward on a device scene (W*H device
// adjoints): zeroes the device gradient arrays, then accumulates.
int32_t dvg_device_render_backward_sdf(dvg_ctx *ctx, dvg_device_scene *ds, const float *d_sdf,
                                       const dvg_grads *grads) {
    if (!ctx || !ds || !d_sdf || !grads) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if (!ds->S.prepared_to_scene || !ds->S.shape_transform)
        return set_err(ctx, DVG_INVALID_SCENE, "device scene created without a backward pack");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    dvgk::DevGrads G;
    grads_to_dev(grads, G);
    dvgk::launch_zero(&G.d_shape_params, ds->grad_len, 14, ctx->stream);
    if (ds->S.num_groups > 0) dvgk::launch_sdf_image_backward(ds->S, d_sdf, G, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    return DVG_OK;
}


int32_t dvg_device_scene_update(dvg_ctx *ctx, dvg_device_scene *ds, const float *path_points,
                                const float *group_colors) {
    if (!ctx || !ds) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if (!ds->plan.curve_pt && !ds->plan.ranges)
        return set_err(ctx, DVG_INVALID_SCENE, "scene was created without the backward tables an update needs");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    dvgk::launch_scene_update(ds->S, ds->plan, path_points, group_colors, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    return DVG_OK;
}

int32_t dvg_device_scene_check(dvg_ctx *ctx, dvg_device_scene *ds) {
    if (!ctx || !ds) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if (!ds->plan.overflow) return DVG_OK;
    uint32_t flag = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&flag, ds->plan.overflow, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (flag) return set_err(ctx, DVG_DEVICE_ERROR, "tile lists outgrew their buffer after an update; recreate the scene");
    return DVG_OK;
}

int32_t dvg_device_scene_set_shapes(dvg_device_scene *ds, uint32_t sblock_log2, uint32_t bnd_log2,
                                    uint32_t unit_log2) {
    if (!ds || sblock_log2 > 4 || bnd_log2 < 1 || bnd_log2 > 4 || unit_log2 > 8) return DVG_INVALID_SCENE;
    dvgk::DevScene &S = ds->S;
    S.sblock_log2 = sblock_log2;
    S.bnd_log2 = bnd_log2;
    uint64_t us = (uint64_t)dvgk::BLOCK << unit_log2, tile = (uint64_t)dvgk::TILE * dvgk::TILE * S.spp;
    S.unit_samples = (uint32_t)std::min(us, tile);
    return DVG_OK;
}

int32_t dvg_device_render_backward(dvg_ctx *ctx, dvg_device_scene *ds, const float *d_render,
                                   const dvg_grads *grads) {
    if (!ctx || !ds || !d_render || !grads) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    const dvgk::DevScene &S = ds->S;
    size_t npix = (size_t)S.W * S.H;
    if (npix == 0 || S.num_groups == 0) return DVG_OK;
    int st = ensure(ctx, ctx->weight, ctx->weight_cap, npix);
    if (st) return st;
    dvgk::DevGrads G;
    grads_to_dev(grads, G);
    // the outputs are accumulations into zeroed arrays
    {
        KernelTimer kt(ctx, DVG_KERNEL_ZERO);
        dvgk::launch_zero(&G.d_shape_params, ds->grad_len, 14, ctx->stream);
    }
    {
        KernelTimer kt(ctx, DVG_KERNEL_WEIGHTS);
        dvgk::launch_weights(S, ctx->weight, true, ctx->stream);
    }
    if (S.prefilter) {
        // analytic coverage: no boundary term (backward/render.rs:192-206)
        KernelTimer kt(ctx, DVG_KERNEL_INTERIOR);
        dvgk::launch_backward_prefiltered(S, ctx->weight, d_render, G, ctx->stream);
        return DVG_OK;
    }
    dvgk::InteriorWork iw = {nullptr, nullptr, 0u, 0u};
    if (ds->gcap > 0) {
        // bit-set slots for the lists the LDS bit set cannot hold
        uint32_t slots = std::max(ctx->n_cu, 1u) * dvgk::INTERIOR_WG_PER_CU;
        size_t need = 1 + (size_t)slots * dvgk::interior_slot_words(ds->gcap);
        if (ctx->imask_cap < need) {
            if (ctx->imask) (void)hipFree(ctx->imask);
            ctx->imask = nullptr;
            ctx->imask_cap = 0;
            HIP_TRY(ctx, hipMalloc(&ctx->imask, need * sizeof(uint32_t)));
            ctx->imask_cap = need;
        }
        iw = {ctx->imask + 1, ctx->imask, ds->gcap, slots};
    }
    // exact boundary mode: the boundary pass first, its uncertain samples
    // again in f64 on the side stream beside the colour adjoint
    bool exact = ctx->exact_boundary && S.n_cdf > 0;
    if (!exact) {
        KernelTimer kt(ctx, DVG_KERNEL_INTERIOR);
        dvgk::launch_backward_interior(S, ctx->weight, d_render, G, iw, ctx->stream);
    }
    bool redo = false;
    if (S.n_cdf > 0) {
        size_t ns = npix * S.spp, nt = dvgk::boundary_buckets(S);
        size_t units = (ns + dvgk::BLOCK - 1) / dvgk::BLOCK;
        size_t nsum = dvgk::scan_chunks((uint32_t)nt + 1);
        size_t need = 3 * ns + 2 * nt + 3 + nsum + (exact ? units * (1 + dvgk::FLAG_UNIT_SLOTS) : 0);
        if (ctx->bwork_cap < need) {
            if (ctx->bwork) (void)hipFree(ctx->bwork);
            ctx->bwork = nullptr;
            ctx->bwork_cap = 0;
            HIP_TRY(ctx, hipMalloc(&ctx->bwork, need * sizeof(uint32_t)));
            ctx->bwork_cap = need;
        }
        dvgk::BoundaryWork bw;
        bw.keys = ctx->bwork;
        bw.ranks = bw.keys + ns;
        bw.order = bw.ranks + ns;
        bw.counts = bw.order + ns;
        bw.offsets = bw.counts + nt + 1;
        bw.chunk_sums = bw.offsets + nt + 2;
        bw.flag_count = exact ? bw.chunk_sums + nsum : nullptr;
        bw.flag_slots = exact ? bw.flag_count + units : nullptr;
        {
            KernelTimer kt(ctx, DVG_KERNEL_BOUNDARY_SORT);
            dvgk::launch_boundary_sort(S, bw, ctx->stream);
        }
        {
            KernelTimer kt(ctx, DVG_KERNEL_BOUNDARY_EVAL);
            dvgk::launch_boundary_eval(S, ctx->weight, d_render, G, bw, ctx->stream);
        }
        if (exact) {
            // a few hundred latency-bound waves: they run beside the colour
            // adjoint instead of after it
            HIP_TRY(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
            HIP_TRY(ctx, hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
            dvgk::launch_boundary_redo(S, ctx->weight, d_render, G, bw, ctx->side);
            redo = true;
            // from here the side stream holds work that writes G and bwork:
            // every return below joins it first
            hipError_t e = hipEventRecord(ctx->ev_join, ctx->side);
            if (e != hipSuccess) {
                (void)hipStreamSynchronize(ctx->side);
                return hip_err(ctx, e, "hipEventRecord(ev_join)");
            }
        }
    }
    if (exact) {
        KernelTimer kt(ctx, DVG_KERNEL_INTERIOR);
        dvgk::launch_backward_interior(S, ctx->weight, d_render, G, iw, ctx->stream);
    }
    if (redo) {
        hipError_t e = hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(ctx->side);
            return hip_err(ctx, e, "hipStreamWaitEvent(ev_join)");
        }
    }
    HIP_TRY(ctx, hipGetLastError());
    return DVG_OK;
}

}  // extern "C"

// ---- synchronous host-pointer entry points ---------------------------------

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct GradLayout {
    size_t off[14], len[14], total = 0;
};

// Lengths of the 14 gradient arrays (backward_gpu.rs:187-200).
GradLayout grad_layout(const dvg_scene_pack *pk, const dvg_backward_pack *bp, const dvg_render_params *prm) {
    GradLayout L;
    size_t S = bp ? bp->num_shapes : 0, G = pk->num_groups;
    size_t npix = (size_t)prm->width * prm->height;
    size_t lens[14] = {8 * S,
                       bp ? bp->path_points_len : 0,
                       bp ? bp->path_thickness_len : 0,
                       S,
                       9 * S,
                       9 * G,
                       16 * G,
                       pk->gradient_data_len,
                       pk->stop_offsets_len,
                       4 * pk->stop_offsets_len,
                       1,
                       4,
                       prm->background_image ? npix * 4 : 0,
                       prm->compute_translation ? npix * 2 : 0};
    for (int k = 0; k < 14; ++k) {
        L.off[k] = L.total;
        L.len[k] = lens[k];
        L.total += (lens[k] + 63) & ~(size_t)63;
    }
    return L;
}

float **grad_slot(const dvg_grads *g, int k) {
    float *const *p = &g->d_shape_params;
    return const_cast<float **>(p + k);
}

int render_backward_host(dvg_ctx *ctx, const dvg_scene_pack *pk, const dvg_backward_pack *bp,
                         const dvg_render_params *prm, const float *d_render, const dvg_grads *grads) {
    size_t npix = (size_t)prm->width * prm->height;
    GradLayout L = grad_layout(pk, bp, prm);
    // zero the caller's arrays first (outputs are accumulated into zeroed buffers)
    for (int k = 0; k < 14; ++k) {
        float *dst = *grad_slot(grads, k);
        if (dst && L.len[k]) std::memset(dst, 0, L.len[k] * sizeof(float));
    }
    if (npix == 0 || pk->num_groups == 0) return DVG_OK;
    dvg_device_scene *ds = nullptr;
    int st = dvg_device_scene_create(ctx, pk, bp, prm, &ds);
    if (st) return st;
    DevBuf gbuf, dr;
    hipError_t e = hipMalloc(&gbuf.p, std::max(L.total, (size_t)1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dr.p, npix * 16);
    if (e == hipSuccess) e = hipMemsetAsync(gbuf.p, 0, std::max(L.total, (size_t)1) * sizeof(float), ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dr.p, d_render, npix * 16, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        dvg_device_scene_destroy(ds);
        return hip_err(ctx, e, "backward buffers");
    }
    dvg_grads dg;
    float *gb = (float *)gbuf.p;
    for (int k = 0; k < 14; ++k) *grad_slot(&dg, k) = L.len[k] ? gb + L.off[k] : nullptr;
    if (!(*grad_slot(grads, 12))) dg.d_background_image = nullptr;
    if (!(*grad_slot(grads, 13))) dg.d_translation = nullptr;
    st = dvg_device_render_backward(ctx, ds, (const float *)dr.p, &dg);
    if (st == DVG_OK) {
        std::vector<float> host(L.total);
        e = hipMemcpyAsync(host.data(), gbuf.p, L.total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = hip_err(ctx, e, "gradient readback");
        else
            for (int k = 0; k < 14; ++k) {
                float *dst = *grad_slot(grads, k);
                if (dst && L.len[k]) std::memcpy(dst, host.data() + L.off[k], L.len[k] * sizeof(float));
            }
    }
    dvg_device_scene_destroy(ds);
    return st;
}

}  // namespace

extern "C" {

int32_t dvg_render_packed(dvg_ctx *ctx, const dvg_scene_pack *pk, const dvg_render_params *prm, float *out) {
    if (!ctx || !pk || !prm || !out) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    size_t npix = (size_t)prm->width * prm->height;
    if (npix == 0) return DVG_OK;
    dvg_device_scene *ds = nullptr;
    int st = dvg_device_scene_create(ctx, pk, nullptr, prm, &ds);
    if (st) return st;
    DevBuf ob;
    hipError_t e = hipMalloc(&ob.p, npix * 16);
    if (e != hipSuccess) {
        dvg_device_scene_destroy(ds);
        return hip_err(ctx, e, "hipMalloc(output)");
    }
    st = dvg_device_render(ctx, ds, (float *)ob.p);
    if (st == DVG_OK) {
        e = hipMemcpyAsync(out, ob.p, npix * 16, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = hip_err(ctx, e, "image readback");
    }
    dvg_device_scene_destroy(ds);
    return st;
}

int32_t dvg_render_backward_packed(dvg_ctx *ctx, const dvg_scene_pack *pk, const dvg_backward_pack *bp,
                                   const dvg_render_params *prm, const float *d_render, const dvg_grads *grads) {
    if (!ctx || !pk || !bp || !prm || !d_render || !grads) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return render_backward_host(ctx, pk, bp, prm, d_render, grads);
}

int32_t dvg_render(dvg_ctx *ctx, const dvg_scene_desc *scene, const dvg_render_options *options, float *out) {
    if (!ctx || !scene || !options || !out) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    // Renderer::render early-outs (renderer.rs:40-66)
    size_t npix = (size_t)scene->width * scene->height;
    if (scene->background_image && scene->background_image_len != npix * 4)
        return set_err(ctx, DVG_INVALID_SCENE, "background image size mismatch");
    if (npix == 0) return DVG_OK;
    if (scene->num_groups == 0) {
        for (size_t i = 0; i < npix; ++i)
            for (int c = 0; c < 4; ++c)
                out[i * 4 + c] = scene->background_image ? scene->background_image[i * 4 + c] : scene->background[c];
        return DVG_OK;
    }
    dvg_prepared P;
    const char *err = "";
    int st = dvgh::prepare(scene, options, P, &err);
    if (st) return set_err(ctx, st, err);
    dvg_scene_pack pk;
    dvg_render_params prm;
    dvgh::pack_views(P, &pk, nullptr, &prm);
    return dvg_render_packed(ctx, &pk, &prm, out);
}

// Renderer::render_sdf (renderer.rs:586-727) and eval_positions
// (renderer.rs:729-848): the scene is packed, uploaded with its backward
// tables, evaluated and read back.
namespace {
int sdf_host(dvg_ctx *ctx, const dvg_scene_desc *scene, const dvg_render_options *options, const float *positions,
             uint64_t count, float *out) {
    dvg_prepared P;
    const char *err = "";
    int st = dvgh::prepare(scene, options, P, &err);
    if (st) return set_err(ctx, st, err);
    dvg_scene_pack pk;
    dvg_backward_pack bp;
    dvg_render_params prm;
    dvgh::pack_views(P, &pk, &bp, &prm);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    dvg_device_scene *ds = nullptr;
    st = dvg_device_scene_create(ctx, &pk, &bp, &prm, &ds);
    if (st) return st;
    size_t n = positions ? (size_t)count : (size_t)scene->width * scene->height;
    DevBuf ob, pb;
    hipError_t e = hipMalloc(&ob.p, n * sizeof(float));
    if (e == hipSuccess && positions) {
        e = hipMalloc(&pb.p, n * 2 * sizeof(float));
        if (e == hipSuccess)
            e = hipMemcpyAsync(pb.p, positions, n * 2 * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    }
    if (e != hipSuccess) {
        dvg_device_scene_destroy(ds);
        return hip_err(ctx, e, "sdf buffers");
    }
    st = positions ? dvg_device_eval_positions(ctx, ds, (const float *)pb.p, count, (float *)ob.p)
                   : dvg_device_render_sdf(ctx, ds, (float *)ob.p);
    if (st == DVG_OK) {
        e = hipMemcpyAsync(out, ob.p, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = hip_err(ctx, e, "sdf readback");
    }
    dvg_device_scene_destroy(ds);
    return st;
}
}  // namespace

// Host form of the signed-distance gradients: at `count` points, or
// (positions null) the d_sdf image term with count = W*H adjoints.
static int sdf_grad_host(dvg_ctx *ctx, const dvg_scene_desc *scene, const dvg_render_options *options,
                         const dvg_backward_options *bopt, const float *positions, uint64_t count,
                         const float *d_sdf, const dvg_grads *grads) {
    dvg_prepared P;
    const char *err = "";
    int st = dvgh::prepare(scene, options, P, &err);
    if (st) return set_err(ctx, st, err);
    dvg_scene_pack pk;
    dvg_backward_pack bp;
    dvg_render_params prm;
    dvgh::pack_views(P, &pk, &bp, &prm);
    prm.compute_translation = (bopt && bopt->compute_translation && grads->d_translation) ? 1u : 0u;
    GradLayout L = grad_layout(&pk, &bp, &prm);
    for (int k = 0; k < 14; ++k) {
        float *dst = *grad_slot(grads, k);
        if (dst && L.len[k]) std::memset(dst, 0, L.len[k] * sizeof(float));
    }
    if (count == 0 || scene->num_groups == 0) return DVG_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    dvg_device_scene *ds = nullptr;
    st = dvg_device_scene_create(ctx, &pk, &bp, &prm, &ds);
    if (st) return st;
    DevBuf gbuf, pb, db;
    hipError_t e = hipMalloc(&gbuf.p, std::max(L.total, (size_t)1) * sizeof(float));
    if (e == hipSuccess && positions) e = hipMalloc(&pb.p, count * 2 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&db.p, count * sizeof(float));
    if (e == hipSuccess && positions)
        e = hipMemcpyAsync(pb.p, positions, count * 2 * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(db.p, d_sdf, count * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        dvg_device_scene_destroy(ds);
        return hip_err(ctx, e, "positions backward buffers");
    }
    dvg_grads dg;
    float *gb = (float *)gbuf.p;
    for (int k = 0; k < 14; ++k) *grad_slot(&dg, k) = L.len[k] ? gb + L.off[k] : nullptr;
    if (!(*grad_slot(grads, 12))) dg.d_background_image = nullptr;
    if (!(bopt && bopt->compute_translation && *grad_slot(grads, 13))) dg.d_translation = nullptr;
    st = positions ? dvg_device_eval_positions_backward(ctx, ds, (const float *)pb.p, count, (const float *)db.p, &dg)
                   : dvg_device_render_backward_sdf(ctx, ds, (const float *)db.p, &dg);
    if (st == DVG_OK) {
        std::vector<float> host(std::max(L.total, (size_t)1));
        e = hipMemcpyAsync(host.data(), gbuf.p, L.total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = hip_err(ctx, e, "gradient readback");
        else
            for (int k = 0; k < 14; ++k) {
                float *dst = *grad_slot(grads, k);
                if (dst && L.len[k] && *grad_slot(&dg, k)) std::memcpy(dst, host.data() + L.off[k], L.len[k] * sizeof(float));
            }
    }
    dvg_device_scene_destroy(ds);
    return st;
}

int32_t dvg_eval_positions_backward(dvg_ctx *ctx, const dvg_scene_desc *scene, const dvg_render_options *options,
                                    const dvg_backward_options *bopt, const float *positions, uint64_t count,
                                    const float *d_sdf, const dvg_grads *grads) {
    if (!ctx || !scene || !options || !grads || (count && (!positions || !d_sdf)))
        return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if (count > 0xffffffffull) return set_err(ctx, DVG_INVALID_SCENE, "too many positions for 1d launch");
    return sdf_grad_host(ctx, scene, options, bopt, positions, count, d_sdf, grads);
}

int32_t dvg_render_backward_sdf(dvg_ctx *ctx, const dvg_scene_desc *scene, const dvg_render_options *options,
                                const dvg_backward_options *bopt, const float *d_sdf, const dvg_grads *grads) {
    if (!ctx || !scene || !options || !d_sdf || !grads) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    uint64_t npix = (uint64_t)scene->width * scene->height;
    if (npix * std::max(options->samples_x, 1u) * std::max(options->samples_y, 1u) > 0xffffffffull)
        return set_err(ctx, DVG_INVALID_SCENE, "too many samples for 1d launch");
    return sdf_grad_host(ctx, scene, options, bopt, nullptr, npix, d_sdf, grads);
}

int32_t dvg_render_sdf(dvg_ctx *ctx, const dvg_scene_desc *scene, const dvg_render_options *options, float *out) {
    if (!ctx || !scene || !options || !out) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    size_t npix = (size_t)scene->width * scene->height;
    if (npix == 0) return DVG_OK;
    if (scene->num_groups == 0) {
        for (size_t i = 0; i < npix; ++i) out[i] = 0.0f;
        return DVG_OK;
    }
    return sdf_host(ctx, scene, options, nullptr, 0, out);
}

int32_t dvg_eval_positions(dvg_ctx *ctx, const dvg_scene_desc *scene, const dvg_render_options *options,
                           const float *positions, uint64_t count, float *out) {
    if (!ctx || !scene || !options || (count && (!positions || !out)))
        return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if (count == 0) return DVG_OK;
    if (count > 0xffffffffull) return set_err(ctx, DVG_INVALID_SCENE, "too many positions for 1d launch");
    if (scene->num_groups == 0) {
        for (uint64_t i = 0; i < count; ++i) out[i] = 0.0f;
        return DVG_OK;
    }
    return sdf_host(ctx, scene, options, positions, count, out);
}

int32_t dvg_render_backward(dvg_ctx *ctx, const dvg_scene_desc *scene, const dvg_render_options *options,
                            const dvg_backward_options *bopt, const float *d_render, const dvg_grads *grads) {
    if (!ctx || !scene || !options || !d_render || !grads) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    size_t npix = (size_t)scene->width * scene->height;
    if (scene->background_image && scene->background_image_len != npix * 4)
        return set_err(ctx, DVG_INVALID_SCENE, "background image size mismatch");
    dvg_prepared P;
    const char *err = "";
    int st = dvgh::prepare(scene, options, P, &err);
    if (st) return set_err(ctx, st, err);
    dvg_scene_pack pk;
    dvg_backward_pack bp;
    dvg_render_params prm;
    dvgh::pack_views(P, &pk, &bp, &prm);
    prm.compute_translation = (bopt && bopt->compute_translation && grads->d_translation) ? 1u : 0u;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    st = render_backward_host(ctx, &pk, &bp, &prm, d_render, grads);
    if (st) return st;
    // finalize_background_gradients (background.rs:29-61)
    if (scene->background_image) {
        if (grads->d_background_image)
            for (size_t i = 0; i < npix; ++i) {
                float *c = grads->d_background_image + i * 4;
                const float *im = scene->background_image + i * 4;
                float dr = c[0] * im[3], dg = c[1] * im[3], db = c[2] * im[3];
                float da = c[0] * im[0] + c[1] * im[1] + c[2] * im[2] + c[3];
                c[0] = dr; c[1] = dg; c[2] = db; c[3] = da;
            }
    } else if (grads->d_background) {
        float *d = grads->d_background;
        const float *b = scene->background;
        float nr = d[0] * b[3], ng = d[1] * b[3], nb = d[2] * b[3], na = d[0] * b[0] + d[1] * b[1] + d[2] * b[2] + d[3];
        d[0] = nr; d[1] = ng; d[2] = nb; d[3] = na;
    }
    return DVG_OK;
}

}  // extern "C"
