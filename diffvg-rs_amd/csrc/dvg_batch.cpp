// Batches of scenes (include/dvg.h, dvg_batch_*): many uneven scenes run
// forward + loss gradient + backward with one persistent launch per phase,
// each draining a device work queue of (scene, unit) items; the step can be
// captured once in a HIP graph and replayed; launch shapes are chosen by a
// measuring autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/dvg.h"
#include "dvg_internal.h"
#include "dvg_kernels.h"

using dvgk::BatchScene;
using dvgk::WorkItem;

struct dvg_batch {
    int device = 0;
    uint32_t n = 0;
    std::vector<dvg_device_scene *> scenes;
    std::vector<BatchScene> bs;          // host copy of the device array
    BatchScene *dev_bs = nullptr;
    void *work = nullptr;                // every per-scene work buffer
    std::vector<const float *> targets;  // caller-owned target images (device)
    uint32_t *counters = nullptr;        // one work-queue counter per phase
    std::vector<WorkItem> items[dvgk::NUM_PHASES];
    WorkItem *dev_items = nullptr;
    size_t item_off[dvgk::NUM_PHASES] = {};
    size_t items_cap = 0;
    std::vector<std::vector<uint32_t>> tile_off;  // tile list offsets per scene
    std::vector<uint32_t> nbmax;                   // bucket capacity per scene
    int n_cu = 256;
    // colour-adjoint bit-set slots for tile lists longer than the LDS bit set
    uint32_t *gmask = nullptr;
    uint32_t gcap = 0;
    dvg_batch_config cfg = {2, 2, 4, 0, 0.f};
    bool use_graph = false;
    hipGraphExec_t exec = nullptr;
    uint32_t exec_flags = 0;
    hipStream_t exec_stream = nullptr;
};

namespace {

void drop_graph(dvg_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->exec_flags = 0;
}

uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Work items of every phase for the current launch shapes.  Tile units are
// ordered heaviest first over the whole batch (cost: group-list length times
// samples in the unit), so the longest units start first.
int build_items(dvg_ctx *ctx, dvg_batch *b) {
    for (auto &v : b->items) v.clear();
    struct Costed {
        uint64_t cost;
        WorkItem w;
    };
    std::vector<Costed> fwd_units, fwd_pf_units, bwd_units, pre_units;
    for (uint32_t i = 0; i < b->n; ++i) {
        const BatchScene &B = b->bs[i];
        const dvgk::DevScene &S = B.S;
        size_t npix = (size_t)S.W * S.H, ns = npix * S.spp;
        if (npix == 0 || S.num_groups == 0) continue;
        for (int k = 0; k < dvgk::BATCH_ZERO_ARRAYS; ++k) {
            uint64_t len = k < 15 ? B.zn[k] : (uint64_t)b->nbmax[i] + (k == 15 ? 1 : 0);
            if (!B.zp[k] || len == 0) continue;
            for (uint32_t c = 0; c < cdiv(len, dvgk::ZERO_CHUNK); ++c)
                b->items[dvgk::PH_ZERO].push_back({i, ((uint32_t)k << 20) | c});
        }
        uint32_t nblk16 = cdiv(S.W, 16) * cdiv(S.H, 16);
        for (uint32_t u = 0; u < nblk16; ++u) {
            b->items[dvgk::PH_WEIGHTS_FWD].push_back({i, u});
            b->items[dvgk::PH_WEIGHTS_BWD].push_back({i, u});
        }
        for (uint32_t u = 0; u < cdiv(npix, dvgk::BLOCK); ++u) b->items[dvgk::PH_RESOLVE].push_back({i, u});
        uint32_t nt = S.tiles_x * S.tiles_y, upt = dvgk::units_per_tile(S);
        uint32_t tile_samples = (uint32_t)dvgk::TILE * dvgk::TILE * S.spp;
        for (uint32_t t = 0; t < nt; ++t) {
            uint64_t len = b->tile_off[i][t + 1] - b->tile_off[i][t] + 1;
            for (uint32_t k = 0; k < upt; ++k) {
                uint32_t ns_unit = std::min(S.unit_samples, tile_samples - k * S.unit_samples);
                Costed c = {len * ns_unit, {i, t * upt + k}};
                (S.prefilter ? fwd_pf_units : fwd_units).push_back(c);
                (S.prefilter ? pre_units : bwd_units).push_back(c);
            }
        }
        if (B.keys) {
            uint32_t nu = cdiv(ns, dvgk::BLOCK);
            for (uint32_t u = 0; u < nu; ++u) {
                b->items[dvgk::PH_BND_KEY].push_back({i, u});
                b->items[dvgk::PH_BND_SCATTER].push_back({i, u});
                b->items[dvgk::PH_BND_EVAL].push_back({i, u});
            }
            b->items[dvgk::PH_SCAN].push_back({i, 0});
        }
    }
    auto heavy_first = [](std::vector<Costed> &v, std::vector<WorkItem> &outv) {
        std::stable_sort(v.begin(), v.end(), [](const Costed &a, const Costed &c) { return a.cost > c.cost; });
        for (const Costed &c : v) outv.push_back(c.w);
    };
    heavy_first(fwd_units, b->items[dvgk::PH_FORWARD]);
    heavy_first(fwd_pf_units, b->items[dvgk::PH_FORWARD_PF]);
    heavy_first(bwd_units, b->items[dvgk::PH_INTERIOR]);
    heavy_first(pre_units, b->items[dvgk::PH_PREFILTERED]);
    size_t n_items = 0;
    for (int ph = 0; ph < dvgk::NUM_PHASES; ++ph) {
        b->item_off[ph] = n_items;
        n_items += b->items[ph].size();
    }
    if (n_items > b->items_cap) {
        if (b->dev_items) (void)hipFree(b->dev_items);
        b->dev_items = nullptr;
        b->items_cap = 0;
        HIP_TRY(ctx, hipMalloc(&b->dev_items, std::max<size_t>(n_items, 1) * sizeof(WorkItem)));
        b->items_cap = n_items;
    }
    for (int ph = 0; ph < dvgk::NUM_PHASES; ++ph) {
        if (b->items[ph].empty()) continue;
        HIP_TRY(ctx, hipMemcpy(b->dev_items + b->item_off[ph], b->items[ph].data(),
                               b->items[ph].size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    }
    return DVG_OK;
}

// Pushes the current launch shapes into every scene, rebuilds the work
// items and uploads both.
int upload_config(dvg_ctx *ctx, dvg_batch *b) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // no step may still read the old arrays
    for (uint32_t i = 0; i < b->n; ++i) {
        BatchScene &B = b->bs[i];
        B.S.sblock_log2 = b->cfg.sblock_log2;
        B.S.bnd_log2 = b->cfg.bnd_log2;
        uint32_t tile_samples = (uint32_t)dvgk::TILE * dvgk::TILE * B.S.spp;
        uint64_t us = (uint64_t)dvgk::BLOCK << b->cfg.unit_log2;
        B.S.unit_samples = (uint32_t)std::min<uint64_t>(us, tile_samples);
        B.nb = dvgk::boundary_buckets(B.S);
        B.zn[15] = B.counts ? B.nb + 1 : 0;
    }
    int st = build_items(ctx, b);
    if (st) return st;
    HIP_TRY(ctx, hipMemcpyAsync(b->dev_bs, b->bs.data(), b->bs.size() * sizeof(BatchScene), hipMemcpyHostToDevice,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(b);
    return DVG_OK;
}

void launch_step(dvg_ctx *ctx, dvg_batch *b, uint32_t flags) {
    hipStream_t st = ctx->stream;
    dvgk::launch_fill_u32(b->counters, dvgk::NUM_PHASES, 0u, st);
    uint32_t cap = (uint32_t)b->n_cu * b->cfg.wg_per_cu;
    // the colour adjoint's workgroups beyond those resident would find the
    // queue drained; each launched one owns a bit-set slot
    uint32_t cap_int = std::min(cap, (uint32_t)b->n_cu * dvgk::INTERIOR_WG_PER_CU);
    auto run = [&](int ph, bool loss = false) {
        uint32_t ni = (uint32_t)b->items[ph].size();
        if (ni == 0) return;
        bool interior = ph == dvgk::PH_INTERIOR;
        dvgk::launch_batch_phase(ph, b->dev_bs, b->dev_items + b->item_off[ph], ni, b->counters + ph,
                                 std::min(ni, interior ? cap_int : cap), loss, interior ? b->gmask : nullptr,
                                 interior ? b->gcap : 0u, st);
    };
    bool fwd = flags & DVG_BATCH_FORWARD, loss = flags & DVG_BATCH_LOSS, bwd = flags & DVG_BATCH_BACKWARD;
    run(dvgk::PH_ZERO);
    if (fwd) {
        run(dvgk::PH_WEIGHTS_FWD);
        run(dvgk::PH_FORWARD);
        run(dvgk::PH_FORWARD_PF);
        run(dvgk::PH_RESOLVE, loss);
    }
    if (bwd) {
        run(dvgk::PH_WEIGHTS_BWD);
        run(dvgk::PH_INTERIOR);
        run(dvgk::PH_PREFILTERED);
        run(dvgk::PH_BND_KEY);
        run(dvgk::PH_SCAN);
        run(dvgk::PH_BND_SCATTER);
        run(dvgk::PH_BND_EVAL);
    }
}

}  // namespace

extern "C" {

int32_t dvg_batch_create(dvg_ctx *ctx, uint32_t n, const dvg_scene_pack *packs, const dvg_backward_pack *bpacks,
                         const dvg_render_params *params, const float *const *targets, dvg_batch **out) {
    if (!ctx || !out || (n > 0 && (!packs || !bpacks || !params))) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    *out = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    dvg_batch *b = new dvg_batch();
    b->device = ctx->device;
    b->n = n;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->n_cu = prop.multiProcessorCount;
    auto fail = [&](int st) {
        dvg_batch_destroy(b);
        return st;
    };
    // scenes: arenas, tile lists (built on the device)
    for (uint32_t i = 0; i < n; ++i) {
        dvg_device_scene *ds = nullptr;
        int st = dvg_device_scene_create(ctx, &packs[i], &bpacks[i], &params[i], &ds);
        if (st) return fail(st);
        b->scenes.push_back(ds);
        b->targets.push_back(targets ? targets[i] : nullptr);
        b->gcap = std::max(b->gcap, ds->gcap);
    }
    if (b->gcap > 0) {
        size_t words = (size_t)b->n_cu * dvgk::INTERIOR_WG_PER_CU * dvgk::interior_slot_words(b->gcap);
        hipError_t e = hipMalloc(&b->gmask, words * sizeof(uint32_t));
        if (e != hipSuccess) {
            b->gmask = nullptr;
            return fail(hip_err(ctx, e, "batch bit-set slots"));
        }
    }
    // one allocation for every work buffer; buckets sized for the finest
    // boundary blocks the autotuner may pick (2x2 pixels: 64 per tile)
    struct Lay {
        size_t weight, accum, img, dr, grads, keys, ranks, order, counts, offsets;
        size_t grad_off[14];
        uint32_t nbmax;
    };
    std::vector<Lay> lay(n);
    size_t total = 0;
    auto take = [&](size_t bytes) {
        size_t o = (total + 255) & ~(size_t)255;
        total = o + std::max(bytes, (size_t)4);
        return o;
    };
    for (uint32_t i = 0; i < n; ++i) {
        const dvgk::DevScene &S = b->scenes[i]->S;
        size_t npix = (size_t)S.W * S.H, ns = npix * S.spp;
        Lay &L = lay[i];
        L.weight = take(npix * 4);
        L.accum = take(npix * 16);
        L.img = take(npix * 16);
        L.dr = take(npix * 16);
        size_t g = 0;
        for (int k = 0; k < 14; ++k) {
            L.grad_off[k] = g;
            g += b->scenes[i]->grad_len[k];
        }
        L.grads = take(g * 4);
        bool bnd = S.n_cdf > 0 && !S.prefilter;
        L.nbmax = S.tiles_x * S.tiles_y * 64u;
        L.keys = bnd ? take(ns * 4) : 0;
        L.ranks = bnd ? take(ns * 4) : 0;
        L.order = bnd ? take(ns * 4) : 0;
        L.counts = bnd ? take(((size_t)L.nbmax + 1) * 4) : 0;
        L.offsets = bnd ? take(((size_t)L.nbmax + 2) * 4) : 0;
    }
    hipError_t e = hipMalloc(&b->work, std::max(total, (size_t)256));
    if (e != hipSuccess) return fail(hip_err(ctx, e, "hipMalloc(batch work)"));
    e = hipMalloc(&b->counters, dvgk::NUM_PHASES * sizeof(uint32_t));
    if (e != hipSuccess) return fail(hip_err(ctx, e, "hipMalloc(batch counters)"));
    uint8_t *base = (uint8_t *)b->work;
    b->bs.resize(n);
    std::vector<std::vector<uint32_t>> tile_off(n);
    for (uint32_t i = 0; i < n; ++i) {
        const dvgk::DevScene &S = b->scenes[i]->S;
        Lay &L = lay[i];
        BatchScene &B = b->bs[i];
        std::memset(&B, 0, sizeof(B));
        B.S = S;
        B.weight = (float *)(base + L.weight);
        B.accum = (float *)(base + L.accum);
        B.img = (float *)(base + L.img);
        B.d_render = (float *)(base + L.dr);
        B.target = const_cast<float *>(b->targets[i]);
        float *gb = (float *)(base + L.grads);
        float **gp = &B.G.d_shape_params;
        for (int k = 0; k < 14; ++k) gp[k] = b->scenes[i]->grad_len[k] ? gb + L.grad_off[k] : nullptr;
        bool bnd = S.n_cdf > 0 && !S.prefilter;
        if (bnd) {
            B.keys = (uint32_t *)(base + L.keys);
            B.ranks = (uint32_t *)(base + L.ranks);
            B.order = (uint32_t *)(base + L.order);
            B.counts = (uint32_t *)(base + L.counts);
            B.offsets = (uint32_t *)(base + L.offsets);
        }
        for (int k = 0; k < 14; ++k) {
            B.zp[k] = gp[k];
            B.zn[k] = gp[k] ? b->scenes[i]->grad_len[k] : 0u;
        }
        B.zp[14] = B.accum;
        B.zn[14] = (uint32_t)((size_t)S.W * S.H * 4);
        B.zp[15] = (float *)B.counts;
        // tile list lengths, for the heaviest-first order of tile work
        tile_off[i].resize((size_t)S.tiles_x * S.tiles_y + 1);
        e = hipMemcpyAsync(tile_off[i].data(), S.tile_offsets, tile_off[i].size() * 4, hipMemcpyDeviceToHost,
                           ctx->stream);
        if (e != hipSuccess) return fail(hip_err(ctx, e, "tile list readback"));
    }
    e = hipMalloc(&b->dev_bs, std::max<size_t>(n, 1) * sizeof(BatchScene));
    if (e != hipSuccess) return fail(hip_err(ctx, e, "hipMalloc(batch scenes)"));
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(hip_err(ctx, e, "batch setup"));
    b->tile_off = std::move(tile_off);
    for (uint32_t i = 0; i < n; ++i) b->nbmax.push_back(lay[i].nbmax);
    int st = upload_config(ctx, b);
    if (st) return fail(st);
    *out = b;
    return DVG_OK;
}

void dvg_batch_destroy(dvg_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    drop_graph(b);
    for (dvg_device_scene *ds : b->scenes) dvg_device_scene_destroy(ds);
    if (b->dev_bs) (void)hipFree(b->dev_bs);
    if (b->work) (void)hipFree(b->work);
    if (b->counters) (void)hipFree(b->counters);
    if (b->dev_items) (void)hipFree(b->dev_items);
    if (b->gmask) (void)hipFree(b->gmask);
    delete b;
}

int32_t dvg_batch_buffers(dvg_batch *b, uint32_t i, float **img, float **d_render, dvg_grads *grads) {
    if (!b || i >= b->n) return DVG_INVALID_SCENE;
    const BatchScene &B = b->bs[i];
    if (img) *img = B.img;
    if (d_render) *d_render = B.d_render;
    if (grads) {
        float **dst = &grads->d_shape_params;
        float *const *src = &B.G.d_shape_params;
        for (int k = 0; k < 14; ++k) dst[k] = src[k];
    }
    return DVG_OK;
}

int32_t dvg_batch_read(dvg_ctx *ctx, dvg_batch *b, uint32_t i, float *img, const dvg_grads *grads) {
    if (!ctx || !b || i >= b->n) return set_err(ctx, DVG_INVALID_SCENE, "batch index out of range");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const BatchScene &B = b->bs[i];
    const dvg_device_scene *ds = b->scenes[i];
    if (img) HIP_TRY(ctx, hipMemcpyAsync(img, B.img, (size_t)B.S.W * B.S.H * 16, hipMemcpyDefault, ctx->stream));
    if (grads) {
        float *const *dst = &grads->d_shape_params;
        float *const *src = &B.G.d_shape_params;
        for (int k = 0; k < 14; ++k)
            if (dst[k] && src[k] && ds->grad_len[k])
                HIP_TRY(ctx, hipMemcpyAsync(dst[k], src[k], (size_t)ds->grad_len[k] * 4, hipMemcpyDefault, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DVG_OK;
}

int32_t dvg_batch_step(dvg_ctx *ctx, dvg_batch *b, uint32_t flags) {
    if (!ctx || !b) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if ((flags & DVG_BATCH_LOSS) && !(flags & DVG_BATCH_FORWARD))
        return set_err(ctx, DVG_INVALID_SCENE, "the loss gradient needs the forward pass");
    if (flags & DVG_BATCH_LOSS)
        for (uint32_t i = 0; i < b->n; ++i)
            if (!b->targets[i]) return set_err(ctx, DVG_INVALID_SCENE, "loss gradient without a target image");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!b->use_graph || ctx->timing) {
        launch_step(ctx, b, flags);
        HIP_TRY(ctx, hipGetLastError());
        return DVG_OK;
    }
    if (!b->exec || b->exec_flags != flags || b->exec_stream != ctx->stream) {
        drop_graph(b);
        hipGraph_t g = nullptr;
        HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        launch_step(ctx, b, flags);
        HIP_TRY(ctx, hipStreamEndCapture(ctx->stream, &g));
        hipError_t e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            b->exec = nullptr;
            return hip_err(ctx, e, "hipGraphInstantiate");
        }
        b->exec_flags = flags;
        b->exec_stream = ctx->stream;
    }
    HIP_TRY(ctx, hipGraphLaunch(b->exec, ctx->stream));
    return DVG_OK;
}

int32_t dvg_batch_set_graph(dvg_batch *b, int enable) {
    if (!b) return DVG_INVALID_SCENE;
    b->use_graph = enable != 0;
    if (!b->use_graph) drop_graph(b);
    return DVG_OK;
}

int32_t dvg_batch_get_config(const dvg_batch *b, dvg_batch_config *cfg) {
    if (!b || !cfg) return DVG_INVALID_SCENE;
    *cfg = b->cfg;
    return DVG_OK;
}

int32_t dvg_batch_set_config(dvg_ctx *ctx, dvg_batch *b, const dvg_batch_config *cfg) {
    if (!ctx || !b || !cfg) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    if (cfg->sblock_log2 > 4 || cfg->bnd_log2 < 1 || cfg->bnd_log2 > 4 || cfg->wg_per_cu < 1 || cfg->wg_per_cu > 64 ||
        cfg->unit_log2 > 8)
        return set_err(ctx, DVG_INVALID_SCENE, "launch shape out of range");
    b->cfg.sblock_log2 = cfg->sblock_log2;
    b->cfg.bnd_log2 = cfg->bnd_log2;
    b->cfg.wg_per_cu = cfg->wg_per_cu;
    b->cfg.unit_log2 = cfg->unit_log2;
    return upload_config(ctx, b);
}

// Coordinate search over the launch shapes: tile unit size, persistent
// workgroups per CU, sample-block edge, boundary-bucket edge; each candidate
// is timed with HIP events over `reps` steps after one warm-up step.
int32_t dvg_batch_autotune(dvg_ctx *ctx, dvg_batch *b, uint32_t flags, int reps, dvg_batch_config *best) {
    if (!ctx || !b) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    reps = std::max(reps, 1);
    hipEvent_t t0, t1;
    HIP_TRY(ctx, hipEventCreate(&t0));
    HIP_TRY(ctx, hipEventCreate(&t1));
    bool graph = b->use_graph;
    b->use_graph = false;
    auto measure = [&](dvg_batch_config c, float &ms) -> int {
        int st = dvg_batch_set_config(ctx, b, &c);
        if (st) return st;
        launch_step(ctx, b, flags);
        HIP_TRY(ctx, hipEventRecord(t0, ctx->stream));
        for (int r = 0; r < reps; ++r) launch_step(ctx, b, flags);
        HIP_TRY(ctx, hipEventRecord(t1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(t1));
        HIP_TRY(ctx, hipEventElapsedTime(&ms, t0, t1));
        ms /= (float)reps;
        return DVG_OK;
    };
    dvg_batch_config cur = b->cfg;
    float cur_ms = 0.f;
    int st = measure(cur, cur_ms);
    const uint32_t ul[] = {0, 1, 2, 8}, wg[] = {2, 4, 8, 16}, sb[] = {1, 2, 3}, bb[] = {1, 2, 3};
    for (int knob = 0; knob < 4 && st == DVG_OK; ++knob) {
        const uint32_t *vals = knob == 0 ? ul : (knob == 1 ? wg : (knob == 2 ? sb : bb));
        int nv = knob < 2 ? 4 : 3;
        for (int k = 0; k < nv && st == DVG_OK; ++k) {
            dvg_batch_config c = cur;
            if (knob == 0) c.unit_log2 = vals[k];
            else if (knob == 1) c.wg_per_cu = vals[k];
            else if (knob == 2) c.sblock_log2 = vals[k];
            else c.bnd_log2 = vals[k];
            if (c.wg_per_cu == cur.wg_per_cu && c.sblock_log2 == cur.sblock_log2 && c.bnd_log2 == cur.bnd_log2 &&
                c.unit_log2 == cur.unit_log2)
                continue;
            float ms = 0.f;
            st = measure(c, ms);
            if (st == DVG_OK && ms < cur_ms) {
                cur = c;
                cur_ms = ms;
            }
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    b->use_graph = graph;
    if (st) return st;
    cur.ms = cur_ms;
    st = dvg_batch_set_config(ctx, b, &cur);
    b->cfg.ms = cur_ms;
    if (best) *best = b->cfg;
    return st;
}

}  // extern "C"
