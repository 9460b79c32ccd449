// Internal host-side types shared by the C ABI translation units (the
// context, device-resident scenes, error helpers, kernel timing).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/dvg.h"
#include "dvg_kernels.h"

// Per-kernel timing (dvg_ctx_set_timing): event pairs recorded around each
// launch on the context's stream, read back by dvg_kernel_times.
struct TimedLaunch {
    int kernel;
    hipEvent_t start, stop;
};

struct dvg_ctx {
    int device = 0;
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> event_pool;
    double kernel_ms[DVG_NUM_KERNELS] = {};
    uint32_t kernel_calls[DVG_NUM_KERNELS] = {};
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    const char *last_error = "";
    std::string err_buf;
    // work buffers, grown on demand
    float *weight = nullptr;
    size_t weight_cap = 0;
    float *accum = nullptr;
    size_t accum_cap = 0;
    // boundary-pass work buffers (u32), grown on demand
    uint32_t *bwork = nullptr;
    size_t bwork_cap = 0;
    // colour-adjoint bit-set slots for long tile lists (u32: counter, then
    // the slots), grown on demand
    uint32_t *imask = nullptr;
    size_t imask_cap = 0;
    // signed-distance work arrays (dvgk::sdf_work_bytes), grown on demand
    float *sdfw = nullptr;
    size_t sdfw_cap = 0;
    uint32_t n_cu = 0;
    // second stream (default priority) for the boundary redo, which runs
    // beside the colour adjoint; fork / join events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool exact_boundary = false;  // dvg_ctx_set_exact_boundary (set on by dvg_ctx_create)
};

struct dvg_device_scene {
    dvgk::DevScene S;
    dvgk::UpdatePlan plan = {};  // device-side updates (plan.curve_pt null: unsupported)
    void *arena = nullptr;
    size_t arena_bytes = 0;
    uint32_t num_groups = 0, num_shapes = 0;
    uint32_t grad_len[14] = {};  // lengths of the 14 gradient arrays
    uint32_t list_max = 0;       // longest tile list at creation
    uint32_t ncurves = 0;        // curve segments of the pack
    uint32_t gcap = 0;           // list capacity of the adjoint's global bit-set slots (0: none needed)
};

// Brackets one kernel launch with timing events when timing is on.
struct KernelTimer {
    dvg_ctx *ctx;
    int id;
    TimedLaunch tl;
    KernelTimer(dvg_ctx *c, int kernel) : ctx(c), id(kernel) {
        if (!ctx->timing) return;
        if (ctx->event_pool.empty()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess) { ctx->timing = false; return; }
            if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); ctx->timing = false; return; }
            ctx->event_pool.push_back({a, b});
        }
        auto ev = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        tl = {id, ev.first, ev.second};
        (void)hipEventRecord(tl.start, ctx->stream);
    }
    ~KernelTimer() {
        if (!ctx->timing) return;
        (void)hipEventRecord(tl.stop, ctx->stream);
        ctx->pending.push_back(tl);
    }
};

inline int set_err(dvg_ctx *ctx, int code, const char *msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

inline int hip_err(dvg_ctx *ctx, hipError_t e, const char *where) {
    if (e == hipSuccess) return DVG_OK;
    if (ctx) {
        ctx->err_buf = std::string(where) + ": " + hipGetErrorString(e);
        ctx->last_error = ctx->err_buf.c_str();
    }
    return DVG_DEVICE_ERROR;
}

namespace dvgn {
// Looks up the no-allocation entry point's device scratch address on
// `device` (which must be current); dvg_noalloc.hip.
hipError_t noalloc_setup(int device);
}  // namespace dvgn

#define HIP_TRY(ctx, expr)                                  \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return hip_err(ctx, _e, #expr); \
    } while (0)

