// dvg_device_render_noalloc: the forward of dvg_device_render with its work
// images (the filter-weight image, W*H floats, and the splat accumulation
// image, W*H*4 floats) in a file-scope device array instead of
// context buffers grown with hipMalloc.  For callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  It calls no hipMalloc, hipFree or other allocating
// routine and launches the same kernels as dvg_device_render.
//
// The array holds DVG_NOALLOC_MAX_PIXELS * 5 floats (20 MiB) on every device
// that loads the library, until the process ends.  Every call on one device
// shares that device's copy, which keeps the last call's contents: the
// accumulation image is cleared by a fill kernel at the start of each call and
// the weight image is overwritten whole.  Calls that overlap on one device (two
// streams, two contexts, two host threads) race on the array.
#include <mutex>

#include "dvg_internal.h"

namespace {

constexpr size_t kMaxPixels = DVG_NOALLOC_MAX_PIXELS;
// the accumulation image starts at a 256-byte boundary after the weight image
constexpr size_t kAccumOffset = (kMaxPixels + 63) / 64 * 64;
constexpr size_t kScratchFloats = kAccumOffset + kMaxPixels * 4;

static __device__ __align__(256) float scratch[kScratchFloats];

// Device address of `scratch` on each device, looked up once per device with
// that device current.  The host-side name `scratch` is not a device address.
constexpr int kMaxDevices = 64;
std::mutex g_lock;
float *g_ptr[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

}  // namespace

namespace dvgn {

// Called by dvg_ctx_create with `device` current.  Returns the lookup's status
// (also kept for the entry point, which reports it).
hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_done[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_status[device] = e;
        g_ptr[device] = (e == hipSuccess) ? static_cast<float *>(p) : nullptr;
        g_done[device] = true;
    }
    return g_status[device];
}

}  // namespace dvgn

int32_t dvg_device_render_noalloc(dvg_ctx *ctx, dvg_device_scene *ds, float *out) {
    if (!ctx || !ds || !out) return set_err(ctx, DVG_INVALID_SCENE, "null argument");
    const dvgk::DevScene &S = ds->S;
    size_t npix = (size_t)S.W * S.H;
    if (npix > kMaxPixels)
        return set_err(ctx, DVG_INVALID_SCENE,
                       "image larger than the no-allocation scratch (DVG_NOALLOC_MAX_PIXELS pixels)");
    if (npix == 0) return DVG_OK;
    float *base = nullptr;
    hipError_t status = hipErrorNotInitialized;
    if (ctx->device >= 0 && ctx->device < kMaxDevices) {
        std::lock_guard<std::mutex> g(g_lock);
        if (g_done[ctx->device]) {
            base = g_ptr[ctx->device];
            status = g_status[ctx->device];
        }
    }
    if (!base) return hip_err(ctx, status == hipSuccess ? hipErrorNotInitialized : status,
                              "no-allocation scratch address");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    float *weight = base;
    float *accum = base + kAccumOffset;
    uint32_t n = (uint32_t)(npix * 4);
    dvgk::launch_zero(&accum, &n, 1, ctx->stream);
    dvgk::launch_weights(S, weight, false, ctx->stream);
    dvgk::launch_forward(S, weight, accum, ctx->stream);
    dvgk::launch_resolve(S, weight, accum, out, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    return DVG_OK;
}
