// xdn_filter_train_noalloc: the fused training filter with its workspace (the per-tile loss
// partial sums that the deterministic loss reduction reads) taken from a fixed device array of
// this library instead of a caller-provided buffer, for callers that cannot allocate device memory.
//
// The array lives once per GPU for the life of the process. Host code never uses its name as an
// address (on the host that name is only a stand-in for the symbol): the device address is looked
// up with hipGetSymbolAddress once per device, with that device current, and cached.
#include <hip/hip_runtime.h>

#include <mutex>

#include "abi_internal.h"

using namespace xdn_abi;

namespace {

constexpr size_t kScratchBytes = (size_t)1 << 20;  // XDN_FILTER_TRAIN_SCRATCH_BYTES
static_assert(kScratchBytes == XDN_FILTER_TRAIN_SCRATCH_BYTES, "header and library disagree");

static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void *g_ptr[kMaxDevices];  // device address of `scratch` on each device (null until looked up)

// The device address of this GPU's copy of the array, looked up on the first call on the device.
int scratch_ptr(void **out)
{
    int dev = -1;
    if (int e = hip_check(hipGetDevice(&dev), "hipGetDevice")) return e;
    if (dev < 0 || dev >= kMaxDevices) return fail(XDN_EINVAL, "filter_train_noalloc: device %d unsupported", dev);
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_ptr[dev])
    {
        void *p = nullptr;
        if (int e = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return e;
        if (!p) return fail(XDN_EHIP, "hipGetSymbolAddress(scratch) returned null");
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return XDN_OK;
}

}  // namespace

extern "C" int xdn_filter_train_noalloc(const float *ols1, const float *ols2, const float *albedo, const float *normal,
                                        const float *param, const float *success, const float *success_prev,
                                        const float *prev_out, const float *opacity, const float *prev_ols1,
                                        const float *prev_ols2, float *out_log, float *out, float *dparam, float *loss,
                                        float *avg_loss, int batch, int height, int width, int win_size, void *stream)
{
    if (int e = check_shape(batch, height, width)) return e;
    const size_t need = xdn_filter_train_workspace_size(batch, height, width);
    if (need > kScratchBytes)
        return fail(XDN_EINVAL,
                    "filter_train_noalloc: [%d, %d, %d] needs %zu bytes of loss partials, more than the library's "
                    "fixed %zu (use xdn_filter_train with a workspace)",
                    batch, height, width, need, kScratchBytes);
    void *ws = nullptr;
    if (batch > 0)  // an empty batch only has its attributes checked: no device call
        if (int e = scratch_ptr(&ws)) return e;
    // every tile writes its own partial before the loss sum reads it: nothing to clear
    return xdn_filter_train(ols1, ols2, albedo, normal, param, success, success_prev, prev_out, opacity, prev_ols1,
                            prev_ols2, out_log, out, dparam, loss, avg_loss, ws, kScratchBytes, batch, height, width,
                            win_size, stream);
}
