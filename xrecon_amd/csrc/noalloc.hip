// An execute entry point that allocates nothing: xr_plan3d_execute with its workspace
// taken from a fixed device array instead of a caller-provided buffer.
//
// For callers that cannot allocate device memory on the call path (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The workspace of the
// gathers (the prefix-summed copy of the input) lives in `scratch`, a file-scope device
// array of XR_NOALLOC_SCRATCH_BYTES.  Each GPU that loads this library holds its own copy
// of the array at its own address until the process ends.  Its device address is looked
// up once per device, with that device current, on the first call on that device; every
// later call passes that address to the kernels as their workspace argument.
//
// All calls on one GPU share that GPU's array: calls that overlap in time on one GPU (on
// different streams, or from different threads) race on it and must be serialised by the
// caller.  Every workspace entry a result depends on is written by the same call before it
// is read (the prefix pass fills the tables, pad entries included), so the contents left
// by an earlier call need no clearing.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "xr_internal.h"

namespace {

constexpr size_t kScratchBytes = XR_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (looked up once per device).
int scratch_on(int device, void **out)
{
    if (device < 0 || device >= kMaxDevices)
        return xr::fail(XR_EINVAL, "device " + std::to_string(device) + " out of range for the static scratch");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch_ptr[device]) {
        int prev = -1;
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != device && hipSetDevice(device) != hipSuccess)
            return xr::fail(XR_EHIP, "cannot make device " + std::to_string(device) + " current");
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
        if (e != hipSuccess || !ptr)
            return xr::fail(XR_EHIP, std::string("hipGetSymbolAddress of the static scratch failed: ") +
                                         hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return XR_OK;
}

}  // namespace

extern "C" int xr_plan3d_execute_noalloc(const xr_plan3d *plan, int op, const float *in, float *out,
                                         int num_batch, void *stream)
{
    xr::set_error("");
    if (!plan) return xr::fail(XR_EINVAL, "null plan");
    if (xr::kernel_kind(plan->p, op) < 0) return xr::fail(XR_EINVAL, "unknown operator " + std::to_string(op));
    if (num_batch < 0) return xr::fail(XR_EINVAL, "negative batch count");
    const size_t need = xr::workspace_bytes(plan->p, op, num_batch);
    if (need > kScratchBytes)
        return xr::fail(XR_EWORKSPACE, "operator " + std::to_string(op) + " on " + std::to_string(num_batch) +
                                           " batch items needs " + std::to_string(need) +
                                           " bytes of workspace; the static scratch holds " +
                                           std::to_string(kScratchBytes) +
                                           " bytes per device (use xr_plan3d_execute with a workspace)");
    void *ws = nullptr;
    const int rc = scratch_on(plan->p.device, &ws);
    if (rc) return rc;
    return xr::execute(plan->p, op, in, out, num_batch, ws, kScratchBytes, 0, 0, static_cast<hipStream_t>(stream));
}
