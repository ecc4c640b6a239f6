/*
 * gm_noalloc.hip -- the device-resident LM state of gm_invert_noalloc.
 *
 * gm_invert keeps its LM state in a per-context allocation that grows with
 * the largest cube it has seen.  Callers that cannot allocate (work under
 * graph capture, tight serving loops) use gm_invert_noalloc instead, whose
 * state lives in one file-scope device array of fixed size: every GPU the
 * library runs on holds its own copy of it until the process ends, and
 * every call on that GPU shares it.
 *
 * In host code the array's name is only a host-side stand-in, not a device
 * address.  The device address is looked up once per device, with that
 * device current, when a context is created there (gm_ctx_create), and kept
 * per device; everything that hands the array to the device uses it.
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "gm_internal.h"

static __device__ __align__(256) char scratch[GM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

} // namespace

gm_status gm_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return GM_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return GM_OK;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) return GM_ERR_HIP;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GM_ERR_HIP;
    g_ptr[device] = p;
    return GM_OK;
}

void *gm_scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}
