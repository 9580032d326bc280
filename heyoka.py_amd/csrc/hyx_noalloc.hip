// hyx_propagate_until_noalloc: propagate_until for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). The inputs and outputs are the caller's device arrays; every scratch buffer
// the kernels need (per-instance bookkeeping, the work-queue counters, default arrays, the
// compact form's Taylor tape) is carved out of one fixed device array of
// HYX_NOALLOC_SCRATCH_BYTES per GPU, `scratch` below, which the code object holds from the
// moment it is loaded on a device until the process ends.
//
// In host code `scratch` names a host-side stand-in, not the device array: its device
// address on each GPU comes from hipGetSymbolAddress, taken once per device (with that
// device current) on the first call there, and only that address is ever handed to a
// kernel. Every call on one GPU shares that GPU's array; calls that overlap on one GPU
// (two streams, two threads) race on it.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/hyx.h"
#include "hyx_internal.h"

static __device__ __align__(256) unsigned char scratch[HYX_NOALLOC_SCRATCH_BYTES];

namespace
{

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
int scratch_on(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= kMaxDevices) return hyx_internal_fail(HYX_EINVAL, "device index beyond the supported 64");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return hyx_internal_fail(HYX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

} // namespace

extern "C" int hyx_propagate_until_noalloc(hyx_sys *s, uint64_t n_inst, double *state, const double *pars,
                                           double *t_hi, double *t_lo, const double *tf_hi, const double *tf_lo,
                                           const double *max_delta_t, uint64_t max_steps, int64_t *outcome,
                                           double *min_h, double *max_h, uint64_t *nsteps, void *stream)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return hyx_internal_fail(HYX_EHIP, "hipGetDevice failed");
    unsigned char *sc = nullptr;
    if (int rc = scratch_on(dev, &sc)) return rc;
    const hyx::NoAllocArgs na{n_inst, state, pars,   t_hi,  t_lo,  tf_hi,  tf_lo, max_delta_t,
                              max_steps, outcome, min_h, max_h, nsteps, stream};
    return hyx_noalloc_propagate(s, na, sc, (uint64_t)HYX_NOALLOC_SCRATCH_BYTES);
}
