// pycbf_das_synthetic_noalloc: the Synthetic DAS on a library-owned scratch
// array, for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The scratch is one file-scope device array per GPU, held until the process
// ends.  Its name is a host-side stand-in in host code, never a device
// address: the device address comes from hipGetSymbolAddress, once per device
// (with that device current) on the first call there, and is kept per device.
// Every call on a GPU uses that GPU's array; calls that overlap on one GPU
// race.  No kernel of the path needs the array cleared: the spline table,
// the tile delay ranges and the partial images are written before they are
// read.
#include <mutex>

#include "common.h"

namespace pycbf {

static __device__ __align__(256) unsigned char scratch[PYCBF_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[NOALLOC_MAX_DEVICES];

// Device address of the current device's scratch array.
static int scratch_of_current_device(void **ptr)
{
    int dev = -1;
    if (int rc = check(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(PYCBF_EINVAL, "device %d out of range (0..%d)", dev, NOALLOC_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        if (int rc = check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                           "hipGetSymbolAddress(scratch)"))
            return rc;
        if (!p) return fail(PYCBF_EINVAL, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_dev[dev] = p;
    }
    *ptr = g_scratch_dev[dev];
    return 0;
}

}  // namespace pycbf

using namespace pycbf;

extern "C" int pycbf_das_synthetic_noalloc(const pycbf_rfinfo *info, const float *rf,
                                           int64_t nfr, const float *txpack,
                                           const float *rxpack, float c0, const float *pvec,
                                           const int32_t *perm, int64_t np, int interp_kind,
                                           int interp_param, const float *S, float thresh,
                                           int sumflag, float *out, void *stream)
{
    void *buf = nullptr;
    if (int rc = scratch_of_current_device(&buf)) return rc;
    return das_synthetic_in(info, rf, nfr, txpack, rxpack, c0, pvec, perm, np, interp_kind,
                            interp_param, S, thresh, sumflag, out, buf,
                            (size_t)PYCBF_NOALLOC_SCRATCH_BYTES, true, (hipStream_t)stream);
}
