// Measured bandwidth peaks of the device: the denominators of the rooflines
// the benchmark reports next to the published HBM figure.
//
//   HBM  one streaming read of a caller buffer larger than the caches (16-byte
//        loads, grid-stride, every byte read once per pass)
//   L2   every workgroup reads one region small enough to stay in each XCD's
//        L2 (larger than a CU's L1), from a rotated start, several passes
//   LDS  every lane reads 16 bytes per step from its workgroup's 16 KB of LDS,
//        consecutive lanes on consecutive addresses (no bank conflicts)
//
// Loaded words are folded into a per-lane sum with three-operand integer adds
// (two loaded words per instruction), so the loads are not removed and VALU
// issue stays well below the memory rates; each lane stores its sum once.
#include "common.h"

namespace pycbf {

__global__ __launch_bounds__(256) void k_peak_stream(const uint4 *__restrict__ src, int64_t n,
                                                     uint32_t *__restrict__ sink)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    uint32_t a = 0, b = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint4 v = src[i];
        a = a + v.x + v.y;
        b = b + v.z + v.w;
    }
    sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = a ^ b;
}

__global__ __launch_bounds__(256) void k_peak_region(const uint4 *__restrict__ src, int64_t n,
                                                     int passes, uint32_t *__restrict__ sink)
{
    uint32_t a = 0, b = 0;
    const int64_t start = ((int64_t)blockIdx.x * 4099) % n;
    for (int p = 0; p < passes; ++p)
        for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
            int64_t i = start + k;
            if (i >= n) i -= n;
            const uint4 v = src[i];
            a = a + v.x + v.y;
            b = b + v.z + v.w;
        }
    sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = a ^ b;
}

constexpr int PEAK_LDS_VEC = 1024;  // 16 KB of uint4 per workgroup: 8 workgroups per CU

// `step` (256 at run time) is a kernel argument so that the compiler cannot
// see the addresses repeat and hoist the loads out of the loop.
__global__ __launch_bounds__(256) void k_peak_lds(int iters, int step, uint32_t *__restrict__ sink)
{
    __shared__ uint4 buf[PEAK_LDS_VEC];
    for (int i = threadIdx.x; i < PEAK_LDS_VEC; i += blockDim.x)
        buf[i] = make_uint4(i, 3 * i, 5 * i, 7 * i);
    __syncthreads();
    uint32_t a = 0, b = 0;
    int i = threadIdx.x;
    for (int k = 0; k < iters; ++k) {
        const uint4 v = buf[i];
        a = a + v.x + v.y;
        b = b + v.z + v.w;
        i = (i + step) & (PEAK_LDS_VEC - 1);
    }
    sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = a ^ b;
}

}  // namespace pycbf

using namespace pycbf;

extern "C" int pycbf_measure_bandwidth(int kind, const void *buf, size_t bytes, int reps,
                                       void *sink, size_t sink_bytes, double *gbps,
                                       void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!gbps || !sink) return fail(PYCBF_EINVAL, "gbps and sink must not be NULL");
    if (reps < 1) return fail(PYCBF_EINVAL, "reps must be >= 1");
    const int blocks = 8192;
    if (sink_bytes < (size_t)blocks * 256 * sizeof(uint32_t))
        return fail(PYCBF_EINVAL, "sink of %zu bytes needed", (size_t)blocks * 256 * 4);
    if ((kind == 0 || kind == 1) && (!buf || bytes < 16 || (bytes % 16) != 0))
        return fail(PYCBF_EINVAL, "buffer must be a multiple of 16 bytes");
    const int64_t n = (int64_t)(bytes / 16);
    const int passes = 8, lds_iters = 4096;
    hipEvent_t e0, e1;
    if (int rc = check(hipEventCreate(&e0), "event")) return rc;
    if (int rc = check(hipEventCreate(&e1), "event")) {
        hipEventDestroy(e0);
        return rc;
    }
    double moved = 0.0;
    int rc = 0;
    for (int r = 0; r <= reps && rc == 0; ++r) {
        if (r == 1) rc = check(hipEventRecord(e0, s), "record");  // first pass warms up
        if (kind == 0) {
            hipLaunchKernelGGL(k_peak_stream, dim3(blocks), dim3(256), 0, s, (const uint4 *)buf, n,
                               (uint32_t *)sink);
            if (r > 0) moved += (double)bytes;
        } else if (kind == 1) {
            hipLaunchKernelGGL(k_peak_region, dim3(blocks), dim3(256), 0, s, (const uint4 *)buf,
                               n, passes, (uint32_t *)sink);
            if (r > 0) moved += (double)bytes * passes * blocks;
        } else if (kind == 2) {
            hipLaunchKernelGGL(k_peak_lds, dim3(blocks), dim3(256), 0, s, lds_iters, 256,
                               (uint32_t *)sink);
            if (r > 0) moved += 16.0 * 256.0 * lds_iters * blocks;
        } else {
            rc = fail(PYCBF_EINVAL, "kind %d is not 0 (HBM), 1 (L2) or 2 (LDS)", kind);
        }
        if (rc == 0) rc = check_launch("bandwidth kernel");
    }
    float ms = 0.0f;
    if (rc == 0) rc = check(hipEventRecord(e1, s), "record");
    if (rc == 0) rc = check(hipEventSynchronize(e1), "synchronize");
    if (rc == 0) rc = check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    *gbps = moved / (ms * 1e-3) / 1e9;
    return 0;
}
