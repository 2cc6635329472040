// Delay-multiply-and-sum over the separated (per-rx or per-tx) axis.
// Formulas of examples/advanced_beamforming.ipynb:296-301 (SSR) and
// examples/develope_gpu_slsc_and_dmas_kernels.ipynb:292-303 (power, P-DMAS-SSR);
// lags come from the caller's array (pycbf/gpu/__advanced_engines__.cu:44-56
// used the loop index instead).  One lane per (frame, pixel); the column
// im[:, p] is read with pixel-coalesced loads and reused across lags from L1.
#include <math.h>

#include "common.h"

namespace pycbf {

__device__ __forceinline__ float ssr(float q)
{
    const float r = sqrtf(fabsf(q));
    return q > 0.0f ? r : (q < 0.0f ? -r : 0.0f);
}

__global__ __launch_bounds__(256) void k_dmas(const float *__restrict__ im, int64_t nfr,
                                              int64_t nsep, int64_t np,
                                              const int32_t *__restrict__ lags, int nlag,
                                              int dmastype, int sumlags, float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nfr * np) return;
    const int64_t f = g / np, p = g - f * np;
    const float *col = im + f * nsep * np + p;
    float total = 0.0f;
    for (int l = 0; l < nlag; ++l) {
        const int m = lags[l];
        float s = 0.0f;
        if (m > 0 && (int64_t)m < nsep) {
            for (int64_t i = 0; i + m < nsep; ++i) {
                const float q = col[i * np] * col[(i + m) * np];
                s += (dmastype == PYCBF_DMAS_SSR) ? ssr(q) : q;
            }
        }
        if (sumlags)
            total += s;
        else
            out[(f * nlag + l) * np + p] = (dmastype == PYCBF_DMAS_PSSR) ? ssr(s) : s;
    }
    if (sumlags) out[f * np + p] = (dmastype == PYCBF_DMAS_PSSR) ? ssr(total) : total;
}

int launch_dmas(const float *im, int64_t nfr, int64_t nsep, int64_t np, const int32_t *lags,
                int nlag, int dmastype, int sumlags, float *out, hipStream_t s)
{
    const int64_t n = nfr * np;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_dmas, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, im, nfr, nsep,
                       np, lags, nlag, dmastype, sumlags, out);
    return check_launch("dmas");
}

}  // namespace pycbf
