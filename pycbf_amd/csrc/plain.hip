// Thread-per-sample DAS: the launch scheme of pycbf's GPU kernels
// (pycbf/gpu/__engines__.cu:620-694, 813-864) on gfx950, kept as the "plain
// GPU" comparator of the staged gather.  One work item per (frame, tx, rx,
// pixel) sample, pixels fastest; each evaluates its delays and weights (from
// the tables, or the synthetic geometry in registers), interpolates from the
// same gather source as the other kernels, and adds w * value to its output
// element with one float atomic.  The output is zero-filled by a kernel first
// (no memset node, so a captured graph replays it correctly).  Sums of one
// output element arrive in scheduling order, so images agree with the
// ordered kernels to float32 rounding, not bit for bit.
#include <stdlib.h>

#include "common.h"
#include "das_dev.h"
#include "prof.h"

namespace pycbf {

__global__ __launch_bounds__(256) void k_plain_zero(float *__restrict__ out, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = 0.0f;
}

template <int GK, int GEO>
__global__ __launch_bounds__(256) void k_das_plain(DasArgs A, int64_t total)
{
    const int64_t np = A.np;
    const int64_t per_tx = (int64_t)A.nrx * np;
    const int64_t per_frame = (int64_t)A.ntx * per_tx;
    const int64_t nchan = (int64_t)A.ntx * A.nrx;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int64_t f = t / per_frame;
        const int64_t r = t - f * per_frame;
        const int tx = (int)(r / per_tx);
        const int rx = (int)((r / np) % A.nrx);
        const int64_t p = r % np;
        float pv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int d = 0; d < GEO; ++d) pv[d] = A.pvec[p * GEO + d];
        float ta = 0.0f, wa = 0.0f, tb = 0.0f, wb = 0.0f;
        elem_eval<GEO>(A, true, tx, p, pv, ta, wa);
        elem_eval<GEO>(A, false, rx, p, pv, tb, wb);
        const float w = wa * wb;
        if (!(fabsf(w) >= A.thresh)) continue;
        float v[1];
        bool ok = false;
        gather<GK, 1>(A, f * nchan + (int64_t)tx * A.nrx + rx, 0, 1, ta + tb, ok, v);
        if (!ok) continue;
        int64_t o;
        if (A.sumflag == PYCBF_SUM_NONE)
            o = ((f * A.ntx + tx) * A.nrx + rx) * np + p;
        else if (A.sumflag == PYCBF_SUM_TX_ONLY)
            o = (f * A.nrx + rx) * np + p;
        else if (A.sumflag == PYCBF_SUM_RX_ONLY)
            o = (f * A.ntx + tx) * np + p;
        else
            o = f * np + p;
        atomicAdd(A.out + o, w * v[0]);
    }
}

bool das_use_plain()
{
    const char *e = getenv("PYCBF_DAS_PLAIN");
    return e && atoi(e) != 0;
}

template <int GK>
static void plain_geo(int geo, const DasArgs &a, int64_t total, hipStream_t s)
{
    const dim3 grid(65536), block(256);
    if (geo == 3)
        hipLaunchKernelGGL((k_das_plain<GK, 3>), grid, block, 0, s, a, total);
    else if (geo == 2)
        hipLaunchKernelGGL((k_das_plain<GK, 2>), grid, block, 0, s, a, total);
    else
        hipLaunchKernelGGL((k_das_plain<GK, 0>), grid, block, 0, s, a, total);
}

int launch_das_plain(int gk, const DasArgs &a, hipStream_t s)
{
    const int64_t nout_per_frame =
        a.sumflag == PYCBF_SUM_NONE      ? (int64_t)a.ntx * a.nrx * a.np
        : a.sumflag == PYCBF_SUM_TX_ONLY ? (int64_t)a.nrx * a.np
        : a.sumflag == PYCBF_SUM_RX_ONLY ? (int64_t)a.ntx * a.np
                                         : a.np;
    const int64_t nout = nout_per_frame * a.nfr;
    const int64_t total = (int64_t)a.nfr * a.ntx * a.nrx * a.np;
    hipLaunchKernelGGL(k_plain_zero, dim3(4096), dim3(256), 0, s, a.out, nout);
    if (int rc = check_launch("plain DAS zero fill")) return rc;
    {
        ProfScope prof("das_plain", s);
        switch (gk) {
        case GK_NEAREST: plain_geo<GK_NEAREST>(a.geo, a, total, s); break;
        case GK_LINEAR: plain_geo<GK_LINEAR>(a.geo, a, total, s); break;
        case GK_CUBIC: plain_geo<GK_CUBIC>(a.geo, a, total, s); break;
        case GK_KORDER: plain_geo<GK_KORDER>(a.geo, a, total, s); break;
        default: plain_geo<GK_AKIMA>(a.geo, a, total, s); break;
        }
    }
    return check_launch("plain DAS");
}

}  // namespace pycbf
