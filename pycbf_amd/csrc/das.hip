// Delay-and-sum gather kernels for gfx950.
//
// Work decomposition: one lane owns one pixel (processing order from an
// optional permutation that groups pixels into spatially compact 64-pixel
// wavefront tiles), and walks every (tx, rx) channel.  The loop over frames is
// innermost, so the delay, weight and sample-index arithmetic of a
// (tx, rx, pixel) triple is done once and reused for up to 16 frames.  Sums
// stay in registers; no atomics.  The outer channel axis may be split across
// blockIdx.z to fill the 256 CUs when the pixel count is small.
//
// Per-sample semantics (pycbf/cpu/pycbf.c:179-271 with the DAS loop of
// pycbf/gpu/__engines__.cu:813-864): w = apodtx*apodrx, x = tautx+taurx; the
// sample contributes w*interp(x) when |w| >= thresh and x lies on the trace.
// With delay residuals (pycbf_das_tabbed_residual) the interpolant is taken
// at x + residuals in double with exact weights (gather_q).
#include <math.h>
#include <stdlib.h>

#include "common.h"
#include "das_dev.h"
#include "prof.h"

namespace pycbf {

template <int GK, int SUM, int FC>
__global__ __launch_bounds__(256) void k_das(DasArgs A)
{
    uint32_t bx, by, bz;
    block_coords(A.xcd_remap, bx, by, bz);
    const int64_t i = (int64_t)bx * blockDim.x + threadIdx.x;
    if (i >= A.np) return;
    const int64_t p = A.perm ? (int64_t)A.perm[i] : i;
    const int f0 = by * FC;
    const int nf = min(FC, A.nfr - f0);
    const int64_t np = A.np;

    // rx_only keeps transmits apart and tx_only receivers, so those are the
    // outer axis there; tx_and_rx and none take the axis the host chose.
    const bool OUTER_TX = (SUM == PYCBF_SUM_RX_ONLY) ||
                          (SUM != PYCBF_SUM_TX_ONLY && A.outer_tx);
    const int na = OUTER_TX ? A.ntx : A.nrx;
    const int nb = OUTER_TX ? A.nrx : A.ntx;
    const float *ta_t = OUTER_TX ? A.ttx : A.trx;
    const float *wa_t = OUTER_TX ? A.atx : A.arx;
    const float *tb_t = OUTER_TX ? A.trx : A.ttx;
    const float *wb_t = OUTER_TX ? A.arx : A.atx;
    const int a0 = bz * A.a_split;
    const int a1 = min(na, a0 + A.a_split);

    const int64_t nchan = (int64_t)A.ntx * A.nrx;
    const int64_t fstride = nchan * A.ch_stride;  // floats between frames in src
    float acc[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) acc[f] = 0.0f;

    // delay residuals (workgroup-uniform): positions in double, exact weights
    const bool fine = A.ttx_lo || A.trx_lo;
    const float *la_t = OUTER_TX ? A.ttx_lo : A.trx_lo;
    const float *lb_t = OUTER_TX ? A.trx_lo : A.ttx_lo;
    // positions in samples from double delays: products with 1 / dt (double)
    const double rdt_d = 1.0 / (double)A.dt;
    for (int a = a0; a < a1; ++a) {
        const float ta = ta_t[(int64_t)a * np + p];
        const float wa = wa_t[(int64_t)a * np + p];
        const double la = (fine && la_t) ? (double)la_t[(int64_t)a * np + p] : 0.0;
        for (int b = 0; b < nb; ++b) {
            const float tb = tb_t[(int64_t)b * np + p];
            const float wb = wb_t[(int64_t)b * np + p];
            const int tx = OUTER_TX ? a : b;
            const int rx = OUTER_TX ? b : a;
            const float w = OUTER_TX ? wa * wb : wb * wa;  // apodtx * apodrx
            const float x = OUTER_TX ? ta + tb : tb + ta;  // tautx + taurx
            const int64_t ch0 = (int64_t)f0 * nchan + (int64_t)tx * A.nrx + rx;
            float v[FC];
            bool ok = false;
            const bool use = fabsf(w) >= A.thresh;
            if (use && fine) {
                const double lb = lb_t ? (double)lb_t[(int64_t)b * np + p] : 0.0;
                const double q = (((double)ta + (double)tb) + (la + lb) - (double)A.t0) * rdt_d;
                gather_q<GK, FC>(A, ch0, fstride, nf, x, q, ok, v);
            } else if (use) {
                gather<GK, FC>(A, ch0, fstride, nf, x, ok, v);
            }
            if (SUM == PYCBF_SUM_NONE) {
                float *o = A.out + ((((int64_t)f0 * A.ntx + tx) * A.nrx + rx) * np + p);
#pragma unroll
                for (int f = 0; f < FC; ++f)
                    if (f < nf) o[(int64_t)f * nchan * np] = (use && ok) ? 0.0f + w * v[f] : 0.0f;
            } else if (use && ok) {
#pragma unroll
                for (int f = 0; f < FC; ++f)
                    if (f < nf) acc[f] = __builtin_fmaf(w, v[f], acc[f]);
            }
        }
        if (SUM == PYCBF_SUM_TX_ONLY || SUM == PYCBF_SUM_RX_ONLY) {
#pragma unroll
            for (int f = 0; f < FC; ++f) {
                if (f < nf) A.out[((int64_t)(f0 + f) * na + a) * np + p] = acc[f];
                acc[f] = 0.0f;
            }
        }
    }
    if (SUM == PYCBF_SUM_TX_AND_RX) {
        // gridDim.z > 1: partial sums (z, nfr, np) reduced by k_reduce_split
        float *o = A.out + (int64_t)bz * A.nfr * np;
#pragma unroll
        for (int f = 0; f < FC; ++f)
            if (f < nf) o[(int64_t)(f0 + f) * np + p] = acc[f];
    }
}

// out[g] = part[0][g] + part[1][g] + ... in split order.  V4: four
// consecutive outputs per lane with 16-byte accesses (n a multiple of 4, both
// buffers 16-byte aligned); the unrolled loop keeps eight loads in flight.
template <bool V4>
__global__ __launch_bounds__(256) void k_reduce_split(const float *__restrict__ part, int nsplit,
                                                      int64_t n, float *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (V4) {
        const int64_t g = 4 * t;
        if (g >= n) return;
        float4 s = *reinterpret_cast<const float4 *>(part + g);
#pragma unroll 8
        for (int z = 1; z < nsplit; ++z) {
            const float4 v = *reinterpret_cast<const float4 *>(part + (int64_t)z * n + g);
            s.x += v.x;
            s.y += v.y;
            s.z += v.z;
            s.w += v.w;
        }
        *reinterpret_cast<float4 *>(out + g) = s;
    } else {
        if (t >= n) return;
        float s = part[t];
#pragma unroll 8
        for (int z = 1; z < nsplit; ++z) s += part[(int64_t)z * n + t];
        out[t] = s;
    }
}

template <int GK, int SUM, int FC>
static void launch_one(const DasArgs &a, dim3 grid, hipStream_t s)
{
    hipLaunchKernelGGL((k_das<GK, SUM, FC>), grid, dim3(256), 0, s, a);
}

template <int GK, int SUM>
static void launch_fc(int fc, const DasArgs &a, dim3 grid, hipStream_t s)
{
    switch (fc) {
    case 1: launch_one<GK, SUM, 1>(a, grid, s); break;
    case 2: launch_one<GK, SUM, 2>(a, grid, s); break;
    case 4: launch_one<GK, SUM, 4>(a, grid, s); break;
    case 8: launch_one<GK, SUM, 8>(a, grid, s); break;
    default: launch_one<GK, SUM, 16>(a, grid, s); break;
    }
}

template <int GK>
static void launch_sum(int sum, int fc, const DasArgs &a, dim3 grid, hipStream_t s)
{
    switch (sum) {
    case PYCBF_SUM_NONE: launch_fc<GK, PYCBF_SUM_NONE>(fc, a, grid, s); break;
    case PYCBF_SUM_TX_ONLY: launch_fc<GK, PYCBF_SUM_TX_ONLY>(fc, a, grid, s); break;
    case PYCBF_SUM_RX_ONLY: launch_fc<GK, PYCBF_SUM_RX_ONLY>(fc, a, grid, s); break;
    default: launch_fc<GK, PYCBF_SUM_TX_AND_RX>(fc, a, grid, s); break;
    }
}

bool das_outer_is_tx(int sumflag, int64_t ntx, int64_t nrx)
{
    if (sumflag == PYCBF_SUM_RX_ONLY) return true;
    if (sumflag == PYCBF_SUM_TX_ONLY) return false;
    return ntx <= nrx;  // fewer outer elements: fewer partial images to reduce
}

int das_frames_per_pass(int64_t nfr)
{
    int cap = 16;
    if (const char *e = getenv("PYCBF_DAS_FC")) cap = atoi(e);
    int fc = 1;
    while (fc * 2 <= cap && fc * 2 <= 16 && fc * 2 <= nfr) fc *= 2;
    return fc;
}

int launch_das(int gk, const DasArgs &a0, int nsplit, float *partial, hipStream_t s)
{
    DasArgs a = a0;
    const int fc = das_frames_per_pass(a.nfr);
    a.xcd_remap = 0;
    if (const char *e = getenv("PYCBF_XCD_REMAP")) a.xcd_remap = atoi(e);
    const int na = das_outer_is_tx(a.sumflag, a.ntx, a.nrx) ? a.ntx : a.nrx;
    a.outer_tx = das_outer_is_tx(a.sumflag, a.ntx, a.nrx) ? 1 : 0;
    if (nsplit < 1) nsplit = 1;
    if (nsplit > na) nsplit = na;
    a.a_split = (na + nsplit - 1) / nsplit;
    nsplit = (na + a.a_split - 1) / a.a_split;
    const bool reduce = (a.sumflag == PYCBF_SUM_TX_AND_RX) && nsplit > 1;
    float *final_out = a.out;
    if (reduce) a.out = partial;
    dim3 grid((unsigned)((a.np + 255) / 256), (unsigned)((a.nfr + fc - 1) / fc), (unsigned)nsplit);
    {
    ProfScope prof("das_gather", s);
    switch (gk) {
    case GK_NEAREST: launch_sum<GK_NEAREST>(a.sumflag, fc, a, grid, s); break;
    case GK_LINEAR: launch_sum<GK_LINEAR>(a.sumflag, fc, a, grid, s); break;
    case GK_CUBIC: launch_sum<GK_CUBIC>(a.sumflag, fc, a, grid, s); break;
    case GK_KORDER: launch_sum<GK_KORDER>(a.sumflag, fc, a, grid, s); break;
    default: launch_sum<GK_AKIMA>(a.sumflag, fc, a, grid, s); break;
    }
    }
    int rc = check_launch("das gather");
    if (rc || !reduce) return rc;
    return launch_reduce_split(partial, nsplit, (int64_t)a.nfr * a.np, final_out, s);
}

int launch_reduce_split(const float *part, int nsplit, int64_t n, float *out, hipStream_t s)
{
    if (n <= 0) return 0;
    ProfScope prof("reduce_split", s);
    const bool v4 = (n & 3) == 0 && (reinterpret_cast<uintptr_t>(part) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    if (v4)
        hipLaunchKernelGGL(k_reduce_split<true>, dim3((unsigned)((n / 4 + 255) / 256)), dim3(256),
                           0, s, part, nsplit, n, out);
    else
        hipLaunchKernelGGL(k_reduce_split<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           s, part, nsplit, n, out);
    return check_launch("das split reduce");
}

template <int NDIM>
__global__ __launch_bounds__(256) void k_synth_tables(int ntx, int nrx,
                                                      const float *__restrict__ txpack,
                                                      const float *__restrict__ rxpack, float c0,
                                                      const float *__restrict__ pvec, int64_t np,
                                                      float *ttx, float *atx, float *trx,
                                                      float *arx)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int e = blockIdx.y;
    float pv[NDIM];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) pv[d] = pvec[p * NDIM + d];
    float tau, apod;
    if (e < ntx) {
        geom_tx<NDIM>(txpack + (int64_t)e * PYCBF_TXPACK_FLOATS, c0, pv, tau, apod);
        ttx[(int64_t)e * np + p] = tau;
        atx[(int64_t)e * np + p] = apod;
    } else {
        const int r = e - ntx;
        geom_rx<NDIM>(rxpack + (int64_t)r * PYCBF_RXPACK_FLOATS, c0, pv, tau, apod);
        trx[(int64_t)r * np + p] = tau;
        arx[(int64_t)r * np + p] = apod;
    }
}

int launch_synth_tables(int ndim, int ntx, int nrx, const float *txpack, const float *rxpack,
                        float c0, const float *pvec, int64_t np, float *ttx, float *atx,
                        float *trx, float *arx, hipStream_t s)
{
    if (np <= 0) return 0;
    dim3 grid((unsigned)((np + 255) / 256), (unsigned)(ntx + nrx));
    ProfScope prof("synthetic_tables", s);
    if (ndim == 3)
        hipLaunchKernelGGL(k_synth_tables<3>, grid, dim3(256), 0, s, ntx, nrx, txpack, rxpack, c0,
                           pvec, np, ttx, atx, trx, arx);
    else
        hipLaunchKernelGGL(k_synth_tables<2>, grid, dim3(256), 0, s, ntx, nrx, txpack, rxpack, c0,
                           pvec, np, ttx, atx, trx, arx);
    return check_launch("synthetic tables");
}

// ---------------------------------------------------------------------------
// index selection exposed for parity tests
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_debug_indices(int ntx, int nrx,
                                                       const float *__restrict__ ttx,
                                                       const float *__restrict__ trx, int64_t np,
                                                       int nearest, float t0, float dt, int n,
                                                       int32_t *__restrict__ idx)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int tx = blockIdx.y / nrx, rx = blockIdx.y % nrx;
    const float x = ttx[(int64_t)tx * np + p] + trx[(int64_t)rx * np + p];
    int j = -1;
    if (nearest) {
        if (!nearest_index(x, t0, dt, n, j)) j = -1;
    } else {
        const float tmax = t0 + dt * (float)(n - 1);
        if (!knot_index(x, t0, dt, tmax, n, j)) j = -1;
    }
    idx[(int64_t)blockIdx.y * np + p] = j;
}

int launch_debug_indices(int ntx, int nrx, const float *ttx, const float *trx, int64_t np,
                         int nearest, float t0, float dt, int n, int32_t *idx, hipStream_t s)
{
    if (np <= 0) return 0;
    dim3 grid((unsigned)((np + 255) / 256), (unsigned)(ntx * nrx));
    hipLaunchKernelGGL(k_debug_indices, grid, dim3(256), 0, s, ntx, nrx, ttx, trx, np, nearest, t0,
                       dt, n, idx);
    return check_launch("debug indices");
}

}  // namespace pycbf
