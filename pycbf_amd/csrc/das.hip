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
#include <math.h>
#include <stdlib.h>

#include "common.h"
#include "prof.h"

namespace pycbf {

// ---------------------------------------------------------------------------
// sample-index selection, shared by the gather kernels and the debug entry
// ---------------------------------------------------------------------------

// Knot index of pycbf/cpu/pycbf.c:75-87: fill outside [t0, tmax] (NaN too).
// The clamp to the last interval only matters when t0 is so large that tmax
// rounds more than one sample past the end (the reference then reads out of
// bounds); for every other input the truncated quotient is already <= n-1.
__device__ __forceinline__ bool knot_index(float x, float t0, float dt, float tmax, int n, int &j)
{
    if (!(x >= t0 && x <= tmax)) return false;
    j = min((int)((x - t0) / dt), n - 1);
    return true;
}

// Nearest index of pycbf/cpu/pycbf.c:264,267: (int)(q + 0.5) with the 0.5 added
// in double; valid when 0 <= it < n.
__device__ __forceinline__ bool nearest_index(float x, float t0, float dt, int n, int &it)
{
    const float q = (x - t0) / dt;
    const double d = (double)q + 0.5;
    if (!(d > -1.0 && d < (double)n)) return false;
    it = (int)d;
    return true;
}

// Two consecutive (y, h*y2) table entries, loaded as one 16-byte access
// from an 8-byte aligned address (gfx950 serves unaligned dwordx4 loads).
struct __attribute__((aligned(8))) Pair2 {
    float y0, z0, y1, z1;
};

// ---------------------------------------------------------------------------
// gather + interpolate for FC frames of one channel
// ---------------------------------------------------------------------------
template <int GK, int FC>
__device__ __forceinline__ void gather(const DasArgs &A, int64_t ch0, int64_t fstride, int nf,
                                       float x, bool &ok, float (&v)[FC])
{
    if (GK == GK_NEAREST) {
        int it = 0;
        ok = nearest_index(x, A.t0, A.dt, A.nsrc, it);
        if (!ok) return;
        const float *s = A.src + ch0 * A.ch_stride + it;
#pragma unroll
        for (int f = 0; f < FC; ++f)
            if (f < nf) v[f] = s[f * fstride];
    } else if (GK == GK_LINEAR) {
        int j = 0;
        ok = knot_index(x, A.t0, A.dt, A.tmax, A.nsrc, j);
        if (!ok) return;
        const float xlead = A.t0 + A.dt * (float)(j + 1);
        const float xlag = A.t0 + A.dt * (float)j;
        const float a = (xlead - x) / A.dt;
        const float b = (x - xlag) / A.dt;
        const bool hi = (j + 1) < A.nsrc;
        const float *s = A.src + ch0 * A.ch_stride + j;
#pragma unroll
        for (int f = 0; f < FC; ++f)
            if (f < nf) {
                const float y0 = s[f * fstride];
                const float y1 = hi ? s[f * fstride + 1] : 0.0f;
                v[f] = a * y0 + b * y1;
            }
    } else if (GK == GK_CUBIC) {
        // value = a*y_j + b*y_j+1 + (a^3-a)*h*y2_j + (b^3-b)*h*y2_j+1 with
        // a, b from separately rounded knot times (pycbf/cpu/pycbf.c:93-103):
        // a + b differs from 1 by up to a few 1e-5, and that is kept
        int j = 0;
        ok = knot_index(x, A.t0, A.dt, A.tmax, A.nsrc, j);
        if (!ok) return;
        const float xlead = A.t0 + A.dt * (float)(j + 1);
        const float xlag = A.t0 + A.dt * (float)j;
        const float a = (xlead - x) * A.rdt;
        const float b = (x - xlag) * A.rdt;
        const float ca = __builtin_fmaf(a * a, a, -a);
        const float cb = __builtin_fmaf(b * b, b, -b);
        // entries j and j+1 of the (y, h*y2) table: one 16-byte load at 8-byte alignment
        const float *s = A.src + ch0 * A.ch_stride + 2 * j;
#pragma unroll
        for (int f = 0; f < FC; ++f)
            if (f < nf) {
                const Pair2 c = *reinterpret_cast<const Pair2 *>(s + f * fstride);
                float r = a * c.y0;
                r = __builtin_fmaf(b, c.y1, r);
                r = __builtin_fmaf(ca, c.z0, r);
                v[f] = __builtin_fmaf(cb, c.z1, r);
            }
    } else {
        int j = 0;
        ok = knot_index(x, A.t0, A.dt, A.tmax, A.nsrc, j);
        if (!ok) return;
        const float t = (GK == GK_KORDER) ? ((x - A.t0) - A.dt * (float)j) * A.rdt
                                          : (x - (A.t0 + A.dt * (float)j)) * A.rdt;
        const float4 *s = reinterpret_cast<const float4 *>(A.src) + ch0 * (A.ch_stride >> 2) + j;
        const int64_t fs4 = fstride >> 2;
#pragma unroll
        for (int f = 0; f < FC; ++f)
            if (f < nf) {
                const float4 c = s[f * fs4];
                v[f] = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c.w, t, c.z), t, c.y), t, c.x);
            }
    }
}

// Workgroups go to the 8 XCDs round robin by linear id.  With remap on, the
// linear id is permuted so each XCD receives one contiguous run of pixel
// tiles: tiles that are neighbours in the (spatially compact) processing order
// read overlapping sample windows of every channel, and now share an L2.
__device__ __forceinline__ void block_coords(int remap, uint32_t &bx, uint32_t &by, uint32_t &bz)
{
    const uint32_t gx = gridDim.x, gy = gridDim.y;
    uint32_t lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (remap) {
        const uint32_t per = (gx * gy * gridDim.z) >> 3;
        if (lin < (per << 3)) lin = (lin & 7u) * per + (lin >> 3);
    }
    bx = lin % gx;
    const uint32_t r = lin / gx;
    by = r % gy;
    bz = r / gy;
}

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

    for (int a = a0; a < a1; ++a) {
        const float ta = ta_t[(int64_t)a * np + p];
        const float wa = wa_t[(int64_t)a * np + p];
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
            if (use) gather<GK, FC>(A, ch0, fstride, nf, x, ok, v);
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

__global__ __launch_bounds__(256) void k_reduce_split(const float *__restrict__ part, int nsplit,
                                                      int64_t n, float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float s = part[g];
    for (int z = 1; z < nsplit; ++z) s += part[(int64_t)z * n + g];
    out[g] = s;
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
    const int64_t n = (int64_t)a.nfr * a.np;
    hipLaunchKernelGGL(k_reduce_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       partial, nsplit, n, final_out);
    return check_launch("das split reduce");
}

// ---------------------------------------------------------------------------
// synthetic-point geometry (pycbf/gpu/__engines__.cu:382-470), float32, unfused
// ---------------------------------------------------------------------------

// "acos(v) > ala" for the T of the element's pack; a NaN acos (|v| > 1) is false.
__device__ __forceinline__ bool acos_gt(float v, float T) { return (v >= -1.0f && v <= 1.0f) && (v < T); }
// "acos(v) <= ala"
__device__ __forceinline__ bool acos_le(float v, float T) { return (v >= -1.0f && v <= 1.0f) && !(v < T); }

template <int NDIM>
__device__ __forceinline__ void geom_tx(const float *__restrict__ pk, float c0, const float *p,
                                        float &tau, float &apod)
{
    float mag = 0.0f, proj = 0.0f;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        const float di = p[d] - pk[d];
        mag += di * di;
        proj += di * pk[3 + d];
    }
    mag = sqrtf(mag);
    const float t0 = pk[6];
    const int mode = (int)pk[7];
    const float T = pk[8];
    tau = 0.0f;
    if (mode == 0) {  // plane wave
        tau = proj / c0;
        apod = 1.0f;
        return;
    }
    const float sgn = (proj > 0.0f) ? 1.0f : -1.0f;
    if (mode == 2) {  // focus with depth of field
        const bool in_cone = ((double)fabsf(proj) > 1e-12) && ((double)mag > 1e-12) &&
                             acos_le(fabsf(proj / mag), T);
        if (in_cone) {
            tau = sgn * (mag / c0) + t0;
            apod = 1.0f;
        } else {
            const float lat = sqrtf(fabsf(mag * mag - proj * proj));
            if ((double)lat < (double)pk[9] / 2.0) {
                tau = proj * pk[10] / c0 + t0;
                apod = 1.0f;
            } else {
                apod = 0.0f;
            }
        }
        return;
    }
    tau = ((double)fabsf(proj) > 1e-12) ? sgn * (mag / c0) + t0 : t0;
    apod = (((double)mag > 1e-12) && acos_gt(fabsf(proj / mag), T)) ? 0.0f : 1.0f;
}

template <int NDIM>
__device__ __forceinline__ void geom_rx(const float *__restrict__ pk, float c0, const float *p,
                                        float &tau, float &apod)
{
    float mag = 0.0f, proj = 0.0f;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        const float di = p[d] - pk[d];
        mag += di * di;
        proj += di * pk[3 + d];
    }
    mag = sqrtf(mag);
    tau = mag / c0;
    apod = (((double)mag > 1e-12) && acos_gt(proj / mag, pk[6])) ? 0.0f : 1.0f;
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
