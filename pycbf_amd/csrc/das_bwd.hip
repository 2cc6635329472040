// Adjoint (backward) kernels of the Tabbed DAS, for gradients of the images
// with respect to the channel data, the delays and the apodisations.
//
// Forward model (pycbf/cpu/pycbf.c:179-207 with the GPU layouts of
// pycbf/gpu/__engines__.cu:813-864):
//   out[f, (tx), (rx), p] += w * v(rf[f, tx, rx, :], x),
//   w = apodtx[tx, p] * apodrx[rx, p] (only where |w| >= thresh),
//   x = tautx[tx, p] + taurx[rx, p],
// with v the nearest / linear / natural-cubic interpolant of the forward
// kernels (same knot index, same separately rounded weights a and b, zero
// outside [t0, t0 + dt (nt-1)]).
//
// Gradients, for an upstream gradient G of the output:
//   d tautx[tx, p]  = sum_f sum_rx G w v'(x)     d apodtx[tx, p] = sum_f sum_rx G apodrx v(x)
//   d taurx[rx, p]  = sum_f sum_tx G w v'(x)     d apodrx[rx, p] = sum_f sum_tx G apodtx v(x)
//   d rf[f, tx, rx, k] = sum_p G w dv(x)/dy_k                       (transposed gather)
// k_das_bwd_tau evaluates the first four with one lane per (element, pixel),
// recomputing x and w from the tables.  The channel-data gradient is a
// scatter into 64-bit fixed-point LDS bins (scale from k_bwd_scale's bound):
// k_das_bwd_rf_ch (traces up to 4096 samples) gives each workgroup one
// channel and writes its rows once; k_das_bwd_rf (longer traces) gives each
// workgroup a (pixel tile, transmit, frame), bins the tile's sample window
// per receive channel (k_tile_bounds) and flushes it with one float atomic
// per non-zero bin.  For the natural
// cubic the bins hold the cotangents of the samples y and of the table's
// e = dt^2/6 y2, and k_spline_adjoint(_seg) adds the transposed spline solve
//   dy += D2(T^-1 de),  T = tridiag(1, 4, 1),  D2 the second difference,
// which is the adjoint of e = T^-1 D2 y (the natural spline with y2 = 0 at
// both ends).  For the Hermite kinds (korder_cubic, akima, makima) the bins
// hold the cotangents of the four coefficients of each interval (the value
// is c.x + c.y t + c.z t^2 + c.w t^3), and k_korder_adjoint /
// k_akima_adjoint transpose their construction from the samples
// (k_korder_coefs: end values and S-derivatives over a k-sample window;
// k_akima_coefs: end values and Akima derivatives).
#include <math.h>

#include <algorithm>

#include "common.h"
#include "das_dev.h"
#include "prof.h"
#include "tiles.h"

namespace pycbf {

namespace {

constexpr int BWD_BINS = 2048;  // LDS bins per array and channel window

// value and derivative of the interpolant at x (0 outside the trace)
// FD: the knot quotient by dt through div_rn (dt accepted by
// recip_division_exact, rdt = 1 / dt rounded): the same index
template <int GK, bool FD = false>
__device__ __forceinline__ bool interp_vd(const float *__restrict__ src, int64_t row, int n,
                                          float x, float t0, float dt, float rdt, float tmax,
                                          float &v, float &d)
{
    auto knot = [&](int &j) {
        return FD ? knot_index_rn(x, t0, dt, rdt, tmax, n, j) : knot_index(x, t0, dt, tmax, n, j);
    };
    if (GK == GK_NEAREST) {
        int it = 0;
        if (!nearest_index(x, t0, dt, n, it)) return false;
        v = src[row * n + it];
        d = 0.0f;
        return true;
    } else if (GK == GK_KORDER || GK == GK_AKIMA) {
        int j = 0;
        if (!knot(j)) return false;
        // the forward's roundings of t (korder_cubic / akima)
        const float t = (GK == GK_KORDER) ? ((x - t0) - dt * (float)j) * rdt
                                          : (x - (t0 + dt * (float)j)) * rdt;
        const float4 c = reinterpret_cast<const float4 *>(src)[row * n + j];
        v = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c.w, t, c.z), t, c.y), t, c.x);
        d = __builtin_fmaf(__builtin_fmaf(3.0f * c.w, t, 2.0f * c.z), t, c.y) * rdt;
        return true;
    } else {
        int j = 0;
        if (!knot(j)) return false;
        const float xlead = t0 + dt * (float)(j + 1);
        const float xlag = t0 + dt * (float)j;
        if (GK == GK_LINEAR) {
            const float a = (xlead - x) / dt;
            const float b = (x - xlag) / dt;
            const float y0 = src[row * n + j];
            const float y1 = (j + 1 < n) ? src[row * n + j + 1] : 0.0f;
            v = a * y0 + b * y1;
            d = (y1 - y0) / dt;
        } else {
            const float a = (xlead - x) * rdt;
            const float b = (x - xlag) * rdt;
            const float ca = __builtin_fmaf(a * a, a, -a);
            const float cb = __builtin_fmaf(b * b, b, -b);
            const Pair2 c = *reinterpret_cast<const Pair2 *>(src + (row * (n + 1) + j) * 2);
            float r = a * c.y0;
            r = __builtin_fmaf(b, c.y1, r);
            r = __builtin_fmaf(ca, c.z0, r);
            v = __builtin_fmaf(cb, c.z1, r);
            const float da = __builtin_fmaf(3.0f * a, a, -1.0f);
            const float db = __builtin_fmaf(3.0f * b, b, -1.0f);
            d = ((c.y1 - c.y0) + (db * c.z1 - da * c.z0)) * rdt;
        }
        return true;
    }
}

// Derivative of the interpolant at the sample position q = (x - t0) / dt held
// in double (x placed to sub-float32 precision by delay residuals): the same
// piece as interp_vd, evaluated at the exact fractional position instead of
// the float32 rounding of x and of the knot times.
template <int GK>
__device__ __forceinline__ float interp_d_at(const float *__restrict__ src, int64_t row, int n,
                                             double q, float rdt)
{
    if (GK == GK_NEAREST) return 0.0f;
    int j = (int)floor(q);
    j = min(max(j, 0), n - 1);
    const double fr = q - (double)j;
    if (GK == GK_KORDER || GK == GK_AKIMA) {
        const float t = (float)fr;
        const float4 c = reinterpret_cast<const float4 *>(src)[row * n + j];
        return __builtin_fmaf(__builtin_fmaf(3.0f * c.w, t, 2.0f * c.z), t, c.y) * rdt;
    }
    if (GK == GK_LINEAR) {
        const float y0 = src[row * n + j];
        const float y1 = (j + 1 < n) ? src[row * n + j + 1] : 0.0f;
        return (y1 - y0) * rdt;
    }
    const float b = (float)fr, a = (float)(1.0 - fr);
    const Pair2 c = *reinterpret_cast<const Pair2 *>(src + (row * (n + 1) + j) * 2);
    const float da = __builtin_fmaf(3.0f * a, a, -1.0f);
    const float db = __builtin_fmaf(3.0f * b, b, -1.0f);
    return ((c.y1 - c.y0) + (db * c.z1 - da * c.z0)) * rdt;
}

__device__ __forceinline__ int64_t gout_index(int sumflag, int f, int tx, int rx, int ntx, int nrx,
                                              int64_t np, int64_t p)
{
    switch (sumflag) {
    case PYCBF_SUM_NONE: return (((int64_t)f * ntx + tx) * nrx + rx) * np + p;
    case PYCBF_SUM_TX_ONLY: return ((int64_t)f * nrx + rx) * np + p;
    case PYCBF_SUM_RX_ONLY: return ((int64_t)f * ntx + tx) * np + p;
    default: return (int64_t)f * np + p;
    }
}

}  // namespace

// pixel position p of the synthetic geometry (GEO = ndim; nothing for tables)
template <int GEO>
__device__ __forceinline__ void bwd_pixel(const BwdArgs &A, int64_t p, float (&pv)[3])
{
    pv[0] = pv[1] = pv[2] = 0.0f;
#pragma unroll
    for (int d = 0; d < GEO; ++d) pv[d] = A.pvec[p * GEO + d];
}

// delay and weight of element e of the transmit (tx) or receive axis at pixel
// p: the tables (GEO 0), or the synthetic geometry evaluated in registers with
// the forward's float32 operations (GEO = ndim)
// knot index of the adjoint's trace (FD: the quotient by dt through div_rn)
template <bool FD>
__device__ __forceinline__ bool knot_at(const BwdArgs &A, float x, int &j)
{
    return FD ? knot_index_rn(x, A.t0, A.dt, A.rdt, A.tmax, A.nt, j)
              : knot_index(x, A.t0, A.dt, A.tmax, A.nt, j);
}

template <int GEO, bool FD = false>
__device__ __forceinline__ void bwd_elem(const BwdArgs &A, bool tx, int e, int64_t p,
                                         const float (&pv)[3], float &tau, float &apod)
{
    if (GEO == 0) {
        const int64_t o = (int64_t)e * A.np + p;
        tau = tx ? A.ttx[o] : A.trx[o];
        apod = tx ? A.atx[o] : A.arx[o];
    } else if (tx) {
        geom_tx<GEO, FD>(A.txpack + (int64_t)e * PYCBF_TXPACK_FLOATS, A.c0, pv, tau, apod, A.rc0);
    } else {
        geom_rx<GEO, FD>(A.rxpack + (int64_t)e * PYCBF_RXPACK_FLOATS, A.c0, pv, tau, apod, A.rc0);
    }
}

// one lane per (outer element, pixel); outer axis tx (OUTER_TX) or rx.  With
// the synthetic geometry (GEO > 0) and A.gc0part set, each workgroup also
// writes its share of d loss / d c0 = sum over (element, pixel) of the delay
// gradient times d tau / d c0 = -tau / c0: every delay of the geometry is a
// path length over c0, the transmit's time intercept included (pycbf's
// helpers compute it as a distance over the sound speed,
// pycbf/helpers/__full_aperture_2D__.py:259,300,346).
template <int GK, int OUTER_TX, int GEO, bool FD>
__global__ __launch_bounds__(256) void k_das_bwd_tau(BwdArgs A)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int o = blockIdx.y;
    const bool act = p < A.np;
    const int ninner = OUTER_TX ? A.nrx : A.ntx;
    float pv[3];
    bwd_pixel<GEO>(A, act ? p : 0, pv);
    float tau_o = 0.0f, ap_o = 0.0f;
    if (act) bwd_elem<GEO, FD>(A, OUTER_TX, o, p, pv, tau_o, ap_o);
    // delay residuals: the derivative is taken at the position in double
    const float *lo_o_t = OUTER_TX ? A.ttx_lo : A.trx_lo;
    const float *lo_i_t = OUTER_TX ? A.trx_lo : A.ttx_lo;
    const bool fine = GEO == 0 && (A.ttx_lo || A.trx_lo) && GK != GK_NEAREST;
    const double lo_o = (fine && lo_o_t && act) ? (double)lo_o_t[(int64_t)o * A.np + p] : 0.0;
    float gt = 0.0f, ga = 0.0f;
    const double rdt_d = 1.0 / (double)A.dt;  // positions in samples: products with 1 / dt
    for (int i = 0; act && i < ninner; ++i) {
        float tau_i = 0.0f, ap_i = 0.0f;
        if (GEO == 0) {
            ap_i = (OUTER_TX ? A.arx : A.atx)[(int64_t)i * A.np + p];
        } else {
            bwd_elem<GEO, FD>(A, !OUTER_TX, i, p, pv, tau_i, ap_i);
        }
        const float w = OUTER_TX ? ap_o * ap_i : ap_i * ap_o;
        if (!(fabsf(w) >= A.thresh)) continue;
        if (GEO == 0) tau_i = (OUTER_TX ? A.trx : A.ttx)[(int64_t)i * A.np + p];
        const float x = OUTER_TX ? tau_o + tau_i : tau_i + tau_o;
        const int tx = OUTER_TX ? o : i, rx = OUTER_TX ? i : o;
        double q = 0.0;
        if (fine) {
            const double lo_i = lo_i_t ? (double)lo_i_t[(int64_t)i * A.np + p] : 0.0;
            const double xd = ((double)tau_o + (double)tau_i) + (lo_o + lo_i);
            q = (xd - (double)A.t0) * rdt_d;
        }
        for (int f = 0; f < A.nfr; ++f) {
            const int64_t row = ((int64_t)f * A.ntx + tx) * A.nrx + rx;
            float v = 0.0f, d = 0.0f;
            if (!interp_vd<GK, FD>(A.src, row, A.nt, x, A.t0, A.dt, A.rdt, A.tmax, v, d)) break;
            if (fine) d = interp_d_at<GK>(A.src, row, A.nt, q, A.rdt);
            const float g = A.gout[gout_index(A.sumflag, f, tx, rx, A.ntx, A.nrx, A.np, p)];
            gt = __builtin_fmaf(g * w, d, gt);
            ga = __builtin_fmaf(g * ap_i, v, ga);
        }
    }
    if (act && A.gtau) A.gtau[(int64_t)o * A.np + p] = gt;
    if (act && A.gapod) A.gapod[(int64_t)o * A.np + p] = ga;
    if (GEO > 0 && A.gc0part) {
        // fixed-order workgroup sum in double: wave butterflies, then the 4 waves
        double c = (act && gt != 0.0f) ? (double)gt * (-(double)tau_o / (double)A.c0) : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
        __shared__ double wsum[4];
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0)
            A.gc0part[A.gc0_off + (int64_t)o * gridDim.x + blockIdx.x] =
                (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
    }
}

// d loss / d c0 from the partials of both delay-gradient launches, summed in a
// fixed order (one workgroup): strided per-lane sums, then a tree
__global__ __launch_bounds__(256) void k_gc0_sum(const double *__restrict__ part, int64_t n,
                                                 float *__restrict__ gc0)
{
    __shared__ double t[256];
    double c = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) c += part[i];
    t[threadIdx.x] = c;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) t[threadIdx.x] += t[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *gc0 = (float)t[0];
}

// max |value| of gout, apodtx and apodrx (blockIdx.y = 0, 1, 2), as bit
// patterns in mbits[blockIdx.y] (non-negative floats order like their bits;
// a NaN or an infinity records the infinity pattern).  |gout * apodtx *
// apodrx| <= m0 m1 m2 bounds every contribution of the transposed gather.
__global__ __launch_bounds__(256) void k_bwd_scale(BwdArgs A, int64_t ngout)
{
    const int which = blockIdx.y;
    const float *x = which == 0 ? A.gout : (which == 1 ? A.atx : A.arx);
    if (!x) {
        // synthetic geometry: its weights are 0 or 1
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(A.mbits + which, __float_as_uint(1.0f));
        return;
    }
    const int64_t n = which == 0 ? ngout : (int64_t)(which == 1 ? A.ntx : A.nrx) * A.np;
    auto fold = [](float m, float v) {
        v = fabsf(v);
        return fmaxf(m, (v <= 3.0e38f) ? v : __int_as_float(0x7f800000));
    };
    float m = 0.0f;
    const int64_t stride = (int64_t)gridDim.x * 256;
    int64_t head = 0;  // scalar elements before the first 16-byte boundary
    if (reinterpret_cast<uintptr_t>(x) & 15)
        head = min((int64_t)((16 - (reinterpret_cast<uintptr_t>(x) & 15)) / 4), n);
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < head; q += stride)
        m = fold(m, x[q]);
    const float4 *x4 = reinterpret_cast<const float4 *>(x + head);
    const int64_t n4 = (n - head) / 4;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n4; q += stride) {
        const float4 v = x4[q];
        m = fold(fold(fold(fold(m, v.x), v.y), v.z), v.w);
    }
    for (int64_t q = head + 4 * n4 + (int64_t)blockIdx.x * 256 + threadIdx.x; q < n; q += stride)
        m = fold(m, x[q]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    // one atomic per workgroup: atomics on one word serialise in L2
    __shared__ float wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(A.mbits + which,
                  __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]))));
}

// Fixed-point scale of the LDS bins, or 0 when a bound is not finite (the
// bins then take float atomics).  A bin receives at most 2^per_bin
// contributions, each below M (1 + 2^-22) with M = m0 m1 m2; the scale
// 2^(59 - per_bin - ilogb M) keeps every bin below 2^61 and quantises a
// contribution to 2^(per_bin - 59) M.
// per_bin: log2 of the most contributions one bin can receive (9 for the
// 2 x 256 of a pixel tile).
__device__ __forceinline__ double bwd_scale(const unsigned *mbits, int per_bin = 9)
{
    const float m0 = __uint_as_float(mbits[0]), m1 = __uint_as_float(mbits[1]),
                m2 = __uint_as_float(mbits[2]);
    if (!(m0 <= 3.0e38f && m1 <= 3.0e38f && m2 <= 3.0e38f)) return 0.0;
    const double M = (double)m0 * (double)m1 * (double)m2;
    if (!(M > 0.0)) return 1.0;  // every contribution is zero
    return ldexp(1.0, 59 - per_bin - ilogb(M));
}

// transposed gather: one workgroup per (pixel tile, transmit, frame).  The
// LDS bins are 64-bit fixed-point integers (ds_add_u64): on gfx950 integer
// LDS atomics run at about 37x the rate of float LDS atomics, and integer
// sums do not depend on the order of the additions.
template <int GK, int GEO, bool FD>
__global__ __launch_bounds__(256) void k_das_bwd_rf(BwdArgs A)
{
    __shared__ unsigned long long by[BWD_BINS];
    __shared__ unsigned long long be[GK == GK_CUBIC ? BWD_BINS : 1];
    float *byf = reinterpret_cast<float *>(by);  // float bins (bound not finite)
    float *bef = reinterpret_cast<float *>(be);
    const int64_t tile = blockIdx.x;
    const int tx = blockIdx.y, f = blockIdx.z;
    const int64_t p = tile * 256 + threadIdx.x;
    const bool act = p < A.np;
    const int nent = A.ntx + A.nrx;
    const float *bt = A.bounds + (tile * nent + tx) * 2;
    const float txlo = bt[0], txhi = bt[1];
    float pv[3];
    bwd_pixel<GEO>(A, act ? p : 0, pv);
    float ttx = 0.0f, atx = 0.0f;
    if (act) bwd_elem<GEO, FD>(A, true, tx, p, pv, ttx, atx);
    const double scale = bwd_scale(A.mbits);
    const bool fix = scale > 0.0;  // uniform
    const double iscale = fix ? 1.0 / scale : 0.0;
    auto bin_add = [&](unsigned long long *bins, float *binsf, int i, float v) {
        if (fix)
            atomicAdd(bins + i, (unsigned long long)(long long)((double)v * scale));
        else
            atomicAdd(binsf + i, v);
    };
    auto bin_take = [&](unsigned long long *bins, float *binsf, int i) -> float {
        float v;
        if (fix) {
            v = (float)((double)(long long)bins[i] * iscale);
            bins[i] = 0ull;
        } else {
            v = binsf[i];
            binsf[i] = 0.0f;
        }
        return v;
    };
    for (int t = threadIdx.x; t < BWD_BINS; t += 256) {
        by[t] = 0ull;
        if (GK == GK_CUBIC) be[t] = 0ull;
    }
    __syncthreads();
    for (int rx = 0; rx < A.nrx; ++rx) {
        const float *br = A.bounds + (tile * nent + A.ntx + rx) * 2;
        const float xlo = txlo + br[0], xhi = txhi + br[1];
        int lo = 0, cnt = 0;
        if (xlo <= xhi) {
            if (GK == GK_NEAREST) {
                const double dlo = (double)((xlo - A.t0) / A.dt) + 0.5;
                const double dhi = (double)((xhi - A.t0) / A.dt) + 0.5;
                if (dhi > -1.0 && dlo < (double)A.nt) {
                    lo = (dlo > -1.0) ? (int)dlo : 0;
                    const int hi = (dhi < (double)A.nt) ? (int)dhi : A.nt - 1;
                    cnt = hi - lo + 1;
                }
            } else if (!(xhi < A.t0 || xlo > A.tmax)) {
                lo = (xlo < A.t0) ? 0 : min((int)((xlo - A.t0) / A.dt), A.nt - 1);
                const int hi = (xhi > A.tmax) ? A.nt - 1 : min((int)((xhi - A.t0) / A.dt), A.nt - 1);
                cnt = hi - lo + 2;
            }
        }
        if (cnt == 0) continue;  // uniform: no pixel of the tile uses this channel
        // Hermite kinds: four cotangents per interval, added with global atomics
        const bool binned = GK != GK_KORDER && GK != GK_AKIMA && cnt <= BWD_BINS;
        const int64_t chan = ((int64_t)f * A.ntx + tx) * A.nrx + rx;
        float *gy = A.gy + chan * A.gstride;
        float *ge = (GK == GK_CUBIC) ? A.ge + chan * A.gstride : nullptr;
        if (act) {
            float trx = 0.0f, arx = 0.0f;
            if (GEO == 0) arx = A.arx[(int64_t)rx * A.np + p];
            else bwd_elem<GEO, FD>(A, false, rx, p, pv, trx, arx);
            const float w = atx * arx;
            if (fabsf(w) >= A.thresh) {
                if (GEO == 0) trx = A.trx[(int64_t)rx * A.np + p];
                const float x = ttx + trx;
                const float g = A.gout[gout_index(A.sumflag, f, tx, rx, A.ntx, A.nrx, A.np, p)] * w;
                if (GK == GK_NEAREST) {
                    int it = 0;
                    if (nearest_index(x, A.t0, A.dt, A.nt, it) && g != 0.0f) {
                        if (binned) bin_add(by, byf, it - lo, g);
                        else unsafeAtomicAdd(gy + it, g);
                    }
                } else if (GK == GK_KORDER || GK == GK_AKIMA) {
                    int j = 0;
                    if (knot_at<FD>(A, x, j) && g != 0.0f) {
                        const float t = (GK == GK_KORDER)
                                            ? ((x - A.t0) - A.dt * (float)j) * A.rdt
                                            : (x - (A.t0 + A.dt * (float)j)) * A.rdt;
                        const int64_t o = chan * A.gstride + j;
                        unsafeAtomicAdd(A.gy + o, g);
                        unsafeAtomicAdd(A.ge + o, g * t);
                        unsafeAtomicAdd(A.gc2 + o, g * (t * t));
                        unsafeAtomicAdd(A.gc3 + o, g * ((t * t) * t));
                    }
                } else {
                    int j = 0;
                    if (knot_at<FD>(A, x, j) && g != 0.0f) {
                        const float xlead = A.t0 + A.dt * (float)(j + 1);
                        const float xlag = A.t0 + A.dt * (float)j;
                        float a, b;
                        if (GK == GK_LINEAR) {
                            a = (xlead - x) / A.dt;
                            b = (x - xlag) / A.dt;
                        } else {
                            a = (xlead - x) * A.rdt;
                            b = (x - xlag) * A.rdt;
                        }
                        // sample j+1 exists for the cubic table (zero entry at nt); for
                        // linear the forward reads 0 past the last sample
                        const bool up = (GK == GK_CUBIC) || (j + 1 < A.nt);
                        if (binned) {
                            bin_add(by, byf, j - lo, g * a);
                            if (up) bin_add(by, byf, j + 1 - lo, g * b);
                        } else {
                            unsafeAtomicAdd(gy + j, g * a);
                            if (up) unsafeAtomicAdd(gy + j + 1, g * b);
                        }
                        if (GK == GK_CUBIC) {
                            const float ca = __builtin_fmaf(a * a, a, -a);
                            const float cb = __builtin_fmaf(b * b, b, -b);
                            if (binned) {
                                bin_add(be, bef, j - lo, g * ca);
                                bin_add(be, bef, j + 1 - lo, g * cb);
                            } else {
                                unsafeAtomicAdd(ge + j, g * ca);
                                unsafeAtomicAdd(ge + j + 1, g * cb);
                            }
                        }
                    }
                }
            }
        }
        if (!binned) continue;  // uniform
        __syncthreads();
        // flush the window and clear the bins for the next channel
        for (int t = threadIdx.x; t < cnt; t += 256) {
            const float vy = bin_take(by, byf, t);
            if (vy != 0.0f && lo + t < A.gstride) unsafeAtomicAdd(gy + lo + t, vy);
            if (GK == GK_CUBIC) {
                const float ve = bin_take(be, bef, t);
                if (ve != 0.0f) unsafeAtomicAdd(ge + lo + t, ve);
            }
        }
        __syncthreads();
    }
}

// Transposed gather with one workgroup per channel (frame, transmit,
// receive) for traces of up to 4096 samples: the workgroup sweeps every pixel
// of the channel into fixed-point LDS bins covering the whole trace (y and,
// for the cubic, e = dt^2/6 y2: 2 x 4097 x 8 bytes) and writes the channel's
// cotangent rows once with plain stores.  No global atomics, no zero fill of
// the rows beforehand, and the integer sums make the result independent of
// the order of the pixels.  16 waves per workgroup sweep 1024 pixels per
// step; a wave skips its 64 pixels when the tile's delay range leaves the
// channel unused.
template <int GK, int GEO, bool FD>
__global__ __launch_bounds__(1024) void k_das_bwd_rf_ch(BwdArgs A)
{
    extern __shared__ unsigned long long bins[];
    const int nb = A.gstride;  // bins per array: the row length
    unsigned long long *by = bins;
    unsigned long long *be = bins + nb;  // cubic only
    float *byf = reinterpret_cast<float *>(by);
    float *bef = reinterpret_cast<float *>(be);
    const int tid = threadIdx.x;
    const int64_t chan = blockIdx.x;
    const int rx = (int)(chan % A.nrx);
    const int tx = (int)((chan / A.nrx) % A.ntx);
    const int f = (int)(chan / ((int64_t)A.nrx * A.ntx));
    const int nent = A.ntx + A.nrx;
    // log2 of the contributions a bin can take: 2 per pixel
    const int per_bin = 64 - __clzll((unsigned long long)(2 * A.np));
    const double scale = bwd_scale(A.mbits, per_bin);
    const bool fix = scale > 0.0;  // uniform
    const double iscale = fix ? 1.0 / scale : 0.0;
    auto bin_add = [&](unsigned long long *b, float *bf, int i, float v) {
        if (fix)
            atomicAdd(b + i, (unsigned long long)(long long)((double)v * scale));
        else
            atomicAdd(bf + i, v);
    };
    constexpr bool HERM = GK == GK_KORDER || GK == GK_AKIMA;
    unsigned long long *b2 = bins + 2 * nb, *b3 = bins + 3 * nb;  // Hermite kinds only
    float *b2f = reinterpret_cast<float *>(b2), *b3f = reinterpret_cast<float *>(b3);
    const int narr = GK == GK_CUBIC ? 2 : (HERM ? 4 : 1);
    for (int t = tid; t < narr * nb; t += blockDim.x) bins[t] = 0ull;
    __syncthreads();

    for (int64_t base = 0; base < A.np; base += blockDim.x) {
        const int64_t p = base + tid;
        // the wave's pixels lie in one tile of 256: skip it when the tile's
        // delay range selects no sample of this channel (wave-uniform)
        const int64_t tile = (base + (tid & ~63)) >> 8;
        if ((base + (tid & ~63)) >= A.np) continue;
        const float *bt = A.bounds + (tile * nent + tx) * 2;
        const float *br = A.bounds + (tile * nent + A.ntx + rx) * 2;
        const float xlo = bt[0] + br[0], xhi = bt[1] + br[1];
        if (!(xlo <= xhi)) continue;
        if (p >= A.np) continue;
        float ttx = 0.0f, atx = 0.0f, trx = 0.0f, arx = 0.0f;
        if (GEO == 0) {
            // tables: the delays are read only where the weight passes
            atx = A.atx[(int64_t)tx * A.np + p];
            arx = A.arx[(int64_t)rx * A.np + p];
        } else {
            float pv[3];
            bwd_pixel<GEO>(A, p, pv);
            bwd_elem<GEO, FD>(A, true, tx, p, pv, ttx, atx);
            bwd_elem<GEO, FD>(A, false, rx, p, pv, trx, arx);
        }
        const float w = atx * arx;
        if (!(fabsf(w) >= A.thresh)) continue;
        if (GEO == 0) {
            ttx = A.ttx[(int64_t)tx * A.np + p];
            trx = A.trx[(int64_t)rx * A.np + p];
        }
        const float x = ttx + trx;
        const float g = A.gout[gout_index(A.sumflag, f, tx, rx, A.ntx, A.nrx, A.np, p)] * w;
        if (g == 0.0f) continue;
        if (GK == GK_NEAREST) {
            int it = 0;
            if (nearest_index(x, A.t0, A.dt, A.nt, it)) bin_add(by, byf, it, g);
        } else if (HERM) {
            int j = 0;
            if (!knot_at<FD>(A, x, j)) continue;
            // |t| <= 1 (to rounding): every contribution stays within the bound
            const float t = (GK == GK_KORDER) ? ((x - A.t0) - A.dt * (float)j) * A.rdt
                                              : (x - (A.t0 + A.dt * (float)j)) * A.rdt;
            bin_add(by, byf, j, g);
            bin_add(be, bef, j, g * t);
            bin_add(b2, b2f, j, g * (t * t));
            bin_add(b3, b3f, j, g * ((t * t) * t));
        } else {
            int j = 0;
            if (!knot_at<FD>(A, x, j)) continue;
            const float xlead = A.t0 + A.dt * (float)(j + 1);
            const float xlag = A.t0 + A.dt * (float)j;
            float a, b;
            if (GK == GK_LINEAR) {
                a = (xlead - x) / A.dt;
                b = (x - xlag) / A.dt;
            } else {
                a = (xlead - x) * A.rdt;
                b = (x - xlag) * A.rdt;
            }
            // sample j+1 exists for the cubic table (zero entry at nt); for
            // linear the forward reads 0 past the last sample
            bin_add(by, byf, j, g * a);
            if (GK == GK_CUBIC || j + 1 < A.nt) bin_add(by, byf, j + 1, g * b);
            if (GK == GK_CUBIC) {
                bin_add(be, bef, j, g * __builtin_fmaf(a * a, a, -a));
                bin_add(be, bef, j + 1, g * __builtin_fmaf(b * b, b, -b));
            }
        }
    }
    __syncthreads();
    float *gy = A.gy + chan * A.gstride;
    float *ge = GK != GK_NEAREST && GK != GK_LINEAR ? A.ge + chan * A.gstride : nullptr;
    auto take = [&](const unsigned long long *b, const float *bf, int t) -> float {
        return fix ? (float)((double)(long long)b[t] * iscale) : bf[t];
    };
    for (int t = tid; t < nb; t += blockDim.x) {
        gy[t] = take(by, byf, t);
        if (GK == GK_CUBIC || HERM) ge[t] = take(be, bef, t);
        if (HERM) {
            A.gc2[chan * A.gstride + t] = take(b2, b2f, t);
            A.gc3[chan * A.gstride + t] = take(b3, b3f, t);
        }
    }
}

// Adjoint of k_korder_coefs: drf[i] of one channel from the cotangents of the
// Hermite coefficients (gc0..gc3 = d/dc.x .. d/dc.w per interval j).  With
// c.x = y_j, c.y = d0, c.z = 3(y_j+1 - y_j) - 2 d0 - d1,
// c.w = 2(y_j - y_j+1) + d0 + d1 and d0, d1 rows ik0, ik0+1 of S applied to
// the window y[iy0 .. iy0+k) (placement of pycbf/gpu/__engines__.cu:314-328),
//   drf[i] = gc0[i] - 3 gc2[i] + 2 gc3[i]  (i < nt-1; gc0 alone at nt-1)
//          + 3 gc2[i-1] - 2 gc3[i-1]
//          + sum over windows holding i of S[ik0][i-iy0] gd0_j + S[ik0+1][i-iy0] gd1_j
// with gd0 = gc1 - 2 gc2 + gc3 and gd1 = gc3 - gc2.  The windows holding i
// belong to intervals j within k of i.  One lane per (channel, sample), S in
// LDS; a gather, so no atomics.
__global__ __launch_bounds__(256) void k_korder_adjoint(const float *__restrict__ gc0,
                                                        const float *__restrict__ gc1,
                                                        const float *__restrict__ gc2,
                                                        const float *__restrict__ gc3, int gstride,
                                                        int64_t nch, int nt,
                                                        const float *__restrict__ S, int k,
                                                        float *__restrict__ grf)
{
    extern __shared__ float Ss[];
    for (int q = threadIdx.x; q < k * k; q += blockDim.x) Ss[q] = S[q];
    __syncthreads();
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nch * nt) return;
    const int64_t ch = g / nt;
    const int i = (int)(g - ch * nt);
    const int64_t rb = ch * (int64_t)gstride;
    float acc = gc0[rb + i];
    if (i < nt - 1) acc += 2.0f * gc3[rb + i] - 3.0f * gc2[rb + i];
    if (i >= 1) acc += 3.0f * gc2[rb + i - 1] - 2.0f * gc3[rb + i - 1];
    const int jlo = max(0, i - k), jhi = min(nt - 2, i + k);
    for (int j = jlo; j <= jhi; ++j) {
        int ik0, iy0;
        if (j < k / 2) {
            ik0 = j;
            iy0 = 0;
        } else if (j > nt - k / 2 - 1) {
            ik0 = k - nt + j;
            iy0 = nt - k;
        } else {
            ik0 = (k - 1) / 2;
            iy0 = j - (k - 1) / 2;
        }
        const int e = i - iy0;
        if (e < 0 || e >= k) continue;
        const float c1 = gc1[rb + j], c2 = gc2[rb + j], c3 = gc3[rb + j];
        const float gd0 = c1 - 2.0f * c2 + c3, gd1 = c3 - c2;
        acc = __builtin_fmaf(Ss[ik0 * k + e], gd0, acc);
        acc = __builtin_fmaf(Ss[(ik0 + 1) * k + e], gd1, acc);
    }
    grf[g] = acc;
}

namespace {
// slope m_v of y (v in [-2, n]), with Akima's two-point extension at each
// end: m_-1 = 2m_0 - m_1, m_-2 = 3m_0 - 2m_1, m_n-1 = 2m_n-2 - m_n-3,
// m_n = 3m_n-2 - 2m_n-3 (the slopes of k_akima_coefs)
__device__ __forceinline__ float ak_slope(const float *y, int n, int v)
{
    if (v >= 0 && v <= n - 2) return y[v + 1] - y[v];
    if (v < 0) {
        const float m0 = y[1] - y[0], m1 = y[2] - y[1];
        return (v == -1) ? (2.0f * m0 - m1) : (3.0f * m0 - 2.0f * m1);
    }
    const float a = y[n - 1] - y[n - 2], b = y[n - 2] - y[n - 3];
    return (v == n - 1) ? (2.0f * a - b) : (3.0f * a - 2.0f * b);
}

__device__ __forceinline__ float sgnf(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }

// partial derivatives of the Akima / makima derivative d_i with respect to
// the slopes m_i-2, m_i-1, m_i, m_i+1 (d[0..3]).  d_i = (w1 m_i-1 + w2 m_i) /
// (w1 + w2) with w1 = |m_i+1 - m_i| (+ |m_i+1 + m_i| / 2 for makima), w2 =
// |m_i-1 - m_i-2| (+ |m_i-1 + m_i-2| / 2), or (m_i-1 + m_i) / 2 when both
// weights vanish; |.| differentiates to sign(.) with sign(0) = 0.
__device__ __forceinline__ void ak_dderiv(const float *y, int n, int i, int modified, float (&d)[4])
{
    const float mm2 = ak_slope(y, n, i - 2), mm1 = ak_slope(y, n, i - 1);
    const float mp0 = ak_slope(y, n, i), mp1 = ak_slope(y, n, i + 1);
    float w1 = fabsf(mp1 - mp0), w2 = fabsf(mm1 - mm2);
    const float s1 = sgnf(mp1 - mp0), s2 = sgnf(mm1 - mm2);
    float h1 = 0.0f, h2 = 0.0f;
    if (modified) {
        w1 += 0.5f * fabsf(mp1 + mp0);
        w2 += 0.5f * fabsf(mm1 + mm2);
        h1 = 0.5f * sgnf(mp1 + mp0);
        h2 = 0.5f * sgnf(mm1 + mm2);
    }
    const float ws = w1 + w2;
    if (!(ws > 0.0f)) {
        d[0] = 0.0f;
        d[1] = 0.5f;
        d[2] = 0.5f;
        d[3] = 0.0f;
        return;
    }
    const float di = (w1 * mm1 + w2 * mp0) / ws;
    const float a1 = (mm1 - di) / ws;  // d di / d w1
    const float a2 = (mp0 - di) / ws;  // d di / d w2
    d[0] = a2 * (h2 - s2);
    d[1] = w1 / ws + a2 * (s2 + h2);
    d[2] = w2 / ws + a1 * (h1 - s1);
    d[3] = a1 * (s1 + h1);
}

// cotangent of the derivative d_i from the interval coefficients: d_i is c.y
// of interval i (c.z, c.w also hold it) and d1 of interval i-1
__device__ __forceinline__ float ak_gd(const float *c1, const float *c2, const float *c3, int n,
                                       int i)
{
    float g = 0.0f;
    if (i <= n - 2) g += c1[i] - 2.0f * c2[i] + c3[i];
    if (i >= 1 && i - 1 <= n - 2) g += c3[i - 1] - c2[i - 1];
    return g;
}

// cotangent of slope m_v (v in [-2, n]) through the derivatives d_v-1 .. d_v+2
__device__ __forceinline__ float ak_gmv(const float *y, const float *c1, const float *c2,
                                        const float *c3, int n, int modified, int v)
{
    float g = 0.0f;
    for (int i = max(v - 1, 0); i <= min(v + 2, n - 1); ++i) {
        float d[4];
        ak_dderiv(y, n, i, modified, d);
        const int e = v - i + 2;  // slope v is argument e of d_i
        g = __builtin_fmaf(ak_gd(c1, c2, c3, n, i), d[e], g);
    }
    return g;
}

// cotangent of the real slope m_k (k in [0, n-2]), with the end extensions
__device__ __forceinline__ float ak_gm(const float *y, const float *c1, const float *c2,
                                       const float *c3, int n, int modified, int k)
{
    float g = ak_gmv(y, c1, c2, c3, n, modified, k);
    if (k == 0)
        g += 2.0f * ak_gmv(y, c1, c2, c3, n, modified, -1) +
             3.0f * ak_gmv(y, c1, c2, c3, n, modified, -2);
    if (k == 1)
        g -= ak_gmv(y, c1, c2, c3, n, modified, -1) +
             2.0f * ak_gmv(y, c1, c2, c3, n, modified, -2);
    if (k == n - 2)
        g += 2.0f * ak_gmv(y, c1, c2, c3, n, modified, n - 1) +
             3.0f * ak_gmv(y, c1, c2, c3, n, modified, n);
    if (k == n - 3)
        g -= ak_gmv(y, c1, c2, c3, n, modified, n - 1) +
             2.0f * ak_gmv(y, c1, c2, c3, n, modified, n);
    return g;
}
}  // namespace

// Adjoint of k_akima_coefs: drf[i] of one channel from the cotangents of the
// Hermite coefficients (gc0..gc3 per interval j; c = hermite(y_j, y_j+1,
// d_j, d_j+1) for j < nt-1, c.x = y_nt-1 on the last interval):
//   drf[i] = gc0[i] - 3 gc2[i] + 2 gc3[i]  (i < nt-1; gc0 alone at nt-1)
//          + 3 gc2[i-1] - 2 gc3[i-1]
//          + gm[i-1] - gm[i]               (m_k = y_k+1 - y_k)
// with gm the slope cotangents through the Akima derivatives.  The Akima
// derivative is evaluated and differentiated in float32 from the samples, as
// the forward evaluates it.  One lane per (channel, sample); a gather.
__global__ __launch_bounds__(256) void k_akima_adjoint(const float *__restrict__ gc0,
                                                       const float *__restrict__ gc1,
                                                       const float *__restrict__ gc2,
                                                       const float *__restrict__ gc3, int gstride,
                                                       int64_t nch, int nt,
                                                       const float *__restrict__ rf, int modified,
                                                       float *__restrict__ grf)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nch * nt) return;
    const int64_t ch = g / nt;
    const int i = (int)(g - ch * nt);
    const int64_t rb = ch * (int64_t)gstride;
    const float *y = rf + ch * (int64_t)nt;
    const float *c1 = gc1 + rb, *c2 = gc2 + rb, *c3 = gc3 + rb;
    float acc = gc0[rb + i];
    if (i < nt - 1) acc += 2.0f * c3[i] - 3.0f * c2[i];
    if (i >= 1) acc += 3.0f * c2[i - 1] - 2.0f * c3[i - 1];
    if (i >= 1) acc += ak_gm(y, c1, c2, c3, nt, modified, i - 1);
    if (i <= nt - 2) acc -= ak_gm(y, c1, c2, c3, nt, modified, i);
    grf[g] = acc;
}

// drf[k] = gy[k] + (u[k-1] - 2u[k] + u[k+1]) with T u = ge[1..n-2], u = 0
// outside 1..n-2; one lane per channel, Thomas algorithm.  T has constant
// diagonals, so its elimination factors c'_i (c'_1 = 1/4, c'_i =
// 1/(4 - c'_{i-1})) are the same for every channel; in float32 they reach
// their fixed point 2 - sqrt(3) within CP_EXACT steps, after which the
// fixed value is used.
constexpr int CP_EXACT = 24;

__device__ __forceinline__ float elim_factor(int i)
{
    float q = 0.25f;
    for (int k = 2; k <= min(i, CP_EXACT); ++k) q = 1.0f / (4.0f - q);
    return q;
}

// One wave per 64 channels, one lane per channel for both recurrences; the
// samples move in coalesced 64 x 32 tiles through one LDS tile (as in the
// spline prefilter).  Tile s of the forward sweep computes d'_i for
// i = s-1 .. s+30 (e[s-1] is carried), stored shifted as D[i+1] in drf, so
// that tile s of the backward sweep reads only D tile s and writes the
// outputs o[k], k = s .. s+31, over it in place:
//   o[k] = y[k] + (u_{k-1} - 2 u_k + u_{k+1}),  u_j = 0 outside 1..n-2.
__global__ __launch_bounds__(64) void k_spline_adjoint(const float *__restrict__ gy,
                                                       const float *__restrict__ ge, int gstride,
                                                       int64_t nch, int n, float *__restrict__ drf)
{
    __shared__ __align__(16) SplRow S[64];
    __shared__ float cpv[32];  // c'_i for i < 32 (c'_i = c'_CP_EXACT beyond)
    const int lane = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int nrow = (int)((nch - c0) < 64 ? (nch - c0) : 64);
    const float *ybase = gy + c0 * (int64_t)gstride;
    const float *ebase = ge + c0 * (int64_t)gstride;
    float *obase = drf + c0 * (int64_t)n;
    if (lane < 32) cpv[lane] = elim_factor(lane < 1 ? 1 : lane);
    const float cinf = elim_factor(CP_EXACT);
    const int nchunk = (n + 31) / 32;

    // ---- forward elimination: d'_i = (e_i - d'_{i-1}) c'_i, d'_0 = 0 ----
    TileRegs pf;
    tile_fetch(ebase, gstride, nrow, 0, n, pf);
    tile_to_lds(pf, S);
    if (nchunk > 1) tile_fetch(ebase, gstride, nrow, 32, n, pf);
    __syncthreads();
    float e_prev = 0.0f, dp = 0.0f;
    for (int cb = 0; cb < nchunk; ++cb) {
        const int s = cb * 32;
        float v[32];
        row_read(S, v);
        float ecar = e_prev;  // e_{s-1+c}
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            const int i = s - 1 + c;
            const float ei = ecar;
            ecar = v[c];  // e_{s+c}, read before v[c] is overwritten
            float d = 0.0f;
            if (i >= 1 && i <= n - 2) {
                const float ci = i < 32 ? cpv[i] : cinf;
                dp = (ei - dp) * ci;
                d = dp;
            }
            v[c] = d;
        }
        e_prev = ecar;
        row_write(S, v);  // the lane's own row
        __syncthreads();
        lds_to_global(S, obase, n, nrow, s, n);
        __syncthreads();
        if (cb + 1 < nchunk) {
            tile_to_lds(pf, S);
            if (cb + 2 < nchunk) tile_fetch(ebase, gstride, nrow, s + 64, n, pf);
        }
        __syncthreads();
    }
    // the D rows are read back below by the lanes that wrote them
    __threadfence();

    // ---- back substitution u_i = d'_i - c'_i u_{i+1} and the outputs ----
    float u1 = 0.0f, u2 = 0.0f;  // u_{k+1}, u_{k+2} at the top of the tile
    const int slast = (nchunk - 1) * 32;
    TileRegs pd, py;
    tile_fetch(obase, n, nrow, slast, n, pd);
    tile_fetch(ybase, gstride, nrow, slast, n, py);
    for (int cb = nchunk - 1; cb >= 0; --cb) {
        const int s = cb * 32;
        float d[32], y[32];
        tile_to_lds(pd, S);
        __syncthreads();
        row_read(S, d);
        __syncthreads();
        tile_to_lds(py, S);
        __syncthreads();
        row_read(S, y);
        if (cb > 0) {
            tile_fetch(obase, n, nrow, s - 32, n, pd);
            tile_fetch(ybase, gstride, nrow, s - 32, n, py);
        }
        // c = 31 .. 0: i = s-1+c, k = i+1 = s+c; d[c] = d'_i
        float o[32];
#pragma unroll
        for (int c = 31; c >= 0; --c) {
            const int i = s - 1 + c;
            float ui = 0.0f;
            if (i >= 1 && i <= n - 2) ui = d[c] - (i < 32 ? cpv[i] : cinf) * u1;
            o[c] = y[c] + (ui - 2.0f * u1 + u2);
            u2 = u1;
            u1 = ui;
        }
        __syncthreads();  // every lane has read its y row
        row_write(S, o);
        __syncthreads();
        lds_to_global(S, obase, n, nrow, s, n);
        __syncthreads();
    }
}

// Same solve for n <= 64*SEG, one wave per channel: lane l sweeps the SEG
// samples i0 = l*SEG .. of the channel in registers, starting each sweep
// from a guessed carry 32 samples outside its segment (both recurrences
// contract by c'_inf = 2 - sqrt(3) per step) that is checked bit for bit
// against the neighbouring lane's value and recomputed on a mismatch, as in
// the spline prefilter: the result equals the one-lane-per-channel sweep.
// Global traffic: ge and gy read once, drf written once, in coalesced
// 64 x 32 tiles through one LDS tile.
template <int SEG>
__device__ __forceinline__ float adj_forward(const float (&e)[SEG], float dp, int i0, int n,
                                             float cinf, const float (&clo)[CP_EXACT],
                                             bool lane0, float (&v)[SEG])
{
#pragma unroll
    for (int c = 0; c < SEG; ++c) {
        const float ci = (c < CP_EXACT && lane0) ? clo[c < CP_EXACT ? c : 0] : cinf;
        float dn = (e[c] - dp) * ci;
        if (c == 0) dn = lane0 ? 0.0f : dn;  // d'_0 = 0
        dp = dn;
        v[c] = (i0 + c <= n - 2) ? dp : 0.0f;
    }
    return dp;
}

template <int SEG>
__device__ __forceinline__ float adj_backward(float (&v)[SEG], float un, float cinf,
                                              const float (&clo)[CP_EXACT], bool lane0)
{
#pragma unroll
    for (int c = SEG - 1; c >= 0; --c) {
        const float ci = (c < CP_EXACT && lane0) ? clo[c < CP_EXACT ? c : 0] : cinf;
        float u = v[c] - ci * un;
        if (c == 0) u = lane0 ? 0.0f : u;  // u_0 = 0
        un = u;
        v[c] = u;
    }
    return un;
}

template <int SEG>
__global__ __launch_bounds__(64) void k_spline_adjoint_seg(const float *__restrict__ gy,
                                                           const float *__restrict__ ge,
                                                           int gstride, int n,
                                                           float *__restrict__ drf)
{
    static_assert(SEG % 32 == 0, "segments are whole 32-sample tiles");
    __shared__ __align__(16) SplRow S[64];
    const int lane = threadIdx.x;
    const bool lane0 = lane == 0;
    const int64_t ch = blockIdx.x;
    const float *erow = ge + ch * (int64_t)gstride;
    const float *yrow = gy + ch * (int64_t)gstride;
    float *orow = drf + ch * (int64_t)n;
    const int i0 = lane * SEG;
    const int c4 = 4 * (lane & 7);

    // c'_i for i < CP_EXACT (c'_0 unused), c'_inf beyond
    float clo[CP_EXACT];
    {
        float q = 0.25f;
        clo[0] = 0.0f;
        clo[1] = q;
#pragma unroll
        for (int i = 2; i < CP_EXACT; ++i) {
            q = 1.0f / (4.0f - q);
            clo[i] = q;
        }
    }
    const float cinf = elim_factor(CP_EXACT);

    // ---- load e: tile t = samples 32t .. 32t+31 of every segment ----------
    float e[SEG];
#pragma unroll
    for (int t = 0; t < SEG / 32; ++t) {
        TileRegs r;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = (8 * q + (lane >> 3)) * SEG + 32 * t + c4;
            F4u w{{0.0f, 0.0f, 0.0f, 0.0f}};
            if (k + 4 <= n) {
                w = *reinterpret_cast<const F4u *>(erow + k);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (k + j < n) w.v[j] = erow[k + j];
            }
            r.q[q] = w;
        }
        if (t > 0) __syncthreads();
        tile_to_lds(r, S);
        __syncthreads();
        float part[32];
        row_read(S, part);
#pragma unroll
        for (int j = 0; j < 32; ++j) e[32 * t + j] = part[j];
    }

    // ---- forward elimination d'_i = (e_i - d'_{i-1}) c'_i -----------------
    // guess of d'_{i0-1}: d' = 0 at i0-32, then the left neighbour's last 31
    // e values (still in the LDS tile)
    float carry = 0.0f;
    if (lane > 0) {
        const bool low = SEG == 32 && lane == 1;  // reaches i < CP_EXACT
        float dp = 0.0f;
#pragma unroll
        for (int c = 1; c < 32; ++c) {
            const float ci = (c < CP_EXACT && low) ? clo[c < CP_EXACT ? c : 0] : cinf;
            dp = (S[lane - 1][c] - dp) * ci;
        }
        carry = dp;
    }
    float v[SEG];
    float last = adj_forward<SEG>(e, carry, i0, n, cinf, clo, lane0, v);
    for (;;) {
        const float want = __shfl_up(last, 1);
        const bool redo = lane > 0 && i0 <= n - 2 &&
                          __float_as_uint(carry) != __float_as_uint(want);
        if (__ballot(redo) == 0) break;
        if (redo) {
            carry = want;
            last = adj_forward<SEG>(e, carry, i0, n, cinf, clo, lane0, v);
        }
    }
    const float carry_fwd = carry;

    // ---- back substitution u_i = d'_i - c'_i u_{i+1} ----------------------
    __syncthreads();  // neighbours' e rows read
#pragma unroll
    for (int j = 0; j < 32; j += 4)  // d'_{i0 .. i0+31}
        *reinterpret_cast<float4 *>(&S[lane][j]) = make_float4(v[j], v[j + 1], v[j + 2], v[j + 3]);
    __syncthreads();
    float bcarry = 0.0f;  // lane 63: u beyond the channel is 0
    if (lane < 63) {
        float un = 0.0f;
#pragma unroll
        for (int c = 31; c >= 0; --c) un = S[lane + 1][c] - cinf * un;
        bcarry = un;
    }
    float first = adj_backward<SEG>(v, bcarry, cinf, clo, lane0);
    for (;;) {
        const float want = __shfl_down(first, 1);
        const bool redo = lane < 63 && i0 < n && __float_as_uint(bcarry) != __float_as_uint(want);
        if (__ballot(redo) == 0) break;
        if (redo) {
            bcarry = want;
            adj_forward<SEG>(e, carry_fwd, i0, n, cinf, clo, lane0, v);
            first = adj_backward<SEG>(v, bcarry, cinf, clo, lane0);
        }
    }

    // ---- output o[k] = y[k] + (u_{k-1} - 2 u_k + u_{k+1}) ------------------
    const float ul = __shfl_up(v[SEG - 1], 1);  // u_{i0-1}
    const float ur = __shfl_down(first, 1);     // u_{i0+SEG}
#pragma unroll
    for (int c = 0; c < SEG; ++c) {
        const float um = c > 0 ? v[c - 1] : (lane0 ? 0.0f : ul);
        const float up = c + 1 < SEG ? v[c + 1] : (lane == 63 ? 0.0f : ur);
        e[c] = um - 2.0f * v[c] + up;  // e is free: holds the second difference
    }
#pragma unroll
    for (int t = 0; t < SEG / 32; ++t) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 32; j += 4)
            *reinterpret_cast<float4 *>(&S[lane][j]) =
                make_float4(e[32 * t + j], e[32 * t + j + 1], e[32 * t + j + 2], e[32 * t + j + 3]);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int rr = 8 * q + (lane >> 3);
            const int k = rr * SEG + 32 * t + c4;
            const float4 w = *reinterpret_cast<const float4 *>(&S[rr][c4]);
            const float d4[4] = {w.x, w.y, w.z, w.w};
            if (k + 4 <= n) {
                const F4u y = *reinterpret_cast<const F4u *>(yrow + k);
                F4u o{{y.v[0] + d4[0], y.v[1] + d4[1], y.v[2] + d4[2], y.v[3] + d4[3]}};
                *reinterpret_cast<F4u *>(orow + k) = o;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (k + j < n) orow[k + j] = yrow[k + j] + d4[j];
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_fill_zero(float *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

static int fill_zero(float *p, int64_t n, hipStream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return check_launch("fill zero");
}

template <int GK, int GEO, bool FD>
static int launch_bwd_t(const BwdArgs &a, float *gtautx, float *gapodtx, float *gtaurx,
                        float *gapodrx, float *grf, float *gc0, hipStream_t s)
{
    const unsigned ptiles = (unsigned)((a.np + 255) / 256);
    const bool want_c0 = GEO > 0 && gc0 && a.gc0part;
    if (gtautx || gapodtx || want_c0) {
        BwdArgs b = a;
        b.gtau = gtautx;
        b.gapod = gapodtx;
        b.gc0part = want_c0 ? a.gc0part : nullptr;
        b.gc0_off = 0;
        ProfScope ps("das_bwd_tau", s);
        hipLaunchKernelGGL((k_das_bwd_tau<GK, 1, GEO, FD>), dim3(ptiles, a.ntx), dim3(256), 0, s, b);
        if (int rc = check_launch("das_bwd_tau")) return rc;
    }
    if (gtaurx || gapodrx || want_c0) {
        BwdArgs b = a;
        b.gtau = gtaurx;
        b.gapod = gapodrx;
        b.gc0part = want_c0 ? a.gc0part : nullptr;
        b.gc0_off = (int)(ptiles * (unsigned)a.ntx);
        ProfScope ps("das_bwd_tau", s);
        hipLaunchKernelGGL((k_das_bwd_tau<GK, 0, GEO, FD>), dim3(ptiles, a.nrx), dim3(256), 0, s, b);
        if (int rc = check_launch("das_bwd_tau")) return rc;
    }
    if (want_c0) {
        hipLaunchKernelGGL(k_gc0_sum, dim3(1), dim3(256), 0, s, (const double *)a.gc0part,
                           (int64_t)ptiles * (a.ntx + a.nrx), gc0);
        if (int rc = check_launch("gc0_sum")) return rc;
    }
    if (grf) {
        const int64_t nch = (int64_t)a.nfr * a.ntx * a.nrx;
        ProfScope ps("das_bwd_rf", s);
        // one workgroup per channel while the whole trace's bins fit in LDS
        const size_t ch_lds =
            (size_t)(GK == GK_CUBIC ? 2 : (GK == GK_KORDER || GK == GK_AKIMA ? 4 : 1)) *
            a.gstride * 8;
        int dev = 0, max_lds = 0;
        if (int rc = check(hipGetDevice(&dev), "device")) return rc;
        if (int rc = check(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                                 dev),
                           "LDS per workgroup"))
            return rc;
        const bool per_channel = a.gstride <= 4097 && ch_lds <= (size_t)max_lds && nch <= 0x7fffffff;
        // bound of the contributions for the fixed-point bins
        if (int rc = fill_zero(reinterpret_cast<float *>(a.mbits), 3, s)) return rc;
        int64_t ngout = (int64_t)a.nfr * a.np;
        if (a.sumflag == PYCBF_SUM_NONE) ngout *= (int64_t)a.ntx * a.nrx;
        else if (a.sumflag == PYCBF_SUM_TX_ONLY) ngout *= a.nrx;
        else if (a.sumflag == PYCBF_SUM_RX_ONLY) ngout *= a.ntx;
        const int64_t nscan = std::max<int64_t>(ngout, (int64_t)std::max(a.ntx, a.nrx) * a.np);
        const unsigned sb = (unsigned)std::min<int64_t>(256, (nscan / 4 + 255) / 256 + 1);
        hipLaunchKernelGGL(k_bwd_scale, dim3(sb, 3), dim3(256), 0, s, a, ngout);
        if (int rc = check_launch("das_bwd_scale")) return rc;
        if (per_channel) {
            hipLaunchKernelGGL((k_das_bwd_rf_ch<GK, GEO, FD>), dim3((unsigned)nch), dim3(1024), ch_lds, s,
                               a);
        } else {
            if (GK == GK_CUBIC || GK == GK_KORDER || GK == GK_AKIMA) {
                if (int rc = fill_zero(a.gy, nch * a.gstride, s)) return rc;
                if (int rc = fill_zero(a.ge, nch * a.gstride, s)) return rc;
            }
            if (GK == GK_KORDER || GK == GK_AKIMA) {
                if (int rc = fill_zero(a.gc2, nch * a.gstride, s)) return rc;
                if (int rc = fill_zero(a.gc3, nch * a.gstride, s)) return rc;
            }
            if (GK == GK_NEAREST || GK == GK_LINEAR) {
                // the cotangent rows (grf itself, or the upsampled grid's rows)
                if (int rc = fill_zero(a.gy, nch * a.gstride, s)) return rc;
            }
            hipLaunchKernelGGL((k_das_bwd_rf<GK, GEO, FD>), dim3(ptiles, a.ntx, a.nfr), dim3(256), 0, s,
                               a);
        }
        if (int rc = check_launch("das_bwd_rf")) return rc;
        if (GK == GK_CUBIC) {
            if (a.nt > SPL_MAX_NT) return fail(PYCBF_EINVAL, "spline adjoint: nt too large");
            const int seg = spline_seg_len(a.nt);
            if (seg == 32)
                hipLaunchKernelGGL(k_spline_adjoint_seg<32>, dim3((unsigned)nch), dim3(64), 0, s,
                                   a.gy, a.ge, a.gstride, a.nt, grf);
            else if (seg == 64)
                hipLaunchKernelGGL(k_spline_adjoint_seg<64>, dim3((unsigned)nch), dim3(64), 0, s,
                                   a.gy, a.ge, a.gstride, a.nt, grf);
            else
                hipLaunchKernelGGL(k_spline_adjoint, dim3((unsigned)((nch + 63) / 64)), dim3(64),
                                   0, s, a.gy, a.ge, a.gstride, nch, a.nt, grf);
            if (int rc = check_launch("spline_adjoint")) return rc;
        }
        if (GK == GK_KORDER) {
            const int64_t n = nch * a.nt;
            hipLaunchKernelGGL(k_korder_adjoint, dim3((unsigned)((n + 255) / 256)), dim3(256),
                               (size_t)a.k * a.k * sizeof(float), s, a.gy, a.ge, a.gc2, a.gc3,
                               a.gstride, nch, a.nt, a.S, a.k, grf);
            if (int rc = check_launch("korder_adjoint")) return rc;
        }
        if (GK == GK_AKIMA) {
            const int64_t n = nch * a.nt;
            hipLaunchKernelGGL(k_akima_adjoint, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                               s, a.gy, a.ge, a.gc2, a.gc3, a.gstride, nch, a.nt, a.rf,
                               a.modified, grf);
            if (int rc = check_launch("akima_adjoint")) return rc;
        }
    }
    return 0;
}

// Adjoint of k_upsample (natural-spline values at usf x the sampling rate,
// pycbf/cpu/pycbf.c:209-271 with usf > 1): the cotangents gu of the upsampled
// trace (nch, usf (nt-1) + 1) go back to the cotangents of the spline table
// entries y_i and e_i = dt^2/6 y2_i (rows of nt + 1, entry nt unused).  Fine
// point it selects interval j with the forward's float32 operations and
// contributes a gu, (a^3 - a) gu to entry j and b gu, (b^3 - b) gu to entry
// j+1 (when j+1 < nt).  One lane per (channel, entry) gathers over the fine
// points that can select it (2 usf + 5 candidates); no atomics.
__global__ __launch_bounds__(256) void k_upsample_adjoint(const float *__restrict__ gu,
                                                          int64_t nch, int nt, int usf, float t0,
                                                          float dt, float *__restrict__ gy,
                                                          float *__restrict__ ge)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int gs = nt + 1;
    if (g >= nch * gs) return;
    const int64_t ch = g / gs;
    const int i = (int)(g - ch * gs);
    const int ntu = usf * (nt - 1) + 1;
    const float dtu = dt / (float)usf;
    const float xmax = t0 + dt * (float)(nt - 1);
    const float *row = gu + ch * (int64_t)ntu;
    float sy = 0.0f, se = 0.0f;
    const int lo = max(0, usf * (i - 1) - 2), hi = min(ntu - 1, usf * (i + 1) + 2);
    for (int it = lo; it <= hi; ++it) {
        const float x = t0 + (float)it * dtu;
        if ((x < t0) || (x > xmax)) continue;
        const int j = min((int)((x - t0) / dt), nt - 1);
        if (j != i && j + 1 != i) continue;
        const float xlead = t0 + dt * (float)(j + 1);
        const float xlag = t0 + dt * (float)j;
        const float v = row[it];
        if (j == i) {
            const float a = (xlead - x) / dt;
            sy = __builtin_fmaf(a, v, sy);
            se = __builtin_fmaf(__builtin_fmaf(a * a, a, -a), v, se);
        } else if (j + 1 < nt) {
            const float b = (x - xlag) / dt;
            sy = __builtin_fmaf(b, v, sy);
            se = __builtin_fmaf(__builtin_fmaf(b * b, b, -b), v, se);
        }
    }
    gy[g] = sy;
    ge[g] = se;
}

int launch_upsample_adjoint(const float *gu, int64_t nch, int32_t nt, int32_t usf, float t0,
                            float dt, float *gy, float *ge, float *grf, hipStream_t s)
{
    if (nch <= 0) return 0;
    if (nt > SPL_MAX_NT) return fail(PYCBF_EINVAL, "spline adjoint: nt too large");
    const int64_t n = nch * ((int64_t)nt + 1);
    hipLaunchKernelGGL(k_upsample_adjoint, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       gu, nch, (int)nt, (int)usf, t0, dt, gy, ge);
    if (int rc = check_launch("upsample_adjoint")) return rc;
    const int seg = spline_seg_len(nt);
    if (seg == 32)
        hipLaunchKernelGGL(k_spline_adjoint_seg<32>, dim3((unsigned)nch), dim3(64), 0, s, gy, ge,
                           nt + 1, (int)nt, grf);
    else if (seg == 64)
        hipLaunchKernelGGL(k_spline_adjoint_seg<64>, dim3((unsigned)nch), dim3(64), 0, s, gy, ge,
                           nt + 1, (int)nt, grf);
    else
        hipLaunchKernelGGL(k_spline_adjoint, dim3((unsigned)((nch + 63) / 64)), dim3(64), 0, s,
                           gy, ge, nt + 1, nch, (int)nt, grf);
    return check_launch("spline_adjoint");
}

template <int GEO, bool FD>
static int launch_das_bwd_fd(int gk, const BwdArgs &a, float *gtautx, float *gapodtx,
                             float *gtaurx, float *gapodrx, float *grf, float *gc0, hipStream_t s)
{
    switch (gk) {
    case GK_NEAREST:
        return launch_bwd_t<GK_NEAREST, GEO, FD>(a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
    case GK_LINEAR:
        return launch_bwd_t<GK_LINEAR, GEO, FD>(a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
    case GK_KORDER:
        return launch_bwd_t<GK_KORDER, GEO, FD>(a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
    case GK_AKIMA:
        return launch_bwd_t<GK_AKIMA, GEO, FD>(a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
    default:
        return launch_bwd_t<GK_CUBIC, GEO, FD>(a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
    }
}

// the quotients by dt (and by c0 with the geometry) as rounded-reciprocal
// products where the host check accepts the divisors (the same indices and
// delays as the divisions)
template <int GEO>
static int launch_das_bwd_geo(int gk, const BwdArgs &a, float *gtautx, float *gapodtx,
                              float *gtaurx, float *gapodrx, float *grf, float *gc0, hipStream_t s)
{
    const bool fd = a.rdt == 1.0f / a.dt && recip_division_exact(a.dt) &&
                    (GEO == 0 || (a.rc0 == 1.0f / a.c0 && recip_division_exact(a.c0)));
    if (fd) return launch_das_bwd_fd<GEO, true>(gk, a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
    return launch_das_bwd_fd<GEO, false>(gk, a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
}

int launch_das_bwd(int gk, const BwdArgs &a, float *gtautx, float *gapodtx, float *gtaurx,
                   float *gapodrx, float *grf, float *gc0, hipStream_t s)
{
    if (a.geo == 2) return launch_das_bwd_geo<2>(gk, a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
    if (a.geo == 3) return launch_das_bwd_geo<3>(gk, a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
    return launch_das_bwd_geo<0>(gk, a, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s);
}

}  // namespace pycbf
