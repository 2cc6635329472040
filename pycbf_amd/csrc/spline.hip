// Per-channel prefilters: turn each RF channel into a table the DAS gather
// reads with one load per sample.
//
//  * k_spline: natural cubic spline (tie_knots1D_fixed, pycbf/cpu/pycbf.c:114-166).
//    One wave owns 64 channels; 32-sample tiles are staged through LDS so that
//    global loads/stores stay coalesced while each lane runs its channel's
//    sequential recurrence in the reference's exact float32 operation order
//    (bit-identical second derivatives).  Output is either y2 itself or the
//    per-interval cubic coefficients used by the DAS gather.
//  * k_upsample: the spline evaluated on a dt/usf grid (pycbf/cpu/pycbf.c:234-253).
//  * k_korder_coefs: Hermite coefficients with k-order derivatives
//    (pycbf/gpu/__engines__.cu:292-355).
//  * k_akima_coefs: Hermite coefficients with Akima / modified-Akima
//    derivatives in consistent per-sample units (the reference mixes units,
//    pycbf/gpu/__engines__.cu:153-160,261-268).
#include <math.h>

#include "common.h"
#include "prof.h"

namespace pycbf {

// First 32 data-independent forward coefficients of the spline sweep,
// y2f[i] = (0.5 - 1) / (0.5*y2f[i-1] + 2).  In float32 the sequence reaches an
// exact fixed point at i = 7, so y2f[i] == y2f[31] for every i >= 31.
__device__ __forceinline__ float y2f_at(const float *tab, int i) { return tab[i < 31 ? i : 31]; }

template <int MODE>  // 0: float4 cubic coefficients, 1: float y2
__global__ __launch_bounds__(64) void k_spline(const float *__restrict__ rf, int64_t nch, int nt,
                                               float dx, float *__restrict__ uscr,
                                               void *__restrict__ outp)
{
    constexpr int CH = SPL_CHUNK;
    __shared__ float ytile[64][CH + 1];  // column CH holds the next tile's first sample
    __shared__ float utile[64][CH + 1];
    __shared__ float y2f[32];

    const int lane = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int64_t ch = c0 + lane;
    const bool active = ch < nch;
    const int nrow = (int)((nch - c0) < 64 ? (nch - c0) : 64);

    if (lane == 0) {
        float v = 0.0f;
        y2f[0] = 0.0f;
        for (int i = 1; i < 32; ++i) {
            float p = 0.5f * v + 2.0f;
            v = (0.5f - 1.0f) / p;
            y2f[i] = v;
        }
    }
    __syncthreads();

    const int nchunk = (nt + CH - 1) / CH;

    // ---- forward sweep: u[i] -------------------------------------------
    float u_prev = 0.0f, y_prev = 0.0f;
    for (int cb = 0; cb < nchunk; ++cb) {
        const int s = cb * CH;
#pragma unroll 8
        for (int r0 = 0; r0 < 64; r0 += 2) {
            const int r = r0 + (lane >> 5), col = lane & 31;
            const int t = s + col;
            float v = 0.0f;
            if (r < nrow && t < nt) v = rf[(c0 + r) * (int64_t)nt + t];
            ytile[r][col] = v;
        }
        {
            const int t = s + CH;
            float v = 0.0f;
            if (active && t < nt) v = rf[ch * (int64_t)nt + t];
            ytile[lane][CH] = v;
        }
        __syncthreads();
#pragma unroll 2
        for (int col = 0; col < CH; ++col) {
            const int i = s + col;
            const float yi = ytile[lane][col];
            float u = 0.0f;
            if (i >= 1 && i <= nt - 2) {
                const float yn = ytile[lane][col + 1];
                const float p = 0.5f * y2f_at(y2f, i - 1) + 2.0f;
                const float jump = (yn - yi) / dx - (yi - y_prev) / dx;
                u = (6.0f * jump / (2.0f * dx) - 0.5f * u_prev) / p;
                u_prev = u;
            }
            y_prev = yi;
            utile[lane][col] = u;
        }
        __syncthreads();
#pragma unroll 8
        for (int r0 = 0; r0 < 64; r0 += 2) {
            const int r = r0 + (lane >> 5), col = lane & 31;
            const int t = s + col;
            if (r < nrow && t < nt) uscr[(c0 + r) * (int64_t)nt + t] = utile[r][col];
        }
        __syncthreads();
    }

    // ---- back substitution: y2[k] = y2f[k]*y2[k+1] + u[k] ----------------
    const float h = (1.0f / 6.0f) * (dx * dx);
    float y2_next = 0.0f;  // y2[k+1]; zero beyond the end (natural end + padding)
    float y_next = 0.0f;   // y[k+1];  zero beyond the end (padding)
    for (int cb = nchunk - 1; cb >= 0; --cb) {
        const int s = cb * CH;
#pragma unroll 8
        for (int r0 = 0; r0 < 64; r0 += 2) {
            const int r = r0 + (lane >> 5), col = lane & 31;
            const int t = s + col;
            float v = 0.0f, u = 0.0f;
            if (r < nrow && t < nt) {
                v = rf[(c0 + r) * (int64_t)nt + t];
                u = uscr[(c0 + r) * (int64_t)nt + t];
            }
            ytile[r][col] = v;
            utile[r][col] = u;
        }
        __syncthreads();
        float4 *coef = static_cast<float4 *>(outp);
#pragma unroll 2
        for (int col = CH - 1; col >= 0; --col) {
            const int k = s + col;
            if (k < nt) {
                const float yk = ytile[lane][col];
                float y2k = 0.0f;
                if (k < nt - 1) y2k = y2f_at(y2f, k) * y2_next + utile[lane][col];
                if (MODE == 1) {
                    utile[lane][col] = y2k;
                } else if (active) {
                    // interval k: value = a*y_k + b*y_k+1 + (a^3-a)*h*y2_k
                    // + (b^3-b)*h*y2_k+1 with h = dx^2/6 (pycbf/cpu/pycbf.c:97-103).
                    // Lane-private rows: eight consecutive intervals complete
                    // each 128-byte line, which L2 merges before write-back.
                    coef[ch * (int64_t)nt + k] = make_float4(yk, y_next, h * y2k, h * y2_next);
                }
                y2_next = y2k;
                y_next = yk;
            }
        }
        if (MODE == 1) {
            __syncthreads();
            float *y2 = static_cast<float *>(outp);
            for (int r0 = 0; r0 < 64; r0 += 2) {
                const int r = r0 + (lane >> 5), col = lane & 31;
                const int t = s + col;
                if (r < nrow && t < nt) y2[(c0 + r) * (int64_t)nt + t] = utile[r][col];
            }
        }
        __syncthreads();
    }
}

int launch_spline_coefs(const float *rf, int64_t nch, int32_t nt, float dt, float *uscr,
                        float4 *coef, hipStream_t s)
{
    if (nch <= 0 || nt <= 0) return 0;
    dim3 grid((unsigned)((nch + 63) / 64));
    ProfScope prof("spline_coefs", s);
    hipLaunchKernelGGL(k_spline<0>, grid, dim3(64), 0, s, rf, nch, (int)nt, dt, uscr,
                       (void *)coef);
    return check_launch("spline coefficients");
}

int launch_spline_y2(const float *rf, int64_t nch, int32_t nt, float dt, float *uscr, float *y2,
                     hipStream_t s)
{
    if (nch <= 0 || nt <= 0) return 0;
    dim3 grid((unsigned)((nch + 63) / 64));
    hipLaunchKernelGGL(k_spline<1>, grid, dim3(64), 0, s, rf, nch, (int)nt, dt, uscr, (void *)y2);
    return check_launch("spline y2");
}

// a^3 rounded once (through double), standing in for powf(a, 3.0f).
__device__ __forceinline__ float cube_rn(float a)
{
    double d = (double)a;
    return (float)(d * d * d);
}

// Natural-spline value at x, same operation order as cubic1D_fixed
// (pycbf/cpu/pycbf.c:27-53), zero padding beyond the last knot.
__global__ __launch_bounds__(256) void k_upsample(const float *__restrict__ rf,
                                                  const float *__restrict__ y2, int64_t nch,
                                                  int nt, float t0, float dt, int usf,
                                                  float *__restrict__ out)
{
    const int64_t ntu = (int64_t)usf * (nt - 1) + 1;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nch * ntu) return;
    const int64_t ch = g / ntu;
    const int it = (int)(g - ch * ntu);
    const float dtu = dt / (float)usf;
    const float x = t0 + (float)it * dtu;
    const float xmax = t0 + dt * (float)(nt - 1);
    float v = 0.0f;
    if (!((x < t0) || (x > xmax))) {
        const int j = min((int)((x - t0) / dt), nt - 1);
        const float *y = rf + ch * (int64_t)nt;
        const float *z = y2 + ch * (int64_t)nt;
        const float xlead = t0 + dt * (float)(j + 1);
        const float xlag = t0 + dt * (float)j;
        const float a = (xlead - x) / dt;
        const float b = (x - xlag) / dt;
        const float dx2 = dt * dt;
        const float c = (1.0f / 6.0f) * (cube_rn(a) - a) * dx2;
        const float d = (1.0f / 6.0f) * (cube_rn(b) - b) * dx2;
        const float y1 = (j + 1 < nt) ? y[j + 1] : 0.0f;
        const float z1 = (j + 1 < nt) ? z[j + 1] : 0.0f;
        v = a * y[j] + b * y1 + c * z[j] + d * z1;
    }
    out[g] = v;
}

int launch_upsample(const float *rf, const float *y2, int64_t nch, int32_t nt, float t0, float dt,
                    int32_t usf, float *out, hipStream_t s)
{
    const int64_t n = nch * ((int64_t)usf * (nt - 1) + 1);
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_upsample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rf, y2,
                       nch, (int)nt, t0, dt, (int)usf, out);
    return check_launch("upsample");
}

// Hermite coefficients on interval j from end values and per-sample derivatives.
__device__ __forceinline__ float4 hermite(float y0, float y1, float d0, float d1)
{
    float4 c;
    c.x = y0;
    c.y = d0;
    c.z = 3.0f * (y1 - y0) - 2.0f * d0 - d1;
    c.w = 2.0f * (y0 - y1) + d0 + d1;
    return c;
}

// k-order cubic: derivatives from a k-sample window through S, window placement
// as pycbf/gpu/__engines__.cu:314-328.
__global__ __launch_bounds__(256) void k_korder_coefs(const float *__restrict__ rf, int64_t nch,
                                                      int nt, const float *__restrict__ S, int k,
                                                      float4 *__restrict__ coef)
{
    extern __shared__ float Ss[];
    for (int i = threadIdx.x; i < k * k; i += blockDim.x) Ss[i] = S[i];
    __syncthreads();
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nch * nt) return;
    const int64_t ch = g / nt;
    const int j = (int)(g - ch * nt);
    const float *y = rf + ch * (int64_t)nt;
    float4 c = make_float4(y[j], 0.0f, 0.0f, 0.0f);
    if (j < nt - 1) {
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
        float d0 = 0.0f, d1 = 0.0f;
        for (int i = 0; i < k; ++i) {
            const float yi = y[iy0 + i];
            d0 = __builtin_fmaf(Ss[ik0 * k + i], yi, d0);
            d1 = __builtin_fmaf(Ss[(ik0 + 1) * k + i], yi, d1);
        }
        c = hermite(y[j], y[j + 1], d0, d1);
    }
    coef[g] = c;
}

int launch_korder_coefs(const float *rf, int64_t nch, int32_t nt, const float *S, int k,
                        float4 *coef, hipStream_t s)
{
    const int64_t n = nch * nt;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_korder_coefs, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       (size_t)k * k * sizeof(float), s, rf, nch, (int)nt, S, k, coef);
    return check_launch("korder coefficients");
}

// Slope m[i] = y[i+1]-y[i] for i in [0, n-2], extended by two on each side as
// in Akima's method: m[-1] = 2m[0]-m[1], m[-2] = 3m[0]-2m[1] (and mirrored).
__device__ __forceinline__ float akima_slope(const float *y, int n, int i)
{
    if (i >= 0 && i <= n - 2) return y[i + 1] - y[i];
    if (i < 0) {
        const float m0 = y[1] - y[0], m1 = y[2] - y[1];
        return (i == -1) ? (2.0f * m0 - m1) : (3.0f * m0 - 2.0f * m1);
    }
    const float a = y[n - 1] - y[n - 2], b = y[n - 2] - y[n - 3];
    return (i == n - 1) ? (2.0f * a - b) : (3.0f * a - 2.0f * b);
}

// Akima (modified = 0) / makima (modified = 1) derivative at sample i.
__device__ __forceinline__ float akima_deriv(const float *y, int n, int i, int modified)
{
    const float mm2 = akima_slope(y, n, i - 2), mm1 = akima_slope(y, n, i - 1);
    const float mp0 = akima_slope(y, n, i), mp1 = akima_slope(y, n, i + 1);
    float w1 = fabsf(mp1 - mp0), w2 = fabsf(mm1 - mm2);
    if (modified) {
        w1 += 0.5f * fabsf(mp1 + mp0);
        w2 += 0.5f * fabsf(mm1 + mm2);
    }
    const float ws = w1 + w2;
    if (!(ws > 0.0f)) return 0.5f * (mm1 + mp0);
    return (w1 * mm1 + w2 * mp0) / ws;
}

__global__ __launch_bounds__(256) void k_akima_coefs(const float *__restrict__ rf, int64_t nch,
                                                     int nt, int modified,
                                                     float4 *__restrict__ coef)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nch * nt) return;
    const int64_t ch = g / nt;
    const int j = (int)(g - ch * nt);
    const float *y = rf + ch * (int64_t)nt;
    float4 c = make_float4(y[j], 0.0f, 0.0f, 0.0f);
    if (j < nt - 1)
        c = hermite(y[j], y[j + 1], akima_deriv(y, nt, j, modified),
                    akima_deriv(y, nt, j + 1, modified));
    coef[g] = c;
}

int launch_akima_coefs(const float *rf, int64_t nch, int32_t nt, int modified, float4 *coef,
                       hipStream_t s)
{
    const int64_t n = nch * nt;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_akima_coefs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rf,
                       nch, (int)nt, modified, coef);
    return check_launch("akima coefficients");
}

}  // namespace pycbf
