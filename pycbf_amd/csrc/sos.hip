// Straight-ray delays through a slowness (1 / speed of sound) map, and their
// adjoint, for speed-of-sound estimation with the differentiable DAS.
//
// The reference's examples/rough_sos_correction.ipynb corrects delays with an
// external eikonal solver; here the delays are straight-ray integrals, the
// model stated for the differentiable workload:
//   plane wave tx (steer theta, origin xo on z = 0): the ray through pixel p
//     along (sin theta, cos theta) enters z = 0 at q = (x_p - z_p tan theta, 0);
//     tau_tx = (q_x - xo) sin theta / c0 + integral of s from q to p
//   receive element at (x_e, 0): tau_rx = integral of s from (x_e, 0) to p
// The forward evaluates every ray in double (sample positions, bilinear weights,
// path length and the fire term from the float32 inputs) and rounds the delay
// to float32; it can also write the float32 residual delay - float32(delay),
// which the DAS adjoint uses to place each sample to sub-float32 precision.
// The integral is the midpoint rule with nsteps samples of the bilinear
// interpolant of the slowness map s (nzm, nxm) whose node (i, k) sits at
// (xm0 + (k + 1/2) dxm, zm0 + (i + 1/2) dzm); outside the node range the map
// is extended by its edge values.  The adjoint scatters
//   d s[node] += sum over rays of g_tau * (L / nsteps) * bilinear weight.
#include <math.h>

#include <algorithm>

#include "common.h"
#include "prof.h"

namespace pycbf {

namespace {

constexpr int SOS_MAXMAP = 8192;    // map nodes held in LDS (64 KB of 64-bit sums in the adjoint)
constexpr int SOS_PARTIALS = 1024;  // workgroups of the adjoint

// bilinear sample positions: node indices and weights of point (x, z)
__device__ __forceinline__ void bilin(const MapGeo &g, float x, float z, int &i0, int &k0,
                                      float &wz, float &wx)
{
    float u = (x - g.xm0) * g.rdxm - 0.5f;
    float v = (z - g.zm0) * g.rdzm - 0.5f;
    u = fminf(fmaxf(u, 0.0f), (float)(g.nxm - 1));
    v = fminf(fmaxf(v, 0.0f), (float)(g.nzm - 1));
    k0 = min((int)u, max(g.nxm - 2, 0));
    i0 = min((int)v, max(g.nzm - 2, 0));
    wx = g.nxm > 1 ? u - (float)k0 : 0.0f;
    wz = g.nzm > 1 ? v - (float)i0 : 0.0f;
}

// ray of element e through pixel p: start (qx, 0), length L, fire delay
__device__ __forceinline__ void ray_of(int e, int ntx, const float *steer, const float *xorig,
                                       const float *xele, float c0, float px, float pz, float &qx,
                                       float &L, float &fire)
{
    if (e < ntx) {
        const float th = steer[e];
        const float sn = sinf(th), cs = cosf(th);
        qx = px - pz * (sn / cs);
        L = pz / cs;
        fire = (qx - xorig[e]) * sn / c0;
    } else {
        qx = xele[e - ntx];
        const float dx = px - qx;
        L = sqrtf(dx * dx + pz * pz);
        fire = 0.0f;
    }
}

}  // namespace

// the same in double, from the map coordinates u = (x - xm0) / dxm - 1/2 and
// v = (z - zm0) / dzm - 1/2 of the point
__device__ __forceinline__ void bilin_d(const RayArgs &a, double u, double v, int &i0, int &k0,
                                        double &wz, double &wx)
{
    const int nxm = a.g.nxm, nzm = a.g.nzm;
    u = fmin(fmax(u, 0.0), (double)(nxm - 1));
    v = fmin(fmax(v, 0.0), (double)(nzm - 1));
    k0 = min((int)u, max(nxm - 2, 0));
    i0 = min((int)v, max(nzm - 2, 0));
    wx = nxm > 1 ? u - (double)k0 : 0.0;
    wz = nzm > 1 ? v - (double)i0 : 0.0;
}

__device__ __forceinline__ void ray_of_d(const RayArgs &a, int e, double px, double pz,
                                         double &qx, double &L, double &fire)
{
    if (e < a.ntx) {
        const double th = (double)a.steer[e];
        const double sn = sin(th), cs = cos(th);
        qx = px - pz * (sn / cs);
        L = pz / cs;
        fire = (qx - (double)a.xorig[e]) * sn / (double)a.c0;
    } else {
        qx = (double)a.xele[e - a.ntx];
        const double dx = px - qx;
        L = sqrt(dx * dx + pz * pz);
        fire = 0.0;
    }
}

// clears the adjoint's max word and its n 64-bit map sums
__global__ void k_ray_clear(unsigned *mbits, unsigned long long *gsum, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *mbits = 0u;
    for (int i = t; i < n; i += gridDim.x * blockDim.x) gsum[i] = 0ull;
}

// one lane per (element, pixel); the map is staged in LDS
__global__ __launch_bounds__(256) void k_ray_delays(RayArgs a)
{
    extern __shared__ float sm[];
    const int nmap = a.g.nzm * a.g.nxm;
    for (int t = threadIdx.x; t < nmap; t += 256) sm[t] = a.slow[t];
    __syncthreads();
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int e = blockIdx.y;
    if (p >= a.np) return;
    const double px = a.pvec[2 * p], pz = a.pvec[2 * p + 1];
    double qx, L, fire;
    ray_of_d(a, e, px, pz, qx, L, fire);
    const int nxm = a.g.nxm, nzm = a.g.nzm;
    // map coordinates along the ray, affine in t: u = u0 + t du, v = v0 + t dv
    const double rn = 1.0 / (double)a.nsteps;
    const double u0 = (qx - a.xm0d) * a.rdxmd - 0.5, du = (px - qx) * a.rdxmd;
    const double v0 = -a.zm0d * a.rdzmd - 0.5, dv = pz * a.rdzmd;
    double acc = 0.0;
    for (int m = 0; m < a.nsteps; ++m) {
        const double t = ((double)m + 0.5) * rn;
        int i0, k0;
        double wz, wx;
        bilin_d(a, fma(t, du, u0), fma(t, dv, v0), i0, k0, wz, wx);
        const int i1 = min(i0 + 1, nzm - 1), k1 = min(k0 + 1, nxm - 1);
        const double s00 = sm[i0 * nxm + k0], s01 = sm[i0 * nxm + k1];
        const double s10 = sm[i1 * nxm + k0], s11 = sm[i1 * nxm + k1];
        const double top = s00 + wx * (s01 - s00), bot = s10 + wx * (s11 - s10);
        acc += top + wz * (bot - top);
    }
    const double tau = fire + L * (acc * rn);
    const float hi = (float)tau;
    const float lo = (float)(tau - (double)hi);
    if (e < a.ntx) {
        a.tautx[(int64_t)e * a.np + p] = hi;
        if (a.tautx_lo) a.tautx_lo[(int64_t)e * a.np + p] = lo;
    } else {
        a.taurx[(int64_t)(e - a.ntx) * a.np + p] = hi;
        if (a.taurx_lo) a.taurx_lo[(int64_t)(e - a.ntx) * a.np + p] = lo;
    }
}

// Adjoint accumulation in 64-bit fixed point.  Float atomic adds in LDS run
// at about 0.3 lanes per clock per CU on gfx950, integer adds at about 12, so
// the map sums are integers: a contribution c is added as round(c * scale).
// k_ray_scale finds M = max over rays of |g| * L, which bounds what one ray
// adds to any node (its bilinear weights sum to at most nsteps per node and
// gl = g L / nsteps), so |sum over all rays| <= total * M, and
// scale = 2^62 / (max(total, 2^12) * M) keeps every partial and final sum
// within int64.  The rounding error per contribution is at most
// 2^-63 * max(total, 2^12) * M, and integer sums do not depend on the order
// of the atomics: the map gradient is deterministic.
__global__ __launch_bounds__(256) void k_ray_scale(RayArgs a, unsigned *mbits)
{
    const int64_t total = (int64_t)(a.ntx + a.nrx) * a.np;
    float m = 0.0f;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * 256) {
        const int e = (int)(q / a.np);
        const int64_t p = q - (int64_t)e * a.np;
        const float g = e < a.ntx ? (a.gtautx ? a.gtautx[q] : 0.0f)
                                  : (a.gtaurx ? a.gtaurx[q - (int64_t)a.ntx * a.np] : 0.0f);
        if (g == 0.0f) continue;
        float qx, L, fire;
        ray_of(e, a.ntx, a.steer, a.xorig, a.xele, a.c0, a.pvec[2 * p], a.pvec[2 * p + 1], qx, L,
               fire);
        const float v = fabsf(g * L);
        // NaN or infinity: the largest pattern, read back as "not finite"
        m = (v <= 3.0e38f) ? fmaxf(m, v) : __int_as_float(0x7f800000);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    // non-negative floats order like their bit patterns
    if ((threadIdx.x & 63) == 0) atomicMax(mbits, __float_as_uint(m));
}

// scale = 2^62 / (max(total, 2^12) * M): every sum fits int64, and one ray's
// contribution to a node stays below 2^51, where fixed_ll converts exactly
__device__ __forceinline__ double ray_scale(const RayArgs &a, const unsigned *mbits)
{
    const float m = __uint_as_float(*mbits);
    const double total = fmax((double)(a.ntx + a.nrx) * (double)a.np, 4096.0);
    return (m > 0.0f && m <= 3.0e38f) ? 4611686018427387904.0 / (total * (double)m) : 0.0;
}

// x rounded to the nearest integer (ties to even) for |x| < 2^51: adding
// 1.5 * 2^52 leaves the integer in the low significand bits
__device__ __forceinline__ long long fixed_ll(double x)
{
    const double magic = 6755399441055744.0;
    return __double_as_longlong(x + magic) - __double_as_longlong(magic);
}

// adjoint: one lane per ray.  Rays are visited in a scrambled order
// (q -> q * stride mod total, stride coprime to total) and each lane starts
// its ray at a different step, so that the lanes of a wave follow unrelated
// rays at different depths; every step adds its four bilinear contributions
// to the LDS map (branch-free across the wave: a run-length accumulation per
// cell flushed when the ray left it diverged on almost every step).  Each
// workgroup adds its LDS map into the global 64-bit sums with integer
// atomics (exact, so the order does not matter), and k_ray_reduce scales them.
__global__ __launch_bounds__(256) void k_ray_delays_bwd(RayArgs a, const unsigned *mbits,
                                                        unsigned long long *gsum, int64_t stride,
                                                        int64_t step)
{
    extern __shared__ unsigned long long gm64[];
    const int nmap = a.g.nzm * a.g.nxm;
    for (int t = threadIdx.x; t < nmap; t += 256) gm64[t] = 0ull;
    __syncthreads();
    const double scale = ray_scale(a, mbits);
    const int64_t total = (int64_t)(a.ntx + a.nrx) * a.np;
    const float rn = 1.0f / (float)a.nsteps;
    // ray r -> q = r * stride mod total; successive r of a lane differ by
    // gridDim * 256, so q advances by step = (gridDim * 256 * stride) mod total
    const int64_t r0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t q = (int64_t)(((unsigned __int128)r0 * (unsigned __int128)stride) %
                          (unsigned __int128)total);
    if (scale > 0.0) {
        for (int64_t r = r0; r < total; r += (int64_t)gridDim.x * 256,
                     q = (q + step >= total) ? q + step - total : q + step) {
            const int e = (int)(q / a.np);
            const int64_t p = q - (int64_t)e * a.np;
            const float g = e < a.ntx ? (a.gtautx ? a.gtautx[q] : 0.0f)
                                      : (a.gtaurx ? a.gtaurx[q - (int64_t)a.ntx * a.np] : 0.0f);
            if (g == 0.0f) continue;
            const float px = a.pvec[2 * p], pz = a.pvec[2 * p + 1];
            float qx, L, fire;
            ray_of(e, a.ntx, a.steer, a.xorig, a.xele, a.c0, px, pz, qx, L, fire);
            const float gl = g * L * rn;
            const double gls = (double)gl * scale;
            const int nxm = a.g.nxm, nzm = a.g.nzm;
            // lane l walks its ray from step l (mod nsteps) on, so the lanes of
            // a wave sit at different depths (and nodes) at any moment
            const int mo = (int)(threadIdx.x & 63) % a.nsteps;
            for (int mm = 0; mm < a.nsteps; ++mm) {
                const int m = mm + mo < a.nsteps ? mm + mo : mm + mo - a.nsteps;
                const float t = ((float)m + 0.5f) * rn;
                const float x = qx + t * (px - qx), z = t * pz;
                int i0, k0;
                float wz, wx;
                bilin(a.g, x, z, i0, k0, wz, wx);
                const int i1 = min(i0 + 1, nzm - 1), k1 = min(k0 + 1, nxm - 1);
                const float uz = 1.0f - wz, ux = 1.0f - wx;
                // weights that are zero (clamped edges) add nothing
                auto add = [&](int idx, float w) {
                    if (w != 0.0f)
                        atomicAdd(&gm64[idx], (unsigned long long)fixed_ll((double)w * gls));
                };
                add(i0 * nxm + k0, uz * ux);
                add(i0 * nxm + k1, uz * wx);
                add(i1 * nxm + k0, wz * ux);
                add(i1 * nxm + k1, wz * wx);
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nmap; t += 256) {
        const unsigned long long v = gm64[t];
        if (v) atomicAdd(gsum + t, v);
    }
}

// gslow[t] = (integer map sum) / scale; integer sums are exact, so the
// result does not depend on the workgroup count or order.  A non-finite
// upstream gradient gives NaN everywhere.
__global__ __launch_bounds__(256) void k_ray_reduce(RayArgs a, const unsigned *mbits,
                                                    const unsigned long long *__restrict__ gsum,
                                                    int nmap, float *__restrict__ gslow)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nmap) return;
    const float m = __uint_as_float(*mbits);
    if (!(m <= 3.0e38f)) {
        gslow[t] = __builtin_nanf("");
        return;
    }
    const double scale = ray_scale(a, mbits);
    if (scale == 0.0) {
        gslow[t] = 0.0f;
        return;
    }
    gslow[t] = (float)((double)(long long)gsum[t] / scale);
}

int launch_ray_delays(const RayArgs &a, int adjoint, hipStream_t s)
{
    const int nmap = a.g.nzm * a.g.nxm;
    if (!adjoint) {
        ProfScope ps("ray_delays", s);
        const dim3 grid((unsigned)((a.np + 255) / 256), (unsigned)(a.ntx + a.nrx));
        hipLaunchKernelGGL(k_ray_delays, grid, dim3(256), (size_t)nmap * 4, s, a);
        return check_launch("ray_delays");
    }
    ProfScope ps("ray_delays_bwd", s);
    const int64_t total = (int64_t)(a.ntx + a.nrx) * a.np;
    const int nb = (int)std::min<int64_t>((total + 255) / 256, SOS_PARTIALS);
    // workspace: the max word, then the map sums (64-bit)
    unsigned *mbits = reinterpret_cast<unsigned *>(a.gpart);
    unsigned long long *gsum =
        reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(a.gpart) + 256);
    hipLaunchKernelGGL(k_ray_clear, dim3((unsigned)((nmap + 255) / 256)), dim3(256), 0, s, mbits,
                       gsum, nmap);
    if (int rc = check_launch("ray adjoint clear")) return rc;
    const unsigned sb = (unsigned)std::min<int64_t>((total + 255) / 256, 2048);
    hipLaunchKernelGGL(k_ray_scale, dim3(sb), dim3(256), 0, s, a, mbits);
    if (int rc = check_launch("ray_scale")) return rc;
    // scramble stride: near total * 0.618, coprime to total
    int64_t stride = std::max<int64_t>(1, (int64_t)((double)total * 0.6180339887498949));
    while (std::__gcd(stride, total) != 1) ++stride;
    const int64_t step = (int64_t)(((unsigned __int128)((int64_t)nb * 256) * stride) % total);
    hipLaunchKernelGGL(k_ray_delays_bwd, dim3((unsigned)nb), dim3(256), (size_t)nmap * 8, s, a,
                       mbits, gsum, stride, step);
    if (int rc = check_launch("ray_delays_bwd")) return rc;
    hipLaunchKernelGGL(k_ray_reduce, dim3((unsigned)((nmap + 255) / 256)), dim3(256), 0, s, a,
                       mbits, gsum, nmap, a.gslow);
    return check_launch("ray_reduce");
}

int sos_max_map() { return SOS_MAXMAP; }
size_t ray_bwd_workspace_bytes(int nzm, int nxm) { return 256 + (size_t)nzm * nxm * 8; }

}  // namespace pycbf
