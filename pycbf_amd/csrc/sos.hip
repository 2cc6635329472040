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
// The integral is the midpoint rule (summed in double) with nsteps samples of the bilinear
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

constexpr int SOS_MAXMAP = 16384;   // map nodes held in LDS (64 KB)
constexpr int SOS_PARTIALS = 1024;  // workgroups (partial maps) of the adjoint

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
    const float px = a.pvec[2 * p], pz = a.pvec[2 * p + 1];
    float qx, L, fire;
    ray_of(e, a.ntx, a.steer, a.xorig, a.xele, a.c0, px, pz, qx, L, fire);
    const float rn = 1.0f / (float)a.nsteps;
    // the ray sum runs in double: a float32 sum of nsteps samples of ~1/c
    // would move the delay by a few thousandths of a sample
    double acc = 0.0;
    for (int m = 0; m < a.nsteps; ++m) {
        const float t = ((float)m + 0.5f) * rn;
        const float x = qx + t * (px - qx), z = t * pz;
        int i0, k0;
        float wz, wx;
        bilin(a.g, x, z, i0, k0, wz, wx);
        const int i1 = min(i0 + 1, a.g.nzm - 1), k1 = min(k0 + 1, a.g.nxm - 1);
        const float s00 = sm[i0 * a.g.nxm + k0], s01 = sm[i0 * a.g.nxm + k1];
        const float s10 = sm[i1 * a.g.nxm + k0], s11 = sm[i1 * a.g.nxm + k1];
        const float top = s00 + wx * (s01 - s00), bot = s10 + wx * (s11 - s10);
        acc += (double)(top + wz * (bot - top));
    }
    const float tau = (float)((double)fire + (double)L * (acc / (double)a.nsteps));
    if (e < a.ntx) a.tautx[(int64_t)e * a.np + p] = tau;
    else a.taurx[(int64_t)(e - a.ntx) * a.np + p] = tau;
}

// adjoint: one lane per ray.  Rays are visited in a scrambled order
// (q -> q * stride mod total, stride coprime to total) so that the lanes of a
// wave follow unrelated rays and rarely touch the same map cell at the same
// step; along its ray a lane accumulates the bilinear weights of the current
// cell in registers and adds them to the LDS map only when the ray leaves the
// cell.  Each workgroup stores its LDS map as a partial map, and k_ray_reduce
// sums the partials (no device-wide atomics on the few map nodes).
__device__ __forceinline__ void flush_cell(float *gm, int nxm, int nzm, int i0, int k0, float gl,
                                           float a00, float a01, float a10, float a11)
{
    if (i0 < 0) return;
    const int i1 = min(i0 + 1, nzm - 1), k1 = min(k0 + 1, nxm - 1);
    atomicAdd(&gm[i0 * nxm + k0], gl * a00);
    atomicAdd(&gm[i0 * nxm + k1], gl * a01);
    atomicAdd(&gm[i1 * nxm + k0], gl * a10);
    atomicAdd(&gm[i1 * nxm + k1], gl * a11);
}

__global__ __launch_bounds__(256) void k_ray_delays_bwd(RayArgs a, int64_t stride, int64_t step)
{
    extern __shared__ float gm[];
    const int nmap = a.g.nzm * a.g.nxm;
    for (int t = threadIdx.x; t < nmap; t += 256) gm[t] = 0.0f;
    __syncthreads();
    const int64_t total = (int64_t)(a.ntx + a.nrx) * a.np;
    const float rn = 1.0f / (float)a.nsteps;
    // ray r -> q = r * stride mod total; successive r of a lane differ by
    // gridDim * 256, so q advances by step = (gridDim * 256 * stride) mod total
    const int64_t r0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t q = (int64_t)(((unsigned __int128)r0 * (unsigned __int128)stride) %
                          (unsigned __int128)total);
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
        int ci = -1, ck = -1;
        float a00 = 0.0f, a01 = 0.0f, a10 = 0.0f, a11 = 0.0f;
        for (int m = 0; m < a.nsteps; ++m) {
            const float t = ((float)m + 0.5f) * rn;
            const float x = qx + t * (px - qx), z = t * pz;
            int i0, k0;
            float wz, wx;
            bilin(a.g, x, z, i0, k0, wz, wx);
            if (i0 != ci || k0 != ck) {
                flush_cell(gm, a.g.nxm, a.g.nzm, ci, ck, gl, a00, a01, a10, a11);
                ci = i0;
                ck = k0;
                a00 = a01 = a10 = a11 = 0.0f;
            }
            const float uz = 1.0f - wz, ux = 1.0f - wx;
            a00 += uz * ux;
            a01 += uz * wx;
            a10 += wz * ux;
            a11 += wz * wx;
        }
        flush_cell(gm, a.g.nxm, a.g.nzm, ci, ck, gl, a00, a01, a10, a11);
    }
    __syncthreads();
    float *part = a.gpart + (int64_t)blockIdx.x * nmap;
    for (int t = threadIdx.x; t < nmap; t += 256) part[t] = gm[t];
}

// gslow[t] = sum over workgroups of their partial maps (fixed order)
__global__ __launch_bounds__(256) void k_ray_reduce(const float *__restrict__ part, int nb, int nmap,
                                                    float *__restrict__ gslow)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nmap) return;
    float acc = 0.0f;
    for (int b = 0; b < nb; ++b) acc += part[(int64_t)b * nmap + t];
    gslow[t] = acc;
}

int launch_ray_delays(const RayArgs &a, int adjoint, hipStream_t s)
{
    const int nmap = a.g.nzm * a.g.nxm;
    const size_t lds = (size_t)nmap * 4;
    if (!adjoint) {
        ProfScope ps("ray_delays", s);
        const dim3 grid((unsigned)((a.np + 255) / 256), (unsigned)(a.ntx + a.nrx));
        hipLaunchKernelGGL(k_ray_delays, grid, dim3(256), lds, s, a);
        return check_launch("ray_delays");
    }
    ProfScope ps("ray_delays_bwd", s);
    const int64_t total = (int64_t)(a.ntx + a.nrx) * a.np;
    const int nb = (int)std::min<int64_t>((total + 255) / 256, SOS_PARTIALS);
    // scramble stride: near total * 0.618, coprime to total
    int64_t stride = std::max<int64_t>(1, (int64_t)((double)total * 0.6180339887498949));
    while (std::__gcd(stride, total) != 1) ++stride;
    const int64_t step = (int64_t)(((unsigned __int128)((int64_t)nb * 256) * stride) % total);
    hipLaunchKernelGGL(k_ray_delays_bwd, dim3((unsigned)nb), dim3(256), lds, s, a, stride, step);
    if (int rc = check_launch("ray_delays_bwd")) return rc;
    hipLaunchKernelGGL(k_ray_reduce, dim3((unsigned)((nmap + 255) / 256)), dim3(256), 0, s, a.gpart,
                       nb, nmap, a.gslow);
    return check_launch("ray_reduce");
}

int sos_max_map() { return SOS_MAXMAP; }
size_t ray_bwd_workspace_bytes(int nzm, int nxm) { return (size_t)SOS_PARTIALS * nzm * nxm * 4; }

}  // namespace pycbf
