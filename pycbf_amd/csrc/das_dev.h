// Device helpers shared by the DAS kernels: sample-index selection, the
// synthetic-point geometry and workgroup-coordinate mapping.
#pragma once
#include <math.h>

#include "common.h"

namespace pycbf {

// x / c for a divisor that recip_division_exact(c) accepts (r = 1 / c
// rounded): the product with the reciprocal and one fused correction, the
// correctly rounded quotient for 2^-60 <= |x| < 2^60.  Its callers' dividends
// lie there or do not matter: a knot offset x - t0 in [0, tmax - t0]
// truncates to index 0 when tiny; a path length of 0 gives a delay of 0 (its
// sign may differ, which changes no sum); an overflowing length gives NaN
// instead of infinity (out of the trace either way).
__device__ __forceinline__ float div_rn(float x, float c, float r)
{
    const float q = x * r;
    const float e = __builtin_fmaf(-q, c, x);
    return __builtin_fmaf(e, r, q);
}

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

// knot_index with the quotient by div_rn (rdt = 1 / dt rounded, dt accepted
// by recip_division_exact): the same index
__device__ __forceinline__ bool knot_index_rn(float x, float t0, float dt, float rdt, float tmax,
                                              int n, int &j)
{
    if (!(x >= t0 && x <= tmax)) return false;
    j = min((int)div_rn(x - t0, dt, rdt), n - 1);
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

// Workgroups go to the 8 XCDs round robin by linear id.  With remap on, the
// linear id is permuted so each XCD receives one contiguous run of the
// tile-fastest (tile, frame chunk, outer element) order: a slab of outer
// elements over every tile, so the channels of an outer element are read
// into one or two XCDs' L2 instead of all eight.
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

// ---------------------------------------------------------------------------
// synthetic-point geometry (pycbf/gpu/__engines__.cu:382-470), float32, unfused
// ---------------------------------------------------------------------------

// An element's geometry pack, read through the constant address space: the
// packs are read-only while a kernel runs, so a pack whose index is uniform
// comes in through the scalar cache (its loads wait on lgkmcnt, not behind
// the vector-memory loads in flight).
template <int N>
__device__ __forceinline__ void load_pack(const float *p, float (&pk)[N])
{
    typedef const __attribute__((address_space(4))) float cfloat;
    const cfloat *cp = (const cfloat *)p;
#pragma unroll
    for (int i = 0; i < N; ++i) pk[i] = cp[i];
}

// "acos(v) > ala" for the T of the element's pack; a NaN acos (|v| > 1) is false.
__device__ __forceinline__ bool acos_gt(float v, float T) { return (v >= -1.0f && v <= 1.0f) && (v < T); }
// "acos(v) <= ala"
__device__ __forceinline__ bool acos_le(float v, float T) { return (v >= -1.0f && v <= 1.0f) && !(v < T); }

// FD: quotients by c0 through div_rn (rc0 = 1 / c0 rounded, c0 accepted by
// recip_division_exact); the same delays
template <int NDIM, bool FD = false>
__device__ __forceinline__ void geom_tx(const float *__restrict__ pk, float c0, const float *p,
                                        float &tau, float &apod, float rc0 = 0.0f)
{
    auto over_c0 = [&](float v) { return FD ? div_rn(v, c0, rc0) : v / c0; };
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
        tau = over_c0(proj);
        apod = 1.0f;
        return;
    }
    const float sgn = (proj > 0.0f) ? 1.0f : -1.0f;
    if (mode == 2) {  // focus with depth of field
        const bool in_cone = ((double)fabsf(proj) > 1e-12) && ((double)mag > 1e-12) &&
                             acos_le(fabsf(proj / mag), T);
        if (in_cone) {
            tau = sgn * over_c0(mag) + t0;
            apod = 1.0f;
        } else {
            const float lat = sqrtf(fabsf(mag * mag - proj * proj));
            if ((double)lat < (double)pk[9] / 2.0) {
                tau = over_c0(proj * pk[10]) + t0;
                apod = 1.0f;
            } else {
                apod = 0.0f;
            }
        }
        return;
    }
    tau = ((double)fabsf(proj) > 1e-12) ? sgn * over_c0(mag) + t0 : t0;
    apod = (((double)mag > 1e-12) && acos_gt(fabsf(proj / mag), T)) ? 0.0f : 1.0f;
}

// sqrtf(x), correctly rounded.  For x >= 2^-96 (+inf included) it is the
// hardware root (within one ulp) moved to the neighbour the residuals
// x - s' * s select: the instructions sqrtf itself runs for such x, without its
// rescaling of tiny inputs and its zero / infinity select.  Smaller x and NaN
// take sqrtf in a branch the compiler cannot flatten (the empty volatile asm
// keeps it from computing both sides and selecting).
__device__ __forceinline__ float sqrt_rn(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    float r = (__builtin_fmaf(-sd, s, x) <= 0.0f) ? sd : s;
    r = (__builtin_fmaf(-su, s, x) > 0.0f) ? su : r;
    if (!(x >= 0x1p-96f)) {
        asm volatile("");
        r = sqrtf(x);
    }
    return r;
}

// The receive weight rule acos(RN(proj / mag)) > ala, decided without the
// division.  m (rx pack floats 8-9, a double) is the midpoint between the
// pack's T and the float below it; for mag > 0:
//   RN(proj / mag) < T   <=>  proj < m * mag
//   RN(proj / mag) >= -1 <=>  proj > -(1 + 2^-24) * mag
// (T <= 1, so RN < T also gives RN <= 1).  Both products are exact in double
// (25 x 24 significant bits) and no float proj equals either one: the midpoints
// have 25 significant bits (an odd numerator; T is never subnormal, as
// |cos(a)| >= 4e-8 for every float angle a in [0, pi]) and a product with a
// float keeps them, where proj has 24.  T = +inf (ala < 0: only |q| <= 1 is
// asked) stores m = 1 + 2^-24, T = -inf stores m = -inf.  The same holds at
// mag = +inf (proj finite: RN(q) = +-0; proj infinite: NaN, no weight) and for
// NaN operands (no weight).
__device__ __forceinline__ bool rx_outside(float proj, float mag, const float *pk)
{
    const double m = __hiloint2double(__float_as_int(pk[9]), __float_as_int(pk[8]));
    const double md = (double)mag, pd = (double)proj;
    return (md > 1e-12) && (pd > md * -0x1.000001p0) && (pd < md * m);
}

template <int NDIM, bool FD = false>
__device__ __forceinline__ void geom_rx(const float *__restrict__ pk, float c0, const float *p,
                                        float &tau, float &apod, float rc0 = 0.0f)
{
    float mag = 0.0f, proj = 0.0f;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        const float di = p[d] - pk[d];
        mag += di * di;
        proj = (d == 0) ? di * pk[3] : proj + di * pk[3 + d];
    }
    mag = sqrt_rn(mag);
    tau = FD ? div_rn(mag, c0, rc0) : mag / c0;
    apod = rx_outside(proj, mag, pk) ? 0.0f : 1.0f;
}

// ---------------------------------------------------------------------------
// gather + interpolate for FC frames of one channel
// ---------------------------------------------------------------------------
// delay and weight of element e of the tx (or rx) axis at pixel p
// FD: quotients by c0 through div_rn (see geom_tx)
template <int GEO, bool FD = false>
__device__ __forceinline__ void elem_eval(const DasArgs &A, bool tx, int e, int64_t p,
                                          const float (&pv)[3], float &tau, float &apod)
{
    if (GEO == 0) {
        const int64_t o = (int64_t)e * A.np + p;
        tau = tx ? A.ttx[o] : A.trx[o];
        apod = tx ? A.atx[o] : A.arx[o];
    } else if (tx) {
        float pk[PYCBF_TXPACK_FLOATS];
        load_pack(A.txpack + (int64_t)e * PYCBF_TXPACK_FLOATS, pk);
        geom_tx<GEO, FD>(pk, A.c0, pv, tau, apod, A.rc0);
    } else {
        float pk[PYCBF_RXPACK_FLOATS];
        load_pack(A.rxpack + (int64_t)e * PYCBF_RXPACK_FLOATS, pk);
        geom_rx<GEO, FD>(pk, A.c0, pv, tau, apod, A.rc0);
    }
}

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

// As gather, at the sample position q = (x - t0) / dt held in double (x
// placed beyond float32 by delay residuals): the piece j = floor(q) with the
// exact weights b = q - j, a = 1 - b (the reference's separately rounded knot
// weights differ from these by up to a few 1e-5); the trace bounds follow the
// float32 x as in gather.  nearest keeps gather's rule.
template <int GK, int FC>
__device__ __forceinline__ void gather_q(const DasArgs &A, int64_t ch0, int64_t fstride, int nf,
                                         float x, double q, bool &ok, float (&v)[FC])
{
    if (GK == GK_NEAREST) {
        gather<GK, FC>(A, ch0, fstride, nf, x, ok, v);
        return;
    }
    ok = x >= A.t0 && x <= A.tmax;
    if (!ok) return;
    int j = (int)floor(q);
    j = min(max(j, 0), A.nsrc - 1);
    const double fr = q - (double)j;
    if (GK == GK_LINEAR) {
        const float b = (float)fr, a = (float)(1.0 - fr);
        const bool hi = (j + 1) < A.nsrc;
        const float *s = A.src + ch0 * A.ch_stride + j;
#pragma unroll
        for (int f = 0; f < FC; ++f)
            if (f < nf) v[f] = a * s[f * fstride] + b * (hi ? s[f * fstride + 1] : 0.0f);
    } else if (GK == GK_CUBIC) {
        const float b = (float)fr, a = (float)(1.0 - fr);
        const float ca = __builtin_fmaf(a * a, a, -a);
        const float cb = __builtin_fmaf(b * b, b, -b);
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
        const float t = (float)fr;
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

}  // namespace pycbf
