// Per-channel prefilters: turn each RF channel into a table the DAS gather
// reads with one load per sample.
//
//  * k_spline_seg (nt <= 4096): natural cubic spline (tie_knots1D_fixed,
//    pycbf/cpu/pycbf.c:114-166), one wave per channel, each lane sweeping a
//    segment from a verified carry; bit-identical second derivatives.
//  * k_spline (longer channels): the same spline, one wave owns 64 channels;
//    32-sample tiles are staged through LDS so that global loads/stores stay
//    coalesced while each lane runs its channel's sequential recurrence in the
//    reference's exact float32 operation order.
//    Output of both is either y2 itself or the (y, dt^2/6*y2) table used by
//    the DAS gather.
//  * k_upsample: the spline evaluated on a dt/usf grid (pycbf/cpu/pycbf.c:234-253).
//  * k_korder_coefs: Hermite coefficients with k-order derivatives
//    (pycbf/gpu/__engines__.cu:292-355).
//  * k_akima_coefs: Hermite coefficients with Akima / modified-Akima
//    derivatives in consistent per-sample units (the reference mixes units,
//    pycbf/gpu/__engines__.cu:153-160,261-268).
#include <math.h>
#include <stdlib.h>

#include "common.h"
#include "prof.h"
#include "tiles.h"

namespace pycbf {

// The data-independent forward coefficients of the spline sweep,
// y2f[i] = (0.5 - 1) / (0.5*y2f[i-1] + 2), reach an exact float32 fixed point
// at i = 7; the kernels keep the first 32 and use y2f[31] for every later i.

// One wave per 64 channels, one lane per channel for the recurrences.  Tiles
// of 32 samples move through LDS (coalesced global traffic, transposed to one
// row per lane) with the next tile prefetched into registers.  Per tile the
// data-parallel part of the forward sweep, g_i = 6*((y[i+1]-y[i])/dx -
// (y[i]-y[i-1])/dx)/(2*dx), is evaluated first (32 independent evaluations
// per lane), so only u_i = (g_i - 0.5*u_{i-1}) / p_i remains on the serial
// chain.  Operations and their order are those of tie_knots1D_fixed
// (pycbf/cpu/pycbf.c:126-141): y2 is bit-identical.
//
// Tile cb of the forward sweep (samples s = 32*cb .. s+31 in LDS) computes
// u_i for i = s-1 .. s+30, whose right neighbour y[i+1] lies in the same
// tile; y[s-1] and the slope ending there are carried in registers.  The
// scratch row therefore holds U[i+1] = u_i (U[0] = 0), so that each tile's
// u values land in one aligned 32-sample tile of the scratch row, and the
// back substitution of tile cb produces y2[s-1 .. s+30].
//
// Resources: 64-thread workgroups, 9.3 KB of LDS and at most 168 VGPRs
// (3 waves per SIMD): the 2400 waves of a 16-frame, 75 x 128-channel batch
// are all resident at once on 1024 SIMDs instead of running in two rounds.
template <int MODE>  // 0: (y, dx^2/6*y2) gather table, 1: y2
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_spline(const float *__restrict__ rf, int64_t nch, int nt,
                                               float dx, float *__restrict__ uscr,
                                               void *__restrict__ outp)
{
    constexpr int CH = SPL_CHUNK;
    static_assert(CH == 32, "tile helpers assume 32-sample tiles");
    __shared__ __align__(16) SplRow S[64];
    __shared__ float y2f[32];  // forward coefficients; y2f[i] = y2f[31] for i >= 7

    const int lane = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int nrow = (int)((nch - c0) < 64 ? (nch - c0) : 64);
    const float *ybase = rf + c0 * (int64_t)nt;
    float *ubase = uscr + c0 * (int64_t)nt;

    float y2f_inf;
    {
        // y2f[i] = (0.5 - 1) / (0.5*y2f[i-1] + 2), y2f[0] = 0 (pycbf.c:127-133)
        float v = 0.0f, mine = 0.0f;
#pragma unroll
        for (int i = 1; i < 32; ++i) {
            const float p = 0.5f * v + 2.0f;
            v = (0.5f - 1.0f) / p;
            if (lane == i) mine = v;
        }
        if (lane < 32) y2f[lane] = mine;
        y2f_inf = v;
    }
    const float p_inf = 0.5f * y2f_inf + 2.0f;  // p_i for every i >= 8
    const int nchunk = (nt + CH - 1) / CH;

    // ---- forward sweep ---------------------------------------------------
    TileRegs pf;
    tile_fetch(ybase, nt, nrow, 0, nt, pf);
    tile_to_lds(pf, S);
    if (nchunk > 1) tile_fetch(ybase, nt, nrow, CH, nt, pf);
    __syncthreads();
    // carried: y[s-1], the slope (y[s-1] - y[s-2])/dx, and u_{s-2}.  The
    // reference evaluates both slopes at every step; the backward one equals
    // the previous step's forward slope (same operands, same rounding).
    float y_prev = 0.0f, d_prev = 0.0f, u_prev = 0.0f;
    for (int cb = 0; cb < nchunk; ++cb) {
        const int s = cb * CH;
        float v[CH];
        row_read(S, v);
        // v[c] <- g_{s-1+c}
        {
            float yl = y_prev;
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const float yr = v[c];
                const float d = (yr - yl) / dx;
                const float jump = d - d_prev;
                v[c] = 6.0f * jump / (2.0f * dx);
                d_prev = d;
                yl = yr;
            }
            y_prev = yl;
        }
        // v[c] <- u_{s-1+c} (zero outside 1 <= i <= nt-2)
        if (s >= 32 && s + CH <= nt) {
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                u_prev = (v[c] - 0.5f * u_prev) / p_inf;
                v[c] = u_prev;
            }
        } else {
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int i = s - 1 + c;
                float ui = 0.0f;
                if (i >= 1 && i <= nt - 2) {
                    const float p = 0.5f * y2f[(i - 1) < 31 ? (i - 1) : 31] + 2.0f;
                    ui = (v[c] - 0.5f * u_prev) / p;
                    u_prev = ui;
                }
                v[c] = ui;
            }
        }
        row_write(S, v);  // the lane's own row: no other lane read it
        __syncthreads();  // u tile complete
        lds_to_global(S, ubase, nt, nrow, s, nt);
        __syncthreads();  // u tile read out
        if (cb + 1 < nchunk) {
            tile_to_lds(pf, S);
            if (cb + 2 < nchunk) tile_fetch(ybase, nt, nrow, s + 2 * CH, nt, pf);
        }
        __syncthreads();  // next y tile complete
    }
    // the scratch rows are read back below by the lanes that wrote them
    __threadfence();

    // ---- back substitution: y2[k] = y2f[k]*y2[k+1] + u[k] ------------------
    const float h = (1.0f / 6.0f) * (dx * dx);
    // y2[s+31] at the top of tile cb: zero for the last tile (s+31 >= nt-1)
    float y2_carry = 0.0f;
    const int slast = (nchunk - 1) * CH;
    TileRegs py, pu;
    if (MODE == 0) tile_fetch(ybase, nt, nrow, slast, nt, py);
    tile_fetch(ubase, nt, nrow, slast, nt, pu);
    if (MODE == 0 && lane < nrow) {
        float2 *tab = static_cast<float2 *>(outp) + (c0 + lane) * (int64_t)(nt + 1);
        tab[nt] = make_float2(0.0f, 0.0f);
    }
    for (int cb = nchunk - 1; cb >= 0; --cb) {
        const int s = cb * CH;
        tile_to_lds(pu, S);  // U[s+c] = u_{s-1+c}
        __syncthreads();
        if (cb > 0) tile_fetch(ubase, nt, nrow, s - CH, nt, pu);
        float v[CH];
        row_read(S, v);
        // k = s-1+c from c = 31 down to 0; v[c] <- output entry s+c, i.e.
        // y2[s+c] (y2[s+31] is the carry, y2[s-1] the next tile's carry)
        float nxt = y2_carry;
        if (s >= 32 && s + CH <= nt) {
#pragma unroll
            for (int c = CH - 1; c >= 0; --c) {
                const float y2k = y2f_inf * nxt + v[c];
                v[c] = nxt;
                nxt = y2k;
            }
        } else {
#pragma unroll
            for (int c = CH - 1; c >= 0; --c) {
                const int k = s - 1 + c;
                float y2k = 0.0f;
                if (k >= 0 && k < nt - 1) y2k = y2f[k < 31 ? k : 31] * nxt + v[c];
                v[c] = nxt;
                nxt = y2k;
            }
        }
        y2_carry = nxt;
        if (MODE == 0) {
#pragma unroll
            for (int c = 0; c < CH; ++c) v[c] = h * v[c];
        }
        row_write(S, v);  // own row, read above by this lane only
        __syncthreads();  // output tile complete
        if (MODE == 0) {
            // entries (y_k, h*y2_k) of rows 8q + lane/8, k = s + 4*(lane%8) .. +3:
            // y from the register tile, h*y2 from LDS, 32 contiguous bytes per lane
            const int cc = 4 * (lane & 7);
            float *tab = static_cast<float *>(outp) + c0 * (int64_t)(2 * (nt + 1));
            const uint32_t o = (uint32_t)(2 * ((lane >> 3) * (nt + 1) + s + cc));
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int r = 8 * q + (lane >> 3);
                if (r >= nrow) continue;
                const float4 ov = *reinterpret_cast<const float4 *>(&S[r][cc]);
                const F4u &y = py.q[q];
                float *dst = at(tab, o + (uint32_t)(16 * q * (nt + 1)));
                if (s + cc + 4 <= nt) {
                    F4u w0{{y.v[0], ov.x, y.v[1], ov.y}};
                    F4u w1{{y.v[2], ov.z, y.v[3], ov.w}};
                    reinterpret_cast<F4u *>(dst)[0] = w0;
                    reinterpret_cast<F4u *>(dst)[1] = w1;
                } else {
                    const float oo[4] = {ov.x, ov.y, ov.z, ov.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (s + cc + e < nt) {
                            dst[2 * e] = y.v[e];
                            dst[2 * e + 1] = oo[e];
                        }
                }
            }
            if (cb > 0) tile_fetch(ybase, nt, nrow, s - CH, nt, py);
        } else {
            float *y2 = static_cast<float *>(outp) + c0 * (int64_t)nt;
            lds_to_global(S, y2, nt, nrow, s, nt);
        }
        __syncthreads();  // staging tile free for the next tile
    }
}

// ---- segment-parallel sweeps: one wave per channel -------------------------
//
// For nt <= 64*SEG, lane l of the wave owns the SEG consecutive samples
// i0 = l*SEG .. i0+SEG-1 of one channel, held in registers.  Both sweeps of
// tie_knots1D_fixed are first-order recurrences that contract by
// |0.5/p| ~ 0.27 per step (p -> 1 + sqrt(3)/2), so a lane can start each
// sweep from a guessed carry a little outside its segment:
//  * forward: u = 0 at i0-32, run over the left neighbour's last 31 samples
//    to get a guess of u_{i0-1};
//  * backward: y2 = 0 at i0+SEG+32, run over the right neighbour's first 32
//    u values to get a guess of y2_{i0+SEG}.
// After each sweep every guess is compared bit for bit with the value the
// neighbouring lane computed.  A lane whose guess differs takes its
// neighbour's value and recomputes its segment, until every carry agrees.
// Lane 0's forward carry and lane 63's backward carry are the exact boundary
// values, and a segment computed from an exact carry is exact, so agreement
// of all carries means every segment equals the sequential sweep bit for bit,
// whatever the data.  A guess misses only when the data's magnitude falls by
// ~1e10 within 32 samples (an echo followed by exact zeros); the recomputation
// costs one segment pass for the few lanes involved.
//
// Global memory is touched once per sample: the channel is read in coalesced
// 64 x 32 tiles (row l = 32 samples of lane l's segment) and transposed
// through one LDS tile; the output goes back the same way.  Compared with
// k_spline there is no scratch round trip and no second read of the channel.
constexpr int SEG_WARM = 32;  // warm-up length of both sweeps (samples)

// Fills u_{i0+c} (c < SEG) from the exact or guessed carry u = u_{i0-1}.
// y = the lane's samples, yl = y_{i0-1}, yr = y_{i0+SEG}.  u values with
// index > nt-2 are stored as 0 (the back substitution starts at y2_{nt-1}=0).
// Returns u_{i0+SEG-1} as seen by the right neighbour.
template <int SEG>
__device__ __forceinline__ float seg_forward(const float (&y)[SEG], float yl, float yr, float dx,
                                             float u, int i0, int nt, float p_inf,
                                             const float (&plo)[9], bool lane0, float (&v)[SEG])
{
    float d_prev = (y[0] - yl) / dx;
#pragma unroll
    for (int c = 0; c < SEG; ++c) {
        const float ynext = (c + 1 < SEG) ? y[c + 1] : yr;
        const float d = (ynext - y[c]) / dx;
        const float g = 6.0f * (d - d_prev) / (2.0f * dx);
        d_prev = d;
        // p_i = 0.5*y2f[i-1] + 2 differs from p_inf only for i < 9 (lane 0)
        const float p = (c < 9 && lane0) ? plo[c < 9 ? c : 0] : p_inf;
        float un = (g - 0.5f * u) / p;
        if (c == 0) un = lane0 ? 0.0f : un;  // u_0 = 0
        u = un;
        v[c] = (i0 + c <= nt - 2) ? u : 0.0f;
    }
    return u;
}

// y2_{i0+c} = y2f[i0+c] * y2_{i0+c+1} + u_{i0+c}, c = SEG-1 .. 0, in place,
// from the carry y2_{i0+SEG}.  Returns y2_{i0}.
template <int SEG>
__device__ __forceinline__ float seg_backward(float (&v)[SEG], float nxt, float f_inf,
                                              const float (&flo)[8], bool lane0)
{
#pragma unroll
    for (int c = SEG - 1; c >= 0; --c) {
        const float f = (c < 8 && lane0) ? flo[c < 8 ? c : 0] : f_inf;
        nxt = f * nxt + v[c];
        v[c] = nxt;
    }
    return nxt;
}

__device__ __forceinline__ bool bits_differ(float a, float b)
{
    return __float_as_uint(a) != __float_as_uint(b);
}

template <int MODE, int SEG>  // MODE 0: (y, dx^2/6*y2) gather table, 1: y2
__global__ __launch_bounds__(64) void k_spline_seg(const float *__restrict__ rf, int nt, float dx,
                                                   void *__restrict__ outp)
{
    static_assert(SEG % 32 == 0 && SEG >= SEG_WARM, "segments are whole 32-sample tiles");
    __shared__ __align__(16) SplRow S[64];
    const int lane = threadIdx.x;
    const bool lane0 = lane == 0;
    const int64_t ch = blockIdx.x;
    const float *yrow = rf + ch * (int64_t)nt;
    const int i0 = lane * SEG;

    // forward coefficients y2f[i] = (0.5 - 1) / (0.5*y2f[i-1] + 2), y2f[0] = 0
    // (pycbf.c:127-133): exact float32 fixed point from i = 7 on
    float flo[8], plo[9];
    float y2f_inf;
    {
        float w = 0.0f;
        flo[0] = 0.0f;
        plo[0] = 0.0f;  // unused: u_0 = 0
#pragma unroll
        for (int i = 1; i < 32; ++i) {
            const float p = 0.5f * w + 2.0f;
            if (i < 9) plo[i] = p;
            w = (0.5f - 1.0f) / p;
            if (i < 8) flo[i] = w;
        }
        y2f_inf = w;
    }
    const float p_inf = 0.5f * y2f_inf + 2.0f;

    // ---- load: tile t holds samples 32t .. 32t+31 of every segment --------
    float y[SEG];
    {
        const int c = 4 * (lane & 7);
#pragma unroll
        for (int t = 0; t < SEG / 32; ++t) {
            TileRegs r;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int k = (8 * q + (lane >> 3)) * SEG + 32 * t + c;
                F4u w{{0.0f, 0.0f, 0.0f, 0.0f}};
                if (k + 4 <= nt) {
                    w = *reinterpret_cast<const F4u *>(yrow + k);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (k + e < nt) w.v[e] = yrow[k + e];
                }
                r.q[q] = w;
            }
            if (t > 0) __syncthreads();  // previous tile read out
            tile_to_lds(r, S);
            __syncthreads();
            float part[32];
            row_read(S, part);
#pragma unroll
            for (int e = 0; e < 32; ++e) y[32 * t + e] = part[e];
        }
    }
    const float yl = __shfl_up(y[SEG - 1], 1);  // y_{i0-1}
    const float yr = __shfl_down(y[0], 1);      // y_{i0+SEG}

    // ---- forward sweep -----------------------------------------------------
    // guess of u_{i0-1}: u = 0 at i0-32, then i = i0-31 .. i0-1 over the left
    // neighbour's last 32 samples (still in the LDS tile)
    float carry = 0.0f;
    if (lane > 0) {
        float w[32];
#pragma unroll
        for (int e = 0; e < 32; e += 4) {
            const float4 q4 = *reinterpret_cast<const float4 *>(&S[lane - 1][e]);
            w[e] = q4.x;
            w[e + 1] = q4.y;
            w[e + 2] = q4.z;
            w[e + 3] = q4.w;
        }
        // only lane 1 of a 32-sample segment reaches i < 9 here
        const bool low = (SEG == 32) && lane == 1;
        float d_prev = (w[1] - w[0]) / dx;
        float u = 0.0f;
#pragma unroll
        for (int c = 1; c < 32; ++c) {
            const float ynext = (c + 1 < 32) ? w[c + 1] : y[0];
            const float d = (ynext - w[c]) / dx;
            const float g = 6.0f * (d - d_prev) / (2.0f * dx);
            d_prev = d;
            const float p = (c < 9 && low) ? plo[c < 9 ? c : 0] : p_inf;
            u = (g - 0.5f * u) / p;
        }
        carry = u;
    }
    float v[SEG];
    float last = seg_forward<SEG>(y, yl, yr, dx, carry, i0, nt, p_inf, plo, lane0, v);
    // lanes with i0 > nt-2 hold no u values; their carries need not agree
    for (;;) {
        const float want = __shfl_up(last, 1);
        const bool redo = lane > 0 && i0 <= nt - 2 && bits_differ(carry, want);
        if (__ballot(redo) == 0) break;
        if (redo) {
            carry = want;
            last = seg_forward<SEG>(y, yl, yr, dx, carry, i0, nt, p_inf, plo, lane0, v);
        }
    }
    const float carry_fwd = carry;

    // ---- back substitution -------------------------------------------------
    __syncthreads();  // every lane has read its neighbour's y row
#pragma unroll
    for (int e = 0; e < 32; e += 4)  // u_{i0 .. i0+31}
        *reinterpret_cast<float4 *>(&S[lane][e]) = make_float4(v[e], v[e + 1], v[e + 2], v[e + 3]);
    __syncthreads();
    float bcarry = 0.0f;  // lane 63: y2 beyond the channel is 0
    if (lane < 63) {
        float nxt = 0.0f;
#pragma unroll
        for (int e = 28; e >= 0; e -= 4) {
            const float4 q4 = *reinterpret_cast<const float4 *>(&S[lane + 1][e]);
            nxt = y2f_inf * nxt + q4.w;
            nxt = y2f_inf * nxt + q4.z;
            nxt = y2f_inf * nxt + q4.y;
            nxt = y2f_inf * nxt + q4.x;
        }
        bcarry = nxt;
    }
    float first = seg_backward<SEG>(v, bcarry, y2f_inf, flo, lane0);
    for (;;) {
        const float want = __shfl_down(first, 1);
        const bool redo = lane < 63 && i0 < nt && bits_differ(bcarry, want);
        if (__ballot(redo) == 0) break;
        if (redo) {
            // rebuild u from the exact forward carry, then substitute again
            bcarry = want;
            seg_forward<SEG>(y, yl, yr, dx, carry_fwd, i0, nt, p_inf, plo, lane0, v);
            first = seg_backward<SEG>(v, bcarry, y2f_inf, flo, lane0);
        }
    }

    // ---- output ------------------------------------------------------------
    const int c = 4 * (lane & 7);
    if (MODE == 0) {
        // 16 samples per pass: lane row = (y_k, h*y2_k) pairs, 32 floats
        const float h = (1.0f / 6.0f) * (dx * dx);
        float *tab = static_cast<float *>(outp) + ch * (int64_t)(2 * (nt + 1));
#pragma unroll
        for (int hh = 0; hh < SEG / 16; ++hh) {
            __syncthreads();  // staging tile free
#pragma unroll
            for (int j = 0; j < 16; j += 2)
                *reinterpret_cast<float4 *>(&S[lane][2 * j]) =
                    make_float4(y[16 * hh + j], h * v[16 * hh + j], y[16 * hh + j + 1],
                                h * v[16 * hh + j + 1]);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int r = 8 * q + (lane >> 3);
                const int f = 2 * (r * SEG + 16 * hh) + c;  // float index in the row
                const float4 w = *reinterpret_cast<const float4 *>(&S[r][c]);
                if (f + 4 <= 2 * nt) {
                    F4u o{{w.x, w.y, w.z, w.w}};
                    *reinterpret_cast<F4u *>(tab + f) = o;
                } else {
                    const float e4[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (f + e < 2 * nt) tab[f + e] = e4[e];
                }
            }
        }
        if (lane0) {
            tab[2 * nt] = 0.0f;
            tab[2 * nt + 1] = 0.0f;
        }
    } else {
        float *y2 = static_cast<float *>(outp) + ch * (int64_t)nt;
#pragma unroll
        for (int t = 0; t < SEG / 32; ++t) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 32; e += 4)
                *reinterpret_cast<float4 *>(&S[lane][e]) =
                    make_float4(v[32 * t + e], v[32 * t + e + 1], v[32 * t + e + 2],
                                v[32 * t + e + 3]);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int r = 8 * q + (lane >> 3);
                const int k = r * SEG + 32 * t + c;
                const float4 w = *reinterpret_cast<const float4 *>(&S[r][c]);
                if (k + 4 <= nt) {
                    F4u o{{w.x, w.y, w.z, w.w}};
                    *reinterpret_cast<F4u *>(y2 + k) = o;
                } else {
                    const float e4[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (k + e < nt) y2[k + e] = e4[e];
                }
            }
        }
    }
}

int spline_seg_len(int32_t nt)
{
    if (const char *e = getenv("PYCBF_SPLINE_SEG"))
        if (atoi(e) == 0) return 0;
    if (nt <= 64 * 32) return 32;
    if (nt <= 64 * 64) return 64;
    return 0;
}

template <int MODE>
static int launch_spline(const float *rf, int64_t nch, int32_t nt, float dt, float *uscr,
                         void *out, hipStream_t s)
{
    if (nch <= 0 || nt <= 0) return 0;
    if (nt > SPL_MAX_NT) return fail(1, "spline: nt %d exceeds %d samples", nt, SPL_MAX_NT);
    switch (spline_seg_len(nt)) {
    case 32:
        hipLaunchKernelGGL((k_spline_seg<MODE, 32>), dim3((unsigned)nch), dim3(64), 0, s, rf,
                           (int)nt, dt, out);
        break;
    case 64:
        hipLaunchKernelGGL((k_spline_seg<MODE, 64>), dim3((unsigned)nch), dim3(64), 0, s, rf,
                           (int)nt, dt, out);
        break;
    default:
        hipLaunchKernelGGL(k_spline<MODE>, dim3((unsigned)((nch + 63) / 64)), dim3(64), 0, s, rf,
                           nch, (int)nt, dt, uscr, out);
    }
    return check_launch(MODE == 0 ? "spline coefficients" : "spline y2");
}

int launch_spline_coefs(const float *rf, int64_t nch, int32_t nt, float dt, float *uscr,
                        float2 *coef, hipStream_t s)
{
    ProfScope prof("spline_coefs", s);
    return launch_spline<0>(rf, nch, nt, dt, uscr, (void *)coef, s);
}

int launch_spline_y2(const float *rf, int64_t nch, int32_t nt, float dt, float *uscr, float *y2,
                     hipStream_t s)
{
    return launch_spline<1>(rf, nch, nt, dt, uscr, (void *)y2, s);
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
