// Coalesced 64-channel x 32-sample tiles for the per-channel recurrences
// (spline prefilter and its adjoint): global memory <-> registers in whole
// cache lines, registers <-> one LDS tile, and a lane's own row of the tile.
#pragma once
#include <stdint.h>

#include "common.h"

namespace pycbf {

// 16-byte unit of a row; rows need only 4-byte alignment (gfx950 serves
// unaligned dwordx4 accesses).
struct __attribute__((aligned(4))) F4u {
    float v[4];
};


// LDS staging tile of the spline sweeps: 64 rows (channels) x 32 floats at a
// pitch of 36 floats, so that rows stay 16-byte aligned and both the
// row-per-lane and the 8-lanes-per-row float4 accesses are free of bank
// conflicts.  The one tile holds, in turn, the y tile and the u tile
// (forward), and the u tile and the output tile (back substitution).
constexpr int SPL_PITCH = 36;
typedef float SplRow[SPL_PITCH];

// A 64-row x 32-column tile moves between global memory and registers in
// whole cache lines: lane l covers row 8q + l/8, columns 4*(l%8)..4*(l%8)+3,
// so every instruction reads or writes eight complete 128-byte lines.
struct TileRegs {
    F4u q[8];
};

// Addresses are a uniform base (the workgroup's first row) plus a 32-bit byte
// offset per lane: 64 table rows of at most SPL_MAX_NT + 1 entries (8 bytes
// each) stay below 2^32 bytes.
constexpr int SPL_MAX_NT = (1 << 22) - 1;
__device__ __forceinline__ const float *at(const float *base, uint32_t elems)
{
    return reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + 4u * elems);
}
__device__ __forceinline__ float *at(float *base, uint32_t elems)
{
    return reinterpret_cast<float *>(reinterpret_cast<char *>(base) + 4u * elems);
}

__device__ __forceinline__ void tile_fetch(const float *base, int ld, int nrow, int s, int ncol,
                                           TileRegs &t)
{
    const int lane = threadIdx.x;
    const int c = 4 * (lane & 7);
    const uint32_t o = (uint32_t)((lane >> 3) * ld + s + c);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int r = 8 * q + (lane >> 3);
        F4u w{{0.0f, 0.0f, 0.0f, 0.0f}};
        if (r < nrow) {
            const float *src = at(base, o + (uint32_t)(8 * q * ld));
            if (s + c + 4 <= ncol) {
                w = *reinterpret_cast<const F4u *>(src);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (s + c + e < ncol) w.v[e] = src[e];
            }
        }
        t.q[q] = w;
    }
}

__device__ __forceinline__ void tile_to_lds(const TileRegs &t, SplRow *S)
{
    const int lane = threadIdx.x;
    const int c = 4 * (lane & 7);
#pragma unroll
    for (int q = 0; q < 8; ++q)
        *reinterpret_cast<float4 *>(&S[8 * q + (lane >> 3)][c]) =
            make_float4(t.q[q].v[0], t.q[q].v[1], t.q[q].v[2], t.q[q].v[3]);
}

__device__ __forceinline__ void lds_to_global(const SplRow *S, float *base, int ld, int nrow, int s,
                                              int ncol)
{
    const int lane = threadIdx.x;
    const int c = 4 * (lane & 7);
    const uint32_t o = (uint32_t)((lane >> 3) * ld + s + c);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int r = 8 * q + (lane >> 3);
        if (r >= nrow) continue;
        const float4 v = *reinterpret_cast<const float4 *>(&S[r][c]);
        float *dst = at(base, o + (uint32_t)(8 * q * ld));
        if (s + c + 4 <= ncol) {
            F4u w{{v.x, v.y, v.z, v.w}};
            *reinterpret_cast<F4u *>(dst) = w;
        } else {
            const float e4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (s + c + e < ncol) dst[e] = e4[e];
        }
    }
}

// lane's row of the staging tile <-> 32 registers
__device__ __forceinline__ void row_read(const SplRow *S, float (&v)[32])
{
#pragma unroll
    for (int c = 0; c < 32; c += 4) {
        const float4 w = *reinterpret_cast<const float4 *>(&S[threadIdx.x][c]);
        v[c] = w.x;
        v[c + 1] = w.y;
        v[c + 2] = w.z;
        v[c + 3] = w.w;
    }
}

__device__ __forceinline__ void row_write(SplRow *S, const float (&v)[32])
{
#pragma unroll
    for (int c = 0; c < 32; c += 4)
        *reinterpret_cast<float4 *>(&S[threadIdx.x][c]) =
            make_float4(v[c], v[c + 1], v[c + 2], v[c + 3]);
}

}  // namespace pycbf
