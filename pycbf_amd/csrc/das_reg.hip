// Single-frame gather for the natural cubic with transmit delays in registers.
//
// With one frame nothing amortises the LDS staging of k_das_lds (a window per
// channel and a barrier per staged channel serve a single use of each
// sample).  Here one wave owns a work item: 64 pixels (positions 64*tile ..
// 64*tile+63 of the processing order, one per lane) and a chunk of receive
// elements.  The lane holds the delays and apodisations of REG_TB transmits
// in registers while it sweeps the receive chunk, and reads each sample's
// table pair (y_j, e_j, y_j+1, e_j+1) with one 16-byte load through L2.
// Items are numbered tile-fastest, so the waves in flight share the table
// rows of one receive chunk.
//
// Sums follow pycbf's layouts for one frame: tx_and_rx adds each item's
// partial image to the zeroed output with float atomics; tx_only writes the
// per-receive images (every (receive element, pixel) belongs to one item).
// Sample selection, weights and the cubic evaluation are those of the
// staged gather (pycbf/cpu/pycbf.c:68-105,179-207).
#include <math.h>
#include <stdlib.h>

#include "common.h"
#include "das_dev.h"
#include "prof.h"

namespace pycbf {

namespace {

// transmits whose delays / weights a lane holds: a block of them is summed in
// registers before the output is touched.  Per-receive images read and write
// their output once per block, so they take 16 (75 x 128 x 256 x 2048: 4.36
// vs 4.96 ms at 8); the summed image adds its blocks in a register and keeps
// 8 (3.88 vs 3.98 ms at 16; 25 costs occupancy for both)
template <int SUM>
struct RegTB {
    static constexpr int value = SUM == PYCBF_SUM_TX_ONLY ? 16 : 8;
};

// natural-cubic value of the table row `row` at x (0 outside the trace); FD:
// the knot quotient by dt through div_rn (dt accepted by recip_division_exact)
template <bool FD>
__device__ __forceinline__ float cubic_row(const float *__restrict__ row, float x, const DasArgs &A)
{
    int j = 0;
    const bool in = FD ? knot_index_rn(x, A.t0, A.dt, A.rdt, A.tmax, A.nsrc, j)
                       : knot_index(x, A.t0, A.dt, A.tmax, A.nsrc, j);
    if (!in) return 0.0f;
    const float xlead = A.t0 + A.dt * (float)(j + 1);
    const float xlag = A.t0 + A.dt * (float)j;
    const float a = (xlead - x) * A.rdt;
    const float b = (x - xlag) * A.rdt;
    const float ca = __builtin_fmaf(a * a, a, -a);
    const float cb = __builtin_fmaf(b * b, b, -b);
    const Pair2 c = *reinterpret_cast<const Pair2 *>(row + 2 * j);
    float v = a * c.y0;
    v = __builtin_fmaf(b, c.y1, v);
    v = __builtin_fmaf(ca, c.z0, v);
    return __builtin_fmaf(cb, c.z1, v);
}

}  // namespace

__global__ __launch_bounds__(256) void k_fill(float *__restrict__ p, int64_t n, float v)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = v;
}

// FD: quotients by dt and c0 through div_rn (both accepted by
// recip_division_exact; the same delays and indices as the divisions)
template <int GEO, int SUM, bool FD>
__global__ __launch_bounds__(256) void k_das_reg(DasArgs A, int64_t ntiles, int rchunk)
{
    constexpr int REG_TB = RegTB<SUM>::value;
    const int lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t tile = item % ntiles;
    const int r0 = (int)(item / ntiles) * rchunk;
    if (r0 >= A.nrx) return;  // wave-uniform
    const int nrc = min(rchunk, A.nrx - r0);
    const int64_t i = tile * 64 + lane;
    const bool act = i < A.np;
    const int64_t p = act ? (A.perm ? (int64_t)A.perm[i] : i) : 0;
    // every lane loads (p is 0 past the end), then a full wait: no load of the
    // position is still counted when the receive loop's gathers are
    float pv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int d = 0; d < GEO; ++d) pv[d] = A.pvec[p * GEO + d];
    __builtin_amdgcn_s_waitcnt(0);
    const int64_t rowlen = A.ch_stride;
    float acc = 0.0f;
#pragma unroll 1
    for (int t0 = 0; t0 < A.ntx; t0 += REG_TB) {
        float ttx[REG_TB], atx[REG_TB];
#pragma unroll
        for (int k = 0; k < REG_TB; ++k) {
            ttx[k] = 0.0f;
            atx[k] = 0.0f;
            if (act && t0 + k < A.ntx)
                elem_eval<GEO, FD>(A, true, t0 + k, p, pv, ttx[k], atx[k]);
        }
#pragma unroll 1
        for (int ii = 0; ii < nrc; ++ii) {
            const int r = r0 + ii;
            float trx = 0.0f, arx = 0.0f;
            if (act) elem_eval<GEO, FD>(A, false, r, p, pv, trx, arx);
            const float *base = A.src + ((int64_t)t0 * A.nrx + r) * rowlen;
            float part = 0.0f;
#pragma unroll
            for (int k = 0; k < REG_TB; ++k) {
                if (t0 + k >= A.ntx) break;
                const float w = atx[k] * arx;          // apodtx * apodrx
                if (act && fabsf(w) >= A.thresh)
                    part = __builtin_fmaf(
                        w, cubic_row<FD>(base + (int64_t)k * A.nrx * rowlen, ttx[k] + trx, A),
                        part);
            }
            if (SUM == PYCBF_SUM_TX_AND_RX) {
                acc += part;
            } else if (act) {
                float *o = A.out + (int64_t)r * A.np + p;
                *o = (t0 == 0) ? part : *o + part;
            }
        }
    }
    if (SUM == PYCBF_SUM_TX_AND_RX && act) unsafeAtomicAdd(A.out + p, acc);
}

bool das_reg_applies(int gk, int64_t nfr, int sumflag)
{
    static int on = -1;
    if (on < 0) {
        const char *e = getenv("PYCBF_DAS_REG");
        on = e ? atoi(e) != 0 : 1;
    }
    return on && gk == GK_CUBIC && nfr == 1 &&
           (sumflag == PYCBF_SUM_TX_AND_RX || sumflag == PYCBF_SUM_TX_ONLY);
}

template <int SUM, bool FD>
static void reg_geo_fd(const DasArgs &a, dim3 grid, int64_t ntiles, int rchunk, hipStream_t s)
{
    if (a.geo == 3)
        hipLaunchKernelGGL((k_das_reg<3, SUM, FD>), grid, dim3(256), 0, s, a, ntiles, rchunk);
    else if (a.geo == 2)
        hipLaunchKernelGGL((k_das_reg<2, SUM, FD>), grid, dim3(256), 0, s, a, ntiles, rchunk);
    else
        hipLaunchKernelGGL((k_das_reg<0, SUM, FD>), grid, dim3(256), 0, s, a, ntiles, rchunk);
}

template <int SUM>
static void reg_geo(const DasArgs &a, dim3 grid, int64_t ntiles, int rchunk, hipStream_t s)
{
    const bool fd = a.rdt == 1.0f / a.dt && recip_division_exact(a.dt) &&
                    (a.geo == 0 || (a.rc0 == 1.0f / a.c0 && recip_division_exact(a.c0)));
    if (fd)
        reg_geo_fd<SUM, true>(a, grid, ntiles, rchunk, s);
    else
        reg_geo_fd<SUM, false>(a, grid, ntiles, rchunk, s);
}

int launch_das_reg(const DasArgs &a, hipStream_t s)
{
    const bool both = a.sumflag == PYCBF_SUM_TX_AND_RX;
    // receive chunk per item (measured on 75 tx x 128 rx: 16 for the summed
    // image; per-receive images with 16 transmits per lane 4.28 / 4.38 /
    // 5.05 ms at 16 / 32 / 64)
    int rchunk = 16;
    if (const char *e = getenv("PYCBF_DAS_REG_RC")) rchunk = std::max(1, atoi(e));
    const int64_t ntiles = (a.np + 63) / 64;
    const int64_t nitems = ntiles * ((a.nrx + rchunk - 1) / rchunk);
    const int64_t nblocks = (nitems + 3) / 4;
    if (nblocks > 0x7fffffffLL) return fail(PYCBF_EINVAL, "single-frame gather grid too large");
    ProfScope prof("das_gather", s);
    if (both) {
        const unsigned fb = (unsigned)std::min<int64_t>((a.np + 255) / 256, 8192);
        hipLaunchKernelGGL(k_fill, dim3(fb), dim3(256), 0, s, a.out, a.np, 0.0f);
        if (int rc = check_launch("gather output clear")) return rc;
        reg_geo<PYCBF_SUM_TX_AND_RX>(a, dim3((unsigned)nblocks), ntiles, rchunk, s);
    } else {
        reg_geo<PYCBF_SUM_TX_ONLY>(a, dim3((unsigned)nblocks), ntiles, rchunk, s);
    }
    return check_launch("das gather (single frame)");
}

}  // namespace pycbf
