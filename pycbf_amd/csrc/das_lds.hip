// LDS-staged delay-and-sum gather for gfx950.
//
// Grid: (pixel tiles of 256, frame chunks, outer channel elements).  Tiles
// are dispatched fastest, so the resident workgroups of one outer element
// sweep the inner channel loop together and each channel's samples pass
// through L2 about once (a channel wavefront).
//
// Per workgroup and channel the samples any of its 256 pixels can touch form
// one contiguous window of the channel's table.  The window comes from the
// tile's delay ranges (k_tile_bounds): the sample index is monotone in the
// delay tau_tx + tau_rx and float addition is monotone, so
//   [index(min_tx + min_rx), index(max_tx + max_rx)]
// contains every index a contributing pixel selects.  The window of channel
// b+1 is fetched into registers (coalesced, all frames of the chunk) while
// channel b is evaluated from LDS, then written to the other LDS buffer; one
// barrier per channel.  A window larger than the LDS slot falls back to
// direct global gathers for that channel (a workgroup-uniform branch).
//
// Delays and weights are read from tables (geo 0) or evaluated in registers
// from the synthetic-point geometry (geo 2 / 3), both with the float32
// operations of the reference (pycbf/gpu/__engines__.cu:382-470).
//
// The natural cubic with sums over 4 or 16 frames per pass runs
// k_das_cubic_dma, which stages the same windows with LDS DMA (below).
#include <math.h>
#include <stdlib.h>

#include <type_traits>

#include "common.h"
#include "das_dev.h"
#include "prof.h"

namespace pycbf {

constexpr int LDS_BLOCK = 256;
constexpr int LDS_BYTES = 40960;  // both staging buffers of one workgroup
constexpr int LDS_NLD = 10;       // staged entries per thread and channel

template <int GK>
struct EntryOf {
    // floats per table entry; interval j also reads entry j+1
    static constexpr int ES = (GK == GK_CUBIC) ? 2 : ((GK == GK_KORDER || GK == GK_AKIMA) ? 4 : 1);
    static constexpr int PAIR = (GK == GK_LINEAR || GK == GK_CUBIC) ? 1 : 0;
};

template <int GK, int FC>
struct Stage {
    static constexpr int ES = EntryOf<GK>::ES;
    static constexpr int PAIR = EntryOf<GK>::PAIR;
    static constexpr int TPF = LDS_BLOCK / FC;  // staging threads per frame
    static constexpr int WIN_LDS = LDS_BYTES / (2 * FC * ES * 4);
    static constexpr int WIN = (WIN_LDS < LDS_NLD * TPF) ? WIN_LDS : LDS_NLD * TPF;
    static constexpr int NLD = (WIN + TPF - 1) / TPF;
};

// ---------------------------------------------------------------------------
// delay range per (tile, element) over the pixels that can contribute; a
// workgroup takes one tile and TB_ELEMS elements (one barrier for all of them)
// ---------------------------------------------------------------------------
constexpr int TB_ELEMS = 8;

template <int GEO>
__global__ __launch_bounds__(256) void k_tile_bounds(DasArgs A)
{
    const int64_t tile = blockIdx.x;
    const int e0 = blockIdx.y * TB_ELEMS;
    const int ne = A.ntx + A.nrx;
    const int64_t i = tile * 256 + threadIdx.x;
    const bool in = i < A.np;
    const int64_t p = in ? (A.perm ? (int64_t)A.perm[i] : i) : 0;
    float pv[3] = {0.0f, 0.0f, 0.0f};
    if (in) {
#pragma unroll
        for (int d = 0; d < GEO; ++d) pv[d] = A.pvec[p * GEO + d];
    }
    __shared__ float wl[TB_ELEMS][4], wh[TB_ELEMS][4];
    const int w = threadIdx.x >> 6;
    for (int k = 0; k < TB_ELEMS && e0 + k < ne; ++k) {
        const int e = e0 + k;  // [0, ntx): tx element, [ntx, ntx+nrx): rx element
        float lo = INFINITY, hi = -INFINITY;
        if (in) {
            const bool tx = e < A.ntx;
            float tau = 0.0f, apod = 0.0f;
            elem_eval<GEO>(A, tx, tx ? e : e - A.ntx, p, pv, tau, apod);
            // with thresh > 0 a zero (or NaN) weight on either axis never contributes
            const bool used =
                (tau == tau) && (!(A.thresh > 0.0f) || (apod != 0.0f && apod == apod));
            if (used) lo = hi = tau;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            lo = fminf(lo, __shfl_xor(lo, off, 64));
            hi = fmaxf(hi, __shfl_xor(hi, off, 64));
        }
        if ((threadIdx.x & 63) == 0) {
            wl[k][w] = lo;
            wh[k][w] = hi;
        }
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k < TB_ELEMS && e0 + k < ne) {
        const float lo = fminf(fminf(wl[k][0], wl[k][1]), fminf(wl[k][2], wl[k][3]));
        const float hi = fmaxf(fmaxf(wh[k][0], wh[k][1]), fmaxf(wh[k][2], wh[k][3]));
        float *b = const_cast<float *>(A.bounds) + (tile * ne + e0 + k) * 2;
        b[0] = lo;
        b[1] = hi;
    }
}

int launch_tile_bounds(const DasArgs &a, float *bounds, hipStream_t s)
{
    DasArgs b = a;
    b.bounds = bounds;
    dim3 grid((unsigned)das_tiles(a.np), (unsigned)((a.ntx + a.nrx + TB_ELEMS - 1) / TB_ELEMS));
    ProfScope prof("tile_bounds", s);
    if (a.geo == 3)
        hipLaunchKernelGGL(k_tile_bounds<3>, grid, dim3(256), 0, s, b);
    else if (a.geo == 2)
        hipLaunchKernelGGL(k_tile_bounds<2>, grid, dim3(256), 0, s, b);
    else
        hipLaunchKernelGGL(k_tile_bounds<0>, grid, dim3(256), 0, s, b);
    return check_launch("tile bounds");
}

// ---------------------------------------------------------------------------
// window of one channel: entries [lo, lo + cnt) of its table row (cnt 0: no
// contributing pixel selects a sample of this channel)
// ---------------------------------------------------------------------------
template <int GK>
__device__ __forceinline__ void channel_window(const DasArgs &A, float xlo, float xhi, int &lo,
                                               int &cnt)
{
    lo = 0;
    cnt = 0;
    if (!(xlo <= xhi)) return;
    const int n = A.nsrc;
    if (GK == GK_NEAREST) {
        const double dlo = (double)((xlo - A.t0) / A.dt) + 0.5;
        const double dhi = (double)((xhi - A.t0) / A.dt) + 0.5;
        if (!(dhi > -1.0) || !(dlo < (double)n)) return;
        const int ilo = (dlo > -1.0) ? (int)dlo : 0;
        const int ihi = (dhi < (double)n) ? (int)dhi : n - 1;
        lo = ilo;
        cnt = ihi - ilo + 1;
    } else {
        if (xhi < A.t0 || xlo > A.tmax) return;
        const int jlo = (xlo < A.t0) ? 0 : min((int)((xlo - A.t0) / A.dt), n - 1);
        const int jhi = (xhi > A.tmax) ? n - 1 : min((int)((xhi - A.t0) / A.dt), n - 1);
        lo = jlo;
        cnt = jhi - jlo + 1 + EntryOf<GK>::PAIR;
    }
}

template <int ES>
struct Ent {
    float v[ES];
};

template <int GK, int SUM, int FC, int GEO>
__global__ __launch_bounds__(LDS_BLOCK) void k_das_lds(DasArgs A)
{
    using St = Stage<GK, FC>;
    constexpr int ES = St::ES, WIN = St::WIN, TPF = St::TPF, NLD = St::NLD;
    __shared__ __align__(16) float lds[2 * FC * WIN * ES];

    uint32_t bx, by, bz;
    block_coords(A.xcd_remap, bx, by, bz);
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)bx * LDS_BLOCK + tid;
    const bool live = i < A.np;
    const int64_t p = live ? (A.perm ? (int64_t)A.perm[i] : i) : 0;
    const int f0 = by * FC;
    const int nf = min(FC, A.nfr - f0);
    const int64_t np = A.np;

    const bool OUTER_TX = (SUM == PYCBF_SUM_RX_ONLY) || (SUM != PYCBF_SUM_TX_ONLY && A.outer_tx);
    const int na = OUTER_TX ? A.ntx : A.nrx;
    const int nb = OUTER_TX ? A.nrx : A.ntx;
    const int a0 = bz * A.a_split;
    const int a1 = min(na, a0 + A.a_split);
    const int64_t nchan = (int64_t)A.ntx * A.nrx;
    const int64_t fstride = nchan * A.ch_stride;
    const float *bnd = A.bounds + (int64_t)bx * (A.ntx + A.nrx) * 2;
    const int boff_a = OUTER_TX ? 0 : A.ntx;  // bounds index offset of the outer axis
    const int boff_b = OUTER_TX ? A.ntx : 0;

    float pv[3] = {0.0f, 0.0f, 0.0f};
    if (GEO > 0 && live) {
#pragma unroll
        for (int d = 0; d < GEO; ++d) pv[d] = A.pvec[p * GEO + d];
    }
    const int sfr = tid / TPF;  // staging role: frame and first entry
    const int se0 = tid % TPF;

    float acc[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) acc[f] = 0.0f;

    for (int a = a0; a < a1; ++a) {
        float ta = 0.0f, wa = 0.0f;
        if (live) elem_eval<GEO>(A, OUTER_TX, a, p, pv, ta, wa);
        const float amin = bnd[2 * (boff_a + a)], amax = bnd[2 * (boff_a + a) + 1];

        Ent<ES> reg[NLD];
        auto window = [&](int b, int &lo, int &cnt) {
            const float bmin = bnd[2 * (boff_b + b)], bmax = bnd[2 * (boff_b + b) + 1];
            channel_window<GK>(A, amin + bmin, amax + bmax, lo, cnt);
        };
        auto fetch = [&](int b, int lo, int cnt) {
            if (cnt == 0 || cnt > WIN || sfr >= nf) return;
            const int tx = OUTER_TX ? a : b, rx = OUTER_TX ? b : a;
            const float *base =
                A.src + ((int64_t)(f0 + sfr) * nchan + (int64_t)tx * A.nrx + rx) * A.ch_stride;
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const int e = se0 + k * TPF;
                if (e < cnt) {
                    const int g = lo + e;
                    if (g < A.nent) {
                        reg[k] = *reinterpret_cast<const Ent<ES> *>(base + (int64_t)g * ES);
                    } else {
#pragma unroll
                        for (int c = 0; c < ES; ++c) reg[k].v[c] = 0.0f;
                    }
                }
            }
        };
        auto store = [&](int buf, int cnt) {
            if (cnt == 0 || cnt > WIN || sfr >= nf) return;
            float *dst = lds + ((buf * FC + sfr) * WIN) * ES;
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const int e = se0 + k * TPF;
                if (e < cnt) *reinterpret_cast<Ent<ES> *>(dst + e * ES) = reg[k];
            }
        };

        int lo_c, cnt_c;
        window(0, lo_c, cnt_c);
        fetch(0, lo_c, cnt_c);
        store(0, cnt_c);
        __syncthreads();
        for (int b = 0; b < nb; ++b) {
            const int buf = b & 1;
            int lo_n = 0, cnt_n = 0;
            if (b + 1 < nb) {
                window(b + 1, lo_n, cnt_n);
                fetch(b + 1, lo_n, cnt_n);
            }
            const int tx = OUTER_TX ? a : b, rx = OUTER_TX ? b : a;
            if (cnt_c > 0) {
                float tb = 0.0f, wb = 0.0f;
                if (live) elem_eval<GEO>(A, !OUTER_TX, b, p, pv, tb, wb);
                const float w = OUTER_TX ? wa * wb : wb * wa;  // apodtx * apodrx
                const float x = OUTER_TX ? ta + tb : tb + ta;  // tautx + taurx
                const bool use = live && cnt_c > 0 && fabsf(w) >= A.thresh;
                float v[FC];
                bool ok = false;
                if (use && cnt_c > WIN) {
                    // window too large for the LDS slot: direct global gather
                    gather<GK, FC>(A, (int64_t)f0 * nchan + (int64_t)tx * A.nrx + rx, fstride, nf, x,
                                   ok, v);
                } else if (use) {
                    const float *L = lds + (buf * FC * WIN) * ES;
                    if (GK == GK_NEAREST) {
                        int it = 0;
                        ok = nearest_index(x, A.t0, A.dt, A.nsrc, it);
                        if (ok) {
                            const int e = it - lo_c;
    #pragma unroll
                            for (int f = 0; f < FC; ++f)
                                if (f < nf) v[f] = L[f * WIN + e];
                        }
                    } else if (GK == GK_LINEAR) {
                        int j = 0;
                        ok = knot_index(x, A.t0, A.dt, A.tmax, A.nsrc, j);
                        if (ok) {
                            const float xlead = A.t0 + A.dt * (float)(j + 1);
                            const float xlag = A.t0 + A.dt * (float)j;
                            const float ca = (xlead - x) / A.dt;
                            const float cb = (x - xlag) / A.dt;
                            const int e = j - lo_c;
    #pragma unroll
                            for (int f = 0; f < FC; ++f)
                                if (f < nf) v[f] = ca * L[f * WIN + e] + cb * L[f * WIN + e + 1];
                        }
                    } else if (GK == GK_CUBIC) {
                        int j = 0;
                        ok = knot_index(x, A.t0, A.dt, A.tmax, A.nsrc, j);
                        if (ok) {
                            const float xlead = A.t0 + A.dt * (float)(j + 1);
                            const float xlag = A.t0 + A.dt * (float)j;
                            float ka = (xlead - x) * A.rdt;
                            float kb = (x - xlag) * A.rdt;
                            float kc = __builtin_fmaf(ka * ka, ka, -ka);
                            float kd = __builtin_fmaf(kb * kb, kb, -kb);
                            const float *E = L + (j - lo_c) * 2;
                            if (SUM == PYCBF_SUM_NONE) {
    #pragma unroll
                                for (int f = 0; f < FC; ++f)
                                    if (f < nf) {
                                        const float *Ef = E + f * WIN * 2;
                                        float r = ka * Ef[0];
                                        r = __builtin_fmaf(kb, Ef[2], r);
                                        r = __builtin_fmaf(kc, Ef[1], r);
                                        v[f] = __builtin_fmaf(kd, Ef[3], r);
                                    }
                            } else {
                                // fold the weight into the four coefficients: four
                                // fused multiply-adds per frame straight into acc
                                ka *= w;
                                kb *= w;
                                kc *= w;
                                kd *= w;
                                // every frame slot, present or not: the slots of
                                // frames past nf lie in the buffer and feed
                                // sums that are never stored, and straight-line
                                // code lets the LDS reads of all frames overlap
    #pragma unroll
                                for (int f = 0; f < FC; ++f) {
                                    const float *Ef = E + f * WIN * 2;
                                    float r = __builtin_fmaf(ka, Ef[0], acc[f]);
                                    r = __builtin_fmaf(kb, Ef[2], r);
                                    r = __builtin_fmaf(kc, Ef[1], r);
                                    acc[f] = __builtin_fmaf(kd, Ef[3], r);
                                }
                                ok = false;  // already accumulated
                            }
                        }
                    } else {
                        int j = 0;
                        ok = knot_index(x, A.t0, A.dt, A.tmax, A.nsrc, j);
                        if (ok) {
                            const float t = (GK == GK_KORDER) ? ((x - A.t0) - A.dt * (float)j) * A.rdt
                                                              : (x - (A.t0 + A.dt * (float)j)) * A.rdt;
                            const float *E = L + (j - lo_c) * 4;
    #pragma unroll
                            for (int f = 0; f < FC; ++f)
                                if (f < nf) {
                                    const float4 c = *reinterpret_cast<const float4 *>(E + f * WIN * 4);
                                    v[f] = __builtin_fmaf(
                                        __builtin_fmaf(__builtin_fmaf(c.w, t, c.z), t, c.y), t, c.x);
                                }
                        }
                    }
                }
                if (SUM == PYCBF_SUM_NONE) {
                    float *o = A.out + ((((int64_t)f0 * A.ntx + tx) * A.nrx + rx) * np + p);
                    if (live) {
    #pragma unroll
                        for (int f = 0; f < FC; ++f)
                            if (f < nf) o[(int64_t)f * nchan * np] = ok ? 0.0f + w * v[f] : 0.0f;
                    }
                } else if (ok) {
    #pragma unroll
                    for (int f = 0; f < FC; ++f)
                        if (f < nf) acc[f] = __builtin_fmaf(w, v[f], acc[f]);
                }
            } else if (SUM == PYCBF_SUM_NONE && live) {
                // no pixel of this tile takes a sample of this channel
                float *o = A.out + ((((int64_t)f0 * A.ntx + tx) * A.nrx + rx) * np + p);
#pragma unroll
                for (int f = 0; f < FC; ++f)
                    if (f < nf) o[(int64_t)f * nchan * np] = 0.0f;
            }
            if (b + 1 < nb) store(buf ^ 1, cnt_n);
            // workgroup-uniform: a barrier is needed only when this channel was
            // read from LDS (its buffer is refilled next) or the next was staged
            const bool staged_c = cnt_c > 0 && cnt_c <= WIN;
            const bool staged_n = cnt_n > 0 && cnt_n <= WIN;
            if (staged_c || staged_n) __syncthreads();
            lo_c = lo_n;
            cnt_c = cnt_n;
        }
        if (SUM == PYCBF_SUM_TX_ONLY || SUM == PYCBF_SUM_RX_ONLY) {
#pragma unroll
            for (int f = 0; f < FC; ++f) {
                if (live && f < nf) A.out[((int64_t)(f0 + f) * na + a) * np + p] = acc[f];
                acc[f] = 0.0f;
            }
        }
    }
    if (SUM == PYCBF_SUM_TX_AND_RX && live) {
        float *o = A.out + (int64_t)bz * A.nfr * np;
#pragma unroll
        for (int f = 0; f < FC; ++f)
            if (f < nf) o[(int64_t)(f0 + f) * np + p] = acc[f];
    }
}

// ---------------------------------------------------------------------------
// natural-cubic gather over FC = 4 or 16 frames with sums (tx_and_rx, tx_only,
// rx_only), staged by LDS DMA.  The windows come to LDS through gfx950's
// global_load_lds_dwordx4: every lane moves one 16-byte chunk (two table
// entries) straight into LDS at base + 16 * lane, without VGPRs and without
// LDS writes by the waves.  A frame row of the window starts at the 16-byte
// aligned entry at or just below its first entry (rows are 8-byte aligned; the
// parity of row * (nt + 1) + lo says which), and each frame owns CH chunk
// slots, so chunk q of the staged block lands at byte 16 q.  Per outer
// element one pass computes every inner channel's window into an LDS table;
// channel b + 1 is in flight while b is evaluated (two LDS buffers, one
// barrier per staged channel).  Nothing in the evaluation of channel b waits
// on the DMA of channel b + 1: the buffers are separate LDS objects (a read
// of one is known not to alias the DMA into the other), geometry packs come
// through the scalar cache and table entries are fetched a channel ahead.
// Windows wider than the slot take direct global gathers, in a loop after
// the staged channels, as in k_das_lds.  The quotients by c0 (geometry) and dt
// (knot index) are products with the rounded reciprocals plus one fused
// correction (div_rn), which the launch takes only for divisors whose every
// quotient it rounds like the division (recip_division_exact): the same
// delays and indices with three VALU instructions instead of a division.
// ---------------------------------------------------------------------------
constexpr int DMA_MAXB = 256;    // inner channels per outer element (window table)
constexpr int DMA_CHUNKS = 768;  // 16-byte chunks per buffer (12 KB), all frames

// one (y, h*y2) table entry as read from a staging buffer
struct __attribute__((aligned(8))) YZ {
    float y, z;
};

typedef __attribute__((address_space(3))) void lds_void;
typedef const __attribute__((address_space(1))) void glb_void;

// FODD: a frame's rows span an odd number of table entries, so the rows of
// consecutive frames start on alternating 8-byte parities
template <int SUM, int FC, int GEO, bool OUTER_TX, bool FODD>
__global__ __launch_bounds__(LDS_BLOCK) void k_das_cubic_dma(DasArgs A)
{
    constexpr int CH = DMA_CHUNKS / FC;     // chunk slots per frame row: 48 (FC 16) or 192
    constexpr int WINMAX = 2 * CH - 1;      // entries a frame row can always hold
    constexpr int NI = DMA_CHUNKS / LDS_BLOCK;  // DMA instructions per wave and channel (3)
    // the two staging buffers are separate objects, so the compiler can tell a
    // read of one from the DMA into the other and does not hold the read for
    // that DMA's count
    __shared__ __align__(16) float lds0[DMA_CHUNKS * 4];
    __shared__ __align__(16) float lds1[DMA_CHUNKS * 4];
    __shared__ int2 wins[DMA_MAXB];

    uint32_t bx, by, bz;
    block_coords(A.xcd_remap, bx, by, bz);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t i = (int64_t)bx * LDS_BLOCK + tid;
    const bool live = i < A.np;
    const int64_t p = live ? (A.perm ? (int64_t)A.perm[i] : i) : 0;
    const int f0 = by * FC;
    const int nf = min(FC, A.nfr - f0);
    const int64_t np = A.np;

    static_assert(SUM != PYCBF_SUM_RX_ONLY || OUTER_TX, "rx_only sums over the inner receives");
    static_assert(SUM != PYCBF_SUM_TX_ONLY || !OUTER_TX, "tx_only sums over the inner transmits");
    const int na = OUTER_TX ? A.ntx : A.nrx;
    const int nb = OUTER_TX ? A.nrx : A.ntx;
    const int a0 = bz * A.a_split;
    const int a1 = min(na, a0 + A.a_split);
    const int64_t nchan = (int64_t)A.ntx * A.nrx;
    const int64_t rowlen = A.ch_stride / 2;  // table entries per channel row (nt + 1)
    const int64_t fstride = nchan * A.ch_stride;
    const float *bnd = A.bounds + (int64_t)bx * (A.ntx + A.nrx) * 2;
    const int boff_a = OUTER_TX ? 0 : A.ntx;
    const int boff_b = OUTER_TX ? A.ntx : 0;
    // A ds_read2_b64 reaches 255 8-byte slots past its base: FG frame rows.
    // The frames' reads take one base per group of FG rows, stepped by
    // fg_step, a value the compiler cannot see (the empty asm), so it adds
    // it once per group; with a known constant it recomputes the address of
    // every row that its offsets cannot reach.
    constexpr int FG = (2032 / (CH * 16) + 1 < FC) ? 2032 / (CH * 16) + 1 : FC;
    int fg_step = FG * CH * 2;  // in (y, h*y2) entries
    asm("" : "+v"(fg_step));

    // pixel position (every lane: p is 0 past the end), and a full wait, so no
    // load of the prologue is still counted when the channel loop's DMAs are
    float pv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int d = 0; d < GEO; ++d) pv[d] = A.pvec[p * GEO + d];
    __builtin_amdgcn_s_waitcnt(0);
    // DMA role of this lane: chunk q = (k * 4 + wave) * 64 + lane of the block,
    // slot dma_c of frame row q / CH; dma_src is that frame's table (channel
    // 0, entry 0) and dma_par its rows' parity relative to frame 0's (FODD);
    // chunks of absent frames reload frame nf - 1
    int dma_c[NI], dma_par[NI];
    const float *dma_src[NI];
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        const int q = (k * (LDS_BLOCK / 64) + wave) * 64 + lane;
        const int f = min(q / CH, nf - 1);
        dma_c[k] = q % CH;
        dma_par[k] = FODD ? (f & 1) : 0;
        dma_src[k] = A.src + (int64_t)(f0 + f) * nchan * rowlen * 2;
    }

    // an inner element's delay and weight at this pixel.  Tables (GEO 0): the
    // entries are per-pixel vector loads, fetched one item ahead (inner_fetch)
    // so that no wait inside an item holds its evaluation behind the next
    // item's DMA (both count in vmcnt).  Geometry: the element's pack comes in
    // through the scalar cache (read-only for the kernel's lifetime), whose
    // loads wait on lgkmcnt, not on the DMA.
    constexpr int NIN = 2;
    auto inner_fetch = [&](int b, float(&d)[NIN]) {
        if (GEO == 0) {
            const int64_t o = (int64_t)b * A.np + p;
            d[0] = OUTER_TX ? A.trx[o] : A.ttx[o];
            d[1] = OUTER_TX ? A.arx[o] : A.atx[o];
        }
    };
    auto inner_eval = [&](int b, const float(&d)[NIN], float &tau, float &apod) {
        if (GEO == 0) {
            tau = d[0];
            apod = d[1];
        } else {
            elem_eval<GEO, true>(A, !OUTER_TX, b, p, pv, tau, apod);
        }
    };

    float acc[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) acc[f] = 0.0f;

    for (int a = a0; a < a1; ++a) {
        float ta = 0.0f, wa = 0.0f;
        if (live) elem_eval<GEO, true>(A, OUTER_TX, a, p, pv, ta, wa);
        const float amin = bnd[2 * (boff_a + a)], amax = bnd[2 * (boff_a + a) + 1];
        for (int b = tid; b < nb; b += LDS_BLOCK) {
            int lo, cnt;
            channel_window<GK_CUBIC>(A, amin + bnd[2 * (boff_b + b)],
                                     amax + bnd[2 * (boff_b + b) + 1], lo, cnt);
            wins[b] = make_int2(lo, cnt);
        }
        __syncthreads();

        // table row of frame f (of this pass) for inner channel b
        auto row_of = [&](int b, int f) {
            const int tx = OUTER_TX ? a : b, rx = OUTER_TX ? b : a;
            return ((int64_t)(f0 + f) * nchan + (int64_t)tx * A.nrx + rx);
        };
        auto staged = [&](int b) {
            if (b >= nb) return false;
            const int2 w = wins[b];
            return w.y > 0 && w.y <= WINMAX;
        };
        // issue the DMA of channel b into buffer buf (workgroup-uniform condition)
        auto dma = [&](int b, auto bufc) {
            constexpr int buf = decltype(bufc)::value;
            if (!staged(b)) return;
            const int2 w = wins[b];
            // the window is the same in every lane (one LDS word): scalar
            // registers, so the channel's offset is scalar arithmetic
            const int wx = __builtin_amdgcn_readfirstlane(w.x);
            const int wy = __builtin_amdgcn_readfirstlane(w.y);
            // entry wx of the channel's row, offset from the frame's table and
            // its parity in frame 0 (workgroup-uniform)
            const int64_t cho = row_of(b, -f0) * rowlen + wx;
            const int cpar = (int)(cho & 1);
#pragma unroll
            for (int k = 0; k < NI; ++k) {
                // the row's first staged entry is wx - par, 16-byte aligned
                // (row r-1's last if -1); its nch = (wy + par + 1) / 2 chunks
                const int par = FODD ? (dma_par[k] ^ cpar) : cpar;
                const uint32_t c = (uint32_t)min(dma_c[k], (wy + par - 1) >> 1);
                const float *src = dma_src[k] + (cho - par) * 2 + 4 * c;
                float *dst = (buf ? lds1 : lds0) + (k * (LDS_BLOCK / 64) + wave) * 64 * 4;
                __builtin_amdgcn_global_load_lds((glb_void *)src, (lds_void *)dst, 16, 0, 0);
            }
        };
        // delay tautx + taurx and weight apodtx * apodrx of channel b at this
        // pixel from the inner element's inputs d
        auto delay = [&](int b, const float(&d)[NIN], float &x, float &wgt) {
            float tb = 0.0f, wb = 0.0f;
            if (live) inner_eval(b, d, tb, wb);
            wgt = OUTER_TX ? wa * wb : wb * wa;
            x = OUTER_TX ? ta + tb : tb + ta;
        };
        // a staged channel (window of 1 .. WINMAX entries) from LDS buffer buf
        auto compute = [&](int b, auto bufc, const float(&d)[NIN]) {
            constexpr int buf = decltype(bufc)::value;
            const int2 w = wins[b];
            if (!(w.y > 0 && w.y <= WINMAX)) return;
            float x, wgt;
            delay(b, d, x, wgt);
            int j = 0;
            if (!(live && fabsf(wgt) >= A.thresh &&
                  knot_index_rn(x, A.t0, A.dt, A.rdt, A.tmax, A.nsrc, j)))
                return;
            const float xlead = A.t0 + A.dt * (float)(j + 1);
            const float xlag = A.t0 + A.dt * (float)j;
            float ka = (xlead - x) * A.rdt;
            float kb = (x - xlag) * A.rdt;
            float kc = __builtin_fmaf(ka * ka, ka, -ka);
            float kd = __builtin_fmaf(kb * kb, kb, -kb);
            ka *= wgt;
            kb *= wgt;
            kc *= wgt;
            kd *= wgt;
            // row parity of frame f: that of frame 0, flipped every frame when a
            // frame's rows span an odd number of entries
            const int par0 = (int)((row_of(b, 0) * rowlen + w.x) & 1);
            // every frame slot: those of frames past nf hold copies of frame
            // nf - 1 and feed sums that are never stored; straight-line code
            // lets the LDS reads of all frames overlap
            const YZ *L = (const YZ *)(buf ? lds1 : lds0) + (j - w.x);
#pragma unroll
            for (int g = 0; g < FC; g += FG) {
                const YZ *Lg = L + (g / FG) * fg_step;
#pragma unroll
                for (int k = 0; k < FG && g + k < FC; ++k) {
                    const int f = g + k;
                    const int par = FODD ? par0 ^ (f & 1) : par0;
                    const YZ *Ef = Lg + k * CH * 2 + par;
                    const YZ e0 = Ef[0], e1 = Ef[1];
                    float r = __builtin_fmaf(ka, e0.y, acc[f]);
                    r = __builtin_fmaf(kb, e1.y, r);
                    r = __builtin_fmaf(kc, e0.z, r);
                    acc[f] = __builtin_fmaf(kd, e1.z, r);
                }
            }
        };

        bool wide = false;  // a channel's window exceeds the slot (workgroup-uniform)
        // channel b: issue the DMA of channel b + 1 and the loads of its inner
        // element's inputs (into dn, when some pixel of the tile takes a
        // sample of it), evaluate channel b from LDS with inputs dc; the wait
        // that ends the step covers both, so no wait inside a step holds its
        // evaluation behind the next channel's DMA
        auto step = [&](int b, auto bufc, const float(&dc)[NIN], float(&dn)[NIN]) {
            constexpr int buf = decltype(bufc)::value;  // b & 1
            dma(b + 1, std::integral_constant<int, buf ^ 1>());
            if (b + 1 < nb && wins[b + 1].y > 0) inner_fetch(b + 1, dn);
            wide |= wins[b].y > WINMAX;
            compute(b, bufc, dc);
            // this wave's DMA of channel b + 1 has landed; the barrier makes
            // every wave's land, and frees buffer buf for channel b + 2
            __builtin_amdgcn_s_waitcnt(0);
            if (staged(b) || staged(b + 1)) __syncthreads();
        };
        float d0[NIN], d1[NIN];
        if (wins[0].y > 0) inner_fetch(0, d0);
        dma(0, std::integral_constant<int, 0>());
        __builtin_amdgcn_s_waitcnt(0);
        if (staged(0)) __syncthreads();
        for (int b = 0; b < nb; b += 2) {
            step(b, std::integral_constant<int, 0>(), d0, d1);
            if (b + 1 < nb) step(b + 1, std::integral_constant<int, 1>(), d1, d0);
        }
        // channels whose window exceeds a staging slot: direct gathers, after
        // the staged ones (a loop of its own keeps the staged loop's registers)
        for (int b = 0; wide && b < nb; ++b) {
            if (wins[b].y <= WINMAX) continue;
            float x, wgt;
            inner_fetch(b, d0);
            delay(b, d0, x, wgt);
            if (!(live && fabsf(wgt) >= A.thresh)) continue;
            const int tx = OUTER_TX ? a : b, rx = OUTER_TX ? b : a;
            float v[FC];
            bool ok = false;
            gather<GK_CUBIC, FC>(A, (int64_t)f0 * nchan + (int64_t)tx * A.nrx + rx, fstride, nf,
                                 x, ok, v);
            if (ok) {
#pragma unroll
                for (int f = 0; f < FC; ++f)
                    if (f < nf) acc[f] = __builtin_fmaf(wgt, v[f], acc[f]);
            }
        }
        if (SUM == PYCBF_SUM_TX_ONLY || SUM == PYCBF_SUM_RX_ONLY) {
#pragma unroll
            for (int f = 0; f < FC; ++f) {
                if (live && f < nf) A.out[((int64_t)(f0 + f) * na + a) * np + p] = acc[f];
                acc[f] = 0.0f;
            }
        }
        __syncthreads();  // the window table is rewritten for the next outer element
    }
    if (SUM == PYCBF_SUM_TX_AND_RX && live) {
        float *o = A.out + (int64_t)bz * A.nfr * np;
#pragma unroll
        for (int f = 0; f < FC; ++f)
            if (f < nf) o[(int64_t)(f0 + f) * np + p] = acc[f];
    }
}

template <int SUM, int FC, bool OT, bool FODD>
static void cubic_dma_launch(int geo, const DasArgs &a, dim3 grid, hipStream_t s)
{
    if (geo == 3)
        hipLaunchKernelGGL((k_das_cubic_dma<SUM, FC, 3, OT, FODD>), grid, dim3(LDS_BLOCK), 0, s, a);
    else if (geo == 2)
        hipLaunchKernelGGL((k_das_cubic_dma<SUM, FC, 2, OT, FODD>), grid, dim3(LDS_BLOCK), 0, s, a);
    else
        hipLaunchKernelGGL((k_das_cubic_dma<SUM, FC, 0, OT, FODD>), grid, dim3(LDS_BLOCK), 0, s, a);
}

template <int SUM, int FC, bool OT>
static void cubic_dma_geo(int geo, const DasArgs &a, dim3 grid, hipStream_t s)
{
    const int64_t frame_entries = (int64_t)a.ntx * a.nrx * (a.ch_stride / 2);
    if (frame_entries & 1)
        cubic_dma_launch<SUM, FC, OT, true>(geo, a, grid, s);
    else
        cubic_dma_launch<SUM, FC, OT, false>(geo, a, grid, s);
}

template <int FC>
static void cubic_dma_sum(const DasArgs &a, dim3 grid, hipStream_t s)
{
    switch (a.sumflag) {
    case PYCBF_SUM_TX_ONLY: cubic_dma_geo<PYCBF_SUM_TX_ONLY, FC, false>(a.geo, a, grid, s); break;
    case PYCBF_SUM_RX_ONLY: cubic_dma_geo<PYCBF_SUM_RX_ONLY, FC, true>(a.geo, a, grid, s); break;
    default:
        if (a.outer_tx)
            cubic_dma_geo<PYCBF_SUM_TX_AND_RX, FC, true>(a.geo, a, grid, s);
        else
            cubic_dma_geo<PYCBF_SUM_TX_AND_RX, FC, false>(a.geo, a, grid, s);
        break;
    }
}

template <int GK, int SUM, int FC>
static void lds_geo(int geo, const DasArgs &a, dim3 grid, hipStream_t s)
{
    if (geo == 3)
        hipLaunchKernelGGL((k_das_lds<GK, SUM, FC, 3>), grid, dim3(LDS_BLOCK), 0, s, a);
    else if (geo == 2)
        hipLaunchKernelGGL((k_das_lds<GK, SUM, FC, 2>), grid, dim3(LDS_BLOCK), 0, s, a);
    else
        hipLaunchKernelGGL((k_das_lds<GK, SUM, FC, 0>), grid, dim3(LDS_BLOCK), 0, s, a);
}

template <int GK, int SUM>
static void lds_fc(int fc, const DasArgs &a, dim3 grid, hipStream_t s)
{
    if (fc >= 16)
        lds_geo<GK, SUM, 16>(a.geo, a, grid, s);
    else if (fc >= 4)
        lds_geo<GK, SUM, 4>(a.geo, a, grid, s);
    else
        lds_geo<GK, SUM, 1>(a.geo, a, grid, s);
}

template <int GK>
static void lds_sum(int fc, const DasArgs &a, dim3 grid, hipStream_t s)
{
    switch (a.sumflag) {
    case PYCBF_SUM_NONE: lds_fc<GK, PYCBF_SUM_NONE>(fc, a, grid, s); break;
    case PYCBF_SUM_TX_ONLY: lds_fc<GK, PYCBF_SUM_TX_ONLY>(fc, a, grid, s); break;
    case PYCBF_SUM_RX_ONLY: lds_fc<GK, PYCBF_SUM_RX_ONLY>(fc, a, grid, s); break;
    default: lds_fc<GK, PYCBF_SUM_TX_AND_RX>(fc, a, grid, s); break;
    }
}

// the natural-cubic DMA-staged gather (PYCBF_DAS_DMA=0 selects k_das_lds)
static bool das_cubic_dma_enabled()
{
    const char *e = getenv("PYCBF_DAS_DMA");
    return !(e && atoi(e) == 0);
}

// frames per workgroup for the staged gather: 16, 4 or 1
int das_lds_frames_per_pass(int64_t nfr)
{
    int cap = 16;
    if (const char *e = getenv("PYCBF_DAS_FC")) cap = atoi(e);
    if (nfr > 8 && cap >= 16) return 16;
    if (nfr > 1 && cap >= 4) return 4;
    return 1;
}

int launch_das_lds(int gk, const DasArgs &a0, int nsplit, float *partial, hipStream_t s)
{
    DasArgs a = a0;
    const int fc = das_lds_frames_per_pass(a.nfr);
    // XCD slabs: each XCD takes a contiguous run of (outer element, tile)
    // work items, so a channel's table rows are pulled into one or two XCDs'
    // L2 instead of all eight (16-frame headline: 13.3 -> 12.5 ms staged,
    // 8.8 -> 8.4 ms DMA-staged); PYCBF_XCD_REMAP=0 restores round-robin
    a.xcd_remap = 1;
    if (const char *e = getenv("PYCBF_XCD_REMAP")) a.xcd_remap = atoi(e);
    a.outer_tx = das_outer_is_tx(a.sumflag, a.ntx, a.nrx) ? 1 : 0;
    const int na = a.outer_tx ? a.ntx : a.nrx;
    const int nb = a.outer_tx ? a.nrx : a.ntx;
    const bool dma = gk == GK_CUBIC && a.sumflag != PYCBF_SUM_NONE && (fc == 4 || fc == 16) &&
                     (a.sumflag == PYCBF_SUM_TX_AND_RX ? nb : (a.sumflag == PYCBF_SUM_TX_ONLY
                                                                   ? a.ntx : a.nrx)) <= DMA_MAXB &&
                     das_cubic_dma_enabled() &&
                     // its quotients by dt and c0 are products with the reciprocals
                     a.rdt == 1.0f / a.dt && recip_division_exact(a.dt) &&
                     (a.geo == 0 || (a.rc0 == 1.0f / a.c0 && recip_division_exact(a.c0)));
    if (nsplit < 1) nsplit = 1;
    if (nsplit > na) nsplit = na;
    a.a_split = (na + nsplit - 1) / nsplit;
    nsplit = (na + a.a_split - 1) / a.a_split;
    const bool reduce = (a.sumflag == PYCBF_SUM_TX_AND_RX) && nsplit > 1;
    float *final_out = a.out;
    if (reduce) a.out = partial;
    dim3 grid((unsigned)das_tiles(a.np), (unsigned)((a.nfr + fc - 1) / fc), (unsigned)nsplit);
    if (dma) {
        ProfScope prof("das_gather", s);
        if (fc == 16)
            cubic_dma_sum<16>(a, grid, s);
        else
            cubic_dma_sum<4>(a, grid, s);
    } else {
        ProfScope prof("das_gather", s);
        switch (gk) {
        case GK_NEAREST: lds_sum<GK_NEAREST>(fc, a, grid, s); break;
        case GK_LINEAR: lds_sum<GK_LINEAR>(fc, a, grid, s); break;
        case GK_CUBIC: lds_sum<GK_CUBIC>(fc, a, grid, s); break;
        case GK_KORDER: lds_sum<GK_KORDER>(fc, a, grid, s); break;
        default: lds_sum<GK_AKIMA>(fc, a, grid, s); break;
        }
    }
    int rc = check_launch("das gather (lds)");
    if (rc || !reduce) return rc;
    return launch_reduce_split(partial, nsplit, (int64_t)a.nfr * a.np, final_out, s);
}

}  // namespace pycbf
