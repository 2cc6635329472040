// Short-lag spatial coherence (SLSC) of per-receive-channel images.
//
// Semantics: calc_lagn_coherence and the pixel loop of
// examples/advanced_beamforming.ipynb:65-72,205-223.  For output pixel
// (ix, iz) = (ixo*dsfx, izo*dsfz) the axial window is
// [iz + kofs, iz + kofs + kwin - 1] clipped to the image; an element is
// selected when its apodisation is non-zero over the whole window; with fewer
// than max(lag)+1 selected elements the output is 0.  Per lag m the
// coherence is the mean, over pairs of selected elements m apart in the
// selected list, of the window cross-covariance over the product of the
// population standard deviations; a deviation below 1e-10 drops the pairs of
// that element (numpy nanmean), and a lag with no pair left gives NaN.  The
// output is the mean over lags.
//
// Both kernels normalise each selected channel once per pixel,
//   z[p][w] = (x[p][w] - mean_p) / (std_p * sqrt(K)),
// so that the lag-m coherence is S_m / C_m with S_m = sum_p <z_p, z_p+m> and
// C_m the number of pairs with two valid channels (invalid channels get
// z = 0).  Normalisation, pair counts and the lag average are FP32.
//
//   PREC 0  S_m on the VALU in FP32 (one lane per channel, wave reductions).
//   PREC 1  z in FP16, S_m from 32x32 Gram blocks on MFMA (FP32 accumulate).
//   PREC 2  z in BF16, same MFMA scheme.
//
// MFMA scheme: with the selected channels in blocks of 32, every pair (p, p+m)
// with m < 32 lies in a diagonal block (b, b) or its right neighbour
// (b, b+1); lags 32..63 use (b, b+1) and (b, b+2).  v_mfma_f32_32x32x8
// leaves G[i][j] of a block in lane (j + 32*h), register r, with
// i = 8*(r/4) + 4*h + r%4.  Taking G0[i][j] if j >= i else G1[i][j] and
// rotating each register row left by i lanes (ds_bpermute) lands every entry
// of lag 32q + m' in lane m', so one add per register accumulates all lags at
// once; no per-pair value leaves the registers.  The extraction is linear and
// the same for every block, so the blocks (b, b+q) and (b, b+q+1) of all b are
// first summed in the MFMA accumulators and extracted once per lag group.
#include <math.h>
#include <stdlib.h>

#include <type_traits>

#include "common.h"
#include "prof.h"

namespace pycbf {

namespace {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// value of lane ^ 1 / lane ^ 2 (DPP quad permutations)
__device__ __forceinline__ float dpp_xor1(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_xor2(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
}

// sum over the wave, in every lane: DPP within each row of 16 lanes
// (quad swaps, half-row and row mirrors), then across the four rows
__device__ __forceinline__ float wave_sum(float v)
{
    v += dpp_xor1(v);
    v += dpp_xor2(v);
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, true));
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, true));
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// Per-lane constants of the band extraction of a 32x32 MFMA accumulator
// block: lane (j + 32h) holds G[i][j] in register r, i = 8(r/4) + 4h + r%4.
// Entry r of a lane is taken from G0 when j >= i (tsel >= 8(r/4) + r%4 with
// tsel = j - 4h) and is fetched by lane 32h + ((j + i) & 31), whose byte
// address for ds_bpermute is baddr(r) = ((4(j + 4h) + 4c_r) & 124) | 128h
// with c_r = 8(r/4) + r%4 (computed at the use: two VALU operations instead
// of sixteen live registers).
struct BandLanes {
    int tsel;
    int base4;  // 4(j + 4h)
    int hb;     // 128h
    __device__ __forceinline__ int baddr(int r) const
    {
        return ((base4 + 4 * (8 * (r >> 2) + (r & 3))) & 124) | hb;
    }
};

__device__ __forceinline__ BandLanes band_lanes()
{
    const int lane = threadIdx.x & 63;
    const int j = lane & 31, h = lane >> 5;
    BandLanes b;
    b.tsel = j - 4 * h;
    b.base4 = 4 * (j + 4 * h);
    b.hb = 128 * h;
    return b;
}

__device__ __forceinline__ unsigned lanes_below(unsigned long long mask)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ unsigned short to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
    return (unsigned short)(u >> 16);
}

// pairs of valid channels m apart (1 <= m <= 63) in a 256-bit mask V0..V3
__device__ __forceinline__ int pair_count(unsigned long long V0, unsigned long long V1,
                                          unsigned long long V2, unsigned long long V3, int m)
{
    const int l = 64 - m;
    return __popcll(V0 & ((V0 >> m) | (V1 << l))) + __popcll(V1 & ((V1 >> m) | (V2 << l))) +
           __popcll(V2 & ((V2 >> m) | (V3 << l))) + __popcll(V3 & (V3 >> m));
}

}  // namespace

struct SlscArgs {
    const float *im;    // (nfr, nrx, nx, nz)
    const unsigned long long *sel;  // (nxo, nzo, 4) usable-element bits, or null
    float *out;         // (nfr, nxo, nzo)
    int64_t nrx, nx, nz;
    int nxo, nzo, kwin, kofs, dsfx, dsfz;
    int pz;        // consecutive output pixels per wave
    int zbytes;    // bytes of one wave's normalised-channel buffer
    int wbytes;    // bytes of one wave's LDS region (channels, flags, reduction rows)
    int nlag, need, ngroups;
    float inv_nlag;
    int lags[64];
};

// 16 bytes at 4-byte alignment (gfx950 serves unaligned dwordx4 loads)
struct __attribute__((aligned(4))) F4s {
    float v[4];
};

// Lag sums and output of one pixel from its normalised, compacted channels:
// n usable channels in zbuf (FP32 rows of KP+1 floats, or FP16 / BF16 rows of
// KP halves), their validity bytes in vflag.  Wave-uniform arguments; every
// lane of the wave calls it.
// One 32x32x8 product of normalised FP16 / BF16 channels.
template <int PREC, typename T>
__device__ __forceinline__ f32x16_t mfma_k8(const T &a, const T &b, const f32x16_t &c)
{
    if constexpr (PREC == 1)
        return __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

// NBMAX: blocks of 32 channels a pixel can have (0: operands read per product)
template <int PREC, int KP, int NBMAX = 0>
__device__ __forceinline__ void slsc_pixel(const SlscArgs &a, unsigned char *zbuf,
                                           const unsigned char *vflag, float *red, int n,
                                           const BandLanes &bl, float *dst)
{
    const int lane = threadIdx.x & 63;
    if (n < a.need) {
        if (lane == 0) *dst = 0.0f;
        return;
    }
    // valid-channel bits in selected order
    unsigned long long V0 = 0ull, V1 = 0ull, V2 = 0ull, V3 = 0ull;
    V0 = __ballot(lane < n && vflag[lane]);
    if (n > 64) V1 = __ballot(64 + lane < n && vflag[64 + lane]);
    if (n > 128) V2 = __ballot(128 + lane < n && vflag[128 + lane]);
    if (n > 192) V3 = __ballot(192 + lane < n && vflag[192 + lane]);

    float S = 0.0f;  // lane li: S of lag lags[li]
    if (PREC == 0) {
        // lags in chunks of 8: per-lane partial sums in registers, one LDS
        // transpose-and-add per chunk gives the 8 wave totals
        const float *zb = reinterpret_cast<const float *>(zbuf);
        for (int lc = 0; lc < a.nlag; lc += 8) {
            int mm[8];
            float acc[8];
#pragma unroll
            for (int l = 0; l < 8; ++l) {
                mm[l] = lc + l < a.nlag ? a.lags[lc + l] : 4096;
                acc[l] = 0.0f;
            }
            for (int pc = 0; pc < n; pc += 64) {
                const int p = pc + lane;
                const bool act = p < n;
                float zr[KP];
#pragma unroll
                for (int w = 0; w < KP; ++w) zr[w] = act ? zb[p * (KP + 1) + w] : 0.0f;
#pragma unroll
                for (int l = 0; l < 8; ++l) {
                    if (act && p + mm[l] < n) {
                        const float *zq = zb + (p + mm[l]) * (KP + 1);
                        float s = 0.0f;
#pragma unroll
                        for (int w = 0; w < KP; ++w) s = __builtin_fmaf(zr[w], zq[w], s);
                        acc[l] += s;
                    }
                }
            }
#pragma unroll
            for (int l = 0; l < 8; ++l) red[l * 65 + lane] = acc[l];
            const int l8 = lane >> 3, part = lane & 7;
            float s = 0.0f;
#pragma unroll
            for (int i = 0; i < 8; ++i) s += red[l8 * 65 + part * 8 + i];
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            s += __shfl_xor(s, 4);
            const float got = __shfl(s, ((lane - lc) & 7) * 8);
            if (lane >= lc && lane < lc + 8 && lane < a.nlag) S = got;
        }
    } else {
        // zero rows n .. 32*nblk-1 of the last block
        const int nblk = (n + 31) >> 5;
        unsigned short *zr16 = reinterpret_cast<unsigned short *>(zbuf);
        for (int t = n * KP + lane; t < nblk * 32 * KP; t += 64) zr16[t] = 0;
        const int j = lane & 31, h = lane >> 5;
        const int mylag = lane < a.nlag ? a.lags[lane] : 0;
        // lags are at most 63: one or two lag groups
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (q >= a.ngroups) break;
            // The band extraction below is linear and the same for every
            // block pair, so the Gram blocks (b, b+q) and (b, b+q+1) of all
            // b are summed in the MFMA accumulators and extracted once.
            // the first product of each accumulator starts from zero (an
            // inline constant operand, no accumulator initialisation)
            const f32x16_t zero = {};
            f32x16_t g0 = zero, g1 = zero;
            if constexpr (KP == 8 && NBMAX > 0) {
                // every block's operand (this lane's row and half-window) is
                // read before the first product, so the products issue back
                // to back instead of waiting on LDS between them
                typedef typename std::conditional<PREC == 1, half4_t, bf16x4_t>::type op_t;
                op_t ops[NBMAX];
#pragma unroll
                for (int b = 0; b < NBMAX; ++b)
                    if (b < nblk) ops[b] = *reinterpret_cast<const op_t *>(zr16 + (32 * b + j) * KP + 4 * h);
#pragma unroll
                for (int b = 0; b < NBMAX; ++b) {
                    if (b + q < nblk) {
                        g0 = mfma_k8<PREC>(ops[b], ops[b + q < NBMAX ? b + q : 0], b == 0 ? zero : g0);
                        if (b + q + 1 < nblk)
                            g1 = mfma_k8<PREC>(ops[b], ops[b + q + 1 < NBMAX ? b + q + 1 : 0],
                                               b == 0 ? zero : g1);
                    }
                }
            } else {
            for (int b = 0; b + q < nblk; ++b) {
                const bool two = b + q + 1 < nblk;
#pragma unroll
                for (int kc = 0; kc < KP / 8; ++kc) {
                    const int ko = kc * 8 + 4 * h;
                    const unsigned short *pa = zr16 + (32 * b + j) * KP + ko;
                    const unsigned short *pb0 = zr16 + (32 * (b + q) + j) * KP + ko;
                    const unsigned short *pb1 = zr16 + (32 * (b + q + 1) + j) * KP + ko;
                    const bool first = b == 0 && kc == 0;
                    if (PREC == 1) {
                        const half4_t va = *reinterpret_cast<const half4_t *>(pa);
                        const half4_t vb0 = *reinterpret_cast<const half4_t *>(pb0);
                        g0 = first ? __builtin_amdgcn_mfma_f32_32x32x8f16(va, vb0, zero, 0, 0, 0)
                                   : __builtin_amdgcn_mfma_f32_32x32x8f16(va, vb0, g0, 0, 0, 0);
                        if (two) {
                            const half4_t vb1 = *reinterpret_cast<const half4_t *>(pb1);
                            g1 = first
                                     ? __builtin_amdgcn_mfma_f32_32x32x8f16(va, vb1, zero, 0, 0, 0)
                                     : __builtin_amdgcn_mfma_f32_32x32x8f16(va, vb1, g1, 0, 0, 0);
                        }
                    } else {
                        const bf16x4_t va = *reinterpret_cast<const bf16x4_t *>(pa);
                        const bf16x4_t vb0 = *reinterpret_cast<const bf16x4_t *>(pb0);
                        g0 = first ? __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(va, vb0, zero, 0,
                                                                               0, 0)
                                   : __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(va, vb0, g0, 0, 0,
                                                                               0);
                        if (two) {
                            const bf16x4_t vb1 = *reinterpret_cast<const bf16x4_t *>(pb1);
                            g1 = first ? __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(va, vb1, zero,
                                                                                   0, 0, 0)
                                       : __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(va, vb1, g1, 0,
                                                                                   0, 0);
                        }
                    }
                }
            }
            }
            float acc = 0.0f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float v = bl.tsel >= 8 * (r >> 2) + (r & 3) ? g0[r] : g1[r];
                acc += __int_as_float(__builtin_amdgcn_ds_bpermute(bl.baddr(r), __float_as_int(v)));
            }
            acc += __shfl_xor(acc, 32);
            const float got = __shfl(acc, mylag & 31);
            if (lane < a.nlag && (mylag >> 5) == q) S = got;
        }
    }

    const bool allvalid = __popcll(V0) + __popcll(V1) + __popcll(V2) + __popcll(V3) == n;
    float contrib = 0.0f;
    if (lane < a.nlag) {
        const int cnt = allvalid ? max(n - a.lags[lane], 0) : pair_count(V0, V1, V2, V3, a.lags[lane]);
        const float rho = cnt > 0 ? S / (float)cnt : __builtin_nanf("");
        contrib = rho * a.inv_nlag;
    }
    contrib = wave_sum(contrib);
    if (lane == 0) *dst = contrib;
}

// normalised sample z in the zbuf element type of PREC
template <int PREC>
__device__ __forceinline__ unsigned zcode(float z)
{
    return PREC == 1 ? (unsigned)__builtin_bit_cast(unsigned short, (_Float16)z)
                     : (unsigned)to_bf16(z);
}

// One wave per run of pz consecutive output pixels of a column; no
// workgroup-level staging or barriers.  Lane l owns channels l, l+64, ...
// (NG groups of 64) and loads each channel's window straight from the image:
// 32 bytes per lane for K = 8, so the windows of consecutive pixels of a
// wave share cache lines, which L2 serves after the first pixel.
template <int PREC, int KP, int NG>
__global__ __launch_bounds__(256) void k_slsc(SlscArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int nrx = (int)a.nrx;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwave = blockDim.x >> 6;
    unsigned char *zbuf = smem + (size_t)wave * a.wbytes;
    unsigned char *vflag = zbuf + a.zbytes;
    float *red = reinterpret_cast<float *>(vflag + 256);

    const int f = blockIdx.z, ixo = blockIdx.y;
    const int ix = ixo * a.dsfx;
    const int64_t rstride = a.nx * a.nz;
    const float *col = a.im + ((int64_t)f * a.nrx * a.nx + ix) * a.nz;  // + r * rstride + z
    const int izo0 = (blockIdx.x * nwave + wave) * a.pz;
    const BandLanes bl = band_lanes();

    for (int k = 0; k < a.pz; ++k) {
        const int izo = izo0 + k;
        if (izo >= a.nzo) break;
        const int iz = izo * a.dsfz;
        const int w0 = max(iz + a.kofs, 0);
        const int w1 = min((int64_t)(iz + a.kofs + a.kwin), a.nz);
        const int keff = w1 - w0;
        const float rk = 1.0f / (float)keff;
        // |std| >= 1e-10  <=>  sum of squared deviations >= 1e-20 * K
        const float vmin = 1e-20f * (float)keff;

        // windows of every usable channel, all loads in flight before use
        const unsigned long long *pbits =
            a.sel ? a.sel + ((int64_t)ixo * a.nzo + izo) * 4 : nullptr;
        float xv[NG][KP];
        bool selg[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int r = 64 * g + lane;
            bool sel = r < nrx;
            if (pbits) sel = sel && ((pbits[g] >> lane) & 1ull);
            selg[g] = sel;
            const float *xr = col + (int64_t)(sel ? r : 0) * rstride + w0;
            if (keff == KP) {
#pragma unroll
                for (int w = 0; w < KP; w += 4) {
                    F4s v{{0.0f, 0.0f, 0.0f, 0.0f}};
                    if (sel) v = *reinterpret_cast<const F4s *>(xr + w);
#pragma unroll
                    for (int e = 0; e < 4; ++e) xv[g][w + e] = v.v[e];
                }
            } else {
#pragma unroll
                for (int w = 0; w < KP; ++w) xv[g][w] = (sel && w < keff) ? xr[w] : 0.0f;
            }
        }

        // normalise and compact the usable channels
        int n = 0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const bool sel = selg[g];
            float mu = 0.0f, rs = 0.0f;
            bool valid = false;
            if (sel) {
                float s = 0.0f;
#pragma unroll
                for (int w = 0; w < KP; ++w) s += xv[g][w];
                mu = s * rk;
                float v = 0.0f;
#pragma unroll
                for (int w = 0; w < KP; ++w) {
                    const float d = w < keff ? xv[g][w] - mu : 0.0f;
                    v = __builtin_fmaf(d, d, v);
                }
                // z = (x - mean) / (std * sqrt(K)) = (x - mean) / sqrt(sum d^2)
                valid = v >= vmin;
                rs = valid ? __frsqrt_rn(v) : 0.0f;
            }
            const unsigned long long sm = __ballot(sel);
            const int p = n + (int)lanes_below(sm);
            if (sel) {
                vflag[p] = valid ? 1 : 0;
                if (PREC == 0) {
                    float *zr = reinterpret_cast<float *>(zbuf) + p * (KP + 1);
#pragma unroll
                    for (int w = 0; w < KP; ++w) zr[w] = w < keff ? (xv[g][w] - mu) * rs : 0.0f;
                } else {
                    unsigned *zr = reinterpret_cast<unsigned *>(zbuf) + p * (KP / 2);
                    unsigned hv[KP];
#pragma unroll
                    for (int w = 0; w < KP; ++w) {
                        const float z = w < keff ? (xv[g][w] - mu) * rs : 0.0f;
                        hv[w] = zcode<PREC>(z);
                    }
#pragma unroll
                    for (int w = 0; w < KP; w += 4)
                        *reinterpret_cast<uint2 *>(zr + w / 2) =
                            make_uint2(hv[w] | (hv[w + 1] << 16), hv[w + 2] | (hv[w + 3] << 16));
                }
            }
            n += __popcll(sm);
        }

        float *dst = a.out + ((int64_t)f * a.nxo + ixo) * a.nzo + izo;
        slsc_pixel<PREC, KP>(a, zbuf, vflag, red, n, bl, dst);
    }
}

// sum over the LPP lanes (aligned groups) that hold one window
template <int LPP>
__device__ __forceinline__ float lpp_sum(float v)
{
    if (LPP >= 2) v += dpp_xor1(v);
    if (LPP >= 4) v += dpp_xor2(v);
    if (LPP >= 8) v += __shfl_xor(v, 4);
    return v;
}

// Fast path of the MFMA precisions for windows that tile the axis
// (dsfz == kwin == KP): a wave takes PZ = RUN/KP consecutive output pixels of
// a column, whose windows are RUN (16 or 32) consecutive samples of every
// channel.  Those are read as channel rows: with S = RUN/4 segments per row,
// lane l holds samples 4*(l%S) .. +3 of channel (64/S)*q + l/S, so one load
// instruction covers 64/S channels x 4*RUN contiguous bytes (lane-per-channel
// loads of 16 bytes make each lane its own L2 request), and a window of pixel
// p is held by the KP/4 lanes l%S = p*KP/4 .. of each row.  For K = 8 the two lanes of a window swap halves (DPP) so that
// each lane normalises one whole window of two rows; for K = 16 / 32 the
// lanes of a window reduce its mean and variance with shuffles.  The
// normalised samples go to the pixel's compacted buffer, and each pixel's lag
// sums then run as in k_slsc with every block operand read before the first
// product.  The wave's selection words are read with scalar loads.
template <int PREC, int KP, int NG, int RUN>
// waves_per_eu(5): a register budget that keeps the MFMA accumulators in
// VGPRs (with the default budget they go to AGPRs and the kernel runs four
// waves per SIMD); the 32-sample runs of more than 128 channels keep the
// default (they would spill)
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(NG == 2 || RUN <= 16 ? 5 : 1))) void k_slsc_rows(SlscArgs a)
{
    static_assert(PREC != 0, "the channel-row path feeds the MFMA kernels");
    static_assert(RUN == 8 || RUN == 16 || RUN == 32, "runs of 8, 16 or 32 samples");
    static_assert(RUN >= KP, "a run holds whole windows");
    // S segments of 4 samples per channel row, CPR channels per load
    constexpr int PZ = RUN / KP, LPP = KP / 4, S = RUN / 4, CPR = 64 / S, NQ = NG * 64 / CPR;
    extern __shared__ __align__(16) unsigned char smem[];
    const int nrx = (int)a.nrx;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwave = blockDim.x >> 6;
    unsigned char *wbase = smem + (size_t)wave * a.wbytes;  // PZ x (zbuf, flags)
    float *red = nullptr;  // FP32 reduction rows: not used by the MFMA kernels

    const int f = blockIdx.z, ixo = blockIdx.y;
    const int ix = ixo * a.dsfx;
    const int64_t rstride = a.nx * a.nz;
    const float *col = a.im + ((int64_t)f * a.nrx * a.nx + ix) * a.nz;  // + r * rstride + z
    const int izo0 = (blockIdx.x * nwave + wave) * PZ;

    const int c8 = lane / S, k8 = lane % S;
    const int pp = k8 / LPP;                  // this lane's pixel
    const int izo = izo0 + pp;
    const bool pix = izo < a.nzo;
    const int zl = izo0 * a.dsfz + a.kofs + 4 * k8;  // first sample of this lane
    // the window of pixel pp clipped to the image: keff samples
    const int wz0 = izo * a.dsfz + a.kofs;
    const int keff = min(wz0 + KP, (int)a.nz) - max(wz0, 0);
    const float rk = 1.0f / (float)keff;
    const float vmin = 1e-20f * (float)keff;
    bool inr[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) inr[e] = zl + e >= 0 && zl + e < a.nz;
    const bool inner = zl >= 0 && zl + 4 <= a.nz;

    // usable-element bits of this lane's pixel: the wave's PZ pixels are
    // consecutive, so their words are read with uniform (scalar) loads
    unsigned long long m[NG];
    {
        const unsigned long long *sw = a.sel ? a.sel + ((int64_t)ixo * a.nzo + izo0) * 4 : nullptr;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int rem = nrx - 64 * g;
            const unsigned long long all =
                rem >= 64 ? ~0ull : (rem <= 0 ? 0ull : ((1ull << rem) - 1ull));
            unsigned long long w = 0ull;
#pragma unroll
            for (int p = 0; p < PZ; ++p) {
                const unsigned long long wp =
                    izo0 + p < a.nzo ? (sw ? sw[p * 4 + g] & all : all) : 0ull;
                w = pp == p ? wp : w;
            }
            m[g] = w;
        }
    }

    // channel rows: all loads in flight before the first use
    float x[NQ][4];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int r = CPR * q + c8;
        const bool sel = (m[(CPR * q) / 64] >> (r & 63)) & 1ull;
        const float *xr = col + (int64_t)(sel ? r : 0) * rstride + zl;
        F4s v{{0.0f, 0.0f, 0.0f, 0.0f}};
        if (sel) {
            if (inner) {
                v = *reinterpret_cast<const F4s *>(xr);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (inr[e]) v.v[e] = xr[e];
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) x[q][e] = v.v[e];
    }

    unsigned char *zb = wbase + pp * (a.zbytes + 256);
    unsigned char *vf = zb + a.zbytes;
    if constexpr (LPP == 2) {
        // Channel rows q0 = 2t and q1 = 2t+1: the two lanes holding the
        // halves of a window swap one half (DPP), so that lane (c8, k8)
        // holds the whole 8-sample window of channel CPR*(2t + k8%2) + c8
        // for its pixel k8/2 and normalises it in registers.
        const int hh = k8 & 1;
        bool inr8[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) inr8[e] = wz0 + e >= 0 && wz0 + e < a.nz;
#pragma unroll
        for (int t = 0; t < NQ / 2; ++t) {
            float xw[8];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float rcv = dpp_xor1(hh ? x[2 * t][e] : x[2 * t + 1][e]);
                xw[e] = hh ? rcv : x[2 * t][e];
                xw[4 + e] = hh ? x[2 * t + 1][e] : rcv;
            }
            const int r = CPR * (2 * t + hh) + c8;
            // rows 2t, 2t+1 lie in one 64-bit word (2*CPR divides 64)
            const int wt = (2 * CPR * t) / 64;
            const bool sel = (m[wt] >> (r & 63)) & 1ull;
            float s8 = 0.0f;
#pragma unroll
            for (int e = 0; e < 8; ++e) s8 += xw[e];
            const float mu = s8 * rk;
            float d[8];
            float v = 0.0f;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                d[e] = inr8[e] ? xw[e] - mu : 0.0f;
                v = __builtin_fmaf(d[e], d[e], v);
            }
            const bool valid = v >= vmin;
            const float rs = valid ? __frsqrt_rn(v) : 0.0f;
            if (sel) {
                int pos = __popcll(m[wt] & ((1ull << (r & 63)) - 1ull));
#pragma unroll
                for (int g = 0; g < wt; ++g) pos += __popcll(m[g]);
                vf[pos] = valid ? 1 : 0;
                unsigned hv[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) hv[e] = zcode<PREC>(d[e] * rs);
                *reinterpret_cast<uint4 *>(zb + (size_t)pos * 16) =
                    make_uint4(hv[0] | (hv[1] << 16), hv[2] | (hv[3] << 16),
                               hv[4] | (hv[5] << 16), hv[6] | (hv[7] << 16));
            }
        }
    } else {
        // normalise each (channel, pixel) window over its LPP lanes and write it
        // to the pixel's compacted buffer
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int r = CPR * q + c8;
            const int wq = (CPR * q) / 64;
            const bool sel = (m[wq] >> (r & 63)) & 1ull;
            float s = (x[q][0] + x[q][1]) + (x[q][2] + x[q][3]);
            s = lpp_sum<LPP>(s);
            const float mu = s * rk;
            float d[4];
            float v = 0.0f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                d[e] = inr[e] ? x[q][e] - mu : 0.0f;
                v = __builtin_fmaf(d[e], d[e], v);
            }
            v = lpp_sum<LPP>(v);
            const bool valid = v >= vmin;
            const float rs = valid ? __frsqrt_rn(v) : 0.0f;
            if (sel) {
                // position among the pixel's usable channels
                int pos = __popcll(m[wq] & ((1ull << (r & 63)) - 1ull));
#pragma unroll
                for (int g = 0; g < wq; ++g) pos += __popcll(m[g]);
                if (k8 % LPP == 0) vf[pos] = valid ? 1 : 0;
                const int w0 = 4 * (k8 % LPP);
                unsigned *zr = reinterpret_cast<unsigned *>(zb) + (pos * KP + w0) / 2;
                *reinterpret_cast<uint2 *>(zr) =
                    make_uint2(zcode<PREC>(d[0] * rs) | (zcode<PREC>(d[1] * rs) << 16),
                               zcode<PREC>(d[2] * rs) | (zcode<PREC>(d[3] * rs) << 16));
            }
        }

    }
    __syncthreads();  // compacted buffers complete (all waves reach this once)
    const BandLanes bl = band_lanes();

#pragma unroll
    for (int p = 0; p < PZ; ++p) {
        const int izp = izo0 + p;
        if (izp >= a.nzo) break;
        // usable channels of pixel p: the bits held by its lanes
        // (readlane yields int: take both halves as unsigned before widening)
        int n = 0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)m[g], p * LPP);
            const unsigned hi =
                (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(m[g] >> 32), p * LPP);
            n += __popc(lo) + __popc(hi);
        }
        unsigned char *zbp = wbase + p * (a.zbytes + 256);
        float *dst = a.out + ((int64_t)f * a.nxo + ixo) * a.nzo + izp;
        slsc_pixel<PREC, KP, 2 * NG>(a, zbp, zbp + a.zbytes, red, n, bl, dst);
    }
}

// Usable-element bits per output pixel: bit r of word r/64 is set when
// mask[r, ix, z] is non-zero over the pixel's whole (clipped) window.  One
// lane per (output pixel, element).
__global__ __launch_bounds__(256) void k_slsc_selection(const unsigned char *__restrict__ mask,
                                                        int nrx, int64_t nx, int64_t nz, int kwin,
                                                        int kofs, int dsfx, int dsfz, int nxo,
                                                        int nzo, unsigned long long *sel)
{
    const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);  // output pixel
    const int lane = threadIdx.x & 63;
    if (g >= (int64_t)nxo * nzo) return;
    const int ixo = (int)(g / nzo), izo = (int)(g - (int64_t)ixo * nzo);
    const int64_t ix = (int64_t)ixo * dsfx, iz = (int64_t)izo * dsfz;
    const int64_t w0 = iz + kofs < 0 ? 0 : iz + kofs;
    const int64_t w1 = iz + kofs + kwin < nz ? iz + kofs + kwin : nz;
    for (int c = 0; c < 4; ++c) {
        const int r = c * 64 + lane;
        bool ok = r < nrx;
        if (ok) {
            const unsigned char *m = mask + ((int64_t)r * nx + ix) * nz;
            for (int64_t z = w0; z < w1; ++z) ok = ok && m[z] != 0;
        }
        const unsigned long long b = __ballot(ok);
        if (lane == 0) sel[g * 4 + c] = b;
    }
}

int launch_slsc_selection(const uint8_t *mask, int64_t nrx, int64_t nx, int64_t nz, int kwin,
                          int kofs, int dsfx, int dsfz, uint64_t *sel, hipStream_t s)
{
    const int nxo = (int)((nx + dsfx - 1) / dsfx), nzo = (int)((nz + dsfz - 1) / dsfz);
    const int64_t npx = (int64_t)nxo * nzo;
    hipLaunchKernelGGL(k_slsc_selection, dim3((unsigned)((npx + 3) / 4)), dim3(256), 0, s, mask,
                       (int)nrx, nx, nz, kwin, kofs, dsfx, dsfz, nxo, nzo,
                       reinterpret_cast<unsigned long long *>(sel));
    return check_launch("slsc_selection");
}

// Per-wave LDS (normalised channels, flags, FP32 reduction rows), waves per
// workgroup and pixels per wave of a launch.  rows: the channel-row fast path
// (PZ = 32/KP pixels per wave, one buffer per pixel).
static void slsc_plan(SlscArgs &a, int prec, int kp, bool rows, int run, int &nwave, size_t &lds)
{
    const int nrx = (int)a.nrx;
    const int npad = (nrx + 31) / 32 * 32;
    a.zbytes = (int)align_up(prec == 0 ? (size_t)npad * (kp + 1) * 4 : (size_t)npad * kp * 2, 16);
    const size_t redb = prec == 0 ? 8 * 65 * 4 : 0;
    if (rows) {
        // MFMA precisions only
        a.pz = run / kp;
        a.wbytes = (int)align_up((size_t)a.pz * (a.zbytes + 256), 16);
        // up to 40 KB per workgroup; waves per workgroup measured on 128
        // channels, 256 x 256 output pixels (FP16, us):
        //   K = 8,  runs of 16: 1 wave 56.8, 2 46.5, 3 41.6, 4 46.1
        //   K = 16, runs of 16: 1 wave 56.8, 2 41.8, 4 42.2
        const int cap = std::max(1, std::min(4, 40960 / a.wbytes));
        nwave = run == 16 ? (kp == 8 ? 3 : 2) : cap;
        if (const char *e = getenv("PYCBF_SLSC_NWAVE")) nwave = atoi(e);
        nwave = std::max(1, std::min(cap, nwave));
    } else {
        a.wbytes = (int)align_up((size_t)a.zbytes + 256 + redb, 16);
        // one pixel per wave: more waves in flight beat longer runs per wave
        // (measured on 128 channels, K = 8, dsfz = 8: 1 px 69 us, 2 px 76 us,
        // 4 px 92 us for the FP16 kernel)
        int pz = 1;
        if (const char *e = getenv("PYCBF_SLSC_PZ")) pz = std::max(1, std::min(64, atoi(e)));
        a.pz = pz;
        nwave = 4;
    }
    lds = (size_t)nwave * a.wbytes;
}

static const char *slsc_prof_name(int prec)
{
    return prec == 0 ? "slsc_fp32" : (prec == 1 ? "slsc_mfma_fp16" : "slsc_mfma_bf16");
}

template <int PREC, int KP, int NG>
static int launch_slsc_t(const SlscArgs &a, bool rows, int nwave, size_t lds, int64_t nfr,
                         hipStream_t s)
{
    const int per_block = a.pz * nwave;
    const dim3 grid((unsigned)((a.nzo + per_block - 1) / per_block), (unsigned)a.nxo, (unsigned)nfr);
    ProfScope ps(slsc_prof_name(PREC), s);
    if constexpr (PREC != 0) {
        if (rows) {
            if constexpr (KP == 8) {
                if (a.pz * KP == 8) {
                    hipLaunchKernelGGL((k_slsc_rows<PREC, KP, NG, 8>), grid, dim3(64 * nwave),
                                       lds, s, a);
                    return check_launch("slsc");
                }
            }
            if constexpr (KP <= 16) {
                if (a.pz * KP == 16) {
                    hipLaunchKernelGGL((k_slsc_rows<PREC, KP, NG, 16>), grid, dim3(64 * nwave),
                                       lds, s, a);
                    return check_launch("slsc");
                }
            }
            hipLaunchKernelGGL((k_slsc_rows<PREC, KP, NG, 32>), grid, dim3(64 * nwave), lds, s, a);
            return check_launch("slsc");
        }
    }
    hipLaunchKernelGGL((k_slsc<PREC, KP, NG>), grid, dim3(64 * nwave), lds, s, a);
    return check_launch("slsc");
}

template <int PREC>
static int launch_slsc_p(const SlscArgs &a, int kp, bool rows, int nwave, size_t lds,
                         int64_t nfr, hipStream_t s)
{
    if (a.nrx <= 128) {
        switch (kp) {
        case 8: return launch_slsc_t<PREC, 8, 2>(a, rows, nwave, lds, nfr, s);
        case 16: return launch_slsc_t<PREC, 16, 2>(a, rows, nwave, lds, nfr, s);
        default: return launch_slsc_t<PREC, 32, 2>(a, rows, nwave, lds, nfr, s);
        }
    }
    switch (kp) {
    case 8: return launch_slsc_t<PREC, 8, 4>(a, rows, nwave, lds, nfr, s);
    case 16: return launch_slsc_t<PREC, 16, 4>(a, rows, nwave, lds, nfr, s);
    default: return launch_slsc_t<PREC, 32, 4>(a, rows, nwave, lds, nfr, s);
    }
}

int launch_slsc(const float *im, int64_t nfr, int64_t nrx, int64_t nx, int64_t nz,
                const uint64_t *sel, int kwin, int kofs, int dsfx, int dsfz, const int32_t *lags,
                int nlag, int precision, float *out, hipStream_t s)
{
    SlscArgs a{};
    a.im = im;
    a.sel = reinterpret_cast<const unsigned long long *>(sel);
    a.out = out;
    a.nrx = nrx;
    a.nx = nx;
    a.nz = nz;
    a.kwin = kwin;
    a.kofs = kofs;
    a.dsfx = dsfx;
    a.dsfz = dsfz;
    a.nxo = (int)((nx + dsfx - 1) / dsfx);
    a.nzo = (int)((nz + dsfz - 1) / dsfz);
    a.nlag = nlag;
    int lagmax = 0;
    for (int i = 0; i < nlag; ++i) {
        a.lags[i] = lags[i];
        lagmax = std::max(lagmax, (int)lags[i]);
    }
    a.need = lagmax + 1;
    a.ngroups = lagmax / 32 + 1;
    a.inv_nlag = 1.0f / (float)nlag;
    const int kp = kwin <= 8 ? 8 : (kwin <= 16 ? 16 : 32);
    // windows that tile the axis take the channel-row path (PYCBF_SLSC_ROWS=0
    // selects the lane-per-channel kernel)
    bool rows = precision != PYCBF_SLSC_FP32 && dsfz == kwin && kwin == kp;
    if (const char *e = getenv("PYCBF_SLSC_ROWS")) rows = rows && atoi(e) != 0;
    // channel-row runs of 16 samples for K <= 16 (K = 8 on 128 channels, 256 x
    // 256 pixels: runs of 8 / 16 / 32 take 54.6 / 41.6 / 55.7 us), 32 for
    // K = 32; PYCBF_SLSC_RUN selects 8, 16 or 32 (at least K)
    int run = kp <= 16 ? 16 : 32;
    if (const char *e = getenv("PYCBF_SLSC_RUN")) {
        const int r = atoi(e);
        run = (r == 8 || r == 16) && r >= kp ? r : 32;
    }
    int nwave = 0;
    size_t lds = 0;
    slsc_plan(a, precision, kp, rows, run, nwave, lds);
    switch (precision) {
    case PYCBF_SLSC_FP32: return launch_slsc_p<0>(a, kp, rows, nwave, lds, nfr, s);
    case PYCBF_SLSC_MFMA_FP16: return launch_slsc_p<1>(a, kp, rows, nwave, lds, nfr, s);
    default: return launch_slsc_p<2>(a, kp, rows, nwave, lds, nfr, s);
    }
}

}  // namespace pycbf
