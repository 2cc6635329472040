// FFT demodulation of beamformed RF lines to IQ (fft_tukey_demod of
// examples/SyntheticVsTabbed.ipynb, cell 1): per line, the real FFT of its n
// samples, the m kept bins [ikmin, ikmin + m) weighted by the Tukey window
// and rotated by the ifftshift, and the m-point inverse DFT of those bins.
//
// One wave per line (up to four per workgroup, which sweeps the lines),
// everything in LDS, one read of the line and one write of its IQ samples
// (the operation's compulsory HBM bytes):
//   1. z_j = x_2j + i x_2j+1 (n/2 complex points), Z = FFT_{n/2}(z) and the
//      real-FFT split X_k = (Z_k + conj Z_{n/2-k}) / 2
//                         - i/2 e^{-2 pi i k / n} (Z_k - conj Z_{n/2-k});
//   2. the m-point inverse DFT of S_k = w_src(k) X_{ikmin + src(k)},
//      src(k) = (k + shift) mod m, by Bluestein's chirp product (Bluestein
//      1970; L. Rabiner, R. Schafer, C. Rader, "The chirp z-transform
//      algorithm", 1969): with k t = (k^2 + t^2 - (t - k)^2) / 2,
//        y_t = e^{i pi t^2 / m} / m * sum_k [S_k e^{i pi k^2 / m}] e^{-i pi (t-k)^2 / m},
//      a linear convolution of length 2m - 1 done as a circular one of
//      length L (a power of two >= 2m - 1): FFT_L, product with the
//      transform of the chirp (precomputed), inverse FFT_L as
//      conj(FFT_L(conj(.))) / L.
// The host precomputes, in float64 rounded to float32: pre_k = w_src(k)
// e^{i pi k^2 / m}, post_t = e^{i pi t^2 / m} / (m L), the chirp transform,
// the split twiddle e^{-2 pi i kb / n} of each kept bin, and one twiddle
// table e^{-2 pi i t / T} (T = max(n / 2, L)) that serves every FFT size.
// The FFTs are Stockham autosort transforms (N. Govindaraju et al., "High
// performance discrete Fourier transforms on graphics processors", SC 2008)
// in radix 16 (a radix-2 and a radix-4 stage first where log2 N needs them),
// in place in the wave's LDS buffer (all of a stage's reads into registers
// before its writes).
#include <algorithm>

#include "common.h"
#include "prof.h"

namespace pycbf {

namespace {

constexpr int DFFT_WAVES = 4;                 // lines in flight per workgroup, one per wave
constexpr int DFFT_THREADS = 64 * DFFT_WAVES;
constexpr int DFFT_MAXN = 2048;               // complex points per FFT
constexpr int DFFT_LDS = 8192;                // float2 entries of LDS per workgroup (64 KB)

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
// -i a
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// The wave's LDS writes are complete and visible to its other lanes (each
// wave works on its own line: no workgroup barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// 4-point DFT (e^{-2 pi i rs / 4}) of (a, b, c, d), in place
__device__ __forceinline__ void dft4(float2 &a, float2 &b, float2 &c, float2 &d)
{
    const float2 a0 = cadd(a, c), a1 = csub(a, c);
    const float2 b0 = cadd(b, d), b1 = mul_mi(csub(b, d));
    a = cadd(a0, b0);
    b = cadd(a1, b1);
    c = csub(a0, b0);
    d = csub(a1, b1);
}

// 16-point DFT in registers, natural order in and out: 4-point DFTs over r of
// x[q + 4r] for each q, the twiddles e^{-2 pi i qs / 16}, 4-point DFTs over q
// (X[s + 4t]).
__device__ __forceinline__ void dft16(float2 (&x)[16])
{
    constexpr float C1 = 0.92387953251128674f;  // cos(pi / 8)
    constexpr float S1 = 0.38268343236508978f;  // sin(pi / 8)
    constexpr float R2 = 0.70710678118654752f;  // sqrt(1/2)
#pragma unroll
    for (int q = 0; q < 4; ++q) dft4(x[q], x[q + 4], x[q + 8], x[q + 12]);
    // x[q + 4s] now holds A[q][s]
    x[5] = cmul(x[5], make_float2(C1, -S1));    // q1 s1: W^1
    x[9] = cmul(x[9], make_float2(R2, -R2));    // q1 s2: W^2
    x[13] = cmul(x[13], make_float2(S1, -C1));  // q1 s3: W^3
    x[6] = cmul(x[6], make_float2(R2, -R2));    // q2 s1: W^2
    x[10] = mul_mi(x[10]);                      // q2 s2: W^4
    x[14] = cmul(x[14], make_float2(-R2, -R2)); // q2 s3: W^6
    x[7] = cmul(x[7], make_float2(S1, -C1));    // q3 s1: W^3
    x[11] = cmul(x[11], make_float2(-R2, -R2)); // q3 s2: W^6
    x[15] = cmul(x[15], make_float2(-C1, S1));  // q3 s3: W^9
    float2 y[16];
#pragma unroll
    for (int sidx = 0; sidx < 4; ++sidx) {
        float2 a = x[4 * sidx], b = x[4 * sidx + 1], c = x[4 * sidx + 2], d = x[4 * sidx + 3];
        dft4(a, b, c, d);
        y[sidx] = a;
        y[sidx + 4] = b;
        y[sidx + 8] = c;
        y[sidx + 12] = d;
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = y[r];
}

// One Stockham stage of radix R (2, 4 or 16) at sub-transform size Ns for an
// N-point FFT in the wave's LDS buffer d: butterfly j reads d[j + r N/R],
// twiddles by e^{-2 pi i r k / (R Ns)} (k = j mod Ns), transforms, and writes
// d[(j - k) R + k + r Ns].  All of the lane's reads precede its writes.
template <int R, int PTS>
__device__ __forceinline__ void fft_stage(float2 *d, int N, int Ns, const float2 *tw, int tstride,
                                          int lane)
{
    constexpr int PER = PTS / R > 0 ? PTS / R : 1;  // butterflies per lane
    const int nb = N / R;
    const int span = (N / (R * Ns)) * tstride;  // table step per (r k)
    float2 v[PER][R];
#pragma unroll
    for (int it = 0; it < PER; ++it) {
        const int j = lane + it * 64;
        if (j < nb) {
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = d[j + r * nb];
            if (Ns > 1) {
                if constexpr (R == 16) {
                    // w^1, w^2, w^4, w^8 from the table, the other powers as
                    // products of at most four of them
                    const int e = k * span;
                    const float2 w1 = tw[e], w2 = tw[2 * e], w4 = tw[4 * e], w8 = tw[8 * e];
                    const float2 w3 = cmul(w1, w2), w5 = cmul(w1, w4), w6 = cmul(w2, w4);
                    const float2 w7 = cmul(w3, w4);
                    v[it][1] = cmul(v[it][1], w1);
                    v[it][2] = cmul(v[it][2], w2);
                    v[it][3] = cmul(v[it][3], w3);
                    v[it][4] = cmul(v[it][4], w4);
                    v[it][5] = cmul(v[it][5], w5);
                    v[it][6] = cmul(v[it][6], w6);
                    v[it][7] = cmul(v[it][7], w7);
                    v[it][8] = cmul(v[it][8], w8);
                    v[it][9] = cmul(v[it][9], cmul(w1, w8));
                    v[it][10] = cmul(v[it][10], cmul(w2, w8));
                    v[it][11] = cmul(v[it][11], cmul(w3, w8));
                    v[it][12] = cmul(v[it][12], cmul(w4, w8));
                    v[it][13] = cmul(v[it][13], cmul(w5, w8));
                    v[it][14] = cmul(v[it][14], cmul(w6, w8));
                    v[it][15] = cmul(v[it][15], cmul(w7, w8));
                } else {
#pragma unroll
                    for (int r = 1; r < R; ++r) v[it][r] = cmul(v[it][r], tw[r * k * span]);
                }
            }
        }
    }
    wave_sync();
#pragma unroll
    for (int it = 0; it < PER; ++it) {
        const int j = lane + it * 64;
        if (j < nb) {
            const int k = j & (Ns - 1);
            if constexpr (R == 2) {
                const float2 a = v[it][0], b = v[it][1];
                v[it][0] = cadd(a, b);
                v[it][1] = csub(a, b);
            } else if constexpr (R == 4) {
                dft4(v[it][0], v[it][1], v[it][2], v[it][3]);
            } else {
                dft16(v[it]);
            }
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) d[base + r * Ns] = v[it][r];
        }
    }
    wave_sync();
}

// Forward FFT (e^{-2 pi i jk / N}) of N = 2^p <= 64 PTS points d[0, N) in LDS
// by one wave, in place, natural order in and out.  tw[t * tstride] =
// e^{-2 pi i t / N} (LDS).  Stages: radix 2 if p is odd, radix 4 if p / 2 is
// odd, then radix 16 (two radix-4 layers in registers per LDS pass).
template <int PTS>
__device__ void fft_wave(float2 *d, int N, int log2n, const float2 *tw, int tstride, int lane)
{
    int Ns = 1;
    if (log2n & 1) {
        fft_stage<2, PTS>(d, N, Ns, tw, tstride, lane);
        Ns = 2;
    }
    if ((log2n >> 1) & 1) {
        fft_stage<4, PTS>(d, N, Ns, tw, tstride, lane);
        Ns <<= 2;
    }
    while (Ns < N) {
        fft_stage<16, PTS>(d, N, Ns, tw, tstride, lane);
        Ns <<= 4;
    }
}

struct DemodFftArgs {
    const float *rf;  // (rows, n)
    float *iq;        // (rows, 2 m): Re, Im
    const float2 *pre, *post, *bhat, *split, *tw;
    int64_t rows;
    int n, half, log2half, ikmin, m, shift, L, log2L, T, buf, nw;
};

// PTS: points per lane of the largest FFT (max(n / 2, L) / 64).  LDS: the
// twiddle table (T entries) and one line buffer of max(n / 2, L)
// points per active wave (a.nw waves, as many as fit in 64 KB); the workgroup
// sweeps the lines a.nw at a time.
template <int PTS>
__global__ __launch_bounds__(DFFT_THREADS) void k_demod_fft(DemodFftArgs a)
{
    // kept bins per lane: m <= (L + 1) / 2 <= 32 PTS + 1
    constexpr int KPL = PTS / 2 + 1;
    extern __shared__ float2 lds[];
    float2 *tw = lds;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float2 *B = lds + a.T + wave * a.buf;
    for (int t = tid; t < a.T; t += DFFT_THREADS) tw[t] = a.tw[t];
    __syncthreads();
    if (wave >= a.nw) return;  // no buffer for this wave
    const int ts_half = a.T / a.half, ts_L = a.T / a.L;
    for (int64_t row = (int64_t)blockIdx.x * a.nw + wave; row < a.rows;
         row += (int64_t)gridDim.x * a.nw) {
        const float2 *x = reinterpret_cast<const float2 *>(a.rf + row * a.n);
        for (int j = lane; j < a.half; j += 64) B[j] = x[j];
        wave_sync();
        fft_wave<PTS>(B, a.half, a.log2half, tw, ts_half, lane);
        // kept bins, window and chirp into registers, then over the spectrum
        // with the zero padding to L
        float2 c[KPL];
#pragma unroll
        for (int it = 0; it < KPL; ++it) {
            const int k = lane + it * 64;
            float2 v = make_float2(0.0f, 0.0f);
            if (k < a.m) {
                int s = k + a.shift;
                s = s >= a.m ? s - a.m : s;
                const int kb = a.ikmin + s;  // rfft bin, 0 .. n/2
                const float2 zk = B[kb == a.half ? 0 : kb];
                const float2 zc = conjf2(B[kb == 0 ? 0 : a.half - kb]);
                const float2 ev = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
                const float2 od = mul_mi(make_float2(0.5f * (zk.x - zc.x), 0.5f * (zk.y - zc.y)));
                const float2 X = cadd(ev, cmul(a.split[k], od));
                v = cmul(X, a.pre[k]);
            }
            c[it] = v;
        }
        wave_sync();
#pragma unroll
        for (int it = 0; it < KPL; ++it) {
            const int k = lane + it * 64;
            if (k < a.L) B[k] = c[it];
        }
        for (int k = lane + KPL * 64; k < a.L; k += 64) B[k] = make_float2(0.0f, 0.0f);
        wave_sync();
        fft_wave<PTS>(B, a.L, a.log2L, tw, ts_L, lane);
        for (int k = lane; k < a.L; k += 64) B[k] = conjf2(cmul(B[k], a.bhat[k]));
        wave_sync();
        fft_wave<PTS>(B, a.L, a.log2L, tw, ts_L, lane);
        float2 *out = reinterpret_cast<float2 *>(a.iq + row * 2 * (int64_t)a.m);
        for (int t = lane; t < a.m; t += 64) out[t] = cmul(conjf2(B[t]), a.post[t]);
        wave_sync();  // the buffer is reloaded for the next line
    }
}

int ilog2_exact(int64_t v)
{
    if (v < 1 || (v & (v - 1))) return -1;
    int p = 0;
    while ((1LL << p) < v) ++p;
    return p;
}

}  // namespace

}  // namespace pycbf

using namespace pycbf;

extern "C" int pycbf_demod_fft(const float *rf, int64_t rows, int64_t n, int64_t ikmin,
                               int64_t nbin, int64_t shift, int64_t nconv, const float *tables,
                               float *iq, void *stream)
{
    const int lh = ilog2_exact(n / 2), lL = ilog2_exact(nconv);
    if (rows < 0) return fail(PYCBF_EINVAL, "rows must be >= 0");
    if (n < 8 || n > 2 * DFFT_MAXN || (n & 1) || lh < 0)
        return fail(PYCBF_EINVAL, "n must be a power of two in 8..%d (got %lld)", 2 * DFFT_MAXN,
                    (long long)n);
    if (nbin < 1 || ikmin < 0 || ikmin + nbin > n / 2 + 1)
        return fail(PYCBF_EINVAL, "bins [%lld, %lld) outside the %lld rfft bins", (long long)ikmin,
                    (long long)(ikmin + nbin), (long long)(n / 2 + 1));
    if (shift < 0 || shift >= nbin) return fail(PYCBF_EINVAL, "shift must be in [0, nbin)");
    if (lL < 0 || nconv < 2 * nbin - 1 || nconv > DFFT_MAXN)
        return fail(PYCBF_EINVAL, "nconv must be a power of two in [2 nbin - 1, %d]", DFFT_MAXN);
    if (rows > 0x7fffffffLL) return fail(PYCBF_EINVAL, "too many rows");
    if (rows == 0) return 0;
    if (!rf || !tables || !iq) return fail(PYCBF_EINVAL, "NULL device pointer");
    if (reinterpret_cast<uintptr_t>(rf) & 7 || reinterpret_cast<uintptr_t>(iq) & 7 ||
        reinterpret_cast<uintptr_t>(tables) & 7)
        return fail(PYCBF_EINVAL, "rf, tables and iq must be 8-byte aligned");
    DemodFftArgs a{};
    a.rf = rf;
    a.iq = iq;
    a.n = (int)n;
    a.half = (int)(n / 2);
    a.log2half = lh;
    a.ikmin = (int)ikmin;
    a.m = (int)nbin;
    a.shift = (int)shift;
    a.L = (int)nconv;
    a.log2L = lL;
    a.T = a.half > a.L ? a.half : a.L;
    const float2 *t2 = reinterpret_cast<const float2 *>(tables);
    a.pre = t2;
    a.post = t2 + nbin;
    a.bhat = t2 + 2 * nbin;
    a.split = t2 + 2 * nbin + nconv;
    a.tw = t2 + 3 * nbin + nconv;
    hipStream_t s = (hipStream_t)stream;
    a.rows = rows;
    a.buf = a.T;  // line buffer
    a.nw = std::min(DFFT_WAVES, (DFFT_LDS - a.T) / a.buf);
    const size_t lds = (size_t)(a.T + a.nw * a.buf) * sizeof(float2);
    int dev = 0, ncu = 0;
    if (int rc = check(hipGetDevice(&dev), "device")) return rc;
    if (int rc = check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev),
                       "CU count"))
        return rc;
    const int64_t groups = (rows + a.nw - 1) / a.nw;
    const unsigned grid = (unsigned)std::min<int64_t>(groups, (int64_t)ncu * 8);
    ProfScope ps("demod_fft", s);
    if (a.T <= 1024)
        hipLaunchKernelGGL(k_demod_fft<16>, dim3(grid), dim3(DFFT_THREADS), lds, s, a);
    else
        hipLaunchKernelGGL(k_demod_fft<32>, dim3(grid), dim3(DFFT_THREADS), lds, s, a);
    return check_launch("demod_fft");
}
