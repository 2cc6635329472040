// FFT-window demodulation of beamformed RF to IQ as one real matrix product.
//
// The notebooks' fft_tukey_demod (examples/SyntheticVsTabbed.ipynb, cell 1:
// rfft, keep a band of bins, Tukey window, ifftshift, ifft) is a fixed linear
// map from the n samples of each line to L complex samples.  The host builds
// that map once in float64 and hands it over transposed, as opT (2L, n):
// row 2m holds Re and row 2m+1 Im of output sample m.  The device computes
//   iq[r][c] = sum_k rf[r][k] * opT[c][k]
// so that iq (rows, 2L) reads as complex64 (rows, L).
//
// Precision: every float32 operand x is split exactly into three BF16 parts,
// x = x0 + x1 + x2 (x0 = RN(x), x1 = RN(x - x0), x2 = x - x0 - x1, each
// remainder exact in float32 and the last one exact in BF16), and the
// product keeps the six cross terms of order >= 2^-16,
//   a0 b0 + a0 b1 + a1 b0 + a0 b2 + a1 b1 + a2 b0,
// each an exact product accumulated in float32 by v_mfma_f32_32x32x16_bf16.
// The dropped terms are below 2^-23 |a||b| per product, so the result is a
// float32-accurate product on the BF16 matrix units (2.5 PFLOP/s dense)
// instead of the FP32 ones (157.3 TFLOP/s).
//
// Tiling: a 256-thread workgroup computes a 128 x 128 tile of iq; each wave
// a 64 x 64 quarter as 2 x 2 blocks of 32 x 32.  k advances in slices of 32:
// the workgroup loads the next slice's rf and opT tiles into registers while
// the current one (three BF16 planes each, in LDS) feeds the MFMAs; edges are
// zero-filled.  48 KB of LDS and 168 registers per thread give three
// workgroups per CU (three waves per SIMD), so one wave's splitting overlaps
// another's MFMAs.
#include <algorithm>

#include "common.h"
#include "prof.h"
#include "tiles.h"

namespace pycbf {

namespace {
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
constexpr int DM = 128;     // workgroup tile rows (lines) and columns
constexpr int DK = 32;      // k slice
constexpr int DP = DK;      // LDS row pitch in BF16 (64 bytes, four 16-byte chunks)

// 16-byte chunk c of LDS row r sits at chunk c ^ ((r >> 1) & 3): the eight
// rows one ds_read_b128 phase reads (8 lanes, rows j .. j+7, one chunk) then
// cover all 32 banks once, where an unswizzled 64-byte pitch puts them 4-way
// on the same banks.  The unpadded pitch keeps a workgroup's planes at 48 KB
// (three workgroups per CU).
__device__ __forceinline__ int lds_off(int row, int k)
{
    return row * DP + ((((k >> 3) ^ (row >> 1)) & 3) << 3) + (k & 7);
}

// rows [row0, row0 + 128) x k [k0, k0 + 32) of a (nrow x ld) float32 matrix,
// zero from column kend on: thread t holds rows t/8 + 32q (q < 4), k 4(t%8) .. +3
__device__ __forceinline__ void load_tile(const float *__restrict__ M, int64_t nrow, int ld,
                                          int kend, int64_t row0, int k0, F4u (&v)[4])
{
    const int t = threadIdx.x;
    const int k = k0 + 4 * (t & 7);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t r = row0 + (t >> 3) + 32 * q;
        v[q] = F4u{{0.0f, 0.0f, 0.0f, 0.0f}};
        if (r < nrow) {
            const float *src = M + r * (int64_t)ld + k;
            if (k + 4 <= kend) {
                v[q] = *reinterpret_cast<const F4u *>(src);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (k + e < kend) v[q].v[e] = src[e];
            }
        }
    }
}

// three-way BF16 split of the thread's tile part into planes s[0..2]
__device__ __forceinline__ void store_split(__bf16 *__restrict__ s, const F4u (&v)[4])
{
    const int t = threadIdx.x;
    const int kc = 4 * (t & 7);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        bf16x4_t p0, p1, p2;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float x = v[q].v[e];
            const __bf16 h0 = (__bf16)x;
            const float r1 = x - (float)h0;
            const __bf16 h1 = (__bf16)r1;
            const float r2 = r1 - (float)h1;
            p0[e] = h0;
            p1[e] = h1;
            p2[e] = (__bf16)r2;
        }
        const int o = lds_off((t >> 3) + 32 * q, kc);
        *reinterpret_cast<bf16x4_t *>(s + o) = p0;
        *reinterpret_cast<bf16x4_t *>(s + DM * DP + o) = p1;
        *reinterpret_cast<bf16x4_t *>(s + 2 * DM * DP + o) = p2;
    }
}
}  // namespace

// blockIdx.z: the slice [z kc, (z+1) kc) of k; with several slices C is the
// (nsplit, R, N) array of partial products
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_demod_gemm(const float *__restrict__ A, int64_t R, int K,
                                                    const float *__restrict__ Bt, int N, int kc,
                                                    float *__restrict__ C)
{
    __shared__ __align__(16) __bf16 As[3 * DM * DP];
    __shared__ __align__(16) __bf16 Bs[3 * DM * DP];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t row0 = (int64_t)blockIdx.x * DM;
    const int col0 = blockIdx.y * DM;
    const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
    const int j = lane & 31, h = lane >> 5;

    f32x16_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x16_t{};

    const int kbeg = blockIdx.z * kc, kend = min(K, kbeg + kc);
    F4u va[4], vb[4];
    load_tile(A, R, K, kend, row0, kbeg, va);
    load_tile(Bt, N, K, kend, col0, kbeg, vb);
    for (int k0 = kbeg; k0 < kend; k0 += DK) {
        store_split(As, va);
        store_split(Bs, vb);
        __syncthreads();
        if (k0 + DK < kend) {
            // next slice in flight while this one multiplies
            load_tile(A, R, K, kend, row0, k0 + DK, va);
            load_tile(Bt, N, K, kend, col0, k0 + DK, vb);
        }
#pragma unroll
        for (int kk = 0; kk < DK; kk += 16) {
            // lane l: row (or column) l%32 of the block, k kk + 8(l/32) .. +7
            bf16x8_t a[2][3], b[2][3];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int p = 0; p < 3; ++p) {
                    const int oa = p * DM * DP + lds_off(wr + 32 * bi + j, kk + 8 * h);
                    const int ob = p * DM * DP + lds_off(wc + 32 * bi + j, kk + 8 * h);
                    a[bi][p] = *reinterpret_cast<const bf16x8_t *>(As + oa);
                    b[bi][p] = *reinterpret_cast<const bf16x8_t *>(Bs + ob);
                }
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) {
                    // small terms first
                    f32x16_t c = acc[bi][bj];
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[bi][2], b[bj][0], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[bi][1], b[bj][1], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[bi][0], b[bj][2], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[bi][1], b[bj][0], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[bi][0], b[bj][1], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[bi][0], b[bj][0], c, 0, 0, 0);
                    acc[bi][bj] = c;
                }
        }
        __syncthreads();
    }
    // D[i][j] of a 32 x 32 block sits in lane j + 32h, register r, with
    // i = 8(r/4) + 4h + r%4
    C += (int64_t)blockIdx.z * R * N;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = 8 * (r >> 2) + 4 * h + (r & 3);
                const int64_t row = row0 + wr + 32 * bi + i;
                const int col = col0 + wc + 32 * bj + j;
                if (row < R && col < N) C[row * (int64_t)N + col] = acc[bi][bj][r];
            }
}

// C[g] = sum over z of P[z][g], z in order (deterministic)
__global__ __launch_bounds__(256) void k_demod_reduce(const float *__restrict__ P, int nsplit,
                                                      int64_t n, float *__restrict__ C)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= n) return;
    float v = P[g];
    for (int z = 1; z < nsplit; ++z) v += P[(int64_t)z * n + g];
    C[g] = v;
}

int demod_split(int64_t R, int K, int N)
{
    // split k while the tiles alone leave the chip short of about 512
    // workgroups, keeping at least 256 k (8 slices) per split
    const int64_t tiles = ((R + DM - 1) / DM) * ((N + DM - 1) / DM);
    if (tiles >= 512 || K < 512) return 1;
    const int64_t want = (512 + tiles - 1) / tiles;
    return (int)std::max<int64_t>(1, std::min<int64_t>({want, (int64_t)K / 256, 64}));
}

int launch_demod_gemm(const float *A, int64_t R, int K, const float *Bt, int N, float *C,
                      float *part, hipStream_t s)
{
    if (R <= 0 || N <= 0) return 0;
    int nsplit = demod_split(R, K, N);
    int kc = K;
    if (nsplit > 1) {
        kc = ((K + nsplit - 1) / nsplit + DK - 1) / DK * DK;
        nsplit = (K + kc - 1) / kc;
    }
    if (nsplit > 1 && !part) return fail(PYCBF_EINVAL, "demodulation workspace missing");
    // row tiles on x (up to 2^31 - 1), column tiles on y, k slices on z
    dim3 grid((unsigned)((R + DM - 1) / DM), (unsigned)((N + DM - 1) / DM), (unsigned)nsplit);
    ProfScope ps("demod_gemm", s);
    hipLaunchKernelGGL(k_demod_gemm, grid, dim3(256), 0, s, A, R, K, Bt, N, kc,
                       nsplit > 1 ? part : C);
    if (int rc = check_launch("demod gemm")) return rc;
    if (nsplit > 1) {
        const int64_t n = R * (int64_t)N;
        hipLaunchKernelGGL(k_demod_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           part, nsplit, n, C);
        return check_launch("demod reduce");
    }
    return 0;
}

}  // namespace pycbf
