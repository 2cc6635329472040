// FFT-window demodulation of beamformed RF to IQ as one real GEMM on MFMA.
//
// The notebooks' fft_tukey_demod (examples/SyntheticVsTabbed.ipynb, cell 1:
// rfft, keep a band of bins, Tukey window, ifftshift, ifft) is a fixed linear
// map from the n samples of each row to L complex samples.  The host builds
// that map once in float64 as the real matrix op (n, 2L), columns (Re, Im)
// interleaved; the device computes iq (rows, 2L) = rf (rows, n) @ op with
// v_mfma_f32_32x32x2_f32 (FP32 inputs, FP32 accumulation), so iq reads as
// complex64 (rows, L).
//
// Tiling: a 256-thread workgroup computes a 64 x 64 tile of iq, one 32 x 32
// quarter per wave; k advances in slices of 16 staged in LDS (rf 64 x 16,
// op 16 x 64), zero-filled past the matrix edges.
#include "common.h"
#include "prof.h"
#include "tiles.h"

namespace pycbf {

namespace {
typedef float f32x16_t __attribute__((ext_vector_type(16)));
constexpr int DT = 64;   // workgroup tile (rows and columns)
constexpr int DK = 16;   // k slice
}  // namespace

__global__ __launch_bounds__(256) void k_demod_gemm(const float *__restrict__ A, int64_t R, int K,
                                                    const float *__restrict__ B, int N,
                                                    float *__restrict__ C)
{
    __shared__ float As[DT][DK + 1];
    __shared__ float Bs[DK][DT + 4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t row0 = (int64_t)blockIdx.x * DT;
    const int col0 = blockIdx.y * DT;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    f32x16_t acc = {};
    // loader roles: rf row tid/4, k 4*(tid%4)..+3; op k tid/16, columns 4*(tid%16)..+3
    const int ar = tid >> 2, ak = 4 * (tid & 3);
    const int bk = tid >> 4, bc = 4 * (tid & 15);
    for (int k0 = 0; k0 < K; k0 += DK) {
        {
            const int64_t r = row0 + ar;
            const int k = k0 + ak;
            F4u v{{0.0f, 0.0f, 0.0f, 0.0f}};
            if (r < R) {
                const float *src = A + r * (int64_t)K + k;
                if (k + 4 <= K) {
                    v = *reinterpret_cast<const F4u *>(src);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (k + e < K) v.v[e] = src[e];
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) As[ar][ak + e] = v.v[e];
        }
        {
            const int k = k0 + bk;
            const int c = col0 + bc;
            F4u v{{0.0f, 0.0f, 0.0f, 0.0f}};
            if (k < K) {
                const float *src = B + (int64_t)k * N + c;
                if (c + 4 <= N) {
                    v = *reinterpret_cast<const F4u *>(src);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (c + e < N) v.v[e] = src[e];
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) Bs[bk][bc + e] = v.v[e];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < DK; kk += 2) {
            // lane l: A[l%32][kk + l/32] and B[kk + l/32][l%32]
            const float a = As[wr + (lane & 31)][kk + (lane >> 5)];
            const float b = Bs[kk + (lane >> 5)][wc + (lane & 31)];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D[i][j] of the 32 x 32 quarter sits in lane j + 32h, register r, with
    // i = 8(r/4) + 4h + r%4
    const int j = lane & 31, h = lane >> 5;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = 8 * (r >> 2) + 4 * h + (r & 3);
        const int64_t row = row0 + wr + i;
        const int col = col0 + wc + j;
        if (row < R && col < N) C[row * (int64_t)N + col] = acc[r];
    }
}

int launch_demod_gemm(const float *A, int64_t R, int K, const float *B, int N, float *C,
                      hipStream_t s)
{
    if (R <= 0 || N <= 0) return 0;
    // row tiles on x (up to 2^31 - 1), column tiles on y (ncol < 2^22)
    dim3 grid((unsigned)((R + DT - 1) / DT), (unsigned)((N + DT - 1) / DT));
    ProfScope ps("demod_gemm", s);
    hipLaunchKernelGGL(k_demod_gemm, grid, dim3(256), 0, s, A, R, K, B, N, C);
    return check_launch("demod gemm");
}

}  // namespace pycbf
