// Shared declarations for the gfx950 beamforming library (libpycbf_hip.so).
// All device code is compiled with -ffp-contract=off: every multiply-add that
// must round like pycbf's float32 C code stays unfused, and the few places that
// may fuse use __builtin_fmaf explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pycbf_hip.h"

namespace pycbf {

// Records a failure for pycbf_last_error() and returns `code`.
int fail(int code, const char *fmt, ...);
// Turns a launch / runtime error into a status; 0 on success.
int check(hipError_t e, const char *what);
inline int check_launch(const char *what) { return check(hipGetLastError(), what); }

__host__ __device__ inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Carves aligned sub-buffers out of a caller-provided workspace.
struct Carver {
    char *base;
    size_t cap;
    size_t used = 0;
    Carver(void *b, size_t c) : base(static_cast<char *>(b)), cap(c) {}
    void *take(size_t bytes) {
        size_t off = align_up(used, 256);
        used = off + bytes;
        return (base && used <= cap) ? base + off : nullptr;
    }
};

// ---------------------------------------------------------------- spline ---
constexpr int SPL_CHUNK = 32;  // samples per LDS tile in the spline sweeps
// Sizes the spline scratch (one float per sample).
inline size_t spline_scratch_bytes(int64_t nch, int32_t nt) {
    return align_up((size_t)nch * (size_t)nt * sizeof(float), 256);
}
// natural-spline table (nch, nt+1) of (y_k, dt^2/6 * y2_k), entry nt zero
int launch_spline_coefs(const float *rf, int64_t nch, int32_t nt, float dt, float *uscr,
                        float2 *coef, hipStream_t s);
int launch_spline_y2(const float *rf, int64_t nch, int32_t nt, float dt, float *uscr, float *y2,
                     hipStream_t s);
// Samples per lane (32 or 64) of the one-wave-per-channel spline sweeps for
// traces of nt samples, or 0 for the one-lane-per-channel sweeps (nt > 4096,
// or PYCBF_SPLINE_SEG=0 in the environment at the call).
int spline_seg_len(int32_t nt);
int launch_upsample(const float *rf, const float *y2, int64_t nch, int32_t nt, float t0, float dt,
                    int32_t usf, float *out, hipStream_t s);
int launch_korder_coefs(const float *rf, int64_t nch, int32_t nt, const float *S, int k,
                        float4 *coef, hipStream_t s);
int launch_akima_coefs(const float *rf, int64_t nch, int32_t nt, int modified, float4 *coef,
                       hipStream_t s);

// ------------------------------------------------------------------- DAS ---
// Gather kinds of the DAS kernel.
//   GK_NEAREST  one sample of the raw or upsampled trace
//   GK_LINEAR   two samples, weights a, b as the reference spline weights
//   GK_CUBIC    natural spline, table (y_j, y_j+1, h*y2_j, h*y2_j+1) evaluated
//               with the reference's separately rounded weights a and b
//   GK_KORDER   Hermite table, t = ((x - t0) - dt*j)/dt  (korder_cubic rounding)
//   GK_AKIMA    Hermite table, t = (x - (t0 + dt*j))/dt  (akima / makima rounding)
enum GatherKind { GK_NEAREST = 0, GK_LINEAR = 1, GK_CUBIC = 2, GK_KORDER = 3, GK_AKIMA = 4 };

struct DasArgs {
    const float *ttx, *atx;  // (ntx, np)
    const float *trx, *arx;  // (nrx, np)
    const int32_t *perm;     // (np) or null
    int64_t np;
    const float *src;        // gather source (raw rf, upsampled rf, or float4 coefficients)
    int64_t ch_stride;       // floats per channel in src
    int ntx, nrx, nfr;
    int nsrc;                // samples (or intervals) per channel
    float t0, dt, rdt, tmax; // time axis of src
    float thresh;
    int sumflag;
    int a_split;             // outer-axis elements per blockIdx.z
    int xcd_remap;           // 1: contiguous tile ranges per XCD
    int outer_tx;            // outer (per-block) axis is tx (else rx) for sums over both
    float *out;              // final output, or partial sums (nsplit, nfr, np)
    // LDS-staged gather only
    int geo;                 // 0: delay tables; 2 / 3: synthetic geometry of that ndim
    const float *txpack, *rxpack, *pvec;  // synthetic geometry (geo > 0)
    float c0;
    const float *bounds;     // (ntiles, ntx + nrx, 2) delay min/max per pixel tile
    int nent;                // table entries per channel row of src
    // direct gather only: float32 delay residuals (may be null); samples are
    // placed at (ttx + ttx_lo) + (trx + trx_lo) in double, exact weights
    const float *ttx_lo, *trx_lo;
    float rc0;  // 1 / c0 rounded (rdt above is 1 / dt rounded)
};

// True when, for every float x with 2^-60 <= |x| < 2^60,
//   q = x r, e = fma(-q, c, x), fma(e, r, q)   (r = 1 / c rounded)
// is the correctly rounded x / c (c itself within 2^-30 .. 2^30).  Decided
// by trying every significand of one binade (in that range no step
// underflows or overflows, so the steps scale exactly by powers of two);
// cached per divisor.  Host only.
bool recip_division_exact(float c);

int launch_das(int gk, const DasArgs &a, int nsplit, float *partial, hipStream_t s);
// LDS-staged gather (delay tables or in-register synthetic geometry); bounds
// must hold the tile delay ranges written by launch_tile_bounds.
int launch_das_lds(int gk, const DasArgs &a, int nsplit, float *partial, hipStream_t s);
// Single-frame natural-cubic gather with transmit delays in registers: used
// for one frame summed over tx and rx, or over tx only (das_reg.hip).
bool das_reg_applies(int gk, int64_t nfr, int sumflag);
int launch_das_reg(const DasArgs &a, hipStream_t s);
// Thread-per-sample gather with one float atomic per contributing sample
// (pycbf's own GPU launch scheme), the comparator selected by PYCBF_DAS_PLAIN=1.
bool das_use_plain();
int launch_das_plain(int gk, const DasArgs &a, hipStream_t s);
// Delay min/max per (pixel tile of 256, element): tx elements then rx.
int launch_tile_bounds(const DasArgs &a, float *bounds, hipStream_t s);
inline int64_t das_tiles(int64_t np) { return (np + 255) / 256; }
// out[g] = sum over z < nsplit of part[z*n + g]
int launch_reduce_split(const float *part, int nsplit, int64_t n, float *out, hipStream_t s);
// Whether the LDS-staged gather is used (PYCBF_DAS_LDS=0 selects the direct one).
bool das_use_lds();
// Frames handled per pass of the gather (1, 2, 4, 8 or 16; PYCBF_DAS_FC caps it).
int das_frames_per_pass(int64_t nfr);
// Whether transmits are the outer (one element per workgroup) axis.
bool das_outer_is_tx(int sumflag, int64_t ntx, int64_t nrx);
// Body of pycbf_das_synthetic on a caller buffer of `cap` bytes; with
// shrink_split the outer split's partial images shrink until the plan fits.
int das_synthetic_in(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                     const float *txpack, const float *rxpack, float c0, const float *pvec,
                     const int32_t *perm, int64_t np, int interp_kind, int interp_param,
                     const float *S, float thresh, int sumflag, float *out, void *buf,
                     size_t cap, bool shrink_split, hipStream_t s);
int launch_synth_tables(int ndim, int ntx, int nrx, const float *txpack, const float *rxpack,
                        float c0, const float *pvec, int64_t np, float *ttx, float *atx,
                        float *trx, float *arx, hipStream_t s);
int launch_debug_indices(int ntx, int nrx, const float *ttx, const float *trx, int64_t np,
                         int nearest, float t0, float dt, int n, int32_t *idx, hipStream_t s);

// -------------------------------------------------------------- backward ---
// Arguments of the Tabbed DAS adjoint kernels (das_bwd.hip).
struct BwdArgs {
    const float *ttx, *atx, *trx, *arx;  // (ntx, np), (nrx, np)
    int64_t np;
    int ntx, nrx, nfr, nt;
    const float *src;  // raw rf (nearest, linear) or the (y, e) table (cubic)
    float t0, dt, rdt, tmax, thresh;
    int sumflag;
    const float *gout;
    const float *bounds;  // (ntiles, ntx + nrx, 2) delay ranges of contributing pixels
    float *gtau, *gapod;  // outer-axis gradients (may be null)
    float *gy, *ge;       // channel-data cotangent bins (ge: cubic only); Hermite kinds
                          // (korder_cubic, akima, makima): cotangents of the
                          // coefficients c.x (gy), c.y (ge),
    float *gc2, *gc3;     // c.z and c.w (Hermite kinds only)
    int gstride;          // floats per channel row of gy / ge / gc2 / gc3
    const float *S;       // korder_cubic: (k, k) derivative operator
    int k;
    int modified;         // akima (0) / makima (1)
    const float *rf;      // akima / makima: the raw channel data (slopes are recomputed)
    unsigned *mbits;      // 3 words of workspace: max |gout|, |apodtx|, |apodrx| (bits)
    const float *ttx_lo, *trx_lo;  // delay residuals (may be null): the delay gradient
                                   // evaluates the interpolant's derivative at
                                   // (tautx + tautx_lo) + (taurx + taurx_lo) in double
    int geo;                       // 0: the tables above; 2 / 3: synthetic geometry of that
                                   // ndim, evaluated in registers (tables unused)
    const float *txpack, *rxpack, *pvec;  // synthetic geometry (geo > 0), pvec (np, geo)
    float c0, rc0;  // rc0 = 1 / c0 rounded
    double *gc0part;  // per-workgroup partial sums of d loss / d c0 (geo > 0), or null
    int gc0_off;      // first partial of this launch
};
// gc0 (geo > 0, may be null): d loss / d c0, one float, reduced in a fixed order
// from a.gc0part ((ntx + nrx) * das_tiles(np) doubles)
int launch_das_bwd(int gk, const BwdArgs &a, float *gtautx, float *gapodtx, float *gtaurx,
                   float *gapodrx, float *grf, float *gc0, hipStream_t s);
// nearest with usf > 1: cotangents gu of the upsampled trace (nch, usf (nt-1)
// + 1) -> spline-table cotangents gy, ge (nch, nt + 1) -> grf (nch, nt)
int launch_upsample_adjoint(const float *gu, int64_t nch, int32_t nt, int32_t usf, float t0,
                            float dt, float *gy, float *ge, float *grf, hipStream_t s);

// ------------------------------------------------------------ delay tabs ---
int launch_tabs_2d(int mode, int ntx, const double *txc, int nrx, const double *xele,
                   const double *xout, int64_t nx, const double *zout, int64_t nz, double cm,
                   double fnum, float *tautx, float *apodtx, float *taurx, float *apodrx,
                   hipStream_t s);

// ------------------------------------------------------------ demodulation ---
// C (R x N) = A (R x K) @ Bt (N x K)^T, row-major float32, float32-accurate
// (three-way BF16 split on the BF16 matrix units); part: demod_split(R, K, N)
// x R x N floats of partial products when that is above 1
int demod_split(int64_t R, int K, int N);
int launch_demod_gemm(const float *A, int64_t R, int K, const float *Bt, int N, float *C,
                      float *part, hipStream_t s);

// ------------------------------------------------------------------ DMAS ---
int launch_dmas(const float *im, int64_t nfr, int64_t nsep, int64_t np, const int32_t *lags,
                int nlag, int dmastype, int sumlags, float *out, hipStream_t s);

// ------------------------------------------------------- speed of sound ---
// slowness map geometry: node (i, k) at (xm0 + (k + 1/2) / rdxm, zm0 + (i + 1/2) / rdzm)
struct MapGeo {
    int nzm, nxm;
    float xm0, zm0, rdxm, rdzm;
};

struct RayArgs {
    const float *slow;  // (nzm, nxm)
    MapGeo g;
    const float *steer, *xorig, *xele;
    int ntx, nrx, nsteps;
    float c0;
    const float *pvec;  // (np, 2)
    int64_t np;
    double xm0d, zm0d, dxmd, dzmd;     // map geometry in double (forward)
    double rdxmd, rdzmd;               // 1 / dxm, 1 / dzm in double
    float *tautx, *taurx;              // forward outputs
    float *tautx_lo, *taurx_lo;        // forward residuals (may be null)
    const float *gtautx, *gtaurx;      // adjoint inputs
    float *gslow;                      // adjoint output
    float *gpart;                      // adjoint scratch: 256 B (max word), then the 64-bit
                                       // fixed-point map sums
};

int launch_ray_delays(const RayArgs &a, int adjoint, hipStream_t s);
int sos_max_map();
size_t ray_bwd_workspace_bytes(int nzm, int nxm);

// ------------------------------------------------------------------ SLSC ---
// lags is a host array; arguments validated by pycbf_slsc.
int launch_slsc_selection(const uint8_t *mask, int64_t nrx, int64_t nx, int64_t nz, int kwin,
                          int kofs, int dsfx, int dsfz, uint64_t *sel, hipStream_t s);
int launch_slsc(const float *im, int64_t nfr, int64_t nrx, int64_t nx, int64_t nz,
                const uint64_t *sel, int kwin, int kofs, int dsfx, int dsfz, const int32_t *lags,
                int nlag, int precision, float *out, hipStream_t s);

}  // namespace pycbf
