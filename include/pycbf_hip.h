/*
 * pycbf_hip.h -- C-ABI of libpycbf_hip.so, the MI355X (gfx950) beamforming
 * engine that replaces pycbf's CuPy-JIT kernel module.
 *
 * Conventions
 *   - Every entry point returns 0 on success, otherwise a nonzero status
 *     (a hipError_t value, or PYCBF_EINVAL for argument errors).  The message
 *     of the last failure on the calling thread is pycbf_last_error().
 *   - Array arguments are caller-owned DEVICE pointers unless the comment says
 *     "host".  Layouts are C-contiguous float32 as in pycbf.
 *   - `stream` is a hipStream_t (NULL = default stream).  Entry points enqueue
 *     work and return without synchronising.
 *   - Outputs are WRITTEN (not accumulated); no zero-filled output is needed.
 *   - `workspace` is caller-owned device scratch of at least the size returned
 *     by the matching *_workspace_bytes query; its contents are not preserved
 *     between calls.  Calls that share a workspace must not overlap.
 *
 * Reference interfaces replaced (paths relative to the wewightman/pycbf tree):
 *   pycbf/gpu/__engines__.py:9-26        CuPy RawModule + kernel handles
 *   pycbf/gpu/__engines__.cu:360-374     struct xInfo / RFInfo (same layout)
 *   pycbf/gpu/__engines__.cu:506-864     das_bmode_{synthetic,tabbed}_{korder_cubic,multi_interp}
 *   pycbf/gpu/__advanced_engines__.cu:17-61  calc_dmas
 *   examples/advanced_beamforming.ipynb:65-72,205-223  SLSC (calc_lagn_coherence + pixel loop)
 *   pycbf/cpu/pycbf.h:23-54              tie_knots1D_fixed / beamform_cubic / beamform_nearest
 */
#ifndef PYCBF_HIP_H
#define PYCBF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYCBF_ABI_VERSION 3
#define PYCBF_EINVAL 1001

/* interp kinds: 0..3 keep pycbf's interp_keys (pycbf/gpu/__core_bf_classes__.py:215,332) */
#define PYCBF_INTERP_NEAREST 0
#define PYCBF_INTERP_LINEAR 1
#define PYCBF_INTERP_AKIMA 2
#define PYCBF_INTERP_MAKIMA 3
#define PYCBF_INTERP_KORDER_CUBIC 4
#define PYCBF_INTERP_CUBIC 5

/* sumtype flags, pycbf/gpu/__core_bf_classes__.py:173,296 */
#define PYCBF_SUM_NONE 0
#define PYCBF_SUM_TX_ONLY 1
#define PYCBF_SUM_RX_ONLY 2
#define PYCBF_SUM_TX_AND_RX 3

/* DMAS kinds: 0 and 1 keep pycbf's dmasflag (pycbf/gpu/__advanced_bf_classes__.py:37-38) */
#define PYCBF_DMAS_SSR 0
#define PYCBF_DMAS_POWER 1
#define PYCBF_DMAS_PSSR 2

/* SLSC arithmetic: FP32 on the VALU, or FP16 / BF16 normalised channels on MFMA
 * with FP32 accumulation */
#define PYCBF_SLSC_FP32 0
#define PYCBF_SLSC_MFMA_FP16 1
#define PYCBF_SLSC_MFMA_BF16 2

/* The `thresh` that reproduces pycbf's GPU weight rule |apodtx*apodrx| > 0.01
 * (pycbf/gpu/__engines__.cu:8,564,678,762,851: a float compared with the
 * double 0.01) under this library's rule |w| >= thresh: the smallest float
 * above 0.01.  pycbf's CPU kernels use |w| >= 1e-2f (pycbf/cpu/pycbf.c:201,267),
 * the default of the Python classes. */
#define PYCBF_GPU_REFERENCE_THRESH 0.010000000707805157f

/* floats per transmit / receive element in the synthetic geometry packs */
#define PYCBF_TXPACK_FLOATS 16
#define PYCBF_RXPACK_FLOATS 12

/* struct xInfo, pycbf/gpu/__engines__.cu:360-364 (12 bytes) */
typedef struct pycbf_xinfo {
    float x0;   /* first sample time t0 */
    float dx;   /* sample period dt */
    int32_t nx; /* samples per channel nt */
} pycbf_xinfo;

/* struct RFInfo, pycbf/gpu/__engines__.cu:369-374 (32 bytes, same offsets) */
typedef struct pycbf_rfinfo {
    int64_t ntx;       /* transmit events */
    int64_t nrx;       /* receive channels */
    int32_t ndim;      /* 2 or 3 for Synthetic, 0 for Tabbed */
    pycbf_xinfo tinfo; /* time axis */
} pycbf_rfinfo;

/* ABI version (PYCBF_ABI_VERSION) and last error of the calling thread. */
int pycbf_abi_version(void);
const char *pycbf_last_error(void);

/* Bytes of workspace the DAS entry points need for this problem.
 * interp_param is usf for nearest, k for korder_cubic, ignored otherwise.
 * synthetic != 0 sizes pycbf_das_synthetic (it adds the delay tables). */
size_t pycbf_das_workspace_bytes(const pycbf_rfinfo *info, int64_t nfr, int64_t np,
                                 int interp_kind, int interp_param, int synthetic);

/* Tabbed DAS.  Replaces das_bmode_tabbed_multi_interp (pycbf/gpu/__engines__.cu:813-864)
 * and das_bmode_tabbed_korder_cubic (:726-778), with the CPU kernels' semantics for
 * cubic / nearest(usf) (pycbf/cpu/pycbf.c:179-271): weight w = apodtx*apodrx, a
 * sample contributes w*interp(rf[f,tx,rx,:], tautx+taurx) when |w| >= thresh.
 *   rf       (nfr, ntx, nrx, nt)
 *   tautx, apodtx (ntx, np); taurx, apodrx (nrx, np)
 *   perm     optional (np) int32 pixel processing order (NULL = identity); it
 *            only changes scheduling, never the result layout
 *   S        (k, k) derivative operator for korder_cubic, else NULL
 *   out      (nfr, ntx, nrx, np) | (nfr, nrx, np) | (nfr, ntx, np) | (nfr, np) by sumflag */
int pycbf_das_tabbed(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                     const float *tautx, const float *apodtx, const float *taurx,
                     const float *apodrx, const int32_t *perm, int64_t np, int interp_kind,
                     int interp_param, const float *S, float thresh, int sumflag, float *out,
                     void *workspace, size_t workspace_bytes, void *stream);

/* pycbf_das_tabbed with float32 residuals of the delays (no reference
 * counterpart: for delays computed beyond float32, as pycbf_ray_delays writes
 * them for the speed-of-sound chain).  Each sample is interpolated at
 * (tautx + tautx_lo) + (taurx + taurx_lo) summed in double, on the piece
 * j = floor(q) with exact weights b = q - j, a = 1 - b; the weight threshold
 * and the trace bounds follow the float32 delays.  Either residual table may
 * be NULL (zero), not both.  Runs the direct-gather kernel; other arguments
 * and the output layout as pycbf_das_tabbed. */
int pycbf_das_tabbed_residual(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                              const float *tautx, const float *apodtx, const float *taurx,
                              const float *apodrx, const float *tautx_lo, const float *taurx_lo,
                              int64_t np, int interp_kind, int interp_param, const float *S,
                              float thresh, int sumflag, float *out, void *workspace,
                              size_t workspace_bytes, void *stream);

/* HOST function.  Builds the per-element geometry packs consumed on the device
 * from pycbf's Synthetic parameters (pycbf/__bf_base_classes__.py:141-183, all
 * host arrays).  The acceptance tests "acos(v) > ala" of
 * pycbf/gpu/__engines__.cu:412,427,468 become exact comparisons v < T with T
 * derived from ala in double precision; each receive pack also carries, as a
 * double in floats 8-9, the midpoint between T and the float below it, with
 * which the receive test runs as proj < m * mag without a division.
 *   txpack: ntx * PYCBF_TXPACK_FLOATS floats, rxpack: nrx * PYCBF_RXPACK_FLOATS. */
int pycbf_synthetic_pack(int ndim, int64_t ntx, const float *ovectx, const float *nvectx,
                         const float *t0tx, const float *alatx, const float *doftx, int64_t nrx,
                         const float *ovecrx, const float *nvecrx, const float *alarx,
                         float *txpack, float *rxpack);

/* Synthetic-point DAS.  Replaces das_bmode_synthetic_multi_interp
 * (pycbf/gpu/__engines__.cu:620-694) and das_bmode_synthetic_korder_cubic
 * (:506-581, whose k is passed as an int here).  Delays and apodisations come
 * from calc_tautx_apodtx / calc_taurx_apodrx (:382-470) evaluated on the device.
 *   txpack, rxpack  device copies of pycbf_synthetic_pack output
 *   pvec            (np, ndim) pixel positions; other arguments as pycbf_das_tabbed */
int pycbf_das_synthetic(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                        const float *txpack, const float *rxpack, float c0, const float *pvec,
                        const int32_t *perm, int64_t np, int interp_kind, int interp_param,
                        const float *S, float thresh, int sumflag, float *out, void *workspace,
                        size_t workspace_bytes, void *stream);

/* No-allocation variant of pycbf_das_synthetic.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as pycbf_das_synthetic without the workspace:
 * the library uses its own device array of PYCBF_NOALLOC_SCRATCH_BYTES per GPU
 * (the calling thread's current device, which the stream must belong to),
 * held until the process ends, and calls no allocator.  When the outer
 * split's partial images do not fit, fewer partials are used (the same sums in
 * another order); a problem whose spline table, tile ranges and tables still
 * exceed the array returns PYCBF_EINVAL with the size it needs.  Example: one
 * frame of 31 x 128 channels x 4096 samples, cubic, fits (132 MB); one frame
 * of 75 x 128 x 4096 cubic does not (315 MB of spline table).  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU race. */
#define PYCBF_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int pycbf_das_synthetic_noalloc(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                                const float *txpack, const float *rxpack, float c0,
                                const float *pvec, const int32_t *perm, int64_t np,
                                int interp_kind, int interp_param, const float *S, float thresh,
                                int sumflag, float *out, void *stream);

/* Delay / apodisation tables from the synthetic geometry (the per-sample
 * calc_tautx_apodtx / calc_taurx_apodrx of pycbf/gpu/__engines__.cu:542-561). */
int pycbf_synthetic_tables(const pycbf_rfinfo *info, const float *txpack, const float *rxpack,
                           float c0, const float *pvec, int64_t np, float *tautx, float *apodtx,
                           float *taurx, float *apodrx, void *stream);

/* Delay tabs and apodisations of a 2-D linear array, written on the device:
 * the per-pixel part of make_tabs_and_apods_2D
 * (pycbf/helpers/__full_aperture_2D__.py:3-203), in float64 with the
 * reference's operation order, rounded to float32.  Pixels p = ix*nz + iz of
 * meshgrid(xout, zout, indexing='ij').  txc (device, ntx x 8 doubles) holds
 * the per-transmit constants the host derives with the reference's NumPy
 * calls:
 *   PYCBF_TABS_PLANE      {ox, oz, nx, nz, nperp_x, nperp_z, xele[0], xele[-1]}
 *                         (steering corrected for ctx, :150-160)
 *   PYCBF_TABS_DIVERGENT  {ox, oz, tau0}  (:62-76)
 *   PYCBF_TABS_FOCUSED    {ox, oz, nx, nz, t0, vstar}: the synthetic point of
 *                         :270-314, and arccos(v) < ala  <=>  vstar <= v <= 1
 * xele (nrx), xout (nx), zout (nz): device doubles.  Outputs (float32,
 * device) tautx / apodtx (ntx, nx*nz), taurx / apodrx (nrx, nx*nz); any may
 * be NULL.  Apodisations are 1.0 / 0.0 (the reference returns bool). */
#define PYCBF_TABS_PLANE 0
#define PYCBF_TABS_DIVERGENT 1
#define PYCBF_TABS_FOCUSED 2
int pycbf_make_tabs_2d(int mode, int64_t ntx, const double *txc, int64_t nrx, const double *xele,
                       const double *xout, int64_t nx, const double *zout, int64_t nz, double cm,
                       double fnum, float *tautx, float *apodtx, float *taurx, float *apodrx,
                       void *stream);

/* FFT-window demodulation of beamformed RF to IQ (fft_tukey_demod of
 * examples/SyntheticVsTabbed.ipynb, cell 1, along the last axis) as one
 * real product on the device: iq[r][c] = sum_k rf[r][k] * opT[c][k], with rf
 * (rows, n), opT (ncol, n) and iq (rows, ncol), float32; float32-accurate
 * (three-way BF16 split of both operands on the BF16 matrix units, float32
 * accumulation).  opT is the transposed demodulation operator the host
 * builds once per (n, band, window): row 2m holds Re and row 2m+1 Im of
 * output sample m, so iq reads as complex64 (rows, ncol/2).  All pointers
 * are device pointers, rows row-major and dense; ncol <= 128 * 65535. */
int pycbf_demod_apply(const float *rf, int64_t rows, int64_t n, const float *opT, int64_t ncol,
                      float *iq, void *workspace, size_t workspace_bytes, void *stream);

/* The same demodulation by FFTs, for n a power of two in 8..4096: one
 * workgroup per line (no reference counterpart beyond the notebook's NumPy
 * chain rfft -> band -> Tukey window -> ifftshift -> ifft, which it follows).
 * Per line: the n-point real FFT (an n/2-point complex FFT and the real
 * split), the nbin kept bins from ikmin (bin ikmin + (k + shift) mod nbin
 * at position k: shift = nbin / 2 for the ifftshift, 0 without), and their
 * nbin-point inverse DFT by Bluestein's chirp product over a circular
 * convolution of nconv points (a power of two >= 2 nbin - 1, <= 2048).
 * tables (device, float32 pairs, built by the host in float64):
 *   pre[nbin]   window[(k + shift) mod nbin] * exp(i pi k^2 / nbin)
 *   post[nbin]  exp(i pi t^2 / nbin) / (nbin * nconv)
 *   bhat[nconv] FFT of the chirp b[d] = exp(-i pi d^2 / nbin), |d| < nbin,
 *               placed at d mod nconv
 *   split[nbin] exp(-2 pi i kb / n) of the bin kb = ikmin + (k + shift) mod nbin
 *   tw[T]       exp(-2 pi i t / T), T = max(n / 2, nconv)
 * rf (rows, n) and iq (rows, 2 nbin) as pycbf_demod_apply (Re, Im pairs);
 * the line is read once and its IQ written once.  8-byte aligned pointers. */
int pycbf_demod_fft(const float *rf, int64_t rows, int64_t n, int64_t ikmin, int64_t nbin,
                    int64_t shift, int64_t nconv, const float *tables, float *iq, void *stream);

/* Bytes of device workspace pycbf_demod_apply needs for rows lines of n
 * samples and ncol outputs: 0 when the output tiles fill the GPU, else the
 * partial products of a split over the samples (few lines, long n).  No
 * reference counterpart (the notebooks' FFT chain needs none). */
size_t pycbf_demod_workspace_bytes(int64_t rows, int64_t n, int64_t ncol);

/* HOST function.  Spatially compact processing order for np pixel positions
 * (host pvec (np, ndim)) written to host perm (np).  No reference counterpart:
 * the reference launches one thread per sample in pixel-index order. */
int pycbf_tile_order(const float *pvec, int64_t np, int ndim, int32_t *perm);

/* Natural cubic spline of each channel (tie_knots1D_fixed, pycbf/cpu/pycbf.c:114-166)
 * as the gather table coef (nch, nt+1, 2): entry k = (y_k, dt^2/6 * y2_k) with y2
 * bit-identical to the reference, entry nt = (0, 0) (zero padding of the
 * neighbour the reference reads out of range).  On interval j the value is
 * a*y_j + b*y_j+1 + (a^3-a)*e_j + (b^3-b)*e_j+1 with e the second component. */
int pycbf_spline_coefs(const float *rf, int64_t nch, int32_t nt, float dt, float *coef,
                       void *workspace, size_t workspace_bytes, void *stream);

/* Natural spline second derivatives y2 (nch, nt), bit-identical to tie_knots1D_fixed. */
int pycbf_spline_y2(const float *rf, int64_t nch, int32_t nt, float dt, float *y2,
                    void *workspace, size_t workspace_bytes, void *stream);

/* Workspace bytes for pycbf_spline_coefs / pycbf_spline_y2. */
size_t pycbf_spline_workspace_bytes(int64_t nch, int32_t nt);

/* Sample index each (tx, rx, px) selects, for parity tests of index selection:
 * idx (ntx, nrx, np) int32, -1 where the sample falls outside the trace.
 * kind nearest uses pycbf/cpu/pycbf.c:264 (with dt/usf), the others the knot
 * index of pycbf/cpu/pycbf.c:81. */
int pycbf_debug_indices(const pycbf_rfinfo *info, const float *tautx, const float *taurx,
                        int64_t np, int interp_kind, int usf, int32_t *idx, void *stream);

/* DMAS over the separated axis.  Replaces calc_dmas (pycbf/gpu/__advanced_engines__.cu:17-61)
 * but uses the lag values in `lags` (the reference used 0..nlag-1).
 *   im  (nfr, nsep, np) delayed, partially summed images
 *   out (nfr, np) if sumlags else (nfr, nlag, np)
 *   dmastype PYCBF_DMAS_SSR: sum sign(q)sqrt|q|; POWER: sum q; PSSR: sign(S)sqrt|S| of the power sum */
int pycbf_dmas(const float *im, int64_t nfr, int64_t nsep, int64_t np, const int32_t *lags,
               int nlag, int dmastype, int sumlags, float *out, void *stream);

/* Bytes of workspace pycbf_das_tabbed_bwd and pycbf_das_synthetic_bwd need
 * (cubic: spline table and
 * cotangent bins of every channel; korder_cubic: Hermite table and the
 * cotangents of its four coefficients, as for akima and makima; nearest with
 * param = usf > 1: the upsampled trace and its cotangents; all kinds:
 * per-tile delay ranges). */
size_t pycbf_das_bwd_workspace_bytes(const pycbf_rfinfo *info, int64_t nfr, int64_t np,
                                     int interp_kind, int param);

/* Adjoint of pycbf_das_tabbed (no reference counterpart: pycbf is forward
 * only; the forward it differentiates is pycbf/cpu/pycbf.c:179-271 with the
 * layouts of pycbf/gpu/__engines__.cu:813-864).  Given gout, the gradient of
 * a loss with respect to out (layout by sumflag as pycbf_das_tabbed), writes
 *   grf      (nfr, ntx, nrx, nt)  d loss / d rf
 *   gtautx, gapodtx (ntx, np); gtaurx, gapodrx (nrx, np)
 * Any output pointer may be NULL (not computed).  interp_kind nearest
 * (param = usf; its delay gradient is 0; with usf > 1 the adjoint of the
 * spline upsampling is applied to grf), linear, cubic (natural spline:
 * the adjoint of the spline prefilter is applied to grf) or korder_cubic
 * (param = k, S the (k, k) operator as for pycbf_das_tabbed; the adjoint of
 * the Hermite coefficient construction is applied to grf), akima or makima
 * (the adjoint through the Akima derivatives, whose weights |.| differentiate
 * to sign(.) with sign(0) = 0).  Samples with |apodtx*apodrx| < thresh
 * contribute nothing, as in the forward.  tautx_lo (ntx, np) and taurx_lo
 * (nrx, np) are optional (NULL) float32 residuals of the delays (as
 * pycbf_ray_delays writes them): with them the delay gradients take the
 * interpolant's derivative at (tautx + tautx_lo) + (taurx + taurx_lo) summed
 * in double with exact weights, as pycbf_das_tabbed_residual's forward
 * (sample selection follows the float32 delays; the channel-data and
 * apodisation gradients are those of pycbf_das_tabbed's float32 evaluation). */
int pycbf_das_tabbed_bwd(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                         const float *tautx, const float *apodtx, const float *taurx,
                         const float *apodrx, const float *tautx_lo, const float *taurx_lo,
                         int64_t np, int interp_kind, int param, const float *S, float thresh,
                         int sumflag, const float *gout, float *grf, float *gtautx,
                         float *gapodtx, float *gtaurx, float *gapodrx, void *workspace,
                         size_t workspace_bytes, void *stream);

/* Adjoint of pycbf_das_synthetic (no reference counterpart; the forward it
 * differentiates is pycbf/gpu/__engines__.cu:382-470 (geometry) with
 * :620-694 (synthetic DAS)).  Same gradients as pycbf_das_tabbed_bwd, but no
 * delay or apodisation table exists: every kernel re-evaluates the delays and
 * weights of its (element, pixel) pairs in registers from the packs
 * (pycbf_synthetic_pack) and the pixel positions pvec (np, ndim), with the
 * forward's float32 operations, so grf, gtautx, gapodtx, gtaurx and gapodrx
 * equal what pycbf_das_tabbed_bwd gives on pycbf_synthetic_tables' tables.
 * gc0 (device, one float; NULL: not computed) receives d loss / d c0 for a
 * geometry whose every delay is a path length over c0, the transmit time
 * intercept t0tx included (pycbf's helpers compute it as a distance over the
 * sound speed, pycbf/helpers/__full_aperture_2D__.py:259,300,346): the sum
 * over (element, pixel) of the delay gradient times d tau / d c0 = -tau / c0,
 * accumulated in double in a fixed order (deterministic).  The weights depend
 * on angles only (their gradient is zero).  Pixels are taken in their natural
 * order (no processing order).  Workspace: pycbf_das_bwd_workspace_bytes. */
int pycbf_das_synthetic_bwd(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                            const float *txpack, const float *rxpack, float c0, const float *pvec,
                            int64_t np, int interp_kind, int param, const float *S, float thresh,
                            int sumflag, const float *gout, float *grf, float *gtautx,
                            float *gapodtx, float *gtaurx, float *gapodrx, float *gc0,
                            void *workspace, size_t workspace_bytes, void *stream);

/* Straight-ray delays through a slowness map (no reference counterpart: the
 * reference's examples/rough_sos_correction.ipynb uses an external eikonal
 * solver).  slow (nzm, nxm) device map of 1/c with node (i, k) at
 * (xm0 + (k+1/2) dxm, zm0 + (i+1/2) dzm), bilinear, edge-extended.
 * Plane-wave transmits (steer radians, origin x xorig, both device (ntx)):
 *   tautx = (q_x - xorig) sin(steer) / c0 + integral of slow along the ray
 *   from q = (x - z tan(steer), 0) to the pixel;
 * receive elements at (xele, 0) (device (nrx)): taurx = integral of slow from
 * the element to the pixel.  Midpoint rule with nsteps samples per ray.
 * pvec (np, 2) device pixel positions (x, z); tautx (ntx, np), taurx (nrx, np).
 * Each ray is evaluated in double from the float32 inputs and rounded to
 * float32; tautx_lo / taurx_lo (same shapes, optional: NULL) receive the
 * residuals delay - float32(delay) for pycbf_das_tabbed_bwd.
 * The map has at most 8192 nodes. */
int pycbf_ray_delays(const float *slow, int nzm, int nxm, float xm0, float zm0, float dxm,
                     float dzm, const float *steer, const float *xorig, int ntx, float c0,
                     const float *xele, int nrx, const float *pvec, int64_t np, int nsteps,
                     float *tautx, float *taurx, float *tautx_lo, float *taurx_lo,
                     void *stream);

/* Workspace bytes of pycbf_ray_delays_bwd (256 bytes, then one 64-bit
 * fixed-point sum per map node). */
size_t pycbf_ray_delays_bwd_workspace_bytes(int nzm, int nxm);

/* Adjoint of pycbf_ray_delays: gslow (nzm, nxm) = d loss / d slow given
 * gtautx (ntx, np) and gtaurx (nrx, np) (either may be NULL).  gslow is
 * written.  The map sums are accumulated in 64-bit fixed point (scaled by the
 * largest |g| * path length), so the result does not depend on scheduling;
 * a non-finite upstream gradient gives NaN in every node. */
int pycbf_ray_delays_bwd(const float *slow, int nzm, int nxm, float xm0, float zm0, float dxm,
                         float dzm, const float *steer, const float *xorig, int ntx, float c0,
                         const float *xele, int nrx, const float *pvec, int64_t np, int nsteps,
                         const float *gtautx, const float *gtaurx, float *gslow, void *workspace,
                         size_t workspace_bytes, void *stream);

/* Short-lag spatial coherence.  Replaces calc_lagn_coherence and the pixel loop
 * of examples/advanced_beamforming.ipynb:65-72,205-223 (the reference has no
 * GPU SLSC).
 *   im    (nfr, nrx, nx, nz) per-receive-channel images (DAS with sumtype
 *         tx_only over an x-major grid of nx * nz pixels)
 *   sel   (ceil(nx/dsfx), ceil(nz/dsfz), 4) words from pycbf_slsc_selection,
 *         or NULL (every element usable at every pixel)
 *   out   (nfr, ceil(nx/dsfx), ceil(nz/dsfz)); output (ixo, izo) is pixel
 *         (ixo*dsfx, izo*dsfz) with axial window [iz+kofs, iz+kofs+kwin-1]
 *         clipped to the image.  The notebook's window iz-nkern..iz+nkern is
 *         kwin = 2*nkern+1, kofs = -nkern.
 * With fewer than max(lags)+1 usable elements the output is 0.  Per lag, the
 * pair-normalised window correlation is averaged over pairs of usable elements
 * that lag apart in the list of usable elements (pairs of an element with
 * standard deviation < 1e-10 are dropped; a lag with no pair left is NaN); the
 * output is the mean over lags.
 *   lags  HOST array of nlag (1..64) lag values in 1..63
 *   limits: 1 <= kwin <= 32, -(kwin-1) <= kofs <= 0, nrx <= 256, dsfx, dsfz >= 1
 *   precision  PYCBF_SLSC_FP32 / _MFMA_FP16 / _MFMA_BF16 */
int pycbf_slsc(const float *im, int64_t nfr, int64_t nrx, int64_t nx, int64_t nz,
               const uint64_t *sel, int kwin, int kofs, int dsfx, int dsfz, const int32_t *lags,
               int nlag, int precision, float *out, void *stream);

/* Usable elements per SLSC output pixel (the notebook's element selection,
 * examples/advanced_beamforming.ipynb:211-214): bit r of word r/64 of
 * sel[ixo, izo] is set when mask[r, ix, z] is non-zero for every z of the
 * pixel's window.  mask (nrx, nx, nz) bytes (non-zero receive apodisation);
 * window arguments as pycbf_slsc.  The selection depends on the geometry
 * only, so it is computed once and reused for every frame. */
int pycbf_slsc_selection(const uint8_t *mask, int64_t nrx, int64_t nx, int64_t nz, int kwin,
                         int kofs, int dsfx, int dsfz, uint64_t *sel, void *stream);

/* ------------------------------------------------------------ batches ---
 * Ragged batches of Synthetic (2-D, plane-wave packs from pycbf_synthetic_pack)
 * cubic DAS and SLSC requests in one persistent gather launch (no reference
 * counterpart: pycbf beamforms one configuration per call,
 * pycbf/gpu/__core_bf_classes__.py:68-118).  All requests share nt and dt.
 * The channel data of request r is the r-th block of one device buffer
 * rf_all (requests in order, each (ntx, nrx, nt)); DAS outputs must be laid
 * out back to back in request order. */
#define PYCBF_BATCH_DAS 0  /* image summed over transmits and receives, out (nx*nz) */
#define PYCBF_BATCH_SLSC 1 /* per-receive images into imrx (nrx, nx*nz), then SLSC into out */

typedef struct pycbf_batch_req {
    int32_t kind, ntx, nrx, nx, nz, nt; /* x-major grid of nx * nz pixels */
    float t0, dt, c0, thresh;
    const float *txpack, *rxpack; /* device packs */
    const float *pvec;            /* device (nx*nz, 2) pixel positions */
    float *out;                   /* device output (DAS image, or SLSC image) */
    float *imrx;                  /* device (nrx, nx*nz) scratch of SLSC requests */
    const uint64_t *sel;          /* SLSC: pycbf_slsc_selection bits or NULL */
    int32_t kwin, kofs, dsfx, dsfz, nlag, precision; /* SLSC arguments as pycbf_slsc */
    int32_t lags[64];
} pycbf_batch_req;

typedef struct pycbf_batch pycbf_batch; /* host-side plan of a batch */

/* Validates the requests and builds the host plan (copied from reqs). */
int pycbf_batch_create(const pycbf_batch_req *reqs, int nreq, pycbf_batch **batch);
void pycbf_batch_destroy(pycbf_batch *batch);
/* Device workspace bytes a batch needs (spline tables of every channel, the
 * request table, the work items of the largest configuration, counters). */
size_t pycbf_batch_workspace_bytes(const pycbf_batch *batch);
/* Picks pixel tile (8, 16 or 32 square) and receive chunk (16, 32, 64 or 128) and
 * uploads the request table and the cost-ordered work items; synchronises. */
int pycbf_batch_configure(pycbf_batch *batch, int tile, int rchunk, void *workspace,
                          size_t workspace_bytes, void *stream);
/* Enqueues one run (kernels only, capturable in a graph): counter / output
 * reset, spline prefilter of all channels, the persistent gather, and the
 * SLSC of the SLSC requests. */
int pycbf_batch_run(pycbf_batch *batch, const float *rf_all, void *workspace,
                    size_t workspace_bytes, void *stream);
/* Times the persistent gather for every tile (8, 16, 32) and receive chunk
 * (16, 32, 64, 128) (reps runs each, after one full run that builds the
 * tables), configures the fastest and returns it; times_ms (12 entries,
 * tile-major) may be NULL.  Synchronises. */
int pycbf_batch_autotune(pycbf_batch *batch, const float *rf_all, int reps, void *workspace,
                         size_t workspace_bytes, void *stream, int *best_tile, int *best_rchunk,
                         float *times_ms);
/* Completion time of every request, in ms after the start of the last run
 * (device wall clock); synchronous readback. */
int pycbf_batch_latencies(const pycbf_batch *batch, const void *workspace,
                          size_t workspace_bytes, double *latency_ms);

/* Measured read bandwidth of the device in GB/s (no reference counterpart:
 * the roofline denominators of the benchmark).  kind 0: one streaming read of
 * buf (bytes, larger than the caches) per rep; kind 1: 8192 workgroups each
 * read the region buf (bytes, small enough for one XCD's L2) 8 times; kind 2:
 * LDS reads of 16 bytes per lane, conflict-free (buf unused).  sink: device
 * scratch of at least 8192 * 256 * 4 bytes.  One untimed warm-up launch, then
 * reps timed launches with HIP events on stream; synchronises. */
int pycbf_measure_bandwidth(int kind, const void *buf, size_t bytes, int reps, void *sink,
                            size_t sink_bytes, double *gbps, void *stream);

/* Per-kernel timing: when enabled, instrumented launches record HIP events on
 * their stream; pycbf_profile_read synchronises them and returns the summed
 * duration (ms) and launch count of kernel `kernel` ("das_gather",
 * "spline_coefs", "synthetic_tables", "slsc_fp32", "slsc_mfma_fp16", ...).  No reference counterpart. */
void pycbf_profile_enable(int on);
void pycbf_profile_reset(void);
int pycbf_profile_read(const char *kernel, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* PYCBF_HIP_H */
