/*
 * pycbf_oracle.c -- CPU restatement of pycbf's beamforming arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (pycbf_amd) never calls it.
 *
 * Every function restates the algorithm of a reference source location
 * (paths relative to the wewightman/pycbf tree) in code written for this
 * project.  Floating point follows the reference's float32 operation order;
 * the file is compiled with -ffp-contract=off so no multiply-add is fused,
 * matching the reference wheel build (x86-64 baseline ISA, no FMA).
 *
 * Deliberate, documented departures from the reference:
 *   - Out-of-range neighbour reads (pycbf/cpu/pycbf.c:52,103 read y[N], y2[N]
 *     when the knot index is N-1) are replaced by zero padding.
 *   - A NaN delay is treated as out of range (the reference converts it to
 *     an int, which is undefined behaviour).
 *   - The synthetic-point acceptance tests "acos(v) > ala"
 *     (pycbf/gpu/__engines__.cu:412,427,468) are evaluated as the exact real
 *     predicate via acos() in double precision; the CUDA float acos is closed
 *     and cannot be reproduced.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_KIND_NEAREST 0
#define ORC_KIND_LINEAR 1
#define ORC_KIND_CUBIC 2
#define ORC_KIND_KORDER 3

#define ORC_SUM_NONE 0
#define ORC_SUM_TX_ONLY 1
#define ORC_SUM_RX_ONLY 2
#define ORC_SUM_TX_AND_RX 3

/* ------------------------------------------------------------------------ */
/* Natural cubic spline (Numerical Recipes "spline" with zero end moments).  */
/* Restates tie_knots1D_fixed, pycbf/cpu/pycbf.c:114-166: forward sweep of   */
/* the tridiagonal system with sig = 0.5, then back substitution.  u is an   */
/* n-element scratch buffer supplied by the caller (the reference mallocs).  */
/* ------------------------------------------------------------------------ */
void orc_spline_y2(const float *y, int n, float dx, float *y2, float *u)
{
    const float sig = 0.5f;
    if (n <= 0) return;
    y2[0] = 0.0f;
    u[0] = 0.0f;
    for (int i = 1; i < n - 1; ++i) {
        float p = sig * y2[i - 1] + 2.0f;
        y2[i] = (sig - 1.0f) / p;
        float slope_jump = (y[i + 1] - y[i]) / dx - (y[i] - y[i - 1]) / dx;
        u[i] = (6.0f * slope_jump / (2.0f * dx) - sig * u[i - 1]) / p;
    }
    y2[n - 1] = 0.0f;
    for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

/* Data-independent forward coefficients of the sweep above (y2 before back
 * substitution).  Exposed so tests can pin the GPU prefilter's shortcut that
 * these reach an exact float32 fixed point. */
void orc_spline_forward_coef(int n, float *y2f)
{
    const float sig = 0.5f;
    if (n <= 0) return;
    y2f[0] = 0.0f;
    for (int i = 1; i < n; ++i) {
        float p = sig * y2f[i - 1] + 2.0f;
        y2f[i] = (sig - 1.0f) / p;
    }
}

/* Knot index used by cubic and linear evaluation: j = (int)((x - x0)/dx) with
 * fill outside [x0, x0 + dx*(n-1)].  pycbf/cpu/pycbf.c:75-87.  Returns -1 when
 * out of range (the caller then uses the fill value 0). */
int orc_index_floor(float x, float x0, float dx, int n)
{
    float xmax = x0 + dx * (float)(n - 1);
    if (!(x == x)) return -1;
    if ((x < x0) || (x > xmax)) return -1;
    return (int)((x - x0) / dx);
}

/* Nearest-sample index, pycbf/cpu/pycbf.c:264,267: the +0.5 is added in
 * double and the result truncated toward zero; valid when 0 <= it < n.
 * Returns -1 when invalid. */
int orc_index_nearest(float x, float x0, float dx, int n)
{
    float q = (x - x0) / dx;
    double d = (double)q + 0.5;
    if (!(d > -1.0) || !(d < (double)n)) return -1;
    return (int)d;
}

/* Cubic spline value at x, pycbf/cpu/pycbf.c:68-105 (powf kept). */
float orc_cubic_point(const float *y, const float *y2, int n, float x0, float dx, float x)
{
    int j = orc_index_floor(x, x0, dx, n);
    if (j < 0) return 0.0f;
    float xlead = x0 + dx * (float)(j + 1);
    float xlag = x0 + dx * (float)j;
    float a = (xlead - x) / dx;
    float b = (x - xlag) / dx;
    float c = (1.0f / 6.0f) * (powf(a, 3.0f) - a) * powf(dx, 2.0f);
    float d = (1.0f / 6.0f) * (powf(b, 3.0f) - b) * powf(dx, 2.0f);
    float y_hi = (j + 1 < n) ? y[j + 1] : 0.0f;
    float y2_hi = (j + 1 < n) ? y2[j + 1] : 0.0f;
    return a * y[j] + b * y_hi + c * y2[j] + d * y2_hi;
}

/* Cubic spline at many points, pycbf/cpu/pycbf.c:14-56. */
void orc_cubic_many(const float *y, const float *y2, int n, float x0, float dx,
                    const float *xin, int nin, float *yout)
{
    for (int i = 0; i < nin; ++i) yout[i] = orc_cubic_point(y, y2, n, x0, dx, xin[i]);
}

/* Project-defined linear interpolation (the reference GPU "linear" returns
 * y[floor], pycbf/gpu/__engines__.cu:52).  Same knot index and fill rule as
 * cubic; a and b exactly as the cubic weights (pycbf/cpu/pycbf.c:93-98). */
float orc_linear_point(const float *y, int n, float x0, float dx, float x)
{
    int j = orc_index_floor(x, x0, dx, n);
    if (j < 0) return 0.0f;
    float xlead = x0 + dx * (float)(j + 1);
    float xlag = x0 + dx * (float)j;
    float a = (xlead - x) / dx;
    float b = (x - xlag) / dx;
    float y_hi = (j + 1 < n) ? y[j + 1] : 0.0f;
    return a * y[j] + b * y_hi;
}

/* k-order cubic (Praesius & Jensen local spline derivatives through a k x k
 * operator S), pycbf/gpu/__engines__.cu:292-355.  Fill 0 outside the trace. */
float orc_korder_point(const float *y, int nx, float x0, float dx,
                       const float *S, int k, float x)
{
    if (!(x == x)) return 0.0f;
    float xmax = x0 + dx * (float)(nx - 1);
    if ((x < x0) || (x > xmax)) return 0.0f;
    int ix = (int)((x - x0) / dx);
    if (ix > nx - 1) return 0.0f;
    if (ix == nx - 1) return y[nx - 1];
    int ik0, iy0;
    if (ix < k / 2) {
        ik0 = ix;
        iy0 = 0;
    } else if (ix > nx - k / 2 - 1) {
        ik0 = k - nx + ix;
        iy0 = nx - k;
    } else {
        ik0 = (k - 1) / 2;
        iy0 = ix - (k - 1) / 2;
    }
    float y0 = y[ix], y1 = y[ix + 1];
    float d0 = 0.0f, d1 = 0.0f;
    for (int i = 0; i < k; ++i) {
        d0 += S[ik0 * k + i] * y[iy0 + i];
        d1 += S[(ik0 + 1) * k + i] * y[iy0 + i];
    }
    float t = (x - x0 - dx * (float)ix) / dx;
    float c2 = 3.0f * (y1 - y0) - 2.0f * d0 - d1;
    float c3 = 2.0f * (y0 - y1) + d0 + d1;
    return y0 + d0 * t + c2 * t * t + c3 * t * t * t;
}

/* ------------------------------------------------------------------------ */
/* One (tx, rx) channel accumulated into out over nout pixels.              */
/* ------------------------------------------------------------------------ */

/* beamform_cubic, pycbf/cpu/pycbf.c:179-207. */
void orc_bf_cubic(float t0, float dt, int nt, const float *sig, int nout, float thresh,
                  const float *tautx, const float *apodtx, const float *taurx,
                  const float *apodrx, float *out)
{
    float *y2 = (float *)malloc(sizeof(float) * (size_t)(nt > 0 ? nt : 1));
    float *u = (float *)malloc(sizeof(float) * (size_t)(nt > 0 ? nt : 1));
    orc_spline_y2(sig, nt, dt, y2, u);
    for (int i = 0; i < nout; ++i) {
        float w = apodtx[i] * apodrx[i];
        if (fabs((double)w) >= (double)thresh)
            out[i] += w * orc_cubic_point(sig, y2, nt, t0, dt, tautx[i] + taurx[i]);
    }
    free(u);
    free(y2);
}

/* Upsample one channel onto a dt/usf grid through the natural spline, the
 * first half of beamform_nearest, pycbf/cpu/pycbf.c:234-253.  out has
 * usf*(nt-1)+1 samples. */
void orc_upsample(float t0, float dt, int nt, const float *sig, int usf, float *out)
{
    int nt_usf = usf * (nt - 1) + 1;
    float dt_usf = dt / (float)usf;
    float *y2 = (float *)malloc(sizeof(float) * (size_t)nt);
    float *u = (float *)malloc(sizeof(float) * (size_t)nt);
    orc_spline_y2(sig, nt, dt, y2, u);
    for (int it = 0; it < nt_usf; ++it) {
        float x = t0 + (float)it * dt_usf;
        out[it] = orc_cubic_point(sig, y2, nt, t0, dt, x);
    }
    free(u);
    free(y2);
}

/* beamform_nearest, pycbf/cpu/pycbf.c:209-271. */
void orc_bf_nearest(float t0, float dt, int nt, const float *sig, int nout, float thresh,
                    const float *tautx, const float *apodtx, const float *taurx,
                    const float *apodrx, float *out, int usf)
{
    const float *s = sig;
    float *tmp = NULL;
    int n = nt;
    float dtu = dt;
    if (usf != 1) {
        n = usf * (nt - 1) + 1;
        dtu = dt / (float)usf;
        tmp = (float *)malloc(sizeof(float) * (size_t)n);
        orc_upsample(t0, dt, nt, sig, usf, tmp);
        s = tmp;
    }
    for (int i = 0; i < nout; ++i) {
        float w = apodtx[i] * apodrx[i];
        int it = orc_index_nearest(tautx[i] + taurx[i], t0, dtu, n);
        if ((fabs((double)w) >= (double)thresh) && (it >= 0)) out[i] += w * s[it];
    }
    free(tmp);
}

/* Linear counterpart of beamform_cubic (project-defined kind). */
void orc_bf_linear(float t0, float dt, int nt, const float *sig, int nout, float thresh,
                   const float *tautx, const float *apodtx, const float *taurx,
                   const float *apodrx, float *out)
{
    for (int i = 0; i < nout; ++i) {
        float w = apodtx[i] * apodrx[i];
        if (fabs((double)w) >= (double)thresh)
            out[i] += w * orc_linear_point(sig, nt, t0, dt, tautx[i] + taurx[i]);
    }
}

/* k-order cubic counterpart (the reference only has it on the GPU,
 * pycbf/gpu/__engines__.cu:726-778), with the CPU threshold rule. */
void orc_bf_korder(float t0, float dt, int nt, const float *sig, int nout, float thresh,
                   const float *tautx, const float *apodtx, const float *taurx,
                   const float *apodrx, float *out, const float *S, int k)
{
    for (int i = 0; i < nout; ++i) {
        float w = apodtx[i] * apodrx[i];
        if (fabs((double)w) >= (double)thresh)
            out[i] += w * orc_korder_point(sig, nt, t0, dt, S, k, tautx[i] + taurx[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* Tabbed driver: loop order and output offsets of the reference CPU class, */
/* pycbf/cpu/__core_bf_classes__.py:50-61 (offsets), :170-180 (pointer      */
/* offsets), :240-245 (tx outer, rx inner, one float32 buffer when          */
/* nwrkr == 1).  rf is (nfr, ntx, nrx, nt); out is (nfr, <sumtype axes>, np) */
/* and is overwritten.                                                      */
/* ------------------------------------------------------------------------ */
static void orc_channel(int kind, int usf, const float *S, int k, float t0, float dt, int nt,
                        const float *sig, int nout, float thresh, const float *ttx,
                        const float *atx, const float *trx, const float *arx, float *out)
{
    switch (kind) {
    case ORC_KIND_NEAREST:
        orc_bf_nearest(t0, dt, nt, sig, nout, thresh, ttx, atx, trx, arx, out, usf);
        break;
    case ORC_KIND_LINEAR:
        orc_bf_linear(t0, dt, nt, sig, nout, thresh, ttx, atx, trx, arx, out);
        break;
    case ORC_KIND_CUBIC:
        orc_bf_cubic(t0, dt, nt, sig, nout, thresh, ttx, atx, trx, arx, out);
        break;
    case ORC_KIND_KORDER:
        orc_bf_korder(t0, dt, nt, sig, nout, thresh, ttx, atx, trx, arx, out, S, k);
        break;
    default:
        break;
    }
}

static int64_t orc_out_offset(int sumflag, int64_t np, int itx, int irx, int nrx)
{
    if (sumflag == ORC_SUM_NONE) return np * ((int64_t)irx + (int64_t)itx * nrx);
    if (sumflag == ORC_SUM_TX_ONLY) return np * (int64_t)irx;
    if (sumflag == ORC_SUM_RX_ONLY) return np * (int64_t)itx;
    return 0;
}

static int64_t orc_out_frame_size(int sumflag, int64_t np, int ntx, int nrx)
{
    if (sumflag == ORC_SUM_NONE) return np * ntx * nrx;
    if (sumflag == ORC_SUM_TX_ONLY) return np * nrx;
    if (sumflag == ORC_SUM_RX_ONLY) return np * ntx;
    return np;
}

/* Pixel-chunked so that np may exceed the reference's int range
 * (pycbf/cpu/__core_bf_classes__.py:20-36 rejects ntx*nrx*np >= 2^31). */
#define ORC_PIX_CHUNK (1 << 26)

void orc_das_tabbed(const float *rf, int nfr, int ntx, int nrx, int nt, float t0, float dt,
                    const float *tautx, const float *apodtx, const float *taurx,
                    const float *apodrx, int64_t np, int kind, int usf, const float *S, int k,
                    float thresh, int sumflag, float *out)
{
    int64_t fsz = orc_out_frame_size(sumflag, np, ntx, nrx);
    memset(out, 0, sizeof(float) * (size_t)(fsz * nfr));
    for (int f = 0; f < nfr; ++f) {
        float *fo = out + (int64_t)f * fsz;
        for (int itx = 0; itx < ntx; ++itx) {
            for (int irx = 0; irx < nrx; ++irx) {
                const float *sig = rf + (((int64_t)f * ntx + itx) * nrx + irx) * nt;
                int64_t ooff = orc_out_offset(sumflag, np, itx, irx, nrx);
                for (int64_t p0 = 0; p0 < np; p0 += ORC_PIX_CHUNK) {
                    int nout = (int)((np - p0) < ORC_PIX_CHUNK ? (np - p0) : ORC_PIX_CHUNK);
                    orc_channel(kind, usf, S, k, t0, dt, nt, sig, nout, thresh,
                                tautx + (int64_t)itx * np + p0, apodtx + (int64_t)itx * np + p0,
                                taurx + (int64_t)irx * np + p0, apodrx + (int64_t)irx * np + p0,
                                fo + ooff + p0);
                }
            }
        }
    }
}

/* Multi-threaded variant used as the CPU baseline: (tx, rx) pairs are split
 * over threads, each with a private float32 accumulator, summed in double at
 * the end -- the reference's nwrkr > 1 pool with per-worker RawArrays summed
 * by numpy in float64 (pycbf/cpu/__core_bf_classes__.py:99-114,240-250).
 * Only sumtype tx_and_rx; one frame. Returns the thread count used. */
int orc_das_tabbed_mt(const float *rf, int ntx, int nrx, int nt, float t0, float dt,
                      const float *tautx, const float *apodtx, const float *taurx,
                      const float *apodrx, int64_t np, int kind, int usf, const float *S, int k,
                      float thresh, float *out, int nthreads)
{
    int used = 1;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = omp_get_max_threads();
    double *total = (double *)calloc((size_t)np, sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
        float *priv = (float *)calloc((size_t)np, sizeof(float));
#pragma omp single
        used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 1)
        for (int pair = 0; pair < ntx * nrx; ++pair) {
            int itx = pair / nrx, irx = pair % nrx;
            const float *sig = rf + ((int64_t)itx * nrx + irx) * nt;
            for (int64_t p0 = 0; p0 < np; p0 += ORC_PIX_CHUNK) {
                int nout = (int)((np - p0) < ORC_PIX_CHUNK ? (np - p0) : ORC_PIX_CHUNK);
                orc_channel(kind, usf, S, k, t0, dt, nt, sig, nout, thresh,
                            tautx + (int64_t)itx * np + p0, apodtx + (int64_t)itx * np + p0,
                            taurx + (int64_t)irx * np + p0, apodrx + (int64_t)irx * np + p0,
                            priv + p0);
            }
        }
#pragma omp critical
        for (int64_t i = 0; i < np; ++i) total[i] += (double)priv[i];
        free(priv);
    }
    for (int64_t i = 0; i < np; ++i) out[i] = (float)total[i];
    free(total);
#else
    (void)nthreads;
    orc_das_tabbed(rf, 1, ntx, nrx, nt, t0, dt, tautx, apodtx, taurx, apodrx, np, kind, usf, S,
                   k, thresh, ORC_SUM_TX_AND_RX, out);
#endif
    return used;
}

/* ------------------------------------------------------------------------ */
/* Synthetic-point geometry, pycbf/gpu/__engines__.cu:382-470, in float32   */
/* without fused multiply-add.  The "acos(v) > ala" acceptance tests are    */
/* evaluated exactly (acos in double); a NaN acos (|v| > 1) compares false, */
/* as in the reference.                                                     */
/* ------------------------------------------------------------------------ */
static int orc_acos_gt(float v, float ala) { return acos((double)v) > (double)ala; }

void orc_geom_tx(int ndim, const float *foc, const float *nvec, float c0, float t0, float ala,
                 float dof, const float *p, float *tau, float *apod)
{
    float mag = 0.0f, proj = 0.0f;
    for (int d = 0; d < ndim; ++d) {
        float di = p[d] - foc[d];
        mag += di * di;
        proj += di * nvec[d];
    }
    mag = sqrtf(mag);
    *tau = 0.0f;
    if (ala != 0.0f) {
        float sgn = (proj > 0.0f) ? 1.0f : -1.0f;
        if (dof != 0.0f) {
            /* hour-glass around the focus (Nguyen et al. 2016), :410-423 */
            int in_cone = (fabs((double)fabsf(proj)) > 1e-12) && (fabs((double)mag) > 1e-12) &&
                          (acos((double)fabsf(proj / mag)) <= (double)ala);
            if (in_cone) {
                *tau = sgn * (mag / c0) + t0;
                *apod = 1.0f;
            } else {
                float sf = (float)sin((double)ala);
                float tf = (float)tan((double)ala);
                float ktan = sqrtf(1.0f + tf * tf);
                float lat = sqrtf(fabsf(mag * mag - proj * proj));
                if ((double)lat < (double)(dof * sf) / 2.0) {
                    *tau = proj * ktan / c0 + t0;
                    *apod = 1.0f;
                } else {
                    *apod = 0.0f;
                }
            }
        } else {
            /* diverging / converging point without depth of field, :424-429 */
            if (fabs((double)fabsf(proj)) > 1e-12)
                *tau = sgn * (mag / c0) + t0;
            else
                *tau = t0;
            if ((fabs((double)mag) > 1e-12) && orc_acos_gt(fabsf(proj / mag), ala))
                *apod = 0.0f;
            else
                *apod = 1.0f;
        }
    } else {
        /* plane wave, :433-436 (t0 is not added, as in the reference) */
        *tau = proj / c0;
        *apod = 1.0f;
    }
}

void orc_geom_rx(int ndim, const float *orig, const float *nvec, float c0, float ala,
                 const float *p, float *tau, float *apod)
{
    float mag = 0.0f, proj = 0.0f;
    for (int d = 0; d < ndim; ++d) {
        float di = p[d] - orig[d];
        mag += di * di;
        proj += di * nvec[d];
    }
    mag = sqrtf(mag);
    *tau = mag / c0;
    if ((fabs((double)mag) > 1e-12) && orc_acos_gt(proj / mag, ala))
        *apod = 0.0f;
    else
        *apod = 1.0f;
}

/* Synthetic parameters -> tabbed tables (ntx, np) and (nrx, np). */
void orc_synthetic_tables(int ndim, int ntx, const float *ovectx, const float *nvectx,
                          const float *t0tx, const float *alatx, const float *doftx, int nrx,
                          const float *ovecrx, const float *nvecrx, const float *alarx, float c0,
                          const float *pvec, int64_t np, float *tautx, float *apodtx,
                          float *taurx, float *apodrx)
{
    for (int t = 0; t < ntx; ++t)
        for (int64_t i = 0; i < np; ++i)
            orc_geom_tx(ndim, ovectx + t * ndim, nvectx + t * ndim, c0, t0tx[t], alatx[t],
                        doftx[t], pvec + i * ndim, tautx + (int64_t)t * np + i,
                        apodtx + (int64_t)t * np + i);
    for (int r = 0; r < nrx; ++r)
        for (int64_t i = 0; i < np; ++i)
            orc_geom_rx(ndim, ovecrx + r * ndim, nvecrx + r * ndim, c0, alarx[r], pvec + i * ndim,
                        taurx + (int64_t)r * np + i, apodrx + (int64_t)r * np + i);
}

/* Sample index selected by every (tx, rx, px): idx (ntx, nrx, np), -1 when out
 * of range.  nearest != 0 uses orc_index_nearest on the dt/usf grid of n
 * samples, otherwise orc_index_floor on n samples. */
void orc_indices(const float *tautx, const float *taurx, int ntx, int nrx, int64_t np, float t0,
                 float dt, int n, int nearest, int32_t *idx)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int t = 0; t < ntx; ++t)
        for (int r = 0; r < nrx; ++r)
            for (int64_t i = 0; i < np; ++i) {
                float x = tautx[(int64_t)t * np + i] + taurx[(int64_t)r * np + i];
                idx[((int64_t)t * nrx + r) * np + i] =
                    nearest ? orc_index_nearest(x, t0, dt, n) : orc_index_floor(x, t0, dt, n);
            }
}
