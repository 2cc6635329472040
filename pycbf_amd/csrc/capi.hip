// extern "C" entry points of libpycbf_hip.so (declared in include/pycbf_hip.h):
// argument validation, workspace planning, prefilter selection and launch.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"

namespace pycbf {

static thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

bool das_use_lds()
{
    const char *e = getenv("PYCBF_DAS_LDS");
    return !(e && atoi(e) == 0);
}

int check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    return fail((int)e, "%s: %s", what, hipGetErrorString(e));
}

// ------------------------------------------------------------------ plans ---
struct DasPlan {
    int gk = GK_LINEAR;
    int64_t nch = 0;       // frames * ntx * nrx channels
    size_t coef = 0;       // interpolation table (float2 or float4 per sample)
    size_t uscr = 0;       // spline sweep scratch
    size_t y2 = 0;         // spline second derivatives (nearest, usf > 1)
    size_t up = 0;         // upsampled channels (nearest, usf > 1)
    size_t tabs = 0;       // synthetic delay / apodisation tables
    size_t part = 0;       // split partial sums
    size_t bounds = 0;     // per-tile delay ranges (staged gather)
    int nsplit = 1;
    size_t total = 0;
};

// Outer-axis groups of the gather grid.  One outer element per workgroup by
// default: the grid is dispatched tile-fastest, so all resident workgroups
// sweep the inner channel loop of the same outer element together and every
// channel row is pulled through L2 about once (a channel wavefront).  For sums
// over the outer axis each group writes a partial image; the group count is
// capped so those partials stay within a memory budget.
static int choose_split(int64_t np, int64_t nfr, int64_t na, bool partials)
{
    int64_t s = na;
    if (partials) {
        const int64_t budget = (int64_t)4 << 30;  // bytes of partial images
        const int64_t per = nfr * np * (int64_t)sizeof(float);
        const int64_t cap = per > 0 ? budget / per : na;
        if (s > cap) s = cap;
    }
    if (const char *e = getenv("PYCBF_DAS_SPLIT")) s = atoll(e);
    if (s < 1) s = 1;
    if (s > na) s = na;
    return (int)s;
}

static DasPlan plan_das(const pycbf_rfinfo *info, int64_t nfr, int64_t np, int kind, int param,
                        int synthetic, int sumflag)
{
    DasPlan p;
    const int64_t nt = info->tinfo.nx;
    p.nch = nfr * info->ntx * info->nrx;
    switch (kind) {
    case PYCBF_INTERP_NEAREST:
        p.gk = GK_NEAREST;
        if (param > 1) {
            p.y2 = align_up((size_t)p.nch * nt * sizeof(float), 256);
            p.uscr = spline_scratch_bytes(p.nch, (int32_t)nt);
            p.up = align_up((size_t)p.nch * ((size_t)param * (nt - 1) + 1) * sizeof(float), 256);
        }
        break;
    case PYCBF_INTERP_LINEAR:
        p.gk = GK_LINEAR;
        break;
    case PYCBF_INTERP_CUBIC:
        p.gk = GK_CUBIC;
        // 256 bytes of slack: the DMA-staged gather moves 16-byte chunks and may
        // read one entry past a row's end
        p.coef = align_up((size_t)p.nch * (nt + 1) * sizeof(float2) + 256, 256);
        p.uscr = spline_scratch_bytes(p.nch, (int32_t)nt);
        break;
    case PYCBF_INTERP_KORDER_CUBIC:
        p.gk = GK_KORDER;
        p.coef = align_up((size_t)p.nch * nt * sizeof(float4), 256);
        break;
    default:  // akima, makima
        p.gk = GK_AKIMA;
        p.coef = align_up((size_t)p.nch * nt * sizeof(float4), 256);
        break;
    }
    if (das_use_lds())
        p.bounds = align_up((size_t)das_tiles(np) * (info->ntx + info->nrx) * 2 * sizeof(float), 256);
    else if (synthetic)
        p.tabs = 4 * align_up((size_t)(info->ntx + info->nrx) * np * sizeof(float), 256);
    // partial sums exist only for tx_and_rx (receivers outer); size them whenever
    // sumflag < 0 (workspace query without a sumtype) or tx_and_rx
    if (sumflag < 0 || sumflag == PYCBF_SUM_TX_AND_RX) {
        const int64_t na = das_outer_is_tx(PYCBF_SUM_TX_AND_RX, info->ntx, info->nrx)
                               ? info->ntx : info->nrx;
        p.nsplit = choose_split(np, nfr, na, true);
        if (p.nsplit > 1) p.part = align_up((size_t)p.nsplit * nfr * np * sizeof(float), 256);
    } else {
        const int64_t na = das_outer_is_tx(sumflag, info->ntx, info->nrx) ? info->ntx : info->nrx;
        p.nsplit = choose_split(np, nfr, na, false);
    }
    p.total = p.coef + p.uscr + p.y2 + p.up + p.tabs + p.part + p.bounds + 2048;
    return p;
}

static int validate_common(const pycbf_rfinfo *info, const float *rf, int64_t nfr, int64_t np,
                           int kind, int param, const float *S, int sumflag, const float *out)
{
    if (!info) return fail(PYCBF_EINVAL, "info is NULL");
    if (info->ntx < 1 || info->nrx < 1) return fail(PYCBF_EINVAL, "ntx and nrx must be >= 1");
    if (info->tinfo.nx < 1) return fail(PYCBF_EINVAL, "nt must be >= 1");
    if (!(info->tinfo.dx > 0.0f) || !isfinite(info->tinfo.dx) || !isfinite(info->tinfo.x0))
        return fail(PYCBF_EINVAL, "dt must be finite and > 0 and t0 finite");
    if (nfr < 1 || np < 1) return fail(PYCBF_EINVAL, "nfr and np must be >= 1");
    if ((info->ntx * info->nrx) > 0x7fffffffLL)
        return fail(PYCBF_EINVAL, "ntx*nrx must fit in int32");
    if (info->ntx + info->nrx > 65535) return fail(PYCBF_EINVAL, "ntx + nrx must be <= 65535");
    if (!rf || !out) return fail(PYCBF_EINVAL, "rf and out must be device pointers");
    if (sumflag < PYCBF_SUM_NONE || sumflag > PYCBF_SUM_TX_AND_RX)
        return fail(PYCBF_EINVAL, "sumflag %d out of range", sumflag);
    const int nt = info->tinfo.nx;
    switch (kind) {
    case PYCBF_INTERP_NEAREST:
        if (param < 1) return fail(PYCBF_EINVAL, "usf must be >= 1");
        if ((int64_t)param * (nt - 1) + 1 > 0x7fffffffLL)
            return fail(PYCBF_EINVAL, "usf*(nt-1)+1 must fit in int32");
        break;
    case PYCBF_INTERP_LINEAR:
    case PYCBF_INTERP_CUBIC:
        break;
    case PYCBF_INTERP_KORDER_CUBIC:
        if (param < 4 || (param % 2) != 0)
            return fail(PYCBF_EINVAL, "'k' for korder_cubic must be at least four and even");
        if (param > nt) return fail(PYCBF_EINVAL, "k (%d) must not exceed nt (%d)", param, nt);
        if (param > 128) return fail(PYCBF_EINVAL, "k (%d) must be <= 128", param);
        if (!S) return fail(PYCBF_EINVAL, "korder_cubic needs the (k, k) S operator");
        break;
    case PYCBF_INTERP_AKIMA:
    case PYCBF_INTERP_MAKIMA:
        if (nt < 5) return fail(PYCBF_EINVAL, "akima / makima need nt >= 5");
        break;
    default:
        return fail(PYCBF_EINVAL, "unknown interp kind %d", kind);
    }
    return 0;
}

// Builds the gather source for `kind` into the workspace and runs the gather.
struct SynGeo {
    const float *txpack = nullptr, *rxpack = nullptr, *pvec = nullptr;
    float c0 = 0.0f;
    int ndim = 0;  // 0: tabbed
};

static int run_das(const pycbf_rfinfo *info, const float *rf, int64_t nfr, const float *ttx,
                   const float *atx, const float *trx, const float *arx, const int32_t *perm,
                   int64_t np, int kind, int param, const float *S, float thresh, int sumflag,
                   float *out, const DasPlan &plan, Carver &cv, hipStream_t s,
                   const SynGeo &geo = SynGeo(), const float *ttx_lo = nullptr,
                   const float *trx_lo = nullptr)
{
    const int nt = info->tinfo.nx;
    const float t0 = info->tinfo.x0, dt = info->tinfo.dx;
    DasArgs a{};
    a.ttx = ttx;
    a.atx = atx;
    a.trx = trx;
    a.arx = arx;
    a.perm = perm;
    a.np = np;
    a.ntx = (int)info->ntx;
    a.nrx = (int)info->nrx;
    a.nfr = (int)nfr;
    a.t0 = t0;
    a.dt = dt;
    a.rdt = 1.0f / dt;
    a.tmax = t0 + dt * (float)(nt - 1);
    a.thresh = thresh;
    a.sumflag = sumflag;
    a.out = out;
    a.src = rf;
    a.ch_stride = nt;
    a.nsrc = nt;
    a.nent = nt;
    a.geo = geo.ndim;
    a.txpack = geo.txpack;
    a.rxpack = geo.rxpack;
    a.pvec = geo.pvec;
    a.c0 = geo.c0;
    a.ttx_lo = ttx_lo;
    a.trx_lo = trx_lo;
    a.rc0 = geo.ndim > 0 ? 1.0f / geo.c0 : 0.0f;
    int rc = 0;
    switch (kind) {
    case PYCBF_INTERP_NEAREST:
        if (param > 1) {
            float *y2 = (float *)cv.take(plan.y2);
            float *u = (float *)cv.take(plan.uscr);
            float *up = (float *)cv.take(plan.up);
            if (!y2 || !u || !up) return fail(PYCBF_EINVAL, "workspace too small");
            if ((rc = launch_spline_y2(rf, plan.nch, nt, dt, u, y2, s))) return rc;
            if ((rc = launch_upsample(rf, y2, plan.nch, nt, t0, dt, param, up, s))) return rc;
            const int ntu = param * (nt - 1) + 1;
            a.src = up;
            a.ch_stride = ntu;
            a.nsrc = ntu;
            a.nent = ntu;
            a.dt = dt / (float)param;
            a.rdt = 1.0f / a.dt;
        }
        break;
    case PYCBF_INTERP_LINEAR:
        break;
    case PYCBF_INTERP_CUBIC: {
        float2 *coef = (float2 *)cv.take(plan.coef);
        float *u = (float *)cv.take(plan.uscr);
        if (!coef || !u) return fail(PYCBF_EINVAL, "workspace too small");
        if ((rc = launch_spline_coefs(rf, plan.nch, nt, dt, u, coef, s))) return rc;
        a.src = (const float *)coef;
        a.ch_stride = 2 * ((int64_t)nt + 1);
        a.nent = nt + 1;
        break;
    }
    case PYCBF_INTERP_KORDER_CUBIC: {
        float4 *coef = (float4 *)cv.take(plan.coef);
        if (!coef) return fail(PYCBF_EINVAL, "workspace too small");
        if ((rc = launch_korder_coefs(rf, plan.nch, nt, S, param, coef, s))) return rc;
        a.src = (const float *)coef;
        a.ch_stride = 4 * (int64_t)nt;
        break;
    }
    default: {  // akima / makima
        float4 *coef = (float4 *)cv.take(plan.coef);
        if (!coef) return fail(PYCBF_EINVAL, "workspace too small");
        if ((rc = launch_akima_coefs(rf, plan.nch, nt, kind == PYCBF_INTERP_MAKIMA, coef, s)))
            return rc;
        a.src = (const float *)coef;
        a.ch_stride = 4 * (int64_t)nt;
        break;
    }
    }
    float *part = nullptr;
    if (plan.part) {
        part = (float *)cv.take(plan.part);
        if (!part) return fail(PYCBF_EINVAL, "workspace too small");
    }
    // delay residuals: the direct gather places samples in double
    if (ttx_lo || trx_lo) return launch_das(plan.gk, a, plan.nsplit, part, s);
    if (das_use_plain()) return launch_das_plain(plan.gk, a, s);
    if (das_reg_applies(plan.gk, nfr, sumflag)) return launch_das_reg(a, s);
    if (plan.bounds) {
        float *bnd = (float *)cv.take(plan.bounds);
        if (!bnd) return fail(PYCBF_EINVAL, "workspace too small");
        if ((rc = launch_tile_bounds(a, bnd, s))) return rc;
        a.bounds = bnd;
        return launch_das_lds(plan.gk, a, plan.nsplit, part, s);
    }
    return launch_das(plan.gk, a, plan.nsplit, part, s);
}

bool recip_division_exact(float c)
{
    if (!(fabsf(c) >= 0x1p-30f && fabsf(c) <= 0x1p+30f)) return false;
    static std::mutex mu;
    static std::vector<std::pair<float, bool>> cache;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (const auto &e : cache)
            if (e.first == c) return e.second;
    }
    const float r = 1.0f / c;
    bool ok = true;
    // every significand of [1, 2); rounding to nearest is symmetric in the sign
    for (uint32_t m = 0; m < (1u << 23) && ok; ++m) {
        float x;
        const uint32_t bits = 0x3f800000u | m;
        memcpy(&x, &bits, 4);
        const float q = x * r;
        const float e = std::fma(-q, c, x);
        ok = (std::fma(e, r, q) == x / c);
    }
    std::lock_guard<std::mutex> lk(mu);
    cache.emplace_back(c, ok);
    return ok;
}

static float acos_threshold(float ala)
{
    // returns T with  acos(v) > ala  <=>  v < T  for every float v in [-1, 1]
    if (ala != ala) return -INFINITY;
    const double a = (double)ala;
    if (a < 0.0) return INFINITY;
    if (a > M_PI) return -INFINITY;
    const double c = cos(a);
    float T = (float)c;
    if ((double)T < c) T = nextafterf(T, INFINITY);
    return T;
}

// The double midpoint between T and the float below it, the bound with which
// the receive kernels decide v < T on proj / mag without dividing
// (RN(proj / mag) < T <=> proj < m * mag); T = +inf stores 1 + 2^-24 (the
// bound of RN(q) <= 1), T = -inf stores -inf.
static double acos_midpoint(float T)
{
    if (T == INFINITY) return 1.0 + 0x1p-24;
    if (T == -INFINITY) return -INFINITY;
    return 0.5 * ((double)nextafterf(T, -INFINITY) + (double)T);
}

// A plan whose partial images do not fit `cap` bytes takes fewer, larger
// outer groups (down to one, no partials); the images are the same sums in
// another order.
static void cap_partials(DasPlan &p, int64_t nfr, int64_t np, size_t cap)
{
    if (p.total <= cap || p.part == 0) return;
    const size_t rest = p.total - p.part;
    const size_t per = (size_t)nfr * (size_t)np * sizeof(float);
    int64_t s = (cap > rest + 256) ? (int64_t)((cap - rest - 256) / per) : 0;
    if (s >= p.nsplit) return;
    if (s <= 1) {
        p.nsplit = 1;
        p.part = 0;
    } else {
        p.nsplit = (int)s;
        p.part = align_up((size_t)s * per, 256);
    }
    p.total = rest + p.part;
}

int das_synthetic_in(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                     const float *txpack, const float *rxpack, float c0, const float *pvec,
                     const int32_t *perm, int64_t np, int interp_kind, int interp_param,
                     const float *S, float thresh, int sumflag, float *out, void *buf,
                     size_t cap, bool shrink_split, hipStream_t s)
{
    int rc = validate_common(info, rf, nfr, np, interp_kind, interp_param, S, sumflag, out);
    if (rc) return rc;
    if (info->ndim != 2 && info->ndim != 3)
        return fail(PYCBF_EINVAL, "synthetic geometry needs ndim 2 or 3");
    if (!txpack || !rxpack || !pvec) return fail(PYCBF_EINVAL, "NULL geometry pointer");
    if (!(c0 > 0.0f) || !isfinite(c0)) return fail(PYCBF_EINVAL, "c0 must be finite and > 0");
    DasPlan plan = plan_das(info, nfr, np, interp_kind, interp_param, 1, sumflag);
    if (shrink_split) cap_partials(plan, nfr, np, cap);
    if (cap < plan.total || !buf)
        return fail(PYCBF_EINVAL, "%s of %zu bytes needed, %zu given",
                    shrink_split ? "scratch" : "workspace", plan.total, cap);
    Carver cv(buf, cap);
    if (plan.bounds) {
        // staged gather: delays and weights are evaluated in registers
        SynGeo g;
        g.txpack = txpack;
        g.rxpack = rxpack;
        g.pvec = pvec;
        g.c0 = c0;
        g.ndim = info->ndim;
        return run_das(info, rf, nfr, nullptr, nullptr, nullptr, nullptr, perm, np, interp_kind,
                       interp_param, S, thresh, sumflag, out, plan, cv, s, g);
    }
    const size_t tb = plan.tabs / 4;
    float *ttx = (float *)cv.take(tb);
    float *atx = (float *)cv.take(tb);
    float *trx = (float *)cv.take(tb);
    float *arx = (float *)cv.take(tb);
    if (!ttx || !atx || !trx || !arx) return fail(PYCBF_EINVAL, "workspace too small");
    rc = launch_synth_tables(info->ndim, (int)info->ntx, (int)info->nrx, txpack, rxpack, c0, pvec,
                             np, ttx, atx, trx, arx, s);
    if (rc) return rc;
    return run_das(info, rf, nfr, ttx, atx, trx, arx, perm, np, interp_kind, interp_param, S,
                   thresh, sumflag, out, plan, cv, s);
}

}  // namespace pycbf

using namespace pycbf;

extern "C" {

int pycbf_abi_version(void) { return PYCBF_ABI_VERSION; }

const char *pycbf_last_error(void) { return g_err; }

size_t pycbf_das_workspace_bytes(const pycbf_rfinfo *info, int64_t nfr, int64_t np,
                                 int interp_kind, int interp_param, int synthetic)
{
    if (!info || nfr < 1 || np < 1) return 0;
    return plan_das(info, nfr, np, interp_kind, interp_param, synthetic, -1).total;
}

int pycbf_das_tabbed(const pycbf_rfinfo *info, const float *rf, int64_t nfr, const float *tautx,
                     const float *apodtx, const float *taurx, const float *apodrx,
                     const int32_t *perm, int64_t np, int interp_kind, int interp_param,
                     const float *S, float thresh, int sumflag, float *out, void *workspace,
                     size_t workspace_bytes, void *stream)
{
    int rc = validate_common(info, rf, nfr, np, interp_kind, interp_param, S, sumflag, out);
    if (rc) return rc;
    if (!tautx || !apodtx || !taurx || !apodrx)
        return fail(PYCBF_EINVAL, "delay / apodisation tables must be device pointers");
    const DasPlan plan = plan_das(info, nfr, np, interp_kind, interp_param, 0, sumflag);
    if (workspace_bytes < plan.total || !workspace)
        return fail(PYCBF_EINVAL, "workspace of %zu bytes needed, %zu given", plan.total,
                    workspace_bytes);
    Carver cv(workspace, workspace_bytes);
    return run_das(info, rf, nfr, tautx, apodtx, taurx, apodrx, perm, np, interp_kind,
                   interp_param, S, thresh, sumflag, out, plan, cv, (hipStream_t)stream);
}

int pycbf_das_tabbed_residual(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                              const float *tautx, const float *apodtx, const float *taurx,
                              const float *apodrx, const float *tautx_lo, const float *taurx_lo,
                              int64_t np, int interp_kind, int interp_param, const float *S,
                              float thresh, int sumflag, float *out, void *workspace,
                              size_t workspace_bytes, void *stream)
{
    int rc = validate_common(info, rf, nfr, np, interp_kind, interp_param, S, sumflag, out);
    if (rc) return rc;
    if (!tautx || !apodtx || !taurx || !apodrx)
        return fail(PYCBF_EINVAL, "delay / apodisation tables must be device pointers");
    if (!tautx_lo && !taurx_lo)
        return fail(PYCBF_EINVAL, "give tautx_lo or taurx_lo (else call pycbf_das_tabbed)");
    const DasPlan plan = plan_das(info, nfr, np, interp_kind, interp_param, 0, sumflag);
    if (workspace_bytes < plan.total || !workspace)
        return fail(PYCBF_EINVAL, "workspace of %zu bytes needed, %zu given", plan.total,
                    workspace_bytes);
    Carver cv(workspace, workspace_bytes);
    return run_das(info, rf, nfr, tautx, apodtx, taurx, apodrx, nullptr, np, interp_kind,
                   interp_param, S, thresh, sumflag, out, plan, cv, (hipStream_t)stream, SynGeo(),
                   tautx_lo, taurx_lo);
}

int pycbf_make_tabs_2d(int mode, int64_t ntx, const double *txc, int64_t nrx, const double *xele,
                       const double *xout, int64_t nx, const double *zout, int64_t nz, double cm,
                       double fnum, float *tautx, float *apodtx, float *taurx, float *apodrx,
                       void *stream)
{
    if (mode != PYCBF_TABS_PLANE && mode != PYCBF_TABS_DIVERGENT && mode != PYCBF_TABS_FOCUSED)
        return fail(PYCBF_EINVAL, "tabs mode %d is not plane (0), divergent (1) or focused (2)",
                    mode);
    if (ntx < 0 || nrx < 0 || ntx + nrx > 65535)
        return fail(PYCBF_EINVAL, "need 0..65535 transmit plus receive elements");
    if (nx < 0 || nz < 0) return fail(PYCBF_EINVAL, "negative grid size");
    if (nx * nz > 0 && (!xout || !zout)) return fail(PYCBF_EINVAL, "xout / zout are NULL");
    if (ntx > 0 && (tautx || apodtx) && !txc) return fail(PYCBF_EINVAL, "txc is NULL");
    if (nrx > 0 && (taurx || apodrx) && !xele) return fail(PYCBF_EINVAL, "xele is NULL");
    if (!(cm > 0.0)) return fail(PYCBF_EINVAL, "cm must be > 0");
    const int nt = (tautx || apodtx) ? (int)ntx : 0;
    const int nr = (taurx || apodrx) ? (int)nrx : 0;
    // transmit rows first; with no transmit outputs the receive rows start at 0
    return launch_tabs_2d(mode, nt, txc, nr, xele, xout, nx, zout, nz, cm, fnum, tautx, apodtx,
                          taurx, apodrx, (hipStream_t)stream);
}

size_t pycbf_demod_workspace_bytes(int64_t rows, int64_t n, int64_t ncol)
{
    if (rows < 1 || n < 1 || ncol < 1 || n > 0x7fffffffLL || ncol > 128LL * 65535) return 0;
    const int nsplit = demod_split(rows, (int)n, (int)ncol);
    return nsplit > 1 ? (size_t)nsplit * (size_t)rows * (size_t)ncol * sizeof(float) : 0;
}

int pycbf_demod_apply(const float *rf, int64_t rows, int64_t n, const float *opT, int64_t ncol,
                      float *iq, void *workspace, size_t workspace_bytes, void *stream)
{
    if (rows < 0 || n < 1 || ncol < 1) return fail(PYCBF_EINVAL, "need rows >= 0, n >= 1, ncol >= 1");
    if (n > 0x7fffffffLL) return fail(PYCBF_EINVAL, "n must fit in int32");
    if (ncol > 128LL * 65535) return fail(PYCBF_EINVAL, "ncol must be <= %d", 128 * 65535);
    if (rows > 0 && (!rf || !opT || !iq)) return fail(PYCBF_EINVAL, "NULL device pointer");
    if ((rows + 127) / 128 > 0x7fffffffLL) return fail(PYCBF_EINVAL, "too many rows");
    const size_t need = pycbf_demod_workspace_bytes(rows, n, ncol);
    if (need > 0 && (!workspace || workspace_bytes < need))
        return fail(PYCBF_EINVAL, "workspace of %zu bytes needed, %zu given", need,
                    workspace_bytes);
    return launch_demod_gemm(rf, rows, (int)n, opT, (int)ncol, iq,
                             need > 0 ? (float *)workspace : nullptr, (hipStream_t)stream);
}

int pycbf_synthetic_pack(int ndim, int64_t ntx, const float *ovectx, const float *nvectx,
                         const float *t0tx, const float *alatx, const float *doftx, int64_t nrx,
                         const float *ovecrx, const float *nvecrx, const float *alarx,
                         float *txpack, float *rxpack)
{
    if (ndim < 1 || ndim > 3) return fail(PYCBF_EINVAL, "ndim must be 1, 2 or 3");
    if (ntx < 0 || nrx < 0) return fail(PYCBF_EINVAL, "negative element count");
    for (int64_t e = 0; e < ntx; ++e) {
        float *pk = txpack + e * PYCBF_TXPACK_FLOATS;
        for (int i = 0; i < PYCBF_TXPACK_FLOATS; ++i) pk[i] = 0.0f;
        for (int d = 0; d < ndim; ++d) {
            pk[d] = ovectx[e * ndim + d];
            pk[3 + d] = nvectx[e * ndim + d];
        }
        const float ala = alatx[e], dof = doftx[e];
        pk[6] = t0tx[e];
        pk[7] = (ala != 0.0f) ? ((dof != 0.0f) ? 2.0f : 1.0f) : 0.0f;
        pk[8] = acos_threshold(ala);
        const float sf = (float)sin((double)ala);
        const float tf = (float)tan((double)ala);
        pk[9] = dof * sf;
        pk[10] = sqrtf(1.0f + tf * tf);
    }
    for (int64_t e = 0; e < nrx; ++e) {
        float *pk = rxpack + e * PYCBF_RXPACK_FLOATS;
        for (int i = 0; i < PYCBF_RXPACK_FLOATS; ++i) pk[i] = 0.0f;
        for (int d = 0; d < ndim; ++d) {
            pk[d] = ovecrx[e * ndim + d];
            pk[3 + d] = nvecrx[e * ndim + d];
        }
        pk[6] = acos_threshold(alarx[e]);
        const double m = acos_midpoint(pk[6]);
        memcpy(pk + 8, &m, sizeof m);
    }
    return 0;
}

int pycbf_synthetic_tables(const pycbf_rfinfo *info, const float *txpack, const float *rxpack,
                           float c0, const float *pvec, int64_t np, float *tautx, float *apodtx,
                           float *taurx, float *apodrx, void *stream)
{
    if (!info || (info->ndim != 2 && info->ndim != 3))
        return fail(PYCBF_EINVAL, "synthetic geometry needs ndim 2 or 3");
    if (!txpack || !rxpack || !pvec || !tautx || !apodtx || !taurx || !apodrx)
        return fail(PYCBF_EINVAL, "NULL device pointer");
    if (info->ntx + info->nrx > 65535) return fail(PYCBF_EINVAL, "ntx + nrx must be <= 65535");
    return launch_synth_tables(info->ndim, (int)info->ntx, (int)info->nrx, txpack, rxpack, c0,
                               pvec, np, tautx, apodtx, taurx, apodrx, (hipStream_t)stream);
}

int pycbf_das_synthetic(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                        const float *txpack, const float *rxpack, float c0, const float *pvec,
                        const int32_t *perm, int64_t np, int interp_kind, int interp_param,
                        const float *S, float thresh, int sumflag, float *out, void *workspace,
                        size_t workspace_bytes, void *stream)
{
    return das_synthetic_in(info, rf, nfr, txpack, rxpack, c0, pvec, perm, np, interp_kind,
                            interp_param, S, thresh, sumflag, out, workspace, workspace_bytes,
                            false, (hipStream_t)stream);
}

int pycbf_tile_order(const float *pvec, int64_t np, int ndim, int32_t *perm)
{
    if (!pvec || !perm || np < 0) return fail(PYCBF_EINVAL, "NULL pointer");
    if (ndim < 1 || ndim > 3) return fail(PYCBF_EINVAL, "ndim must be 1, 2 or 3");
    if (np > 0x7fffffffLL) return fail(PYCBF_EINVAL, "np must fit in int32");
    // grid coordinate per axis: dense rank of the value among the axis' unique
    // values (exact for regular grids), quantised to 21 bits otherwise
    std::vector<uint64_t> q((size_t)np * ndim);
    for (int d = 0; d < ndim; ++d) {
        std::vector<float> vals((size_t)np);
        for (int64_t i = 0; i < np; ++i) vals[i] = pvec[i * ndim + d];
        std::vector<float> u = vals;
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        const double scale = u.size() > (1u << 21) ? (double)((1u << 21) - 1) / (double)(u.size() - 1) : 1.0;
        for (int64_t i = 0; i < np; ++i) {
            const size_t r = (size_t)(std::lower_bound(u.begin(), u.end(), vals[i]) - u.begin());
            q[i * ndim + d] = (uint64_t)((double)r * scale);
        }
    }
    std::vector<std::pair<uint64_t, int32_t>> key((size_t)np);
    for (int64_t i = 0; i < np; ++i) {
        uint64_t code = 0;
        for (int b = 0; b < 21; ++b)
            for (int d = 0; d < ndim; ++d)
                code |= ((q[i * ndim + d] >> b) & 1ull) << (b * ndim + d);
        key[i] = {code, (int32_t)i};
    }
    std::sort(key.begin(), key.end());
    for (int64_t i = 0; i < np; ++i) perm[i] = key[i].second;
    return 0;
}

size_t pycbf_spline_workspace_bytes(int64_t nch, int32_t nt)
{
    return spline_scratch_bytes(nch, nt) + 256;
}

int pycbf_spline_coefs(const float *rf, int64_t nch, int32_t nt, float dt, float *coef,
                       void *workspace, size_t workspace_bytes, void *stream)
{
    if (!rf || !coef || nch < 0 || nt < 1) return fail(PYCBF_EINVAL, "bad spline arguments");
    Carver cv(workspace, workspace_bytes);
    float *u = (float *)cv.take(spline_scratch_bytes(nch, nt));
    if (!u) return fail(PYCBF_EINVAL, "workspace too small");
    return launch_spline_coefs(rf, nch, nt, dt, u, (float2 *)coef, (hipStream_t)stream);
}

int pycbf_spline_y2(const float *rf, int64_t nch, int32_t nt, float dt, float *y2,
                    void *workspace, size_t workspace_bytes, void *stream)
{
    if (!rf || !y2 || nch < 0 || nt < 1) return fail(PYCBF_EINVAL, "bad spline arguments");
    Carver cv(workspace, workspace_bytes);
    float *u = (float *)cv.take(spline_scratch_bytes(nch, nt));
    if (!u) return fail(PYCBF_EINVAL, "workspace too small");
    return launch_spline_y2(rf, nch, nt, dt, u, y2, (hipStream_t)stream);
}

int pycbf_debug_indices(const pycbf_rfinfo *info, const float *tautx, const float *taurx,
                        int64_t np, int interp_kind, int usf, int32_t *idx, void *stream)
{
    if (!info || !tautx || !taurx || !idx) return fail(PYCBF_EINVAL, "NULL pointer");
    if (info->ntx * info->nrx > 65535) return fail(PYCBF_EINVAL, "ntx*nrx must be <= 65535");
    const int nt = info->tinfo.nx;
    const bool nearest = interp_kind == PYCBF_INTERP_NEAREST;
    if (nearest && usf < 1) return fail(PYCBF_EINVAL, "usf must be >= 1");
    const int n = nearest ? usf * (nt - 1) + 1 : nt;
    const float dt = nearest ? info->tinfo.dx / (float)usf : info->tinfo.dx;
    return launch_debug_indices((int)info->ntx, (int)info->nrx, tautx, taurx, np, nearest ? 1 : 0,
                                info->tinfo.x0, dt, n, idx, (hipStream_t)stream);
}

// ------------------------------------------------------------- backward ---
struct BwdPlan {
    size_t bounds = 0, coef = 0, uscr = 0, gy = 0, ge = 0, gc2 = 0, gc3 = 0, mbits = 256;
    size_t y2 = 0, up = 0, gu = 0;  // nearest, usf > 1
    size_t gc0 = 0;                 // partial sums of d loss / d c0 (synthetic geometry)
    size_t total = 0;
};

static BwdPlan plan_bwd(const pycbf_rfinfo *info, int64_t nfr, int64_t np, int kind, int param)
{
    BwdPlan b;
    const int64_t nt = info->tinfo.nx;
    const int64_t nch = nfr * info->ntx * info->nrx;
    b.bounds = align_up((size_t)das_tiles(np) * (size_t)(info->ntx + info->nrx) * 2 * 4, 256);
    b.gc0 = align_up((size_t)das_tiles(np) * (size_t)(info->ntx + info->nrx) * 8, 256);
    if (kind == PYCBF_INTERP_CUBIC) {
        b.coef = align_up((size_t)nch * (size_t)(nt + 1) * 8, 256);
        b.uscr = spline_scratch_bytes(nch, (int32_t)nt);
        b.gy = b.ge = align_up((size_t)nch * (size_t)(nt + 1) * 4, 256);
    } else if (kind == PYCBF_INTERP_KORDER_CUBIC || kind == PYCBF_INTERP_AKIMA ||
               kind == PYCBF_INTERP_MAKIMA) {
        b.coef = align_up((size_t)nch * (size_t)nt * 16, 256);
        b.gy = b.ge = b.gc2 = b.gc3 = align_up((size_t)nch * (size_t)nt * 4, 256);
    } else if (kind == PYCBF_INTERP_NEAREST && param > 1) {
        const size_t ntu = (size_t)param * (size_t)(nt - 1) + 1;
        b.y2 = align_up((size_t)nch * (size_t)nt * 4, 256);
        b.uscr = spline_scratch_bytes(nch, (int32_t)nt);
        b.up = b.gu = align_up((size_t)nch * ntu * 4, 256);
        b.gy = b.ge = align_up((size_t)nch * (size_t)(nt + 1) * 4, 256);
    }
    b.total = b.bounds + b.coef + b.uscr + b.gy + b.ge + b.gc2 + b.gc3 + b.mbits + b.y2 + b.up +
              b.gu + b.gc0 + 12 * 256;
    return b;
}

size_t pycbf_das_bwd_workspace_bytes(const pycbf_rfinfo *info, int64_t nfr, int64_t np,
                                     int interp_kind, int param)
{
    if (!info || nfr < 1 || np < 1) return 0;
    return plan_bwd(info, nfr, np, interp_kind, param).total;
}

// Argument checks shared by the two adjoints; returns the kernel parameter
// (k, usf or 1) through kparam.
static int bwd_validate(const pycbf_rfinfo *info, const float *rf, int64_t nfr, int64_t np,
                        int interp_kind, int param, const float *S, int sumflag,
                        const float *gout, int &kparam)
{
    if (interp_kind != PYCBF_INTERP_NEAREST && interp_kind != PYCBF_INTERP_LINEAR &&
        interp_kind != PYCBF_INTERP_CUBIC && interp_kind != PYCBF_INTERP_KORDER_CUBIC &&
        interp_kind != PYCBF_INTERP_AKIMA && interp_kind != PYCBF_INTERP_MAKIMA)
        return fail(PYCBF_EINVAL,
                    "the adjoint supports nearest, linear, cubic, korder_cubic, akima and makima "
                    "(got %d)",
                    interp_kind);
    kparam =
        (interp_kind == PYCBF_INTERP_KORDER_CUBIC || interp_kind == PYCBF_INTERP_NEAREST) ? param
                                                                                          : 1;
    int rc = validate_common(info, rf, nfr, np, interp_kind, kparam,
                             interp_kind == PYCBF_INTERP_KORDER_CUBIC ? S : nullptr, sumflag, gout);
    if (rc) return rc;
    if (nfr > 65535 || info->ntx > 65535) return fail(PYCBF_EINVAL, "nfr and ntx must be <= 65535");
    return 0;
}

// The adjoint pipeline once the delays and weights are fixed: b carries the
// tables (b.geo 0) or the synthetic geometry (b.geo 2 / 3); tb the same
// source for the per-tile delay ranges.  Prefilters the channel data as the
// forward does, runs the adjoint kernels and transposes the prefilter.
static int bwd_run(const pycbf_rfinfo *info, const float *rf, int64_t nfr, int64_t np,
                   int interp_kind, int param, const float *S, float thresh, int sumflag,
                   const float *gout, BwdArgs b, DasArgs tb, float *grf, float *gtautx,
                   float *gapodtx, float *gtaurx, float *gapodrx, float *gc0, void *workspace,
                   size_t workspace_bytes, hipStream_t s)
{
    const BwdPlan plan = plan_bwd(info, nfr, np, interp_kind, param);
    if (!workspace || workspace_bytes < plan.total)
        return fail(PYCBF_EINVAL, "workspace of %zu bytes needed, %zu given", plan.total,
                    workspace_bytes);
    int rc = 0;
    Carver cv(workspace, workspace_bytes);
    const int nt = info->tinfo.nx;
    const float t0 = info->tinfo.x0, dt = info->tinfo.dx;
    const int64_t nch = nfr * info->ntx * info->nrx;

    b.np = np;
    b.ntx = (int)info->ntx;
    b.nrx = (int)info->nrx;
    b.nfr = (int)nfr;
    b.nt = nt;
    b.src = rf;
    b.t0 = t0;
    b.dt = dt;
    b.rdt = 1.0f / dt;
    b.tmax = t0 + dt * (float)(nt - 1);
    b.thresh = thresh;
    b.sumflag = sumflag;
    b.gout = gout;
    float *bounds = (float *)cv.take(plan.bounds);
    b.bounds = bounds;
    b.mbits = (unsigned *)cv.take(plan.mbits);
    b.gc0part = (double *)cv.take(plan.gc0);
    if (interp_kind == PYCBF_INTERP_CUBIC) {
        float2 *coef = (float2 *)cv.take(plan.coef);
        float *u = (float *)cv.take(plan.uscr);
        b.gy = (float *)cv.take(plan.gy);
        b.ge = (float *)cv.take(plan.ge);
        b.gstride = nt + 1;
        if ((rc = launch_spline_coefs(rf, nch, nt, dt, u, coef, s))) return rc;
        b.src = (const float *)coef;
    } else if (interp_kind == PYCBF_INTERP_KORDER_CUBIC) {
        float4 *coef = (float4 *)cv.take(plan.coef);
        b.gy = (float *)cv.take(plan.gy);
        b.ge = (float *)cv.take(plan.ge);
        b.gc2 = (float *)cv.take(plan.gc2);
        b.gc3 = (float *)cv.take(plan.gc3);
        b.gstride = nt;
        b.S = S;
        b.k = param;
        if ((rc = launch_korder_coefs(rf, nch, nt, S, param, coef, s))) return rc;
        b.src = (const float *)coef;
    } else if (interp_kind == PYCBF_INTERP_AKIMA || interp_kind == PYCBF_INTERP_MAKIMA) {
        float4 *coef = (float4 *)cv.take(plan.coef);
        b.gy = (float *)cv.take(plan.gy);
        b.ge = (float *)cv.take(plan.ge);
        b.gc2 = (float *)cv.take(plan.gc2);
        b.gc3 = (float *)cv.take(plan.gc3);
        b.gstride = nt;
        b.rf = rf;
        b.modified = interp_kind == PYCBF_INTERP_MAKIMA;
        if ((rc = launch_akima_coefs(rf, nch, nt, b.modified, coef, s))) return rc;
        b.src = (const float *)coef;
    } else if (interp_kind == PYCBF_INTERP_NEAREST && param > 1) {
        // nearest on the upsampled trace, as the forward gathers it
        float *y2 = (float *)cv.take(plan.y2);
        float *u = (float *)cv.take(plan.uscr);
        float *up = (float *)cv.take(plan.up);
        if ((rc = launch_spline_y2(rf, nch, nt, dt, u, y2, s))) return rc;
        if ((rc = launch_upsample(rf, y2, nch, nt, t0, dt, param, up, s))) return rc;
        const int ntu = param * (nt - 1) + 1;
        b.src = up;
        b.nt = ntu;
        b.dt = dt / (float)param;
        b.rdt = 1.0f / b.dt;
        b.tmax = t0 + b.dt * (float)(ntu - 1);
        b.gy = (float *)cv.take(plan.gu);
        b.gstride = ntu;
    } else {
        b.gy = grf;
        b.gstride = nt;
    }
    if (grf) {
        tb.np = np;
        tb.ntx = b.ntx;
        tb.nrx = b.nrx;
        tb.thresh = thresh;
        tb.perm = nullptr;
        if ((rc = launch_tile_bounds(tb, bounds, s))) return rc;
    }
    const int gk = interp_kind == PYCBF_INTERP_NEAREST  ? GK_NEAREST
                   : interp_kind == PYCBF_INTERP_LINEAR ? GK_LINEAR
                   : interp_kind == PYCBF_INTERP_CUBIC  ? GK_CUBIC
                   : interp_kind == PYCBF_INTERP_KORDER_CUBIC ? GK_KORDER
                                                              : GK_AKIMA;
    if ((rc = launch_das_bwd(gk, b, gtautx, gapodtx, gtaurx, gapodrx, grf, gc0, s))) return rc;
    if (grf && interp_kind == PYCBF_INTERP_NEAREST && param > 1) {
        float *gy = (float *)cv.take(plan.gy);
        float *ge = (float *)cv.take(plan.ge);
        return launch_upsample_adjoint(b.gy, nch, nt, param, t0, dt, gy, ge, grf, s);
    }
    return 0;
}

int pycbf_das_tabbed_bwd(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                         const float *tautx, const float *apodtx, const float *taurx,
                         const float *apodrx, const float *tautx_lo, const float *taurx_lo,
                         int64_t np, int interp_kind, int param, const float *S, float thresh,
                         int sumflag, const float *gout, float *grf, float *gtautx,
                         float *gapodtx, float *gtaurx, float *gapodrx, void *workspace,
                         size_t workspace_bytes, void *stream)
{
    int kparam = 1;
    int rc = bwd_validate(info, rf, nfr, np, interp_kind, param, S, sumflag, gout, kparam);
    if (rc) return rc;
    if (!tautx || !apodtx || !taurx || !apodrx)
        return fail(PYCBF_EINVAL, "delay / apodisation tables must be device pointers");
    BwdArgs b{};
    b.ttx = tautx;
    b.atx = apodtx;
    b.trx = taurx;
    b.arx = apodrx;
    b.ttx_lo = tautx_lo;
    b.trx_lo = taurx_lo;
    b.geo = 0;
    DasArgs tb{};
    tb.ttx = tautx;
    tb.atx = apodtx;
    tb.trx = taurx;
    tb.arx = apodrx;
    tb.geo = 0;
    return bwd_run(info, rf, nfr, np, interp_kind, kparam, S, thresh, sumflag, gout, b, tb, grf,
                   gtautx, gapodtx, gtaurx, gapodrx, nullptr, workspace, workspace_bytes,
                   (hipStream_t)stream);
}

int pycbf_das_synthetic_bwd(const pycbf_rfinfo *info, const float *rf, int64_t nfr,
                            const float *txpack, const float *rxpack, float c0, const float *pvec,
                            int64_t np, int interp_kind, int param, const float *S, float thresh,
                            int sumflag, const float *gout, float *grf, float *gtautx,
                            float *gapodtx, float *gtaurx, float *gapodrx, float *gc0,
                            void *workspace, size_t workspace_bytes, void *stream)
{
    int kparam = 1;
    int rc = bwd_validate(info, rf, nfr, np, interp_kind, param, S, sumflag, gout, kparam);
    if (rc) return rc;
    if (info->ndim != 2 && info->ndim != 3)
        return fail(PYCBF_EINVAL, "synthetic geometry needs ndim 2 or 3");
    if (!txpack || !rxpack || !pvec) return fail(PYCBF_EINVAL, "NULL geometry pointer");
    if (!(c0 > 0.0f) || !isfinite(c0)) return fail(PYCBF_EINVAL, "c0 must be finite and > 0");
    BwdArgs b{};
    b.geo = info->ndim;
    b.txpack = txpack;
    b.rxpack = rxpack;
    b.pvec = pvec;
    b.c0 = c0;
    b.rc0 = 1.0f / c0;
    DasArgs tb{};
    tb.geo = info->ndim;
    tb.txpack = txpack;
    tb.rxpack = rxpack;
    tb.pvec = pvec;
    tb.c0 = c0;
    return bwd_run(info, rf, nfr, np, interp_kind, kparam, S, thresh, sumflag, gout, b, tb, grf,
                   gtautx, gapodtx, gtaurx, gapodrx, gc0, workspace, workspace_bytes,
                   (hipStream_t)stream);
}

// ------------------------------------------------------- speed of sound ---
static int ray_args(const float *slow, int nzm, int nxm, float xm0, float zm0, float dxm,
                    float dzm, const float *steer, const float *xorig, int ntx, float c0,
                    const float *xele, int nrx, const float *pvec, int64_t np, int nsteps,
                    RayArgs &a)
{
    if (!slow || !pvec) return fail(PYCBF_EINVAL, "NULL pointer");
    if (nzm < 1 || nxm < 1 || (int64_t)nzm * nxm > sos_max_map())
        return fail(PYCBF_EINVAL, "slowness map must have 1..%d nodes", sos_max_map());
    if (!(dxm > 0.0f) || !(dzm > 0.0f)) return fail(PYCBF_EINVAL, "map spacing must be > 0");
    if (ntx < 0 || nrx < 0 || ntx + nrx < 1 || ntx + nrx > 65535)
        return fail(PYCBF_EINVAL, "need 1..65535 elements");
    if (ntx > 0 && (!steer || !xorig)) return fail(PYCBF_EINVAL, "steer / xorig are NULL");
    if (nrx > 0 && !xele) return fail(PYCBF_EINVAL, "xele is NULL");
    if (np < 1) return fail(PYCBF_EINVAL, "np must be >= 1");
    if (nsteps < 1 || nsteps > 4096) return fail(PYCBF_EINVAL, "nsteps must be in 1..4096");
    if (!(c0 > 0.0f)) return fail(PYCBF_EINVAL, "c0 must be > 0");
    a.slow = slow;
    a.g.nzm = nzm;
    a.g.nxm = nxm;
    a.g.xm0 = xm0;
    a.g.zm0 = zm0;
    a.g.rdxm = 1.0f / dxm;
    a.g.rdzm = 1.0f / dzm;
    a.xm0d = (double)xm0;
    a.zm0d = (double)zm0;
    a.dxmd = (double)dxm;
    a.dzmd = (double)dzm;
    a.rdxmd = 1.0 / (double)dxm;
    a.rdzmd = 1.0 / (double)dzm;
    a.steer = steer;
    a.xorig = xorig;
    a.xele = xele;
    a.ntx = ntx;
    a.nrx = nrx;
    a.nsteps = nsteps;
    a.c0 = c0;
    a.pvec = pvec;
    a.np = np;
    return 0;
}

int pycbf_ray_delays(const float *slow, int nzm, int nxm, float xm0, float zm0, float dxm,
                     float dzm, const float *steer, const float *xorig, int ntx, float c0,
                     const float *xele, int nrx, const float *pvec, int64_t np, int nsteps,
                     float *tautx, float *taurx, float *tautx_lo, float *taurx_lo, void *stream)
{
    RayArgs a{};
    if (int rc = ray_args(slow, nzm, nxm, xm0, zm0, dxm, dzm, steer, xorig, ntx, c0, xele, nrx,
                          pvec, np, nsteps, a))
        return rc;
    if ((ntx > 0 && !tautx) || (nrx > 0 && !taurx)) return fail(PYCBF_EINVAL, "NULL output");
    a.tautx = tautx;
    a.taurx = taurx;
    a.tautx_lo = tautx_lo;
    a.taurx_lo = taurx_lo;
    return launch_ray_delays(a, 0, (hipStream_t)stream);
}

size_t pycbf_ray_delays_bwd_workspace_bytes(int nzm, int nxm)
{
    if (nzm < 1 || nxm < 1) return 0;
    return ray_bwd_workspace_bytes(nzm, nxm);
}

int pycbf_ray_delays_bwd(const float *slow, int nzm, int nxm, float xm0, float zm0, float dxm,
                         float dzm, const float *steer, const float *xorig, int ntx, float c0,
                         const float *xele, int nrx, const float *pvec, int64_t np, int nsteps,
                         const float *gtautx, const float *gtaurx, float *gslow, void *workspace,
                         size_t workspace_bytes, void *stream)
{
    RayArgs a{};
    if (int rc = ray_args(slow, nzm, nxm, xm0, zm0, dxm, dzm, steer, xorig, ntx, c0, xele, nrx,
                          pvec, np, nsteps, a))
        return rc;
    if (!gslow) return fail(PYCBF_EINVAL, "gslow is NULL");
    const size_t need = ray_bwd_workspace_bytes(nzm, nxm);
    if (!workspace || workspace_bytes < need)
        return fail(PYCBF_EINVAL, "workspace of %zu bytes needed, %zu given", need,
                    workspace_bytes);
    a.gtautx = gtautx;
    a.gtaurx = gtaurx;
    a.gslow = gslow;
    a.gpart = (float *)workspace;
    return launch_ray_delays(a, 1, (hipStream_t)stream);
}

int pycbf_dmas(const float *im, int64_t nfr, int64_t nsep, int64_t np, const int32_t *lags,
               int nlag, int dmastype, int sumlags, float *out, void *stream)
{
    if (!im || !lags || !out) return fail(PYCBF_EINVAL, "NULL pointer");
    if (nfr < 1 || nsep < 1 || np < 1 || nlag < 1) return fail(PYCBF_EINVAL, "empty DMAS problem");
    if (dmastype < PYCBF_DMAS_SSR || dmastype > PYCBF_DMAS_PSSR)
        return fail(PYCBF_EINVAL, "dmastype %d out of range", dmastype);
    return launch_dmas(im, nfr, nsep, np, lags, nlag, dmastype, sumlags, out, (hipStream_t)stream);
}

static int slsc_check_window(int64_t nrx, int64_t nx, int64_t nz, int kwin, int kofs, int dsfx,
                             int dsfz)
{
    if (nrx < 1 || nx < 1 || nz < 1) return fail(PYCBF_EINVAL, "empty SLSC problem");
    if (nrx > 256) return fail(PYCBF_EINVAL, "SLSC supports at most 256 channels (got %lld)",
                               (long long)nrx);
    if (kwin < 1 || kwin > 32) return fail(PYCBF_EINVAL, "kwin must be in 1..32 (got %d)", kwin);
    if (kofs > 0 || kofs < -(kwin - 1))
        return fail(PYCBF_EINVAL, "kofs must be in -(kwin-1)..0 (got %d)", kofs);
    if (dsfx < 1 || dsfz < 1) return fail(PYCBF_EINVAL, "dsfx and dsfz must be >= 1");
    if ((nx + dsfx - 1) / dsfx > 65535)
        return fail(PYCBF_EINVAL, "SLSC grid too large (nx/dsfx must be <= 65535)");
    if ((nz + dsfz - 1) / dsfz > 0x7fffffffLL || nx * nz > (1LL << 40))
        return fail(PYCBF_EINVAL, "SLSC image too large");
    return 0;
}

int pycbf_slsc_selection(const uint8_t *mask, int64_t nrx, int64_t nx, int64_t nz, int kwin,
                         int kofs, int dsfx, int dsfz, uint64_t *sel, void *stream)
{
    if (!mask || !sel) return fail(PYCBF_EINVAL, "NULL pointer");
    if (int rc = slsc_check_window(nrx, nx, nz, kwin, kofs, dsfx, dsfz)) return rc;
    return launch_slsc_selection(mask, nrx, nx, nz, kwin, kofs, dsfx, dsfz, sel,
                                 (hipStream_t)stream);
}

int pycbf_slsc(const float *im, int64_t nfr, int64_t nrx, int64_t nx, int64_t nz,
               const uint64_t *sel, int kwin, int kofs, int dsfx, int dsfz, const int32_t *lags,
               int nlag, int precision, float *out, void *stream)
{
    if (!im || !lags || !out) return fail(PYCBF_EINVAL, "NULL pointer");
    if (nfr < 1) return fail(PYCBF_EINVAL, "empty SLSC problem");
    if (int rc = slsc_check_window(nrx, nx, nz, kwin, kofs, dsfx, dsfz)) return rc;
    if (nlag < 1 || nlag > 64) return fail(PYCBF_EINVAL, "nlag must be in 1..64 (got %d)", nlag);
    for (int i = 0; i < nlag; ++i)
        if (lags[i] < 1 || lags[i] > 63)
            return fail(PYCBF_EINVAL, "lag %d out of range 1..63", (int)lags[i]);
    if (precision < PYCBF_SLSC_FP32 || precision > PYCBF_SLSC_MFMA_BF16)
        return fail(PYCBF_EINVAL, "precision %d out of range", precision);
    if (nfr > 65535) return fail(PYCBF_EINVAL, "nfr must be <= 65535");
    return launch_slsc(im, nfr, nrx, nx, nz, sel, kwin, kofs, dsfx, dsfz, lags, nlag, precision,
                       out, (hipStream_t)stream);
}

}  // extern "C"
