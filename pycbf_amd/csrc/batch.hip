// Ragged batches of Synthetic cubic DAS / SLSC requests in one persistent
// launch (no reference counterpart: pycbf beamforms one configuration per
// call, pycbf/gpu/__core_bf_classes__.py:68-118).
//
// A batch is a list of requests, each with its own plane-wave / element
// geometry (pycbf_synthetic_pack), x-major pixel grid, transmit / receive
// counts and output.  All requests share nt, so their channels are
// prefiltered by one launch of the natural-spline kernel over the
// concatenated channel data.  The gather is one persistent kernel: every wave
// pulls work items (request, TILE x TILE pixel tile, chunk of RC receive
// elements) from a queue ordered by cost (largest first) through one atomic
// counter, and exits when the queue is empty.  Within an item a lane owns a
// pixel; it holds the delays / apodisations of 8 transmits (16 for SLSC
// requests) in registers while it sweeps the receive chunk (receive geometry
// evaluated in registers) and gathers the cubic table.  DAS requests add
// their partial sums into the zeroed image with float atomics; SLSC requests
// write per-receive images (every (receive element, pixel) is owned by one
// item), which the per-request SLSC launches then reduce.  Each item stamps
// its request's completion time (wall clock) for per-request latencies.
// Every launch of a run is a kernel (no memset), so a run can be captured in
// a graph.
#include <math.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "das_dev.h"
#include "prof.h"

namespace pycbf {

struct BatchReqDev {
    const float *txpack, *rxpack, *pvec;
    const float *coef;  // (ntx*nrx, nt+1) (y, e) pairs of this request
    float *out;         // DAS: (np); SLSC: per-receive images (nrx, np)
    int ntx, nrx, nx, nz, kind, nt;
    float c0, t0, dt, rdt, tmax, thresh;
};

struct BatchItem {
    int req, tile, r0, pad;
};

namespace {

// cubic value of channel row `row` at x (0 outside the trace)
__device__ __forceinline__ float cubic_at(const float *__restrict__ row, float x, const BatchReqDev &R)
{
    int j = 0;
    if (!knot_index(x, R.t0, R.dt, R.tmax, R.nt, j)) return 0.0f;
    const float xlead = R.t0 + R.dt * (float)(j + 1);
    const float xlag = R.t0 + R.dt * (float)j;
    const float a = (xlead - x) * R.rdt;
    const float b = (x - xlag) * R.rdt;
    const float ca = __builtin_fmaf(a * a, a, -a);
    const float cb = __builtin_fmaf(b * b, b, -b);
    const Pair2 c = *reinterpret_cast<const Pair2 *>(row + 2 * j);
    float v = a * c.y0;
    v = __builtin_fmaf(b, c.y1, v);
    v = __builtin_fmaf(ca, c.z0, v);
    return __builtin_fmaf(cb, c.z1, v);
}

// one work item: TILE x TILE pixels (64 per pass, one per lane), receive
// elements r0 .. r0+RC-1, all transmits.  Transmits are taken in blocks of
// BATCH_TB whose delays / apodisations stay in registers while the receive
// chunk is swept; per-receive images (KIND 1) accumulate in place, since each
// (receive element, pixel) belongs to this item alone, once per block: they
// take blocks of 16 (half the read-modify-write passes; the bench batch's
// best gather 20.46 -> 17.74 ms), summed images 8.
template <int KIND, int TILE, int RC>
__device__ __forceinline__ void batch_item(const BatchReqDev &R, int tile, int r0, int lane)
{
    constexpr int BATCH_TB = KIND == 1 ? 16 : 8;  // transmits a lane holds in registers
    const int tiles_z = (R.nz + TILE - 1) / TILE;
    const int tx0 = (tile / tiles_z) * TILE, tz0 = (tile % tiles_z) * TILE;
    const int nrc = min(RC, R.nrx - r0);
    const int64_t np = (int64_t)R.nx * R.nz;
    const int64_t rowlen = 2 * ((int64_t)R.nt + 1);
    constexpr int CPP = 64 / TILE;  // tile columns per 64-pixel pass
#pragma unroll 1
    for (int pass = 0; pass < TILE / CPP; ++pass) {
        const int ix = tx0 + pass * CPP + lane / TILE, iz = tz0 + lane % TILE;
        if (tx0 + pass * CPP >= R.nx) break;  // uniform: the rest of the tile is outside
        const bool act = ix < R.nx && iz < R.nz;
        const int64_t p = act ? (int64_t)ix * R.nz + iz : 0;
        // every lane loads (pixel 0 past the grid), then a full wait: no load
        // of the position is still counted when the receive loop's gathers are
        float pv[3] = {R.pvec[2 * p], R.pvec[2 * p + 1], 0.0f};
        __builtin_amdgcn_s_waitcnt(0);
        float acc = 0.0f;
#pragma unroll 1
        for (int t0 = 0; t0 < R.ntx; t0 += BATCH_TB) {
            float ttx[BATCH_TB], atx[BATCH_TB];
#pragma unroll
            for (int k = 0; k < BATCH_TB; ++k) {
                ttx[k] = 0.0f;
                atx[k] = 0.0f;
                if (t0 + k < R.ntx) {
                    float pk[PYCBF_TXPACK_FLOATS];
                    load_pack(R.txpack + (int64_t)(t0 + k) * PYCBF_TXPACK_FLOATS, pk);
                    geom_tx<2>(pk, R.c0, pv, ttx[k], atx[k]);
                }
            }
#pragma unroll 1
            for (int i = 0; i < nrc; ++i) {
                float trx = 0.0f, arx = 0.0f;
                float pk[PYCBF_RXPACK_FLOATS];
                load_pack(R.rxpack + (int64_t)(r0 + i) * PYCBF_RXPACK_FLOATS, pk);
                geom_rx<2>(pk, R.c0, pv, trx, arx);
                const float *base = R.coef + ((int64_t)t0 * R.nrx + r0 + i) * rowlen;
                float part = 0.0f;
#pragma unroll
                for (int k = 0; k < BATCH_TB; ++k) {
                    if (t0 + k >= R.ntx) break;
                    const float w = atx[k] * arx;
                    if (act && fabsf(w) >= R.thresh)
                        part = __builtin_fmaf(
                            w, cubic_at(base + (int64_t)k * R.nrx * rowlen, ttx[k] + trx, R),
                            part);
                }
                if (KIND == 0) {
                    acc += part;
                } else if (act) {
                    float *o = R.out + (int64_t)(r0 + i) * np + p;
                    *o = (t0 == 0) ? part : *o + part;
                }
            }
        }
        if (KIND == 0 && act) unsafeAtomicAdd(R.out + p, acc);
    }
}

}  // namespace

template <int TILE, int RC>
__global__ __launch_bounds__(256) void k_batch_das(const BatchReqDev *__restrict__ reqs,
                                                   const BatchItem *__restrict__ items, int nitems,
                                                   unsigned *counter,
                                                   unsigned long long *finish)
{
    const int lane = threadIdx.x & 63;
    for (;;) {
        // every lane executes the atomic, lane 0 adding 1 and the others 0, and
        // readfirstlane takes lane 0's result into a scalar register: the item
        // index and the exit test are wave-uniform by construction.  (Pulling
        // under `if (lane == 0)` let the compiler split the loop on that
        // loop-invariant lane test, parking lane 0 outside the loop the other
        // lanes kept running on a stale index.)
        const unsigned v = atomicAdd(counter, lane == 0 ? 1u : 0u);
        const int it = (int)__builtin_amdgcn_readfirstlane(v);
        if (it >= nitems) break;  // every wave leaves once the queue is drained
        const BatchItem w = items[it];
        const BatchReqDev R = reqs[w.req];
        if (R.kind == 0) batch_item<0, TILE, RC>(R, w.tile, w.r0, lane);
        else batch_item<1, TILE, RC>(R, w.tile, w.r0, lane);
        if (lane == 0) atomicMax(finish + w.req, (unsigned long long)wall_clock64());
    }
}

// clears the queue counter, the completion stamps and the DAS images, and
// records the batch start time
__global__ __launch_bounds__(256) void k_batch_reset(unsigned *counter, unsigned long long *finish,
                                                     int nreq, float *zero, int64_t nzero,
                                                     unsigned long long *start)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g == 0) {
        *counter = 0u;
        *start = (unsigned long long)wall_clock64();
    }
    if (g < nreq) finish[g] = 0ull;
    for (int64_t i = g; i < nzero; i += (int64_t)gridDim.x * 256) zero[i] = 0.0f;
}

__global__ void k_batch_stamp(unsigned long long *slot)
{
    if (threadIdx.x == 0) *slot = (unsigned long long)wall_clock64();
}

// ---------------------------------------------------------------- host side
struct BatchPlan {
    std::vector<pycbf_batch_req> reqs;
    std::vector<int64_t> ch_off;  // first channel of each request in the concatenation
    int64_t nch = 0, npix_das = 0;
    int nt = 0;
    int tile = 16, rchunk = 32;
    int nitems = 0, max_items = 0;
    float *das_out_base = nullptr;  // DAS outputs are zeroed as one range when contiguous
};

static int64_t items_for(const std::vector<pycbf_batch_req> &reqs, int tile, int rchunk,
                         std::vector<BatchItem> *out)
{
    struct Cost {
        double c;
        BatchItem it;
    };
    std::vector<Cost> v;
    for (size_t r = 0; r < reqs.size(); ++r) {
        const pycbf_batch_req &q = reqs[r];
        const int tx = (q.nx + tile - 1) / tile, tz = (q.nz + tile - 1) / tile;
        for (int r0 = 0; r0 < q.nrx; r0 += rchunk)
            for (int t = 0; t < tx * tz; ++t) {
                const double c = (double)q.ntx * std::min(rchunk, q.nrx - r0) * tile * tile;
                if (out) v.push_back({c, BatchItem{(int)r, t, r0, 0}});
                else v.push_back({c, BatchItem{}});
            }
    }
    if (out) {
        std::stable_sort(v.begin(), v.end(), [](const Cost &a, const Cost &b) { return a.c > b.c; });
        out->clear();
        for (auto &c : v) out->push_back(c.it);
    }
    return (int64_t)v.size();
}

struct BatchLayout {
    size_t coef = 0, uscr = 0, reqs = 0, items = 0, ctrl = 0, total = 0;
};

static BatchLayout batch_layout(const BatchPlan &P)
{
    BatchLayout L;
    L.coef = align_up((size_t)P.nch * (size_t)(P.nt + 1) * 8, 256);
    L.uscr = spline_scratch_bytes(P.nch, P.nt);
    L.reqs = align_up(P.reqs.size() * sizeof(BatchReqDev), 256);
    L.items = align_up((size_t)P.max_items * sizeof(BatchItem), 256);
    L.ctrl = align_up(256 + (P.reqs.size() + 2) * 8, 256);
    L.total = L.coef + L.uscr + L.reqs + L.items + L.ctrl + 6 * 256;
    return L;
}

struct BatchPtrs {
    float2 *coef;
    float *uscr;
    BatchReqDev *reqs;
    BatchItem *items;
    unsigned *counter;
    unsigned long long *start, *finish;
};

static BatchPtrs carve(const BatchPlan &P, void *ws, size_t bytes)
{
    const BatchLayout L = batch_layout(P);
    Carver cv(ws, bytes);
    BatchPtrs b;
    b.coef = (float2 *)cv.take(L.coef);
    b.uscr = (float *)cv.take(L.uscr);
    b.reqs = (BatchReqDev *)cv.take(L.reqs);
    b.items = (BatchItem *)cv.take(L.items);
    char *ctrl = (char *)cv.take(L.ctrl);
    b.counter = (unsigned *)ctrl;
    b.start = (unsigned long long *)(ctrl + 256);
    b.finish = b.start + 1;
    return b;
}

// uploads the request table and the item list for (tile, rchunk)
static int upload(BatchPlan &P, const BatchPtrs &b, hipStream_t s)
{
    std::vector<BatchReqDev> dev(P.reqs.size());
    for (size_t r = 0; r < P.reqs.size(); ++r) {
        const pycbf_batch_req &q = P.reqs[r];
        BatchReqDev &d = dev[r];
        d.txpack = q.txpack;
        d.rxpack = q.rxpack;
        d.pvec = q.pvec;
        d.coef = (const float *)(b.coef + P.ch_off[r] * (int64_t)(P.nt + 1));
        d.out = q.kind == PYCBF_BATCH_DAS ? q.out : q.imrx;
        d.ntx = q.ntx;
        d.nrx = q.nrx;
        d.nx = q.nx;
        d.nz = q.nz;
        d.kind = q.kind == PYCBF_BATCH_DAS ? 0 : 1;
        d.nt = P.nt;
        d.c0 = q.c0;
        d.t0 = q.t0;
        d.dt = q.dt;
        d.rdt = 1.0f / q.dt;
        d.tmax = q.t0 + q.dt * (float)(P.nt - 1);
        d.thresh = q.thresh;
    }
    std::vector<BatchItem> items;
    P.nitems = (int)items_for(P.reqs, P.tile, P.rchunk, &items);
    if (int rc = check(hipMemcpyAsync(b.reqs, dev.data(), dev.size() * sizeof(BatchReqDev),
                                      hipMemcpyHostToDevice, s),
                       "batch request upload"))
        return rc;
    if (int rc = check(hipMemcpyAsync(b.items, items.data(), items.size() * sizeof(BatchItem),
                                      hipMemcpyHostToDevice, s),
                       "batch item upload"))
        return rc;
    // the host vectors go out of scope: wait for the copies
    return check(hipStreamSynchronize(s), "batch upload");
}

// Workgroups of a persistent launch: as many as can be resident at once
// (every CU times the occupancy the kernel's registers allow), so that the
// queue is drained by the full complement of waves.
template <typename K>
static int persistent_blocks(K kernel)
{
    int dev = 0, cus = 256, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 4;
    return cus * per_cu;
}

template <int TILE, int RC>
static void launch_gather_t(const BatchPlan &P, const BatchPtrs &b, hipStream_t s)
{
    static const int blocks = persistent_blocks(k_batch_das<TILE, RC>);
    hipLaunchKernelGGL((k_batch_das<TILE, RC>), dim3(blocks), dim3(256), 0, s, b.reqs, b.items,
                       P.nitems, b.counter, b.finish);
}

static int launch_gather(const BatchPlan &P, const BatchPtrs &b, hipStream_t s)
{
    ProfScope ps("batch_gather", s);
    switch (P.tile * 100 + P.rchunk) {
    case 816: launch_gather_t<8, 16>(P, b, s); break;
    case 832: launch_gather_t<8, 32>(P, b, s); break;
    case 864: launch_gather_t<8, 64>(P, b, s); break;
    case 1616: launch_gather_t<16, 16>(P, b, s); break;
    case 1632: launch_gather_t<16, 32>(P, b, s); break;
    case 1664: launch_gather_t<16, 64>(P, b, s); break;
    case 3216: launch_gather_t<32, 16>(P, b, s); break;
    case 3232: launch_gather_t<32, 32>(P, b, s); break;
    case 3264: launch_gather_t<32, 64>(P, b, s); break;
    case 928: launch_gather_t<8, 128>(P, b, s); break;     // 8 * 100 + 128
    case 1728: launch_gather_t<16, 128>(P, b, s); break;   // 16 * 100 + 128
    case 3328: launch_gather_t<32, 128>(P, b, s); break;   // 32 * 100 + 128
    default: return fail(PYCBF_EINVAL, "unknown gather configuration");
    }
    return check_launch("batch gather");
}

static int run_batch(BatchPlan &P, const float *rf_all, void *ws, size_t bytes, hipStream_t s,
                     bool gather_only)
{
    const BatchPtrs b = carve(P, ws, bytes);
    const int64_t nzero = P.das_out_base ? P.npix_das : 0;
    const int zb = (int)std::min<int64_t>(std::max<int64_t>((nzero + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(k_batch_reset, dim3((unsigned)std::max<int64_t>(zb, (P.reqs.size() + 255) / 256)),
                       dim3(256), 0, s, b.counter, b.finish, (int)P.reqs.size(), P.das_out_base,
                       nzero, b.start);
    if (int rc = check_launch("batch reset")) return rc;
    if (!gather_only)
        if (int rc = launch_spline_coefs(rf_all, P.nch, P.nt, P.reqs[0].dt, b.uscr, b.coef, s))
            return rc;
    if (int rc = launch_gather(P, b, s)) return rc;
    if (gather_only) return 0;
    for (size_t r = 0; r < P.reqs.size(); ++r) {
        const pycbf_batch_req &q = P.reqs[r];
        if (q.kind != PYCBF_BATCH_SLSC) continue;
        if (int rc = launch_slsc(q.imrx, 1, q.nrx, q.nx, q.nz, q.sel, q.kwin, q.kofs, q.dsfx,
                                 q.dsfz, q.lags, q.nlag, q.precision, q.out, s))
            return rc;
        hipLaunchKernelGGL(k_batch_stamp, dim3(1), dim3(64), 0, s, b.finish + r);
        if (int rc = check_launch("batch stamp")) return rc;
    }
    return 0;
}

}  // namespace pycbf

using namespace pycbf;

struct pycbf_batch {
    BatchPlan plan;
};

extern "C" {

int pycbf_batch_create(const pycbf_batch_req *reqs, int nreq, pycbf_batch **out)
{
    if (!reqs || nreq < 1 || !out) return fail(PYCBF_EINVAL, "empty batch");
    BatchPlan P;
    P.nt = reqs[0].nt;
    int64_t chans = 0, das_pix = 0;
    float *das_next = nullptr;
    bool das_contig = true;
    for (int r = 0; r < nreq; ++r) {
        const pycbf_batch_req &q = reqs[r];
        if (q.kind != PYCBF_BATCH_DAS && q.kind != PYCBF_BATCH_SLSC)
            return fail(PYCBF_EINVAL, "request %d: kind %d", r, q.kind);
        if (q.ntx < 1 || q.nrx < 1 || q.nx < 1 || q.nz < 1 || q.nt != P.nt || q.nt < 1)
            return fail(PYCBF_EINVAL, "request %d: sizes (all requests share nt)", r);
        if (!q.txpack || !q.rxpack || !q.pvec || !q.out)
            return fail(PYCBF_EINVAL, "request %d: NULL pointer", r);
        if (!(q.dt > 0.0f) || q.dt != reqs[0].dt)
            return fail(PYCBF_EINVAL, "request %d: all requests share dt > 0", r);
        if (q.kind == PYCBF_BATCH_SLSC) {
            if (!q.imrx) return fail(PYCBF_EINVAL, "request %d: SLSC needs imrx scratch", r);
            if (q.nrx > 256 || q.kwin < 1 || q.kwin > 32 || q.kofs > 0 || q.kofs < -(q.kwin - 1) ||
                q.dsfx < 1 || q.dsfz < 1 || q.nlag < 1 || q.nlag > 64 ||
                q.precision < PYCBF_SLSC_FP32 || q.precision > PYCBF_SLSC_MFMA_BF16)
                return fail(PYCBF_EINVAL, "request %d: SLSC arguments", r);
            for (int i = 0; i < q.nlag; ++i)
                if (q.lags[i] < 1 || q.lags[i] > 63)
                    return fail(PYCBF_EINVAL, "request %d: lag out of range", r);
        } else {
            // DAS images are cleared by one range when they are laid out back to back
            if (das_next && q.out != das_next) das_contig = false;
            if (!P.das_out_base) P.das_out_base = q.out;
            das_next = q.out + (int64_t)q.nx * q.nz;
            das_pix += (int64_t)q.nx * q.nz;
        }
        P.ch_off.push_back(chans);
        chans += (int64_t)q.ntx * q.nrx;
        P.reqs.push_back(q);
    }
    if (!das_contig)
        return fail(PYCBF_EINVAL, "DAS outputs must be contiguous, in request order");
    P.nch = chans;
    P.npix_das = das_pix;
    for (int t : {8, 16, 32})
        for (int c : {16, 32, 64, 128})
            P.max_items = (int)std::max<int64_t>(P.max_items, items_for(P.reqs, t, c, nullptr));
    *out = new pycbf_batch{P};
    return 0;
}

void pycbf_batch_destroy(pycbf_batch *b) { delete b; }

size_t pycbf_batch_workspace_bytes(const pycbf_batch *b)
{
    return b ? batch_layout(b->plan).total : 0;
}

int pycbf_batch_configure(pycbf_batch *b, int tile, int rchunk, void *workspace,
                          size_t workspace_bytes, void *stream)
{
    if (!b) return fail(PYCBF_EINVAL, "NULL batch");
    if ((tile != 8 && tile != 16 && tile != 32) ||
        (rchunk != 16 && rchunk != 32 && rchunk != 64 && rchunk != 128))
        return fail(PYCBF_EINVAL, "tile must be 8, 16 or 32 and rchunk 16, 32, 64 or 128");
    if (!workspace || workspace_bytes < batch_layout(b->plan).total)
        return fail(PYCBF_EINVAL, "workspace too small");
    b->plan.tile = tile;
    b->plan.rchunk = rchunk;
    return upload(b->plan, carve(b->plan, workspace, workspace_bytes), (hipStream_t)stream);
}

int pycbf_batch_run(pycbf_batch *b, const float *rf_all, void *workspace, size_t workspace_bytes,
                    void *stream)
{
    if (!b || !rf_all) return fail(PYCBF_EINVAL, "NULL argument");
    if (!workspace || workspace_bytes < batch_layout(b->plan).total)
        return fail(PYCBF_EINVAL, "workspace too small");
    if (b->plan.nitems == 0) return fail(PYCBF_EINVAL, "batch not configured");
    return run_batch(b->plan, rf_all, workspace, workspace_bytes, (hipStream_t)stream, false);
}

int pycbf_batch_autotune(pycbf_batch *b, const float *rf_all, int reps, void *workspace,
                         size_t workspace_bytes, void *stream, int *best_tile, int *best_rchunk,
                         float *times_ms)
{
    if (!b || !rf_all || reps < 1) return fail(PYCBF_EINVAL, "bad autotune arguments");
    hipStream_t s = (hipStream_t)stream;
    const int tiles[3] = {8, 16, 32}, chunks[4] = {16, 32, 64, 128};
    float best = 1e30f;
    int bt = 16, bc = 32;
    hipEvent_t e0, e1;
    if (int rc = check(hipEventCreate(&e0), "event")) return rc;
    if (int rc = check(hipEventCreate(&e1), "event")) return rc;
    int rc = pycbf_batch_configure(b, 16, 32, workspace, workspace_bytes, stream);
    if (!rc) rc = run_batch(b->plan, rf_all, workspace, workspace_bytes, s, false);  // tables
    for (int ti = 0; ti < 3 && !rc; ++ti)
        for (int ci = 0; ci < 4 && !rc; ++ci) {
            rc = pycbf_batch_configure(b, tiles[ti], chunks[ci], workspace, workspace_bytes, stream);
            if (rc) break;
            rc = run_batch(b->plan, rf_all, workspace, workspace_bytes, s, true);  // warm
            if (rc) break;
            hipEventRecord(e0, s);
            for (int k = 0; k < reps && !rc; ++k)
                rc = run_batch(b->plan, rf_all, workspace, workspace_bytes, s, true);
            hipEventRecord(e1, s);
            if (!rc) rc = check(hipEventSynchronize(e1), "autotune sync");
            float ms = 0.0f;
            hipEventElapsedTime(&ms, e0, e1);
            ms /= (float)reps;
            if (times_ms) times_ms[ti * 4 + ci] = ms;
            if (ms < best) {
                best = ms;
                bt = tiles[ti];
                bc = chunks[ci];
            }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    if (best_tile) *best_tile = bt;
    if (best_rchunk) *best_rchunk = bc;
    return pycbf_batch_configure(b, bt, bc, workspace, workspace_bytes, stream);
}

int pycbf_batch_latencies(const pycbf_batch *b, const void *workspace, size_t workspace_bytes,
                          double *latency_ms)
{
    if (!b || !workspace || !latency_ms) return fail(PYCBF_EINVAL, "NULL argument");
    const BatchPtrs p = carve(b->plan, const_cast<void *>(workspace), workspace_bytes);
    const size_t n = b->plan.reqs.size();
    std::vector<unsigned long long> v(n + 1);
    if (int rc = check(hipMemcpy(v.data(), p.start, (n + 1) * 8, hipMemcpyDeviceToHost),
                       "latency readback"))
        return rc;
    int dev = 0, khz = 0;
    hipGetDevice(&dev);
    if (int rc = check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev),
                       "wall clock rate"))
        return rc;
    for (size_t r = 0; r < n; ++r)
        latency_ms[r] = (double)(long long)(v[r + 1] - v[0]) / (double)khz;
    return 0;
}

}  // extern "C"
