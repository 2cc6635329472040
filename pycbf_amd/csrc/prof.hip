// Optional per-kernel timing with HIP events recorded on the launch stream.
// Off by default; when enabled each instrumented launch records a start/stop
// event pair, and pycbf_profile_read() synchronises the pairs and reports the
// summed duration per kernel name.
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "prof.h"

namespace pycbf {

namespace {
struct Pair {
    const char *name;
    hipEvent_t a, b;
};
struct Stat {
    std::string name;
    double ms = 0.0;
    int64_t n = 0;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Pair> g_pending;
std::vector<Stat> g_stats;

void drain_locked()
{
    for (auto &p : g_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Stat *s = nullptr;
            for (auto &st : g_stats)
                if (st.name == p.name) s = &st;
            if (!s) {
                g_stats.push_back(Stat{p.name, 0.0, 0});
                s = &g_stats.back();
            }
            s->ms += ms;
            s->n += 1;
        }
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    g_pending.clear();
}
}  // namespace

ProfScope::ProfScope(const char *name, hipStream_t s) : name_(name), s_(s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_on) return;
    if (hipEventCreate(&a_) != hipSuccess) return;
    if (hipEventCreate(&b_) != hipSuccess) {
        hipEventDestroy(a_);
        return;
    }
    hipEventRecord(a_, s_);
    live_ = true;
}

ProfScope::~ProfScope()
{
    if (!live_) return;
    hipEventRecord(b_, s_);
    std::lock_guard<std::mutex> lk(g_mu);
    g_pending.push_back(Pair{name_, a_, b_});
}

}  // namespace pycbf

using namespace pycbf;

extern "C" {

void pycbf_profile_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = on != 0;
}

void pycbf_profile_reset(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    g_stats.clear();
}

int pycbf_profile_read(const char *kernel, double *total_ms, int64_t *launches)
{
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    for (auto &st : g_stats)
        if (st.name == kernel) {
            if (total_ms) *total_ms = st.ms;
            if (launches) *launches = st.n;
            return 0;
        }
    if (total_ms) *total_ms = 0.0;
    if (launches) *launches = 0;
    return 0;
}

}  // extern "C"
