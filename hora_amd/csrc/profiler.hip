// profiler.hip -- per-kernel HIP-event timing (off unless hm_profile_enable(1)).
#include "profiler.h"

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace hm {

namespace {
std::atomic<bool> g_on{false};
std::mutex g_mu;
struct Pending {
    std::string name;
    hipEvent_t a, b;
};
std::vector<Pending> g_pending;
struct Total {
    double ms = 0;
    uint64_t n = 0;
};
std::map<std::string, Total> g_totals;
std::string g_only;  // when set, only this family is timed

void drain_locked()
{
    for (auto &p : g_pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto &t = g_totals[p.name];
            t.ms += ms;
            t.n += 1;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    g_pending.clear();
}
}  // namespace

bool profiling_enabled() { return g_on.load(std::memory_order_relaxed); }

ProfScope::ProfScope(const char *name, hipStream_t s) : name_(name), s_(s)
{
    if (!profiling_enabled()) return;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_only.empty() && g_only != name) return;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    if (hipEventCreate(&start_) != hipSuccess || hipEventCreate(&stop_) != hipSuccess) {
        start_ = stop_ = nullptr;
        return;
    }
    (void)hipEventRecord(start_, s_);
}

ProfScope::~ProfScope()
{
    if (!start_) return;
    (void)hipEventRecord(stop_, s_);
    std::lock_guard<std::mutex> lk(g_mu);
    g_pending.push_back({name_, start_, stop_});
}

void profile_reset()
{
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    g_totals.clear();
}

bool profile_query(const char *name, double *total_ms, uint64_t *launches)
{
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    auto it = g_totals.find(name);
    if (it == g_totals.end()) {
        *total_ms = 0;
        *launches = 0;
        return false;
    }
    *total_ms = it->second.ms;
    *launches = it->second.n;
    return true;
}

}  // namespace hm

extern "C" {
HM_API int32_t hm_profile_enable(int32_t on)
{
    hm::g_on.store(on != 0);
    return HM_OK;
}
HM_API int32_t hm_profile_only(const char *kernel)
{
    std::lock_guard<std::mutex> lk(hm::g_mu);
    hm::g_only = kernel ? kernel : "";
    return HM_OK;
}
HM_API int32_t hm_profile_reset(void)
{
    hm::profile_reset();
    return HM_OK;
}
HM_API int32_t hm_profile_query(const char *kernel, double *total_ms, uint64_t *launches)
{
    if (!kernel || !total_ms || !launches) return HM_ERR_INVALID_ARG;
    hm::profile_query(kernel, total_ms, launches);
    return HM_OK;
}
}
