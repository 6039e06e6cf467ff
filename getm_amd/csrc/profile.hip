// Optional per-kernel timing with HIP events recorded on the launch stream.
// Disabled by default (one branch per launch); when enabled, every kernel
// launch of the library is bracketed by two events on the stream it is
// launched on, and getm_profile_read() accumulates their elapsed times.
// This is instrumentation state, separate from the operator objects.
#include <mutex>
#include <vector>

#include "internal.h"

namespace getm {
namespace {
struct Pending {
    int id;
    hipEvent_t a, b;
};
std::mutex g_mu;
bool g_enabled = false;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_free;
double g_ms[GETM_NUM_KERNELS];
long g_count[GETM_NUM_KERNELS];

hipEvent_t take_event()
{
    if (!g_free.empty()) {
        hipEvent_t e = g_free.back();
        g_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void drain_locked()
{
    for (auto& p : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            g_ms[p.id] += ms;
            g_count[p.id] += 1;
        }
        g_free.push_back(p.a);
        g_free.push_back(p.b);
    }
    g_pending.clear();
}
}  // namespace

ProfileScope::ProfileScope(int id, hipStream_t s) : id_(id), s_(s)
{
    if (!g_enabled) return;
    std::lock_guard<std::mutex> lk(g_mu);
    a_ = take_event();
    b_ = take_event();
    if (a_ && b_) hipEventRecord(static_cast<hipEvent_t>(a_), s_);
}

ProfileScope::~ProfileScope()
{
    if (!a_ || !b_) return;
    std::lock_guard<std::mutex> lk(g_mu);
    hipEventRecord(static_cast<hipEvent_t>(b_), s_);
    g_pending.push_back({id_, static_cast<hipEvent_t>(a_), static_cast<hipEvent_t>(b_)});
    if (g_pending.size() > 4096) drain_locked();
}

}  // namespace getm

extern "C" {

int getm_profile_enable(int enable)
{
    std::lock_guard<std::mutex> lk(getm::g_mu);
    getm::drain_locked();
    getm::g_enabled = enable != 0;
    return GETM_OK;
}

int getm_profile_read(int kernel, double* total_ms, long* launches, int reset)
{
    if (kernel < 0 || kernel >= GETM_NUM_KERNELS) return GETM_ERR_ARG;
    std::lock_guard<std::mutex> lk(getm::g_mu);
    getm::drain_locked();
    if (total_ms) *total_ms = getm::g_ms[kernel];
    if (launches) *launches = getm::g_count[kernel];
    if (reset) {
        getm::g_ms[kernel] = 0.0;
        getm::g_count[kernel] = 0;
    }
    return GETM_OK;
}

}  // extern "C"
