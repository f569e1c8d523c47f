// Optional per-kernel timing with HIP events recorded on the launch stream.
// Off by default; enabled through omh_profile_enable for benchmarks.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace omh {

namespace {
struct Rec {
    std::string name;
    double work;
    hipEvent_t a, b;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
struct Agg {
    int64_t calls = 0;
    double seconds = 0.0, work = 0.0;
};
std::map<std::string, Agg> g_agg;

void drain_locked() {
    for (auto& r : g_recs) {
        float ms = 0.f;
        (void)hipEventSynchronize(r.b);
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        Agg& a = g_agg[r.name];
        a.calls += 1;
        a.seconds += ms * 1e-3;
        a.work += r.work;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_recs.clear();
}
}  // namespace

bool prof_enabled() { return g_on; }

ProfScope::ProfScope(const char* name, double work, hipStream_t st) : st_(st) {
    if (!g_on) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return;   // not in graphs
    active_ = true;
    name_ = name;
    work_ = work;
    OMH_HIP(hipEventCreate(&a_));
    OMH_HIP(hipEventCreate(&b_));
    OMH_HIP(hipEventRecord(a_, st_));
}

void ProfScope::cancel() {
    if (!active_) return;
    active_ = false;
    (void)hipEventDestroy(a_);
    (void)hipEventDestroy(b_);
}

ProfScope::~ProfScope() {
    if (!active_) return;
    (void)hipEventRecord(b_, st_);
    std::lock_guard<std::mutex> lk(g_mu);
    g_recs.push_back({name_, work_, a_, b_});
}

void prof_enable(bool on) {
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    g_agg.clear();
    g_on = on;
}

bool prof_read(const std::string& name, int64_t& calls, double& seconds, double& work) {
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    auto it = g_agg.find(name);
    if (it == g_agg.end()) {
        calls = 0;
        seconds = work = 0.0;
        return false;
    }
    calls = it->second.calls;
    seconds = it->second.seconds;
    work = it->second.work;
    return true;
}

}  // namespace omh
