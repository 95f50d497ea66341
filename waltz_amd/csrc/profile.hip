// Optional per-stage timing with HIP events recorded on the stream the stage's
// kernels run on. Off by default (no events recorded); when enabled, each
// stage is bracketed by two events and the durations are accumulated after the
// pipeline's own synchronisation point.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/waltz_amd.h"
#include "profile.hpp"

namespace wz {
namespace {
std::atomic<bool> g_enabled{false};
std::mutex g_lock;
struct Pending {
    const char* name;
    hipEvent_t a, b;
};
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_pool;
std::map<std::string, std::pair<double, long>> g_totals;

hipEvent_t take() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

StageClock::StageClock(const char* name, hipStream_t s) : name_(name), s_(s) {
    if (!g_enabled.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> g(g_lock);
    a_ = take();
    b_ = take();
    if (a_ && b_) hipEventRecord(static_cast<hipEvent_t>(a_), s_);
}

StageClock::~StageClock() {
    if (!a_ || !b_) return;
    hipEventRecord(static_cast<hipEvent_t>(b_), s_);
    std::lock_guard<std::mutex> g(g_lock);
    g_pending.push_back({name_, static_cast<hipEvent_t>(a_), static_cast<hipEvent_t>(b_)});
}

void profile_collect() {
    if (!g_enabled.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> g(g_lock);
    for (auto& p : g_pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto& t = g_totals[p.name];
            t.first += ms;
            t.second += 1;
        }
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
}

}  // namespace wz

extern "C" {

/* Enable (1) or disable (0) per-stage event timing; returns the old state. */
int waltz_amd_profile_enable(int on) { return wz::g_enabled.exchange(on != 0) ? 1 : 0; }

/* JSON object {"stage": [total_ms, launches], ...} accumulated since the last
 * report; resets the totals. Returns the length written (or needed). */
size_t waltz_amd_profile_report(char* buf, size_t cap) {
    wz::profile_collect();
    std::string out = "{";
    {
        std::lock_guard<std::mutex> g(wz::g_lock);
        bool first = true;
        for (auto& kv : wz::g_totals) {
            char tmp[160];
            std::snprintf(tmp, sizeof tmp, "%s\"%s\": [%.6f, %ld]", first ? "" : ", ", kv.first.c_str(),
                          kv.second.first, kv.second.second);
            out += tmp;
            first = false;
        }
        wz::g_totals.clear();
    }
    out += "}";
    if (buf && cap) {
        const size_t k = out.size() < cap - 1 ? out.size() : cap - 1;
        std::memcpy(buf, out.data(), k);
        buf[k] = 0;
    }
    return out.size();
}

}  // extern "C"
