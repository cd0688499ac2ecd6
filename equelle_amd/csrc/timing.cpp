// Live kernel timing (see timing.hpp).
#include "timing.hpp"

#include <map>
#include <mutex>
#include <vector>

#include "core.hpp"

namespace eqhip {

namespace {
struct Timer {
    std::mutex mu;
    bool on = false;
    std::vector<hipEvent_t> pool;   // free events
    struct Rec {
        std::vector<std::pair<hipEvent_t, hipEvent_t>> done;
        hipEvent_t open = nullptr;
    };
    std::map<std::string, Rec> recs;
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        EQ_HIP(hipEventCreate(&e));
        return e;
    }
};
Timer& T() {
    static Timer* t = new Timer();
    return *t;
}
}  // namespace

void timing_enable(bool on) {
    std::lock_guard<std::mutex> lk(T().mu);
    T().on = on;
}
bool timing_enabled() { return T().on; }

void timing_reset() {
    Device::get().sync();
    Timer& t = T();
    std::lock_guard<std::mutex> lk(t.mu);
    for (auto& kv : t.recs) {
        for (auto& pr : kv.second.done) {
            t.pool.push_back(pr.first);
            t.pool.push_back(pr.second);
        }
        if (kv.second.open) t.pool.push_back(kv.second.open);
    }
    t.recs.clear();
}

void timing_begin(const char* name) {
    Timer& t = T();
    std::lock_guard<std::mutex> lk(t.mu);
    hipEvent_t e = t.get();
    EQ_HIP(hipEventRecord(e, Device::get().stream()));
    t.recs[name].open = e;
}

void timing_end(const char* name) {
    Timer& t = T();
    std::lock_guard<std::mutex> lk(t.mu);
    auto& r = t.recs[name];
    if (!r.open) return;
    hipEvent_t e = t.get();
    EQ_HIP(hipEventRecord(e, Device::get().stream()));
    r.done.emplace_back(r.open, e);
    r.open = nullptr;
}

void timing_query(const std::string& name, int* count, double* total_ms) {
    Device::get().sync();
    Timer& t = T();
    std::lock_guard<std::mutex> lk(t.mu);
    *count = 0;
    *total_ms = 0.0;
    auto it = t.recs.find(name);
    if (it == t.recs.end()) return;
    for (auto& pr : it->second.done) {
        float ms = 0.f;
        EQ_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
        *total_ms += ms;
        ++*count;
    }
}

}  // namespace eqhip
