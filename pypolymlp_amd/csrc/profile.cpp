// Kernel timing: event pairs are kept until queried, then resolved with
// hipEventElapsedTime and accumulated per name.
#include "profile.hpp"

#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/polymlp_amd.h"
#include "capi_internal.hpp"

namespace plm {

namespace {
struct Profiler {
    bool on = false;
    std::mutex mu;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::map<std::string, std::pair<double, long>> totals;
    void resolve() {
        for (auto& p : pending) {
            (void)hipEventSynchronize(p.second.second);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, p.second.first, p.second.second);
            auto& t = totals[p.first];
            t.first += ms;
            t.second += 1;
            (void)hipEventDestroy(p.second.first);
            (void)hipEventDestroy(p.second.second);
        }
        pending.clear();
    }
};
Profiler& prof() {
    static Profiler p;
    return p;
}
}  // namespace

bool profiling_on() { return prof().on; }

ProfPause::ProfPause() {
    std::lock_guard<std::mutex> g(prof().mu);
    was_ = prof().on;
    prof().on = false;
}

ProfPause::~ProfPause() {
    std::lock_guard<std::mutex> g(prof().mu);
    prof().on = was_;
}

ProfScope::ProfScope(const char* name, hipStream_t st) : name_(name), st_(st) {
    if (!prof().on) return;
    (void)hipEventCreate(&a_);
    (void)hipEventCreate(&b_);
    (void)hipEventRecord(a_, st_);
}

ProfScope::~ProfScope() {
    if (!a_) return;
    (void)hipEventRecord(b_, st_);
    std::lock_guard<std::mutex> g(prof().mu);
    prof().pending.push_back({name_, {a_, b_}});
}

}  // namespace plm

using namespace plm;

extern "C" {

int plm_profile_enable(int on) {
    std::lock_guard<std::mutex> g(prof().mu);
    prof().on = on != 0;
    return PLM_OK;
}

int plm_profile_reset(void) {
    std::lock_guard<std::mutex> g(prof().mu);
    prof().resolve();
    prof().totals.clear();
    return PLM_OK;
}

int plm_profile_get(const char* name, double* total_ms, long* launches) {
    std::lock_guard<std::mutex> g(prof().mu);
    prof().resolve();
    auto it = prof().totals.find(name);
    *total_ms = it == prof().totals.end() ? 0.0 : it->second.first;
    *launches = it == prof().totals.end() ? 0 : it->second.second;
    return PLM_OK;
}

}  // extern "C"
