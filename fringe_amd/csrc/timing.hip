// timing.hip -- implementation of the per-stage event timing and its C ABI.
#include <mutex>
#include <vector>

#include "timing.hpp"
#include "fringe_amd.h"

namespace fringe {

namespace {
struct Pending {
    int stage;
    hipEvent_t a, b;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_pool;
double g_ms[STAGE_COUNT] = {};
long long g_calls[STAGE_COUNT] = {};
thread_local hipEvent_t t_open[STAGE_COUNT] = {};

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
const char* kNames[STAGE_COUNT] = {"pack", "gram", "solve", "nmap_prep", "nmap_ks", "resolve"};
}  // namespace

void timing_begin(int stage, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_on) return;
    hipEvent_t e = take_event();
    if (e && hipEventRecord(e, s) == hipSuccess) t_open[stage] = e;
}

void timing_end(int stage, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t a = t_open[stage];
    t_open[stage] = nullptr;
    if (!g_on || !a) return;
    hipEvent_t b = take_event();
    if (b && hipEventRecord(b, s) == hipSuccess) g_pending.push_back({stage, a, b});
}

}  // namespace fringe

using namespace fringe;

extern "C" {

void fringe_timing_enable(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = on != 0;
}

void fringe_timing_reset(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (Pending& p : g_pending) {
        hipEventSynchronize(p.b);
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
    for (int i = 0; i < STAGE_COUNT; ++i) {
        g_ms[i] = 0.0;
        g_calls[i] = 0;
    }
}

int fringe_timing_read(double* ms, long long* calls, int max_stages) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (Pending& p : g_pending) {
        float t = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
            g_ms[p.stage] += t;
            g_calls[p.stage] += 1;
        }
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
    const int k = max_stages < STAGE_COUNT ? max_stages : STAGE_COUNT;
    for (int i = 0; i < k; ++i) {
        if (ms) ms[i] = g_ms[i];
        if (calls) calls[i] = g_calls[i];
    }
    return STAGE_COUNT;
}

const char* fringe_timing_stage_name(int stage) {
    return (stage >= 0 && stage < STAGE_COUNT) ? kNames[stage] : "";
}

}  // extern "C"
