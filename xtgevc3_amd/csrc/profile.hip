// Live per-kernel timing with HIP events on the launch stream (off by default).
// bench.py turns it on around its timed region to report each kernel class's average
// launch duration and algorithmic work next to the whole-call throughput.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/xtgevc3.h"
#include "driver.h"

namespace xtg {

bool debug_sync()
{
    static int v = -1;
    if (v < 0) {
        const char *e = std::getenv("XTGEVC3_DEBUG_SYNC");
        v = (e && e[0] == '1') ? 1 : 0;
    }
    return v == 1;
}

namespace {
struct Rec {
    int kind;
    hipEvent_t a, b;
    double flops, bytes;
};
std::mutex g_pm;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    return e;
}
} // namespace

int prof_begin(hipStream_t st)
{
    std::lock_guard<std::mutex> l(g_pm);
    if (!g_on) {
        return -1;
    }
    Rec r{};
    r.a = take();
    r.b = take();
    if (!r.a || !r.b || hipEventRecord(r.a, st) != hipSuccess) {
        return -1;
    }
    g_recs.push_back(r);
    return (int)g_recs.size() - 1;
}

void prof_end(int idx, hipStream_t st, int kind, double flops, double bytes)
{
    if (idx < 0) {
        return;
    }
    std::lock_guard<std::mutex> l(g_pm);
    Rec &r = g_recs[idx];
    r.kind = kind;
    r.flops = flops;
    r.bytes = bytes;
    (void)hipEventRecord(r.b, st);
}

} // namespace xtg

extern "C" {

void xtgevc3_profile_enable(int on)
{
    std::lock_guard<std::mutex> l(xtg::g_pm);
    xtg::g_on = on != 0;
    // create the events up front: hipEventCreate inside the timed launches would stall them
    while (xtg::g_on && xtg::g_pool.size() < 16384) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) {
            break;
        }
        xtg::g_pool.push_back(e);
    }
}

int xtgevc3_profile_read(double *ms, long *launches, double *flops, double *bytes)
{
    std::lock_guard<std::mutex> l(xtg::g_pm);
    for (int k = 0; k < XTGEVC3_PROF_KINDS; ++k) {
        ms[k] = 0.0;
        launches[k] = 0;
        flops[k] = 0.0;
        bytes[k] = 0.0;
    }
    int rc = 0;
    for (auto &r : xtg::g_recs) {
        float t = 0.0f;
        if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) {
            rc = -1;
        }
        if (r.kind >= 0 && r.kind < XTGEVC3_PROF_KINDS) {
            ms[r.kind] += t;
            launches[r.kind] += 1;
            flops[r.kind] += r.flops;
            bytes[r.kind] += r.bytes;
        }
        xtg::g_pool.push_back(r.a);
        xtg::g_pool.push_back(r.b);
    }
    xtg::g_recs.clear();
    return rc;
}

} // extern "C"
