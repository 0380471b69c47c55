// Per-kernel event timing for measurements (fc3_profile_enable / fc3_profile_query), and
// roctx ranges around the C-ABI calls and each kernel kind's launches.
#include <rocprofiler-sdk-roctx/roctx.h>

#include "fc3_internal.hpp"

namespace fc3 {

namespace {
thread_local fc3_context* tl_current = nullptr;

hipEvent_t take_event(Profiler& p)
{
    if (!p.pool.empty()) {
        hipEvent_t e = p.pool.back();
        p.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

void set_current(fc3_context* h) { tl_current = h; }

const char* kind_range_name(int kind)
{
    switch (kind) {
    case FC3_KERNEL_IFC3_WRITE: return "fc3 IFC3 writer";
    case FC3_KERNEL_GEMM_MODE3: return "fc3 mcc3 mode 3";
    case FC3_KERNEL_GEMM_MODE12: return "fc3 mcc3 mode 1";
    case FC3_KERNEL_PAIR_BLOCKS: return "fc3 pair blocks";
    case FC3_KERNEL_NEIGHBORS: return "fc3 neighbour lists";
    case FC3_KERNEL_GEMM_MODE2: return "fc3 mcc3 mode 2";
    default: return "fc3 kernel";
    }
}

TraceRange::TraceRange(const char* name) { roctxRangePushA(name); }
TraceRange::~TraceRange() { roctxRangePop(); }

ProfScope::ProfScope(int kind, hipStream_t stream)
{
    roctxRangePushA(kind_range_name(kind));
    if (!tl_current || !tl_current->prof.enabled) return;
    h = tl_current;
    s = stream;
    ProfRecord r{kind, take_event(h->prof), take_event(h->prof)};
    if (!r.start || !r.stop) { h = nullptr; return; }
    hipEventRecord(r.start, s);
    h->prof.records.push_back(r);
    rec = (int)h->prof.records.size() - 1;
}

ProfScope::~ProfScope()
{
    if (h && rec >= 0) hipEventRecord(h->prof.records[rec].stop, s);
    roctxRangePop();
}

}  // namespace fc3

using namespace fc3;

extern "C" {

int fc3_profile_enable(fc3_handle h, int enable)
{
    if (!h) return set_error(FC3_ERR_ARGUMENT, "null fc3 handle");
    FC3_HIP(hipSetDevice(h->device));
    FC3_HIP(hipStreamSynchronize(h->stream));
    for (auto& r : h->prof.records) {
        h->prof.pool.push_back(r.start);
        h->prof.pool.push_back(r.stop);
    }
    h->prof.records.clear();
    h->prof.enabled = enable != 0;
    return FC3_OK;
}

int fc3_profile_query(fc3_handle h, int kind, double* total_ms, int* launches)
{
    if (!h) return set_error(FC3_ERR_ARGUMENT, "null fc3 handle");
    if (kind < 0 || kind >= FC3_KERNEL_COUNT) return set_error(FC3_ERR_ARGUMENT, "unknown kind");
    FC3_HIP(hipSetDevice(h->device));
    FC3_HIP(hipStreamSynchronize(h->stream));
    double t = 0.0;
    int c = 0;
    for (auto& r : h->prof.records) {
        if (r.kind != kind) continue;
        float ms = 0.0f;
        FC3_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
        t += ms;
        ++c;
    }
    if (total_ms) *total_ms = t;
    if (launches) *launches = c;
    return FC3_OK;
}

}  // extern "C"
