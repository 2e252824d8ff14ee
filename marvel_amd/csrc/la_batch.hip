// Batched Local_Alignment on gfx950 (kernels and level loop: la_batch_core.hpp).
// The library keeps one grown workspace per device for the device batch and
// device buffers for the host entry point, so repeated calls do not allocate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "la_batch_core.hpp"

namespace mi {

// ---- the library's grown workspace (per device) ----------------------------

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool grow(size_t need)
    {
        if (need <= bytes) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t b = std::max(need, (size_t)1 << 20);
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = b;
        return true;
    }
};

struct BatchWorkspace {
    DevBuf scratch;
    // inputs and outputs of the host entry point
    DevBuf packed, tasks, res, arena, top, check;
};

static std::mutex g_ws_lock;
static std::vector<BatchWorkspace*> g_ws;   // per device

static BatchWorkspace* workspace()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1, nullptr);
    if (g_ws[dev] == nullptr) g_ws[dev] = new BatchWorkspace;
    return g_ws[dev];
}

static size_t full_scratch_bytes(int ntasks, int maxlen, int tspace)
{
    const BatchPlan plan = make_plan(maxlen, tspace, mi_num_cus());
    size_t lv = 0;
    for (const Level& L : plan.lv)
        lv = std::max(lv, carve_bytes(std::min(L.waves, ntasks), L.pool_cap, plan.pts_cap, plan.latr_cap));
    return kCounterBytes + 2 * align_up((size_t)ntasks * sizeof(int)) + lv;
}

static int device_batch_locked(BatchWorkspace* W, const MiAlignParams* prm, const uint64_t* d_packed,
                               long long npacked, const MiLaTask* d_tasks, int ntasks, const int* d_order,
                               MiLaResult* d_results, uint16_t* d_traces, long long trace_cap,
                               unsigned long long* d_trace_used, long long* work)
{
    DevSpecHost* hs = mi_spec_device(prm);
    if (hs == nullptr) return MI_ERR_DEVICE;
    const int* flag = mi_abort_flag_device();
    if (flag == nullptr) return MI_ERR_DEVICE;
    if (!W->check.grow(64)) return MI_ERR_DEVICE;
    int maxlen = 0;
    int rc = check_tasks(d_tasks, ntasks, npacked, (int*)W->check.p, &maxlen);
    if (rc != 0) return rc;
    if (!W->scratch.grow(full_scratch_bytes(ntasks, maxlen, prm->trace_space))) {
        mi_set_error("batch: device memory for the workspace");
        return MI_ERR_DEVICE;
    }
    return run_batch(hs->spec, d_packed, d_tasks, ntasks, maxlen, d_order, d_results, d_traces, trace_cap,
                     d_trace_used, ScratchArea{(char*)W->scratch.p, W->scratch.bytes}, flag, true, work,
                     "mi_la_batch");
}

}  // namespace mi

using namespace mi;

// ---- C-ABI -----------------------------------------------------------------

extern "C" {

int mi_la_batch_device(const MiAlignParams* prm, const uint64_t* d_packed, long long npacked, const MiLaTask* d_tasks,
                       int ntasks, const int* d_order, MiLaResult* d_results, uint16_t* d_traces, long long trace_cap,
                       unsigned long long* d_trace_used, long long* work)
{
    if (ntasks <= 0) return 0;
    if (prm == nullptr || d_packed == nullptr || d_tasks == nullptr || d_results == nullptr ||
        d_trace_used == nullptr || (d_traces == nullptr && trace_cap > 0))
        return MI_ERR_ARG;
    std::lock_guard<std::mutex> g(g_ws_lock);
    BatchWorkspace* W = workspace();
    if (W == nullptr) return MI_ERR_DEVICE;
    return device_batch_locked(W, prm, d_packed, npacked, d_tasks, ntasks, d_order, d_results, d_traces, trace_cap,
                               d_trace_used, work);
}

// Batched Local_Alignment over host arrays (see include/marvl_align.h): the
// inputs go to device buffers kept between calls, the device batch runs, and
// the compact trace arena comes back and is laid out in task order.
int mi_la_batch_host(const MiAlignParams* prm, const uint64_t* packed, long long npacked, const MiLaTask* tasks,
                     int ntasks, MiLaResult* results, uint16_t* traces, long long ntrace)
{
    if (ntasks <= 0) return 0;
    long long bound = 0;   // traces a task can produce: maxtp values on each side of both origins
    for (int i = 0; i < ntasks; i++) {
        const MiLaTask& s = tasks[i];
        if (s.aword < 1 || s.bword < 1 || s.alen < 0 || s.blen < 0 || s.hgh < s.low ||
            s.aword + (s.alen + 31) / 32 + 2 > npacked || s.bword + (s.blen + 31) / 32 + 2 > npacked)
            return MI_ERR_ARG;
        bound += 2LL * (2 * (std::max(s.alen, s.blen) / prm->trace_space + 2));
    }
    std::lock_guard<std::mutex> g(g_ws_lock);
    BatchWorkspace* W = workspace();
    if (W == nullptr) return MI_ERR_DEVICE;
    if (!W->packed.grow(sizeof(uint64_t) * npacked) || !W->tasks.grow(sizeof(MiLaTask) * ntasks) ||
        !W->res.grow(sizeof(MiLaResult) * ntasks) || !W->arena.grow(sizeof(uint16_t) * (bound + 64)) ||
        !W->top.grow(64)) {
        mi_set_error("mi_la_batch_host: device memory");
        return MI_ERR_DEVICE;
    }
    uint64_t* dpk = (uint64_t*)W->packed.p;
    MiLaTask* dtk = (MiLaTask*)W->tasks.p;
    MiLaResult* dres = (MiLaResult*)W->res.p;
    uint16_t* dar = (uint16_t*)W->arena.p;
    unsigned long long* dtop = (unsigned long long*)W->top.p;
    if (hipMemcpy(dpk, packed, sizeof(uint64_t) * npacked, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dtk, tasks, sizeof(MiLaTask) * ntasks, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(dtop, 0, sizeof(unsigned long long)) != hipSuccess) {
        mi_set_error("mi_la_batch_host: HIP error");
        return MI_ERR_DEVICE;
    }
    const int rc = device_batch_locked(W, prm, dpk, npacked, dtk, ntasks, nullptr, dres, dar, bound + 64, dtop,
                                       nullptr);
    if (rc != 0) return rc;
    unsigned long long used = 0;
    std::vector<MiLaResult> hr(ntasks);
    if (hipMemcpy(&used, dtop, sizeof(used), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hr.data(), dres, sizeof(MiLaResult) * ntasks, hipMemcpyDeviceToHost) != hipSuccess) {
        mi_set_error("mi_la_batch_host: HIP error");
        return MI_ERR_DEVICE;
    }
    if ((long long)used > bound + 64) {
        mi_set_error("mi_la_batch_host: trace arena overflow");
        return MI_ERR_DEVICE;
    }
    std::vector<uint16_t> ha(used > 0 ? used : 1);
    if (used > 0 && hipMemcpy(ha.data(), dar, sizeof(uint16_t) * used, hipMemcpyDeviceToHost) != hipSuccess) {
        mi_set_error("mi_la_batch_host: HIP error");
        return MI_ERR_DEVICE;
    }
    long long o = 0;
    for (int i = 0; i < ntasks; i++) {
        MiLaResult r = hr[i];
        if (r.status == LA_UNRUN || r.status == LA_SPACE) {
            mi_set_error("mi_la_batch_host: a task was not completed");
            return MI_ERR_DEVICE;
        }
        const long long at = r.atrace, bt = r.btrace;
        if (r.status == LA_OK &&
            (at < 0 || bt < 0 || at + r.atlen > (long long)used || bt + r.btlen > (long long)used)) {
            mi_set_error("mi_la_batch_host: trace offsets outside the arena");
            return MI_ERR_DEVICE;
        }
        if (o + r.atlen + r.btlen > ntrace) return MI_ERR_ARG;
        if (r.atlen) memcpy(traces + o, ha.data() + at, sizeof(uint16_t) * r.atlen);
        if (r.btlen) memcpy(traces + o + r.atlen, ha.data() + bt, sizeof(uint16_t) * r.btlen);
        r.atrace = o;
        r.btrace = o + r.atlen;
        o += r.atlen + r.btlen;
        results[i] = r;
    }
    return 0;
}

}  // extern "C"
