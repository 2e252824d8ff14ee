// Batched Local_Alignment without device allocation (kernels and level loop:
// la_batch_core.hpp).  The per-call scratch -- alignment tables, counters,
// retry lists and per-wave pebble pools -- lives in one file-scope device
// array of fixed size; the call allocates nothing.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "la_batch_core.hpp"

namespace mi {

// Layout: the two alignment tables (2 x 32768 int16), 256 bytes of flag and
// task-check words, then the batch area (counters, two retry lists of at most
// kNoAllocMaxTasks ints, per-level wave scratch).

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kNoAllocMaxTasks = 1 << 20;
constexpr size_t kTableBytes = 2 * 32768 * sizeof(int16_t);

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

struct NoAllocDev {
    char* ptr = nullptr;   // device address of `scratch` on this device
    bool spec_valid = false;
    MiAlignParams spec_prm;
    int ave_path = 0;
};
static std::mutex g_na_lock;
static NoAllocDev g_na[64];
static HostSpec g_na_host;   // host staging of the tables (guarded by g_na_lock)

}  // namespace mi

using namespace mi;

extern "C" {

int mi_la_batch_device_noalloc(const MiAlignParams* prm, const uint64_t* d_packed, long long npacked,
                               const MiLaTask* d_tasks, int ntasks, MiLaResult* d_results, uint16_t* d_traces,
                               long long trace_cap, unsigned long long* d_trace_used)
{
    if (ntasks <= 0) return 0;
    if (ntasks > kNoAllocMaxTasks) {
        mi_set_error("mi_la_batch_device_noalloc: more than 1048576 tasks in one call");
        return MI_ERR_ARG;
    }
    if (prm == nullptr || d_packed == nullptr || d_tasks == nullptr || d_results == nullptr ||
        d_trace_used == nullptr || (d_traces == nullptr && trace_cap > 0))
        return MI_ERR_ARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return MI_ERR_DEVICE;
    std::lock_guard<std::mutex> g(g_na_lock);
    NoAllocDev& D = g_na[dev];
    if (D.ptr == nullptr) {
        // the array's device address on this device; the symbol itself is only a host-side name
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) {
            mi_set_error("mi_la_batch_device_noalloc: no device address for the scratch array");
            return MI_ERR_DEVICE;
        }
        D.ptr = (char*)p;
    }
    // tables of this parameter set at the front (reloaded when the parameters change)
    if (!D.spec_valid || memcmp(&D.spec_prm, prm, sizeof(MiAlignParams)) != 0) {
        D.spec_valid = false;
        spec_build(&g_na_host, prm->ave_corr, prm->trace_space, prm->freq);
        if (hipMemcpy(D.ptr, g_na_host.table, 32768 * sizeof(int16_t), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(D.ptr + 32768 * sizeof(int16_t), g_na_host.score, 32768 * sizeof(int16_t),
                      hipMemcpyHostToDevice) != hipSuccess)
            return MI_ERR_DEVICE;
        D.spec_prm = *prm;
        D.ave_path = g_na_host.ave_path;
        D.spec_valid = true;
    }
    DevSpec sp;
    sp.table = (const int16_t*)D.ptr;
    sp.score = sp.table + 32768;
    sp.tspace = prm->trace_space;
    sp.ave_path = D.ave_path;
    // flag word (never raised: the kernels end by their own bounds) and check words
    int* words = (int*)(D.ptr + kTableBytes);
    if (hipMemset(words, 0, 256) != hipSuccess) return MI_ERR_DEVICE;
    int maxlen = 0;
    int rc = check_tasks(d_tasks, ntasks, npacked, words + 16, &maxlen);
    if (rc != 0) return rc;
    const ScratchArea area{D.ptr + kTableBytes + 256, kNoAllocScratchBytes - kTableBytes - 256};
    return run_batch(sp, d_packed, d_tasks, ntasks, maxlen, nullptr, d_results, d_traces, trace_cap, d_trace_used,
                     area, words, false, nullptr, "mi_la_batch_device_noalloc");
}

}  // extern "C"
