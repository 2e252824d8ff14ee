// Alignment constants and library-wide helpers (host side of libmarvl_align).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "mi_common.h"

namespace mi {

// Bias factors of New_Align_Spec (dalign/align.c:161-162); the same table of
// numbers is needed to derive the same integer parameters.
static const double kBias[10] = {.690, .690, .690, .690, .780, .850, .900, .933, .966, 1.000};

// dalign/align.c:214-279.  The trim tables are built iteratively: for a
// 15-column history x (first column in the most significant bit) score[x]
// is the running sum (+mscore per match, -dscore per mismatch) after all
// columns and table[x] = score[x] - (largest running sum before a column,
// counting the empty prefix).  The float/double expressions are written in
// the reference's order so the truncated integers agree.
void spec_build(HostSpec* s, double ave_corr, int tspace, const float* freq)
{
    s->ave_corr = ave_corr;
    s->tspace = tspace;
    for (int i = 0; i < 4; i++) s->freq[i] = freq[i];
    double match = (float)(freq[0] + freq[3]);
    if (match > .5) match = 1. - match;
    int bias = (int)((match + .025) * 20. - 1.);
    if (match < .2) bias = 3;
    s->ave_path = (int)(60 * (1. - kBias[bias] * (1. - ave_corr)));
    s->mscore = (int)(1000 * kBias[bias] * (1. - ave_corr));
    s->dscore = 1000 - s->mscore;
    for (int x = 0; x < 32768; x++) {
        int run = 0, top = 0;
        for (int i = 14; i >= 0; i--) {
            top = run > top ? run : top;
            run += ((x >> i) & 1) ? s->mscore : -s->dscore;
        }
        s->table[x] = (int16_t)(run - top);
        s->score[x] = (int16_t)run;
    }
}

static std::mutex g_lock;
static std::vector<DevSpecHost*> g_specs;

DevSpecHost* mi_spec_device(const MiAlignParams* prm)
{
    std::lock_guard<std::mutex> g(g_lock);
    for (DevSpecHost* d : g_specs)
        if (d->host.ave_corr == prm->ave_corr && d->host.tspace == prm->trace_space &&
            memcmp(d->host.freq, prm->freq, sizeof(float) * 4) == 0)
            return d;
    DevSpecHost* d = new DevSpecHost;
    spec_build(&d->host, prm->ave_corr, prm->trace_space, prm->freq);
    int16_t* dt = nullptr;
    if (hipMalloc(&dt, 2 * 32768 * sizeof(int16_t)) != hipSuccess) {
        delete d;
        return nullptr;
    }
    if (hipMemcpy(dt, d->host.table, 32768 * sizeof(int16_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt + 32768, d->host.score, 32768 * sizeof(int16_t), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(dt);
        delete d;
        return nullptr;
    }
    d->spec.table = dt;
    d->spec.score = dt + 32768;
    d->spec.tspace = d->host.tspace;
    d->spec.ave_path = d->host.ave_path;
    g_specs.push_back(d);
    return d;
}

int mi_num_cus()
{
    static int n = 0;
    if (n == 0) {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
            n = p.multiProcessorCount;
        if (n <= 0) n = 256;
    }
    return n;
}

static int* g_abort_host = nullptr;
static int* g_abort_dev = nullptr;

const int* mi_abort_flag_device()
{
    std::lock_guard<std::mutex> g(g_lock);
    if (g_abort_host == nullptr) {
        if (hipHostMalloc((void**)&g_abort_host, sizeof(int), hipHostMallocMapped) != hipSuccess) return nullptr;
        *g_abort_host = 0;
        if (hipHostGetDevicePointer((void**)&g_abort_dev, g_abort_host, 0) != hipSuccess) return nullptr;
    }
    return g_abort_dev;
}

int mi_wait_watchdog(hipStream_t st, const char* what)
{
    double limit = 120.0;
    if (const char* e = getenv("MI_WATCHDOG_S")) limit = atof(e);
    auto t0 = std::chrono::steady_clock::now();
    bool raised = false;
    for (;;) {
        hipError_t q = hipStreamQuery(st);
        if (q == hipSuccess) break;
        if (q != hipErrorNotReady) {
            if (raised) __atomic_store_n(g_abort_host, 0, __ATOMIC_SEQ_CST);
            return -1;
        }
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (!raised && el > limit) {
            fprintf(stderr, "%s: watchdog deadline %.0f s passed, stopping the kernel\n", what, limit);
            __atomic_store_n(g_abort_host, 1, __ATOMIC_SEQ_CST);
            raised = true;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    if (raised) {
        __atomic_store_n(g_abort_host, 0, __ATOMIC_SEQ_CST);
        return 1;
    }
    return 0;
}

// Launch shapes picked by the tuner (0: built-in default); see MiTuning.
static MiTuning g_tuning = {0, 0, 0, 0};
const MiTuning& mi_tuning() { return g_tuning; }

static thread_local std::string g_err;
void mi_set_error(const char* msg) { g_err = msg; }

}  // namespace mi

extern "C" {

const char* mi_last_error(void) { return mi::g_err.c_str(); }

int mi_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void* mi_device_alloc(long long bytes)
{
    void* p = nullptr;
    if (bytes <= 0 || hipMalloc(&p, (size_t)bytes) != hipSuccess) {
        mi::mi_set_error("mi_device_alloc: device allocation failed");
        return nullptr;
    }
    return p;
}

int mi_device_free(void* p)
{
    if (p && hipFree(p) != hipSuccess) return MI_ERR_DEVICE;
    return 0;
}

int mi_device_copy(void* dst, const void* src, long long bytes, int to_device)
{
    if (bytes <= 0) return 0;
    if (hipMemcpy(dst, src, (size_t)bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost) != hipSuccess) {
        mi::mi_set_error("mi_device_copy: copy failed");
        return MI_ERR_DEVICE;
    }
    return 0;
}

int mi_device_memset(void* p, int value, long long bytes)
{
    if (bytes <= 0) return 0;
    if (hipMemset(p, value, (size_t)bytes) != hipSuccess) {
        mi::mi_set_error("mi_device_memset: failed");
        return MI_ERR_DEVICE;
    }
    return 0;
}

int mi_set_tuning(const MiTuning* t)
{
    if (t == nullptr) {
        mi::g_tuning = MiTuning{0, 0, 0, 0, 0};
        return 0;
    }
    const bool ok = (t->report_waves_per_cu == 0 || (t->report_waves_per_cu >= 4 && t->report_waves_per_cu <= 32)) &&
                    (t->report_pool_log2 == 0 || (t->report_pool_log2 >= 12 && t->report_pool_log2 <= 18)) &&
                    (t->report_l1_waves_per_cu == 0 || (t->report_l1_waves_per_cu >= 2 && t->report_l1_waves_per_cu <= 32)) &&
                    (t->batch_waves_per_cu == 0 || (t->batch_waves_per_cu >= 4 && t->batch_waves_per_cu <= 32)) &&
                    (t->batch_wide_waves_per_cu == 0 ||
                     (t->batch_wide_waves_per_cu >= 2 && t->batch_wide_waves_per_cu <= 32));
    if (!ok) {
        mi::mi_set_error("mi_set_tuning: value out of range");
        return MI_ERR_ARG;
    }
    mi::g_tuning = *t;
    return 0;
}

int mi_get_tuning(MiTuning* t)
{
    if (t == nullptr) return MI_ERR_ARG;
    *t = mi::g_tuning;
    return 0;
}

int mi_device_info(char* name, int cap, int* cus)
{
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return MI_ERR_DEVICE;
    if (name && cap > 0) {
        strncpy(name, p.gcnArchName, cap - 1);
        name[cap - 1] = 0;
    }
    if (cus) *cus = p.multiProcessorCount;
    return 0;
}

int mi_abort_device_work(int raise)
{
    if (mi::mi_abort_flag_device() == nullptr) return MI_ERR_DEVICE;
    __atomic_store_n(mi::g_abort_host, raise ? 1 : 0, __ATOMIC_SEQ_CST);
    return 0;
}

int mi_align_params(const MiAlignParams* prm, int* ave_path, int* mscore, int* dscore)
{
    mi::HostSpec* s = new mi::HostSpec;
    mi::spec_build(s, prm->ave_corr, prm->trace_space, prm->freq);
    *ave_path = s->ave_path;
    *mscore = s->mscore;
    *dscore = s->dscore;
    delete s;
    return 0;
}

}  // extern "C"
