// MARVEL daligner's k-mer match filter (dalign/filter.c) on gfx950, behind
// the reference's Set_Filter_Params / Sort_Kmers / Match_Filter interface.
//
// Sort_Kmers   : reads -> 2-bit packed words in HBM -> one 64-bit key per
//                k-mer -> stable LSD radix sort on the code bits (keys are
//                generated in (read, rpos) order, so sorting the code alone
//                yields the reference's (code, read, rpos) order)
// Match_Filter : merge-join of the A and B indexes by code (binary-searched
//                group ranges, per-entry counts, prefix sums), the MEM_LIMIT
//                cap from the count histogram, seed emission in merge order,
//                stable radix sort of the seeds by (bread, aread, apos),
//                split into read pairs, the report_thread chunk rule, then
//                one wave per read pair (report.hpp): band scores, gating,
//                Local_Alignment, redundancy fusion, records.
// Records come back to the host and are appended to the Align_Spec's
// overlap buffer in the order report_thread would have produced for equal
// sort keys.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/marvl_align.h"
#include "../../include/marvl_filter.h"
#include "filter_kernels.hpp"
#include "mi_common.h"
#include "radix.hpp"
#include "report.hpp"
#include "scorer.hpp"

extern "C" {
int BIASED = 0;
int VERBOSE = 0;
int MINOVER = 2000;
int HGAP_MIN = 0;
int SYMMETRIC = 1;
int IDENTITY = 0;
uint64_t MEM_LIMIT = 0;
uint64_t MEM_PHYSICAL = 0;
int NTHREADS = 4;
int NSHIFT = 2;
}

namespace mi {

typedef unsigned long long u64;

struct Fail : std::runtime_error {
    int code;
    Fail(const std::string& m, int c = MI_ERR_DEVICE) : std::runtime_error(m), code(c) {}
};

#define HC(x)                                                                                   \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) throw Fail(std::string(#x) + ": " + hipGetErrorString(e_));       \
    } while (0)

// Device memory cache: released blocks go to a free list and are handed out
// again for requests of up to their size (and at least half of it), so the
// multi-GB buffers of a pass are mapped once per process, not once per call.
// Sizes are rounded up to 2 MiB.
struct DevCache {
    std::mutex m;
    std::multimap<size_t, void*> free_blocks;
    void* get(size_t bytes, size_t* got)
    {
        const size_t r = (bytes + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
        {
            std::lock_guard<std::mutex> g(m);
            auto it = free_blocks.lower_bound(r);
            if (it != free_blocks.end() && it->first <= 2 * r) {
                void* p = it->second;
                *got = it->first;
                free_blocks.erase(it);
                return p;
            }
        }
        void* p = nullptr;
        if (hipMalloc(&p, r) != hipSuccess) {
            trim();   // give cached blocks back and try once more
            HC(hipMalloc(&p, r));
        }
        *got = r;
        return p;
    }
    void put(void* p, size_t bytes)
    {
        std::lock_guard<std::mutex> g(m);
        free_blocks.emplace(bytes, p);
    }
    void trim()
    {
        std::lock_guard<std::mutex> g(m);
        for (auto& kv : free_blocks) (void)hipFree(kv.second);
        free_blocks.clear();
    }
};
static DevCache g_cache;

// Grow-only pinned host staging buffer (records and traces coming back from
// the device); kept for the life of the process.
struct PinnedBuf {
    void* p = nullptr;
    size_t n = 0;
    void* get(size_t need)
    {
        if (need > n) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            n = 0;
            const size_t want = need + need / 4;
            if (hipHostMalloc(&p, want) != hipSuccess) throw Fail("pinned host allocation failed");
            n = want;
        }
        return p;
    }
};
static PinnedBuf g_pin_rec, g_pin_tr;

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;        // elements usable
    size_t bytes = 0;    // size of the cached block
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) g_cache.put(p, bytes);
        p = nullptr;
        n = 0;
        bytes = 0;
    }
    T* ensure(size_t m)
    {
        if (m == 0) m = 1;
        if (m <= n) return p;
        release();
        p = (T*)g_cache.get(m * sizeof(T), &bytes);
        n = bytes / sizeof(T);
        return p;
    }
    void swap(DBuf& o)
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(bytes, o.bytes);
    }
};

static hipStream_t fstream()
{
    static hipStream_t s = nullptr;
    if (s == nullptr) HC(hipStreamCreate(&s));
    return s;
}

static unsigned nblk(long long n, int b = 256) { return (unsigned)std::max<long long>(1, (n + b - 1) / b); }

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int bits_for(long long v)   // bits to hold 0..v
{
    int b = 1;
    while (b < 62 && (1LL << b) <= v) b++;
    return b;
}

template <class T>
static T d2h1(const T* dp)
{
    T v;
    HC(hipMemcpyAsync(&v, dp, sizeof(T), hipMemcpyDeviceToHost, fstream()));
    HC(hipStreamSynchronize(fstream()));
    return v;
}

// ---- hipGraph replay of fixed launch sequences ----------------------------------------
// A radix sort or a scan is a fixed sequence of launches on the filter stream
// with no host synchronisation and no memset, determined by its arguments and
// the addresses of its buffers.  The first call with a given key records the
// sequence by stream capture and instantiates it; a later call with the same
// key (the same block pair compared again, or the next B block of the same
// size through the caching allocator's reused buffers) replays the
// instantiated graph with one launch.  MI_GRAPHS=0 turns replay off.

struct GraphKey {
    uint64_t w[12];
    bool operator<(const GraphKey& o) const { return memcmp(w, o.w, sizeof w) < 0; }
};

struct GraphCache {
    std::map<GraphKey, hipGraphExec_t> execs;
    std::vector<GraphKey> order;   // insertion order, for eviction
    long long captured = 0, replayed = 0, launches_replayed = 0;
    static constexpr size_t kMax = 96;
    bool enabled()
    {
        static int on = -1;
        if (on < 0) {
            const char* e = getenv("MI_GRAPHS");
            on = (e && atoi(e) == 0) ? 0 : 1;
        }
        return on == 1;
    }
    // runs `seq` (which enqueues `nlaunch` kernels on st) through the cache
    template <class F>
    void run(const GraphKey& k, hipStream_t st, int nlaunch, F seq)
    {
        if (!enabled()) {
            seq();
            return;
        }
        auto it = execs.find(k);
        if (it != execs.end()) {
            HC(hipGraphLaunch(it->second, st));
            replayed += 1;
            launches_replayed += nlaunch;
            return;
        }
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        HC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            seq();
        } catch (...) {
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HC(hipStreamEndCapture(st, &g));
        HC(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        HC(hipGraphLaunch(ex, st));
        captured += 1;
        if (execs.size() >= kMax) {   // drop the oldest
            auto old = execs.find(order.front());
            if (old != execs.end()) {
                HC(hipStreamSynchronize(st));
                (void)hipGraphExecDestroy(old->second);
                execs.erase(old);
            }
            order.erase(order.begin());
        }
        execs.emplace(k, ex);
        order.push_back(k);
    }
};
static GraphCache g_graphs;

// ---- exclusive scan: out[0..n] with out[n] = total ---------------------------------

__global__ void k_scan_total(const u64* in, u64* out, long long n)
{
    out[n] = (n > 0) ? out[n - 1] + in[n - 1] : 0;
}

struct Scanner {
    std::vector<std::unique_ptr<DBuf<u64>>> lv;
    DBuf<u64>& tmp(int i)
    {
        while ((int)lv.size() <= i) lv.emplace_back(new DBuf<u64>);
        return *lv[i];
    }
    // the scan through the graph cache
    void run(const u64* in, u64* out, long long n)
    {
        // the temporary levels this n needs, allocated before any capture
        int nl = 0;
        for (long long m = n; m > 0;) {
            const long long nb = (m + 1023) / 1024;
            tmp(2 * nl).ensure(nb + 1);
            if (nb > 1) tmp(2 * nl + 1).ensure(nb + 1);
            m = nb > 1 ? nb : 0;
            nl++;
        }
        GraphKey k{};
        k.w[0] = 0x5343414eull;   // scan
        k.w[1] = (uint64_t)in;
        k.w[2] = (uint64_t)out;
        k.w[3] = (uint64_t)n;
        for (int i = 0; i < nl && i < 4; i++) {
            k.w[4 + 2 * i] = (uint64_t)tmp(2 * i).p;
            k.w[5 + 2 * i] = (uint64_t)tmp(2 * i + 1).p;
        }
        if (nl > 4) {   // (2^60 elements): no cache
            run_level(in, out, n, 0);
            return;
        }
        g_graphs.run(k, fstream(), 3 * nl + 1, [&] { run_level(in, out, n, 0); });
    }
    void run_level(const u64* in, u64* out, long long n, int level)
    {
        hipStream_t st = fstream();
        const long long nb = (n + 1023) / 1024;
        if (n > 0) {
            u64* sums = tmp(2 * level).ensure(nb + 1);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(scan_block<u64>), dim3(nb), dim3(256), 0, st, in, out, n, sums);
            HC(hipGetLastError());
            if (nb > 1) {
                u64* ss = tmp(2 * level + 1).ensure(nb + 1);
                run_level(sums, ss, nb, level + 1);
                hipLaunchKernelGGL(HIP_KERNEL_NAME(scan_add<u64>), dim3(nb), dim3(256), 0, st, out, n, ss);
                HC(hipGetLastError());
            }
        }
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_scan_total), dim3(1), dim3(1), 0, st, in, out, n);
        HC(hipGetLastError());
    }
};
static Scanner g_scan;

// ---- stable LSD radix sort over bits [lo, hi) ---------------------------------------

static DBuf<unsigned> g_rcounts;
static DBuf<u64> g_roffs, g_rtot;

template <class Rec, class KeyOf>
static Rec* radix_sort(Rec* a, Rec* b, long long n, int lo, int hi, KeyOf key)
{
    if (n <= 1) return a;
    hipStream_t st = fstream();
    const int ntiles = (int)((n + kTile - 1) / kTile);
    unsigned* counts = g_rcounts.ensure((size_t)256 * ntiles);
    u64* offs = g_roffs.ensure((size_t)256 * ntiles);
    u64* tot = g_rtot.ensure(256);
    static const int kind = 0;   // one address per Rec / KeyOf instantiation
    static_assert(sizeof(KeyOf) <= 16, "sort key functor fits the graph key");
    GraphKey k{};
    k.w[0] = (uint64_t)&kind;
    k.w[1] = (uint64_t)a;
    k.w[2] = (uint64_t)b;
    k.w[3] = (uint64_t)n;
    k.w[4] = (uint64_t)lo | ((uint64_t)hi << 32);
    memcpy(&k.w[5], &key, sizeof(KeyOf));
    k.w[7] = (uint64_t)counts;
    k.w[8] = (uint64_t)offs;
    k.w[9] = (uint64_t)tot;
    int passes = 0;
    for (int sh = lo; sh < hi; sh += 8) passes++;
    Rec* pa = a;
    Rec* pb = b;
    g_graphs.run(k, st, 4 * passes, [&] {
        for (int sh = lo; sh < hi; sh += 8) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(rs_hist<Rec, KeyOf>), dim3(ntiles), dim3(256), 0, st, pa, n, sh, key,
                               counts, ntiles);
            HC(hipGetLastError());
            hipLaunchKernelGGL(HIP_KERNEL_NAME(rs_digit_totals), dim3(256), dim3(256), 0, st, counts, ntiles, tot);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(rs_digit_offsets), dim3(256), dim3(256), 0, st, counts, ntiles, tot,
                               offs);
            HC(hipGetLastError());
            hipLaunchKernelGGL(HIP_KERNEL_NAME(rs_scatter<Rec, KeyOf>), dim3(ntiles), dim3(256), 0, st, pa, pb, n, sh,
                               key, offs, ntiles);
            HC(hipGetLastError());
            std::swap(pa, pb);
        }
    });
    return (passes & 1) ? b : a;
}

// ---- k-mer index -----------------------------------------------------------------------

constexpr uint32_t kIdxMagic = 0x4d49584bu;   // "KXIM"

struct KIdx {
    uint32_t magic = kIdxMagic;
    int nreads = 0, maxlen = 0, ufirst = 0, part = 0, K = 0;
    float freq[4] = {.25f, .25f, .25f, .25f};   // the block's base frequencies (-b weights)
    KeyFmt fmt{1, 1};
    long long nk = 0;
    long long nwords = 0;
    std::vector<int> rlen;
    std::vector<long long> w0;
    // soft mask (-m): read i's points at mpts[manno[i] .. manno[i+1]); empty when unmasked
    std::vector<long long> manno;
    std::vector<int> mpts;
    DBuf<uint64_t> keys, pk;
    DBuf<long long> dw0;
    DBuf<int> dlen;
    // group structure (built when the index is the A side)
    long long ng = -1;
    DBuf<long long> gidx, gstart;
};

static int g_kmer = 14, g_binshift = 6, g_suppress = 0, g_hitmin = 35;

static void layout_reads(KIdx* x, const HITS_DB* blk)
{
    x->nreads = blk->nreads;
    x->maxlen = blk->maxlen;
    x->ufirst = blk->ufirst;
    x->part = blk->part;
    for (int i = 0; i < 4; i++) x->freq[i] = blk->freq[i];
    x->rlen.resize(x->nreads);
    x->w0.resize(x->nreads);
    long long wpos = 1;
    int mx = 1;
    for (int i = 0; i < x->nreads; i++) {
        const int L = blk->reads[i].rlen;
        if (L < 0) throw Fail("negative read length", MI_ERR_ARG);
        x->rlen[i] = L;
        x->w0[i] = wpos;
        wpos += (L + 31) / 32 + 2;
        mx = std::max(mx, L);
    }
    x->nwords = wpos + 1;
    // band arrays are sized from maxlen: never smaller than the longest read
    x->maxlen = std::max(x->maxlen, mx);
    x->fmt.pbits = bits_for(mx);
    x->fmt.rbits = bits_for(std::max(1, x->nreads - 1));
}

// The block's mask track (the first of blk->tracks, as read_DB leaves it:
// offsets in points) copied into the index, checked to be increasing pairs
// within the reads (other tracks make tuple_thread read outside the read,
// dalign/filter.c:424-452).
static void take_mask(KIdx* x, const HITS_DB* blk)
{
    x->manno.clear();
    x->mpts.clear();
    const HITS_TRACK* t = (const HITS_TRACK*)blk->tracks;
    if (t == nullptr) return;
    if (t->anno == nullptr || (t->data == nullptr && ((const int64_t*)t->anno)[x->nreads] > ((const int64_t*)t->anno)[0]))
        throw Fail("mask track without offsets or points", MI_ERR_ARG);
    const int64_t* an = (const int64_t*)t->anno;
    const int* pt = (const int*)t->data;
    x->manno.resize(x->nreads + 1);
    for (int i = 0; i <= x->nreads; i++) x->manno[i] = an[i] - an[0];
    x->mpts.assign(pt ? pt + an[0] : nullptr, pt ? pt + an[x->nreads] : nullptr);
    for (int i = 0; i < x->nreads; i++) {
        const long long b = x->manno[i], f = x->manno[i + 1];
        if (f < b || ((f - b) & 1)) throw Fail("mask track: odd number of points for a read", MI_ERR_ARG);
        int prev = 0;
        for (long long k = b; k < f; k++) {
            if (x->mpts[k] < prev || x->mpts[k] > x->rlen[i])
                throw Fail("mask track: points not increasing within the read", MI_ERR_ARG);
            prev = x->mpts[k];
        }
    }
}

static void suppress_frequent(KIdx* x, DBuf<uint64_t>& k2);

// -b weights: ceil(-10000 log4 freq) per base, fixed by the first biased
// index of the process (dalign/filter.c:718-732 computes them once, from the
// block Sort_Kmers first sees, and keeps them for every later block)
static bool g_bias_set = false;
static int g_logbase[4];

// the -b k-mer list of x into x->keys (unsorted, in read / position order); returns the count
static long long biased_keys(KIdx* x, int K)
{
    hipStream_t st = fstream();
    if (!g_bias_set) {
        const double scale = -10000. / std::log(4.);
        for (int i = 0; i < 4; i++) {
            if (!(x->freq[i] > 0.f && x->freq[i] < 1.f))
                throw Fail("-b needs every base frequency of the block in (0, 1)", MI_ERR_ARG);
            g_logbase[i] = (int)std::ceil(scale * std::log((double)x->freq[i]));
        }
        g_bias_set = true;
    }
    // runs: whole reads, or with a mask the gaps that hold a K-mer
    const bool masked = !x->manno.empty();
    std::vector<int> rr, rp, rq;
    for (int i = 0; i < x->nreads; i++) {
        if (!masked) {
            rr.push_back(i);
            rp.push_back(0);
            rq.push_back(x->rlen[i]);
            continue;
        }
        const long long b = x->manno[i], f = x->manno[i + 1];
        for (long long a = b; a <= f; a += 2) {
            const int p = (a == b) ? 0 : x->mpts[a - 1];
            const int q = (a == f) ? x->rlen[i] : x->mpts[a];
            if (p + K > q) continue;
            rr.push_back(i);
            rp.push_back(p);
            rq.push_back(q);
        }
    }
    const int nrun = (int)rr.size();
    if (nrun == 0) return 0;
    BiasParams bp;
    for (int i = 0; i < 4; i++) bp.lb[i] = g_logbase[i];
    bp.lognorm = 10000 * K;
    bp.logthresh = 10000 * (K - 2);   // MAX_BIAS 2 (dalign/filter.c:18)
    bp.masked = masked ? 1 : 0;
    DBuf<int> drr, drp, drq, dcnt;
    DBuf<long long> dout;
    drr.ensure(nrun);
    drp.ensure(nrun);
    drq.ensure(nrun);
    dcnt.ensure(nrun);
    dout.ensure(nrun);
    HC(hipMemcpyAsync(drr.p, rr.data(), sizeof(int) * nrun, hipMemcpyHostToDevice, st));
    HC(hipMemcpyAsync(drp.p, rp.data(), sizeof(int) * nrun, hipMemcpyHostToDevice, st));
    HC(hipMemcpyAsync(drq.p, rq.data(), sizeof(int) * nrun, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_biased<false>), dim3(nblk(nrun, 64)), dim3(64), 0, st, x->pk.p, x->dw0.p, drr.p,
                       drp.p, drq.p, nrun, K, bp, x->fmt, nullptr, dcnt.p, nullptr);
    HC(hipGetLastError());
    std::vector<int> cnt(nrun);
    HC(hipMemcpyAsync(cnt.data(), dcnt.p, sizeof(int) * nrun, hipMemcpyDeviceToHost, st));
    HC(hipStreamSynchronize(st));
    std::vector<long long> ofs(nrun);
    long long nk = 0;
    for (int s = 0; s < nrun; s++) {
        ofs[s] = nk;
        nk += cnt[s];
    }
    x->keys.ensure(nk + 1);
    if (nk > 0) {
        HC(hipMemcpyAsync(dout.p, ofs.data(), sizeof(long long) * nrun, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_biased<true>), dim3(nblk(nrun, 64)), dim3(64), 0, st, x->pk.p, x->dw0.p,
                           drr.p, drp.p, drq.p, nrun, K, bp, x->fmt, dout.p, nullptr, x->keys.p);
        HC(hipGetLastError());
    }
    return nk;
}

static void build_keys(KIdx* x)
{
    hipStream_t st = fstream();
    const int K = g_kmer;
    x->K = K;
    if (2 * K + x->fmt.rbits + x->fmt.pbits > 64)
        throw Fail("block too large for 64-bit k-mer keys (2k + read bits + position bits > 64)", MI_ERR_ARG);
    if (BIASED) {
        const long long nk = biased_keys(x, K);
        DBuf<uint64_t> k2;
        k2.ensure(nk + 1);
        const int cs = x->fmt.cshift();
        uint64_t* sorted = radix_sort(x->keys.p, k2.p, nk, cs, cs + 2 * K, KeyU64{});
        if (sorted != x->keys.p) x->keys.swap(k2);
        x->nk = nk;
        suppress_frequent(x, k2);
        HC(hipStreamSynchronize(st));
        return;
    }
    // runs of consecutive k-mers: one per read, or with a mask one per gap
    // between masked intervals that holds a k-mer, [0, pt0), [pt1, pt2), ...,
    // [ptlast, rlen) (dalign/filter.c:416-452); kofs[s] = k-mers before run s
    std::vector<long long> kofs;
    std::vector<int> sread, spos0;
    long long nk = 0;
    const bool masked = !x->manno.empty();
    if (!masked) {
        kofs.resize(x->nreads + 1);
        for (int i = 0; i < x->nreads; i++) {
            kofs[i] = nk;
            nk += std::max(0, x->rlen[i] - K + 1);
        }
    } else {
        for (int i = 0; i < x->nreads; i++) {
            const long long b = x->manno[i], f = x->manno[i + 1];
            for (long long a = b; a <= f; a += 2) {
                const int p = (a == b) ? 0 : x->mpts[a - 1];
                const int q = (a == f) ? x->rlen[i] : x->mpts[a];
                if (p + K > q) continue;
                kofs.push_back(nk);
                sread.push_back(i);
                spos0.push_back(p + K - 1);
                nk += q - p - K + 1;
            }
        }
        kofs.push_back(nk);
    }
    const int nrun = (int)kofs.size() - 1;
    kofs.back() = nk;
    DBuf<long long> dk;
    DBuf<int> dsr, dsp;
    dk.ensure(nrun + 1);
    HC(hipMemcpyAsync(dk.p, kofs.data(), sizeof(long long) * (nrun + 1), hipMemcpyHostToDevice, st));
    if (masked && nrun > 0) {
        dsr.ensure(nrun);
        dsp.ensure(nrun);
        HC(hipMemcpyAsync(dsr.p, sread.data(), sizeof(int) * nrun, hipMemcpyHostToDevice, st));
        HC(hipMemcpyAsync(dsp.p, spos0.data(), sizeof(int) * nrun, hipMemcpyHostToDevice, st));
    }
    DBuf<uint64_t> k2;
    x->keys.ensure(nk + 1);
    k2.ensure(nk + 1);
    if (nk > 0) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_kmers), dim3(nblk(nk)), dim3(256), 0, st, x->pk.p, x->dw0.p, dk.p, nrun,
                           masked ? dsr.p : nullptr, masked ? dsp.p : nullptr, nk, K, x->fmt, x->keys.p);
        HC(hipGetLastError());
    }
    const int cs = x->fmt.cshift();
    uint64_t* sorted = radix_sort(x->keys.p, k2.p, nk, cs, cs + 2 * K, KeyU64{});
    if (sorted != x->keys.p) x->keys.swap(k2);
    x->nk = nk;
    suppress_frequent(x, k2);
    HC(hipStreamSynchronize(st));
}

// -t: drop the k-mer groups of >= t entries from the sorted keys (dalign/filter.c:645-696);
// k2 is scratch of at least nk + 1 keys
static void suppress_frequent(KIdx* x, DBuf<uint64_t>& k2)
{
    hipStream_t st = fstream();
    const long long nk = x->nk;
    if (g_suppress > 0 && nk > 0) {
        DBuf<u64> flag, scan;
        DBuf<long long> gidx, gstart;
        flag.ensure(nk);
        scan.ensure(nk + 1);
        gidx.ensure(nk);
        gstart.ensure(nk + 1);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_flags), dim3(nblk(nk)), dim3(256), 0, st, x->keys.p, nk, x->fmt, flag.p);
        g_scan.run(flag.p, scan.p, nk);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_index), dim3(nblk(nk)), dim3(256), 0, st, flag.p, scan.p, nk, gidx.p, gstart.p);
        const long long ng = (long long)d2h1(scan.p + nk);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_set_ll), dim3(1), dim3(1), 0, st, gstart.p + ng, nk);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_keep_small), dim3(nblk(nk)), dim3(256), 0, st, gidx.p, gstart.p, nk, g_suppress, flag.p);
        g_scan.run(flag.p, scan.p, nk);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_compact_u64), dim3(nblk(nk)), dim3(256), 0, st, x->keys.p, flag.p, scan.p, nk, k2.p);
        HC(hipGetLastError());
        x->keys.swap(k2);
        x->nk = (long long)d2h1(scan.p + nk);
    }
    HC(hipStreamSynchronize(st));
}

static void upload_layout(KIdx* x)
{
    hipStream_t st = fstream();
    x->dw0.ensure(x->nreads + 1);
    x->dlen.ensure(x->nreads + 1);
    x->pk.ensure(x->nwords);
    HC(hipMemsetAsync(x->pk.p, 0, sizeof(uint64_t) * x->nwords, st));
    if (x->nreads > 0) {
        HC(hipMemcpyAsync(x->dw0.p, x->w0.data(), sizeof(long long) * x->nreads, hipMemcpyHostToDevice, st));
        HC(hipMemcpyAsync(x->dlen.p, x->rlen.data(), sizeof(int) * x->nreads, hipMemcpyHostToDevice, st));
    }
}

static KIdx* index_from_block(HITS_DB* blk, bool keys = true)
{
    std::unique_ptr<KIdx> x(new KIdx);
    layout_reads(x.get(), blk);
    upload_layout(x.get());
    hipStream_t st = fstream();
    if (x->nreads > 0) {
        // bytes of all reads, one upload of the span they occupy
        long long lo = INT64_MAX, hi = 0;
        for (int i = 0; i < x->nreads; i++) {
            lo = std::min<long long>(lo, blk->reads[i].boff);
            hi = std::max<long long>(hi, blk->reads[i].boff + x->rlen[i]);
        }
        std::vector<long long> boff(x->nreads);
        for (int i = 0; i < x->nreads; i++) boff[i] = blk->reads[i].boff - lo;
        DBuf<uint8_t> bytes;
        DBuf<long long> dboff;
        bytes.ensure(hi - lo);
        dboff.ensure(x->nreads);
        HC(hipMemcpyAsync(bytes.p, (const char*)blk->bases + lo, hi - lo, hipMemcpyHostToDevice, st));
        HC(hipMemcpyAsync(dboff.p, boff.data(), sizeof(long long) * x->nreads, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pack), dim3(x->nreads), dim3(256), 0, st, bytes.p, dboff.p, x->dw0.p, x->dlen.p, x->nreads, x->pk.p);
        HC(hipGetLastError());
        HC(hipStreamSynchronize(st));
    }
    take_mask(x.get(), blk);
    if (keys) build_keys(x.get());
    return x.release();
}

// k-mer index of reads already packed in HBM (the handle of mi_block_device)
static KIdx* index_from_device(const KIdx* src)
{
    std::unique_ptr<KIdx> x(new KIdx);
    x->nreads = src->nreads;
    x->maxlen = src->maxlen;
    x->ufirst = src->ufirst;
    x->part = src->part;
    for (int i = 0; i < 4; i++) x->freq[i] = src->freq[i];
    x->fmt = src->fmt;
    x->rlen = src->rlen;
    x->w0 = src->w0;
    x->manno = src->manno;
    x->mpts = src->mpts;
    x->nwords = src->nwords;
    hipStream_t st = fstream();
    x->pk.ensure(x->nwords);
    x->dw0.ensure(x->nreads + 1);
    x->dlen.ensure(x->nreads + 1);
    HC(hipMemcpyAsync(x->pk.p, src->pk.p, sizeof(uint64_t) * x->nwords, hipMemcpyDeviceToDevice, st));
    HC(hipMemcpyAsync(x->dw0.p, src->dw0.p, sizeof(long long) * (x->nreads + 1), hipMemcpyDeviceToDevice, st));
    HC(hipMemcpyAsync(x->dlen.p, src->dlen.p, sizeof(int) * (x->nreads + 1), hipMemcpyDeviceToDevice, st));
    build_keys(x.get());
    return x.release();
}

static KIdx* index_complement(const KIdx* src, HITS_DB* cblk)
{
    if (cblk->nreads != src->nreads) throw Fail("complement block has a different read count", MI_ERR_ARG);
    for (int i = 0; i < src->nreads; i++)
        if (cblk->reads[i].rlen != src->rlen[i]) throw Fail("complement block has different read lengths", MI_ERR_ARG);
    std::unique_ptr<KIdx> x(new KIdx);
    layout_reads(x.get(), cblk);
    upload_layout(x.get());
    hipStream_t st = fstream();
    if (x->nreads > 0) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_revcomp), dim3(x->nreads), dim3(256), 0, st, src->pk.p, x->pk.p, x->dw0.p, x->dlen.p, x->nreads);
        HC(hipGetLastError());
    }
    take_mask(x.get(), cblk);
    build_keys(x.get());
    return x.release();
}

static void ensure_groups(KIdx* x)
{
    if (x->ng >= 0) return;
    hipStream_t st = fstream();
    const long long n = x->nk;
    x->gidx.ensure(n);
    x->gstart.ensure(n + 1);
    DBuf<u64> flag, scan;
    flag.ensure(n);
    scan.ensure(n + 1);
    if (n > 0) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_flags), dim3(nblk(n)), dim3(256), 0, st, x->keys.p, n, x->fmt, flag.p);
        g_scan.run(flag.p, scan.p, n);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_index), dim3(nblk(n)), dim3(256), 0, st, flag.p, scan.p, n, x->gidx.p, x->gstart.p);
        HC(hipGetLastError());
    }
    const long long ng = n > 0 ? (long long)d2h1(scan.p + n) : 0;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_set_ll), dim3(1), dim3(1), 0, st, x->gstart.p + ng, n);
    HC(hipStreamSynchronize(st));
    x->ng = ng;
}

static KIdx* as_index(void* t)
{
    KIdx* x = (KIdx*)t;
    if (x == nullptr || x->magic != kIdxMagic) throw Fail("not a k-mer index of this library", MI_ERR_ARG);
    return x;
}

// ---- one Match_Filter pass ----------------------------------------------------------

static MiFilterStats g_stats;

// HIP events on the filter stream around the kernels whose rates are reported
struct EvPair {
    hipEvent_t a = nullptr, b = nullptr;
    void record_start()
    {
        if (a == nullptr) {
            HC(hipEventCreate(&a));
            HC(hipEventCreate(&b));
        }
        HC(hipEventRecord(a, fstream()));
    }
    void record_stop() { HC(hipEventRecord(b, fstream())); }
    double seconds()
    {
        HC(hipEventSynchronize(b));
        float ms = 0;
        HC(hipEventElapsedTime(&ms, a, b));
        return ms * 1e-3;
    }
};
static EvPair g_ev_report, g_ev_sort;

// The report kernel's first window for the next pass.  Every window computes
// the same alignments, so this only decides speed: a pass that starts at 64
// diagonals re-runs the pairs that outgrow them at 128 after a partial run,
// which pays on blocks where few pairs do; on repeat-rich blocks, where the
// 128-diagonal re-runs take longer than the whole 64-diagonal launch, the
// next passes start at 128 (kReprobe of them, then one starts at 64 again to
// re-measure).  Match_Filter runs one call at a time, as the reference's.
static int g_report_first = 0, g_report_reprobe = 0;
constexpr double kWideOverNarrow = 1.0;
constexpr int kReprobe = 8;

struct Level {
    int nch, pool_cap, mcap;
    long long mtcap;
    int waves;
};

__global__ void k_gather_start(const int* jobs, const long long* pstart, int n, long long* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pstart[jobs[i]];
}

// launch order of the jobs: most seeds first (the pairs with long true
// overlaps), so the longest alignments do not start at the end of the queue.
// key = (2^20 - 1 - min(seeds, 2^20 - 1)) << 32 | job, sorted on bits 32..51
__global__ void k_job_keys(const int* jobs, const long long* pstart, const unsigned long long* keep, int n,
                           unsigned long long* keys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long c = pstart[jobs[i] + 1] - pstart[jobs[i]];
    // kept jobs have c >= minhit >= 1, so their keys sort before the skipped ones (r = 2^20 - 1)
    const unsigned long long r = keep[i] ? 0xFFFFFull - (unsigned long long)(c < 0xFFFFE ? c : 0xFFFFE) : 0xFFFFFull;
    keys[i] = (r << 32) | (unsigned)i;
}

__global__ void k_low32(const unsigned long long* keys, int n, int* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int)(keys[i] & 0xffffffffu);
}

// (job, status) of the jobs of a launch that did not finish
__global__ void k_collect_failed(const int* __restrict__ order, long long n, const int* __restrict__ status,
                                 int* __restrict__ flist, int* __restrict__ fcount)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = order[i];
    const int s = status[j];
    if (s != 0) {
        const int q = atomicAdd(fcount, 1);
        flist[2 * q] = j;
        flist[2 * q + 1] = s;
    }
}

__global__ void k_flag_nonzero(const int* __restrict__ v, long long n, unsigned long long* __restrict__ flag)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = v[i] != 0 ? 1 : 0;
}

__global__ void k_touched_info(const int* __restrict__ tj, long long n, const int* __restrict__ jobs,
                               const long long* __restrict__ pstart, const int* __restrict__ edge_out,
                               int* __restrict__ te, long long* __restrict__ ts)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = tj[i];
    te[i] = edge_out[j];
    ts[i] = pstart[jobs[j]];
}

__global__ void k_scatter_int(const int* __restrict__ idx, const int* __restrict__ v, long long n, int* __restrict__ dst)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = v[i];
}

__global__ void k_gather_int(const int* __restrict__ idx, const int* __restrict__ src, long long n, int* __restrict__ dst)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// ---- record ordering on the GPU ----
struct RecKey {
    unsigned long long key;
    long long idx;
};
struct RecKeyOf {
    __device__ __forceinline__ unsigned long long operator()(const RecKey& r) const { return r.key; }
};
struct RecKeyFmt {
    int rfirst, rb, pb, lb;   // read id base and bit widths of read ids, positions, list positions
};

// key (aread, bread, abpos, local) of the records a job's final launch wrote; all ones otherwise
__global__ void k_rec_keys(const OutRec* __restrict__ r, long long n, int njobs, const int* __restrict__ final_gen,
                           RecKeyFmt f, RecKey* __restrict__ keys, unsigned long long* __restrict__ flag)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const OutRec x = r[i];
    const bool ok = x.job >= 0 && x.job < njobs && final_gen[x.job] == x.gen;
    unsigned long long k = ~0ull;
    if (ok)
        k = ((unsigned long long)(x.aread - f.rfirst) << (f.rb + f.pb + f.lb)) |
            ((unsigned long long)(x.bread - f.rfirst) << (f.pb + f.lb)) | ((unsigned long long)x.abpos << f.lb) |
            (unsigned long long)x.local;
    keys[i] = RecKey{k, i};
    flag[i] = ok ? 1 : 0;
}

__global__ void k_rec_gather(const OutRec* __restrict__ r, const RecKey* __restrict__ sorted, long long n, int tbytes,
                             OutRec* __restrict__ out, unsigned long long* __restrict__ tb)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const OutRec x = r[sorted[i].idx];
    out[i] = x;
    tb[i] = (unsigned long long)x.tlen * tbytes;
}

// one wave per record: its trace bytes to their place in the compacted array;
// the record's toff becomes that place
__global__ __launch_bounds__(64) void k_trace_gather(OutRec* __restrict__ recs, const uint8_t* __restrict__ src,
                                                     const unsigned long long* __restrict__ pos, long long n,
                                                     int tbytes, uint8_t* __restrict__ dst)
{
    const long long i = blockIdx.x;
    if (i >= n) return;
    const long long from = recs[i].toff;
    const long long len = (long long)recs[i].tlen * tbytes;
    uint8_t* d = dst + pos[i];
    for (long long q = threadIdx.x; q < len; q += 64) d[q] = src[from + q];
    __syncthreads();
    if (threadIdx.x == 0) recs[i].toff = (long long)pos[i];
}

// candidate scorer state (scorer.hip): mode 0 exact, 1 capture, 2 gate
constexpr int SC_EXACT = 0, SC_CAPTURE = 1, SC_GATE = 2;
int scorer_mode();
bool scorer_weights(MlpW* w, float* thr);
void scorer_logits(const float* dfeat, long long n, const MlpW& w, float* dout, hipStream_t st);
void scorer_capture_append(const float* feat, const int* nrec, const unsigned long long* work, long long n);

// records of each job's final launch, counted per job
__global__ void k_job_records(const OutRec* __restrict__ r, long long n, int njobs, const int* __restrict__ final_gen,
                              int* __restrict__ cnt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const OutRec x = r[i];
    if (x.job >= 0 && x.job < njobs && final_gen[x.job] == x.gen) atomicAdd(&cnt[x.job], 1);
}

template <int NCH>
static void launch_report(const ReportArgs& a, int grid)
{
    if (a.job_work || a.redund)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_report<NCH, true>), dim3(grid), dim3(64), 0, fstream(), a);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_report<NCH, false>), dim3(grid), dim3(64), 0, fstream(), a);
}

static void match_pass(KIdx* A, KIdx* B, bool self, int comp, bool same_index, SpecObj* spec)
{
    MiFilterStats& S = g_stats;
    const double t0 = now_s();
    hipStream_t st = fstream();
    const int K = g_kmer, BS = g_binshift;
    S.nkmers_a = A->nk;
    S.nkmers_b = B->nk;
    if (A->nk == 0 || B->nk == 0) return;
    if (A->K == 0 || B->K == 0) throw Fail("device block without a k-mer index (use mi_sort_kmers_device)", MI_ERR_ARG);
    if (A->K != K || B->K != K) throw Fail("index built with another k-mer length", MI_ERR_ARG);
    ensure_groups(A);
    const long long na = A->nk, nb = B->nk, ng = A->ng;

    // merge-join counts (count_thread, dalign/filter.c:948-1056)
    // B range of every A group: with B the A index itself, the group's own
    // entries (gstart[g], gstart[g + 1]) -- no copy
    DBuf<long long> bs, be;
    DBuf<u64> cnt, off, gtot, hist;
    if (!same_index) {
        bs.ensure(ng);
        be.ensure(ng);
    }
    const long long* bsp = same_index ? A->gstart.p : bs.p;
    const long long* bep = same_index ? A->gstart.p + 1 : be.p;
    cnt.ensure(na);
    off.ensure(na + 1);
    gtot.ensure(ng);
    const int MAXGRAM = 10000;
    hist.ensure(MAXGRAM);
    HC(hipMemsetAsync(hist.p, 0, sizeof(u64) * MAXGRAM, st));
    DBuf<long long> bound;
    if (!same_index) {
        const int nblock = (int)nblk(ng);
        bound.ensure(2 * (size_t)nblock);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_brange_bounds), dim3(nblk(nblock)), dim3(256), 0, st, A->keys.p, A->fmt,
                           A->gstart.p, ng, B->keys.p, B->fmt, nb, nblock, bound.p);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_brange), dim3(nblk(ng)), dim3(256), 0, st, A->keys.p, A->fmt,
                           A->gstart.p, ng, B->keys.p, B->fmt, nb, bound.p, bs.p, be.p);
    }
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_counts), dim3(nblk(na)), dim3(256), 0, st, A->keys.p, A->fmt, na, A->gidx.p, bsp, bep, B->keys.p, B->fmt,
                       self ? (IDENTITY ? (comp ? 3 : 2) : 1) : 0, cnt.p);
    g_scan.run(cnt.p, off.p, na);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_totals), dim3(nblk(ng, 256 * 16)), dim3(256), 0, st, A->gstart.p, ng, off.p, gtot.p, hist.p, MAXGRAM);
    HC(hipGetLastError());
    std::vector<u64> h(MAXGRAM);
    HC(hipMemcpyAsync(h.data(), hist.p, sizeof(u64) * MAXGRAM, hipMemcpyDeviceToHost, st));
    const u64 total = d2h1(off.p + na);

    // cap from MEM_LIMIT (dalign/filter.c:2107-2174)
    long long limit = INT32_MAX;
    if (MEM_LIMIT > 0) {
        const uint64_t M16 = MEM_LIMIT / 16;
        long long avail;
        if (same_index || (long long)M16 > na + 2 * nb) avail = (long long)((M16 - (uint64_t)na) / 2);
        else avail = (long long)(M16 - (uint64_t)(na + nb));
        avail = (long long)(avail * .98);
        long long tom = 0;
        int j;
        for (j = 0; j < MAXGRAM; j++) {
            tom += (long long)j * (long long)h[j];
            if (tom > avail) break;
        }
        limit = j;
        if (limit == 1) throw Fail("insufficient memory limit (-M) for the k-mer matches", MI_ERR_ARG);
    }
    S.limit = (int)std::min<long long>(limit, INT32_MAX);
    long long nh = (long long)total;
    if (limit < INT32_MAX) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_apply_limit), dim3(nblk(na)), dim3(256), 0, st, A->gidx.p, gtot.p, na, (u64)limit, cnt.p);
        g_scan.run(cnt.p, off.p, na);
        nh = (long long)d2h1(off.p + na);
        long long expect = 0;
        for (int j = 1; j < limit && j < MAXGRAM; j++) expect += (long long)j * (long long)h[j];
        if (expect != nh) throw Fail("seed count disagrees with the count histogram");
    }
    S.nhits += nh;
    if (nh == 0) return;

    // seeds, then the pair sort (dalign/filter.c:1210-1222, :2030-2054)
    DBuf<Seed> s1, s2;
    s1.ensure(nh + 1);
    s2.ensure(nh + 1);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_emit), dim3(nblk(na)), dim3(256), 0, st, A->keys.p, A->fmt, na, A->gidx.p,
                       bsp, B->keys.p, B->fmt, off.p, s1.p);
    HC(hipGetLastError());
    const int pa = A->fmt.pbits, ra = A->fmt.rbits, rb = B->fmt.rbits;
    if (pa + ra + rb > 64) throw Fail("blocks too large for 64-bit seed keys", MI_ERR_ARG);
    g_ev_sort.record_start();
    Seed* hits = radix_sort(s1.p, s2.p, nh, 0, pa + ra + rb, KeySeed{pa, pa + ra});
    g_ev_sort.record_stop();
    S.t_seed_sort += g_ev_sort.seconds();
    S.seed_sort_passes += (pa + ra + rb + 7) / 8;
    Seed* spare = (hits == s1.p) ? s2.p : s1.p;
    (void)spare;

    // read pairs
    DBuf<u64> pflag, pscan;
    pflag.ensure(nh);
    pscan.ensure(nh + 1);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_flags), dim3(nblk(nh)), dim3(256), 0, st, hits, nh, pflag.p);
    g_scan.run(pflag.p, pscan.p, nh);
    const long long np = (long long)d2h1(pscan.p + nh);
    DBuf<long long> pstart;
    pstart.ensure(np + 1);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_group_index), dim3(nblk(nh)), dim3(256), 0, st, pflag.p, pscan.p, nh, nullptr, pstart.p);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_set_ll), dim3(1), dim3(1), 0, st, pstart.p + np, nh);
    S.npairs += np;

    // report chunks and the pairs they scan (dalign/filter.c:1699-1731, :2282-2294)
    const int nthreads = std::max(1, NTHREADS);
    if (nthreads > 1023 || (1 << NSHIFT) != nthreads) throw Fail("NTHREADS must be a power of 2 below 1024", MI_ERR_ARG);
    DBuf<long long> cb;
    cb.ensure(nthreads + 1);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_chunk_bounds), dim3(1), dim3(1024), 0, st, hits, nh, nthreads, NSHIFT, cb.p);
    std::vector<long long> hcb(nthreads + 1);
    HC(hipMemcpyAsync(hcb.data(), cb.p, sizeof(long long) * (nthreads + 1), hipMemcpyDeviceToHost, st));
    const int minhit = (g_hitmin - 1) / K + 1;
    DBuf<u64> sel, sscan;
    sel.ensure(np);
    sscan.ensure(np + 1);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_select), dim3(nblk(np)), dim3(256), 0, st, pstart.p, np, hits, A->dlen.p, B->dlen.p, cb.p, nthreads, minhit, HGAP_MIN,
                                        sel.p);
    g_scan.run(sel.p, sscan.p, np);
    const long long njobs = (long long)d2h1(sscan.p + np);
    S.njobs += njobs;
    if (njobs == 0) return;
    if (njobs > INT32_MAX) throw Fail("too many read pairs in one pass");
    DBuf<int> jobs;
    jobs.ensure(njobs);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_compact_idx), dim3(nblk(np)), dim3(256), 0, st, sel.p, sscan.p, np, jobs.p);
    HC(hipStreamSynchronize(st));
    const double t1 = now_s();
    S.t_filter += t1 - t0;

    // ---- report: one wave per pair ----
    DevSpecHost* ds = mi_spec_device(&spec->prm);
    if (ds == nullptr) throw Fail("cannot upload alignment tables");
    const int TS = spec->hs.tspace;
    const int maxdiag = A->maxlen >> BS, mindiag = (-B->maxlen) >> BS;
    const int w = maxdiag - mindiag + 1;
    const int maxtp_all = 2 * (std::max(A->maxlen, B->maxlen) / TS + 2);
    const int cus = mi_num_cus();
    // waves per level: up to 8 per SIMD for the 64-window kernel; the per-wave
    // pools are sized for typical pairs, outliers go to the larger levels
    // aligner windows of 64, 128, 256 and 1,024 diagonals; the last level
    // also has the large pebble pools and match arenas
    constexpr int kLevels = 5;
    const MiTuning& tu = mi_tuning();
    const int l0_waves = cus * (tu.report_waves_per_cu ? tu.report_waves_per_cu : 32);
    const int l0_pool = 1 << (tu.report_pool_log2 ? tu.report_pool_log2 : 15);
    const int l1_waves = cus * (tu.report_l1_waves_per_cu ? tu.report_l1_waves_per_cu : 16);
    const Level levels[kLevels] = {{1, l0_pool, 128, 1 << 16, l0_waves},
                                   {2, 1 << 16, 256, 1 << 18, l1_waves},
                                   {4, 1 << 17, 256, 1 << 18, cus * 8},
                                   {16, 1 << 17, 256, 1 << 18, cus * 4},
                                   {16, 1 << 20, 4096, 1LL << 22, std::min(cus, 256)}};
    const int small_trace = TS <= TRACE_XOVR ? 1 : 0;
    const int tbytes = small_trace ? 1 : 2;

    DBuf<int> bands, status, final_gen, edge_in, edge_out, touch, order, order_all, counter, flist, fcount;
    DBuf<Peb> pools, pts;
    DBuf<uint16_t> latr, mtrace;
    DBuf<MPath> mpaths;
    DBuf<u64> ocount, stats;
    status.ensure(njobs);
    final_gen.ensure(njobs);
    edge_in.ensure(njobs);
    edge_out.ensure(njobs);
    touch.ensure(njobs);
    order.ensure(njobs);
    order_all.ensure(njobs);
    flist.ensure(njobs);
    counter.ensure(1);
    fcount.ensure(1);
    ocount.ensure(2);
    stats.ensure(7);
    HC(hipMemsetAsync(ocount.p, 0, 2 * sizeof(u64), st));
    HC(hipMemsetAsync(stats.p, 0, 7 * sizeof(u64), st));
    HC(hipMemsetAsync(edge_in.p, 0, sizeof(int) * njobs, st));
    HC(hipMemsetAsync(edge_out.p, 0, sizeof(int) * njobs, st));
    HC(hipMemsetAsync(status.p, 0, sizeof(int) * njobs, st));
    HC(hipMemsetAsync(final_gen.p, 0xff, sizeof(int) * njobs, st));

    // exact gate pre-check: pairs whose seeds cannot pass report_thread's
    // gate are finished here (no alignment, no record); the rest run
    // longest-first
    long long nfirst = 0;
    const int sc_mode = scorer_mode();
    DBuf<float> jfeat;   // scorer features per job (capture / gate)
    DBuf<u64> jwork;     // capture: aligner work per job
    DBuf<u64> kflag;     // capture: the jobs past the pre-check
    {
        DBuf<u64> keep, kscan, k1, k2;
        keep.ensure(njobs);
        kscan.ensure(njobs + 1);
        k1.ensure(njobs);
        k2.ensure(njobs);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_prefilter), dim3(nblk(njobs)), dim3(256), 0, st, jobs.p, (int)njobs,
                           pstart.p, hits, K, BS, g_hitmin, minhit, mindiag, maxdiag, keep.p, touch.p);
        HC(hipGetLastError());
        // candidate scorer: features of every pair past the pre-check; gating
        // drops the pairs whose logit is below the threshold before any
        // alignment, capture keeps the features for the outcome labels
        if (sc_mode != SC_EXACT) {
            jfeat.ensure((size_t)njobs * kJobFeat);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_job_features), dim3((njobs + 127) / 128), dim3(128), 0, st, jobs.p,
                               (int)njobs, keep.p, pstart.p, hits, A->pk.p, A->dw0.p, A->dlen.p, B->dlen.p, K, BS,
                               minhit, comp, jfeat.p);
            HC(hipGetLastError());
            if (sc_mode == SC_GATE) {
                MlpW w;
                float thr;
                if (!scorer_weights(&w, &thr)) throw Fail("candidate gating requested without scorer weights", MI_ERR_ARG);
                DBuf<float> lg;
                DBuf<u64> gc;
                lg.ensure(njobs);
                gc.ensure(2);
                HC(hipMemsetAsync(gc.p, 0, 2 * sizeof(u64), st));
                scorer_logits(jfeat.p, njobs, w, lg.p, st);
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gate_jobs), dim3(nblk(njobs)), dim3(256), 0, st, keep.p, lg.p,
                                   (int)njobs, thr, gc.p);
                HC(hipGetLastError());
                u64 hc[2];
                HC(hipMemcpy(hc, gc.p, sizeof hc, hipMemcpyDeviceToHost));
                S.cand_scored += (long long)hc[0];
                S.cand_skipped += (long long)hc[1];
            } else {
                jwork.ensure(njobs);
                kflag.ensure(njobs);
                HC(hipMemsetAsync(jwork.p, 0, sizeof(u64) * njobs, st));
                HC(hipMemcpyAsync(kflag.p, keep.p, sizeof(u64) * njobs, hipMemcpyDeviceToDevice, st));
            }
        }
        g_scan.run(keep.p, kscan.p, njobs);
        nfirst = (long long)d2h1(kscan.p + njobs);
        S.njobs_kept += nfirst;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_job_keys), dim3(nblk(njobs)), dim3(256), 0, st, jobs.p, pstart.p, keep.p,
                           (int)njobs, k1.p);
        HC(hipGetLastError());
        u64* sorted = radix_sort(k1.p, k2.p, njobs, 32, 52, KeyU64{});
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_low32), dim3(nblk(njobs)), dim3(256), 0, st, sorted, (int)njobs,
                           order_all.p);
        HC(hipGetLastError());
    }

    // output arenas: about two records per kept pair; overflows grow them and
    // re-run the pairs that did not fit.  MI_REPORT_ARENA (records) sets the
    // initial size instead, e.g. a tiny one to exercise that path.
    long long ocap = std::max<long long>(4096, 4 * nfirst);
    if (const char* e = getenv("MI_REPORT_ARENA")) ocap = std::max(16LL, atoll(e));
    long long otcap = ocap * 160;
    DBuf<OutRec> orecs;
    DBuf<uint8_t> otr;
    orecs.ensure(ocap);
    otr.ensure(otcap);
    HC(hipMemsetAsync(orecs.p, 0xff, sizeof(OutRec) * ocap, st));

    ReportArgs a;
    memset(&a, 0, sizeof a);
    a.hits = hits;
    a.pstart = pstart.p;
    a.jobs = jobs.p;
    a.pkA = A->pk.p;
    a.w0A = A->dw0.p;
    a.lenA = A->dlen.p;
    a.pkB = B->pk.p;
    a.w0B = B->dw0.p;
    a.lenB = B->dlen.p;
    a.afirst = A->ufirst;
    a.bfirst = B->ufirst;
    a.comp = comp;
    a.self = self ? 1 : 0;
    a.symmetric = SYMMETRIC;
    a.hgap_min = HGAP_MIN;
    a.K = K;
    a.binshift = BS;
    a.hitmin = g_hitmin;
    a.minhit = minhit;
    a.minover = MINOVER;
    a.tspace = TS;
    a.small_trace = small_trace;
    a.mindiag = mindiag;
    a.maxdiag = maxdiag;
    a.w = w;
    a.sp = ds->spec;
    a.maxtp_all = maxtp_all;
    a.edge_in = edge_in.p;
    a.edge_out = edge_out.p;
    a.touch = touch.p;
    a.status = status.p;
    a.final_gen = final_gen.p;
    a.counter = counter.p;
    a.abort_flag = mi_abort_flag_device();
    a.stats = stats.p;
    a.ocount = ocount.p;
    a.job_work = sc_mode == SC_CAPTURE ? jwork.p : nullptr;
    // MI_REDUND_STATS=1: count the branches handle_redundancies takes (final run of every job)
    DBuf<unsigned> redund;
    const bool want_redund = getenv("MI_REDUND_STATS") != nullptr && atoi(getenv("MI_REDUND_STATS")) != 0;
    if (want_redund) {
        redund.ensure((size_t)njobs * kRedundBranches);
        HC(hipMemsetAsync(redund.p, 0, sizeof(unsigned) * njobs * kRedundBranches, st));
        a.redund = redund.p;
    }

    int cur_level = -1, gen = 0;
    double t_level[kLevels] = {0., 0., 0., 0., 0.};   // report kernel seconds per window of this pass

    // runs the n jobs listed at dorder (device) at `level`; returns the
    // (job, status) pairs that need another launch
    auto run = [&](const int* dorder, long long n, int level) {
        std::vector<std::pair<int, int>> again;
        if (n == 0) return again;
        const Level& L = levels[level];
        const int grid = (int)std::min<long long>(L.waves, n);
        if (level != cur_level) {
            bands.release();
            pools.release();
            pts.release();
            latr.release();
            mpaths.release();
            mtrace.release();
            cur_level = level;
        }
        const int gmax = L.waves;
        bands.ensure((size_t)(3 * w + 2) * gmax);
        pools.ensure((size_t)L.pool_cap * gmax);
        pts.ensure((size_t)2 * 8192 * gmax);
        latr.ensure((size_t)4 * maxtp_all * gmax);
        mpaths.ensure((size_t)2 * L.mcap * gmax);
        mtrace.ensure((size_t)L.mtcap * gmax);
        HC(hipMemsetAsync(bands.p, 0, sizeof(int) * (size_t)(3 * w + 2) * gmax, st));
        HC(hipMemsetAsync(counter.p, 0, sizeof(int), st));
        HC(hipMemsetAsync(fcount.p, 0, sizeof(int), st));
        ReportArgs b = a;
        b.order = dorder;
        b.norder = (int)n;
        b.gen = gen++;
        b.bands = bands.p;
        b.pools = pools.p;
        b.pool_cap = L.pool_cap;
        b.pts = pts.p;
        b.pts_cap = 8192;
        b.latr = latr.p;
        b.mpaths = mpaths.p;
        b.mcap = L.mcap;
        b.mtrace = mtrace.p;
        b.mtcap = L.mtcap;
        b.orecs = orecs.p;
        b.ocap = ocap;
        b.otr = otr.p;
        b.otcap = otcap;
        g_ev_report.record_start();
        switch (L.nch) {
        case 1: launch_report<1>(b, grid); break;
        case 2: launch_report<2>(b, grid); break;
        case 4: launch_report<4>(b, grid); break;
        default: launch_report<16>(b, grid); break;
        }
        HC(hipGetLastError());
        g_ev_report.record_stop();
        // jobs of this launch that did not finish
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_collect_failed), dim3(nblk(n)), dim3(256), 0, st, dorder, n, status.p,
                           flist.p, fcount.p);
        const int wd = mi_wait_watchdog(st, "Match_Filter report kernel");
        if (wd > 0) throw Fail("report kernel exceeded the watchdog deadline and was stopped", MI_ERR_TIMEOUT);
        if (wd < 0) throw Fail("report kernel failed on the device");
        HC(hipGetLastError());
        const double sec = g_ev_report.seconds();
        S.t_report_kernel += sec;
        S.report_launches += 1;
        t_level[level] += sec;
        u64 oc2[2];
        int nf = 0;
        HC(hipMemcpy(oc2, ocount.p, sizeof oc2, hipMemcpyDeviceToHost));
        HC(hipMemcpy(&nf, fcount.p, sizeof(int), hipMemcpyDeviceToHost));
        std::vector<int> fl(2 * (size_t)nf);
        if (nf > 0) HC(hipMemcpy(fl.data(), flist.p, sizeof(int) * 2 * nf, hipMemcpyDeviceToHost));
        bool grow = false;
        for (int i = 0; i < nf; i++) {
            const int j = fl[2 * i], st_ = fl[2 * i + 1];
            if (st_ == RP_UNDEF) {
                char m[160];
                snprintf(m, sizeof m, "pair job %d: an alignment reached a wave state the reference leaves undefined", j);
                throw Fail(m, MI_ERR_UNDEF);
            }
            if (st_ == RP_OUT) grow = true;
            again.push_back({j, st_});
        }
        std::sort(again.begin(), again.end());
        if (grow || (long long)oc2[0] > ocap || (long long)oc2[1] > otcap) {
            // grow the output arenas past what has been reserved so far
            const long long nrec = std::max<long long>(ocap * 2, (long long)oc2[0] * 3 / 2 + 4096);
            const long long ntr = std::max<long long>(otcap * 2, (long long)oc2[1] * 3 / 2 + 65536);
            DBuf<OutRec> r2;
            DBuf<uint8_t> t2;
            r2.ensure(nrec);
            t2.ensure(ntr);
            HC(hipMemsetAsync(r2.p, 0xff, sizeof(OutRec) * nrec, st));
            HC(hipMemcpyAsync(r2.p, orecs.p, sizeof(OutRec) * ocap, hipMemcpyDeviceToDevice, st));
            HC(hipMemcpyAsync(t2.p, otr.p, otcap, hipMemcpyDeviceToDevice, st));
            HC(hipStreamSynchronize(st));
            orecs.swap(r2);
            otr.swap(t2);
            ocap = nrec;
            otcap = ntr;
        }
        return again;
    };

    // runs a job list until every job is done: output-arena overflows are
    // re-run at the same window, wide bands / full pools at the next one.
    // Jobs start at g_report_first; MI_REPORT_FIRST_LEVEL starts every job at
    // the window it names instead (and keeps the choice fixed).
    int first_level = g_report_first;
    const char* fixed_first = getenv("MI_REPORT_FIRST_LEVEL");
    if (fixed_first) first_level = std::min(std::max(atoi(fixed_first), 0), kLevels - 1);
    auto run_all = [&](const int* dorder, long long n) {
        int level = first_level;
        std::vector<std::pair<int, int>> again = run(dorder, n, level);
        while (!again.empty()) {
            S.retries += (long long)again.size();
            bool up = false, arena = false;
            std::vector<int> todo;
            for (auto& f : again) {
                up |= f.second != RP_OUT;
                arena |= f.second == RP_ARENA;
                todo.push_back(f.first);
            }
            if (up) {
                if (level == kLevels - 1) throw Fail("pairs exceed the largest aligner window / arena");
                level = arena ? kLevels - 1 : level + 1;
            }
            HC(hipMemcpyAsync(order.p, todo.data(), sizeof(int) * todo.size(), hipMemcpyHostToDevice, st));
            again = run(order.p, (long long)todo.size(), level);
        }
    };
    run_all(order_all.p, nfirst);
    if (!fixed_first) {
        if (first_level == 0) {
            if (t_level[0] > 0. && t_level[1] > kWideOverNarrow * t_level[0]) {
                g_report_first = 1;
                g_report_reprobe = kReprobe;
            }
        } else if (--g_report_reprobe <= 0) {
            g_report_first = 0;
        }
    }

    // the cell shared by lastp[maxdiag] and lasta[mindiag-1] carries from
    // one pair to the next within a report chunk: re-run the pairs that
    // would have seen a non-zero value (rare; needs an alignment reaching
    // the lowest band).  Only pairs that touch the cell take part: the
    // others pass it on unchanged.
    {
        DBuf<u64> tflag, tpos;
        tflag.ensure(njobs);
        tpos.ensure(njobs + 1);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_flag_nonzero), dim3(nblk(njobs)), dim3(256), 0, st, touch.p, njobs,
                           tflag.p);
        g_scan.run(tflag.p, tpos.p, njobs);
        const long long nt = (long long)d2h1(tpos.p + njobs);
        if (nt > 0) {
            DBuf<int> tj, te;
            DBuf<long long> ts;
            tj.ensure(nt);
            te.ensure(nt);
            ts.ensure(nt);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_compact_idx), dim3(nblk(njobs)), dim3(256), 0, st, tflag.p, tpos.p,
                               njobs, tj.p);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_touched_info), dim3(nblk(nt)), dim3(256), 0, st, tj.p, nt, jobs.p,
                               pstart.p, edge_out.p, te.p, ts.p);
            std::vector<int> hj(nt), he(nt);
            std::vector<long long> hs(nt);
            HC(hipMemcpyAsync(hj.data(), tj.p, sizeof(int) * nt, hipMemcpyDeviceToHost, st));
            HC(hipMemcpyAsync(he.data(), te.p, sizeof(int) * nt, hipMemcpyDeviceToHost, st));
            HC(hipMemcpyAsync(hs.data(), ts.p, sizeof(long long) * nt, hipMemcpyDeviceToHost, st));
            HC(hipStreamSynchronize(st));
            std::vector<int> chunk(nt), ein(nt, 0);
            for (long long i = 0; i < nt; i++) {
                int c = 0;
                while (c + 1 < nthreads && hcb[c + 1] <= hs[i]) c++;
                chunk[i] = c;
            }
            // walk the touched jobs of every chunk in order with the chunk's carry
            std::vector<long long> pos(nthreads, -1);
            std::vector<int> carry(nthreads, 0);
            for (;;) {
                std::vector<int> rerun;
                std::vector<long long> ridx;
                for (long long i = 0; i < nt; i++) {
                    const int c = chunk[i];
                    if (i <= pos[c] || pos[c] == -2) continue;
                    if (ein[i] != carry[c]) {   // first job that saw another value: re-run it
                        ein[i] = carry[c];
                        rerun.push_back(hj[i]);
                        ridx.push_back(i);
                        pos[c] = -2;   // this chunk waits for the re-run
                        continue;
                    }
                    carry[c] = he[i];
                    pos[c] = i;
                }
                if (rerun.empty()) break;
                S.edge_fixups += (long long)rerun.size();
                // new inputs, re-run, and read back the new outputs
                std::vector<int> vals;
                for (long long i : ridx) vals.push_back(ein[i]);
                DBuf<int> rj, rv;
                rj.ensure(rerun.size());
                rv.ensure(rerun.size());
                HC(hipMemcpyAsync(rj.p, rerun.data(), sizeof(int) * rerun.size(), hipMemcpyHostToDevice, st));
                HC(hipMemcpyAsync(rv.p, vals.data(), sizeof(int) * vals.size(), hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_scatter_int), dim3(nblk(rerun.size())), dim3(256), 0, st, rj.p,
                                   rv.p, (long long)rerun.size(), edge_in.p);
                HC(hipMemcpyAsync(order.p, rerun.data(), sizeof(int) * rerun.size(), hipMemcpyHostToDevice, st));
                run_all(order.p, (long long)rerun.size());
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gather_int), dim3(nblk(rerun.size())), dim3(256), 0, st, rj.p,
                                   edge_out.p, (long long)rerun.size(), rv.p);
                HC(hipMemcpy(vals.data(), rv.p, sizeof(int) * vals.size(), hipMemcpyDeviceToHost));
                for (size_t q = 0; q < ridx.size(); q++) {
                    const long long i = ridx[q];
                    he[i] = vals[q];
                    const int c = chunk[i];
                    carry[c] = he[i];
                    pos[c] = i;
                }
            }
        }
    }
    const double t2 = now_s();
    S.t_report += t2 - t1;

    // ---- records back to the host ----
    // The arena holds the records of every launch; a record is kept when it
    // comes from its job's final launch.  The kept ones are ordered on the
    // GPU by (aread, bread, abpos, position in the pair's list) -- the order
    // report_thread's buffers give equal (aread, bread, COMP, abpos) keys --
    // compacted with their traces, and copied back in one piece.
    u64 oc[2], sv[7];
    HC(hipMemcpy(oc, ocount.p, sizeof oc, hipMemcpyDeviceToHost));
    HC(hipMemcpy(sv, stats.p, sizeof sv, hipMemcpyDeviceToHost));
    S.naligned += (long long)sv[0];
    S.cells += (long long)sv[1];
    S.snake += (long long)sv[2];
    S.edge_events += (long long)sv[3];
    S.wave_steps += (long long)sv[4];
    const long long nrec = std::min<long long>((long long)oc[0], ocap);
    if (want_redund) {
        std::vector<unsigned> hr((size_t)njobs * kRedundBranches);
        HC(hipMemcpy(hr.data(), redund.p, sizeof(unsigned) * hr.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < hr.size(); i++) S.redund[i % kRedundBranches] += hr[i];
    }
    if (sc_mode == SC_CAPTURE) {
        // every pair past the pre-check: its features, the records of its final
        // run and the aligner work that run did
        DBuf<int> jrec;
        jrec.ensure(njobs);
        HC(hipMemsetAsync(jrec.p, 0, sizeof(int) * njobs, st));
        if (nrec > 0)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_job_records), dim3(nblk(nrec)), dim3(256), 0, st, orecs.p, nrec,
                               (int)njobs, final_gen.p, jrec.p);
        std::vector<float> hf((size_t)njobs * kJobFeat);
        std::vector<int> hr(njobs);
        std::vector<u64> hw(njobs), hk(njobs);
        HC(hipMemcpy(hf.data(), jfeat.p, sizeof(float) * hf.size(), hipMemcpyDeviceToHost));
        HC(hipMemcpy(hr.data(), jrec.p, sizeof(int) * njobs, hipMemcpyDeviceToHost));
        HC(hipMemcpy(hw.data(), jwork.p, sizeof(u64) * njobs, hipMemcpyDeviceToHost));
        HC(hipMemcpy(hk.data(), kflag.p, sizeof(u64) * njobs, hipMemcpyDeviceToHost));
        std::vector<float> kf;
        std::vector<int> kr;
        std::vector<u64> kw;
        for (long long j = 0; j < njobs; j++) {
            if (!hk[j]) continue;
            kf.insert(kf.end(), hf.begin() + j * kJobFeat, hf.begin() + (j + 1) * kJobFeat);
            kr.push_back(hr[j]);
            kw.push_back(hw[j]);
        }
        scorer_capture_append(kf.data(), kr.data(), kw.data(), (long long)kr.size());
    }
    long long nkeep = 0, tbytes_all = 0;
    DBuf<OutRec> srec;
    DBuf<uint8_t> str;
    if (nrec > 0) {
        const int rfirst = std::min(A->ufirst, B->ufirst);
        const int rlast = std::max(A->ufirst + A->nreads, B->ufirst + B->nreads);
        const int rb = bits_for(std::max(1, rlast - rfirst));
        const int pb = bits_for(std::max(A->maxlen, B->maxlen));
        const int lb = bits_for(std::max(levels[0].mcap, levels[kLevels - 1].mcap));
        if (2 * rb + pb + lb > 64) throw Fail("record keys exceed 64 bits", MI_ERR_ARG);
        DBuf<RecKey> k1, k2;
        DBuf<u64> flag, pos;
        k1.ensure(nrec);
        k2.ensure(nrec);
        flag.ensure(nrec);
        pos.ensure(nrec + 1);
        const RecKeyFmt kf{rfirst, rb, pb, lb};
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_rec_keys), dim3(nblk(nrec)), dim3(256), 0, st, orecs.p, nrec, (int)njobs,
                           final_gen.p, kf, k1.p, flag.p);
        HC(hipGetLastError());
        g_scan.run(flag.p, pos.p, nrec);
        nkeep = (long long)d2h1(pos.p + nrec);
        RecKey* sorted = radix_sort(k1.p, k2.p, nrec, 0, 2 * rb + pb + lb, RecKeyOf{});
        // sorted: the kept records first (the others have the all-ones key)
        DBuf<u64> tb, tpos;
        tb.ensure(std::max(nkeep, 1LL));
        tpos.ensure(nkeep + 1);
        srec.ensure(std::max(nkeep, 1LL));
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_rec_gather), dim3(nblk(nkeep)), dim3(256), 0, st, orecs.p, sorted, nkeep,
                           tbytes, srec.p, tb.p);
        HC(hipGetLastError());
        g_scan.run(tb.p, tpos.p, nkeep);
        tbytes_all = (long long)d2h1(tpos.p + nkeep);
        str.ensure(std::max(tbytes_all, 1LL));
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace_gather), dim3(std::max<long long>(nkeep, 1)), dim3(64), 0, st, srec.p,
                           otr.p, tpos.p, nkeep, tbytes, str.p);
        HC(hipGetLastError());
    }
    OutRec* recs = (OutRec*)g_pin_rec.get(sizeof(OutRec) * std::max(nkeep, 1LL));
    uint8_t* tr = (uint8_t*)g_pin_tr.get(std::max(tbytes_all, 1LL));
    if (nkeep > 0) HC(hipMemcpyAsync(recs, srec.p, sizeof(OutRec) * nkeep, hipMemcpyDeviceToHost, st));
    if (tbytes_all > 0) HC(hipMemcpyAsync(tr, str.p, tbytes_all, hipMemcpyDeviceToHost, st));
    HC(hipStreamSynchronize(st));
    // srec[i].toff now holds the offset in tr
    if (mi_append_records(spec->iobuf, recs, nkeep, tr, tbytes_all, tbytes) != 0)
        throw Fail("cannot grow the overlap buffer");
    S.nrecords += nkeep;
    S.t_copy += now_s() - t2;
}

}  // namespace mi

using namespace mi;

// ---- C-ABI ------------------------------------------------------------------------------

extern "C" {

int Set_Filter_Params(int kmer, int binshift, int suppress, int hitmin)
{
    if (kmer <= 1) return 1;
    if (kmer > 31) {   // the reference's 1 << 2k is undefined from 32 on (dalign/filter.c:129-131)
        mi_set_error("k-mer length above 31 is not supported");
        return 1;
    }
    g_kmer = kmer;
    g_binshift = binshift;
    g_suppress = suppress;
    g_hitmin = hitmin;
    return 0;
}

static void* sort_kmers_impl(HITS_DB* block, int* len, const KIdx* comp_of)
{
    try {
        KIdx* x = comp_of ? index_complement(comp_of, block) : index_from_block(block);
        if (len) *len = (int)std::min<long long>(x->nk, INT32_MAX);
        return x;
    } catch (const Fail& e) {
        mi_set_error(e.what());
        return nullptr;
    } catch (const std::bad_alloc&) {
        mi_set_error("host out of memory");
        return nullptr;
    }
}

void* Sort_Kmers(HITS_DB* block, int* len)
{
    void* r = sort_kmers_impl(block, len, nullptr);
    if (r == nullptr) {
        fprintf(stderr, "Sort_Kmers: %s\n", mi_last_error());
        exit(1);
    }
    return r;
}

void* mi_block_device(HITS_DB* block)
{
    try {
        return index_from_block(block, false);
    } catch (const Fail& e) {
        mi_set_error(e.what());
        return nullptr;
    }
}

void* mi_sort_kmers_device(void* dblock, int* len)
{
    try {
        KIdx* x = index_from_device(as_index(dblock));
        if (len) *len = (int)std::min<long long>(x->nk, INT32_MAX);
        return x;
    } catch (const Fail& e) {
        mi_set_error(e.what());
        return nullptr;
    }
}

int mi_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : MI_ERR_DEVICE; }

void* mi_sort_kmers_complement(void* table, HITS_DB* cblock, int* len)
{
    try {
        return sort_kmers_impl(cblock, len, as_index(table));
    } catch (const Fail& e) {
        mi_set_error(e.what());
        return nullptr;
    }
}

void mi_free_kmer_index(void* table)
{
    KIdx* x = (KIdx*)table;
    if (x && x->magic == kIdxMagic) {
        x->magic = 0;
        delete x;
    }
}

int mi_match_filter(char* aname, HITS_DB* ablock, char* bname, HITS_DB* bblock, void* atable, void* btable, int comp,
                    Align_Spec* asettings)
{
    (void)ablock;
    (void)bblock;
    memset(&g_stats, 0, sizeof g_stats);
    const double t0 = now_s();
    const long long gc0 = g_graphs.captured, gr0 = g_graphs.replayed, gl0 = g_graphs.launches_replayed;
    int rc = 0;
    try {
        KIdx* A = as_index(atable);
        KIdx* B = as_index(btable);
        const bool self = (aname == bname);
        match_pass(A, B, self, comp ? 1 : 0, atable == btable, (SpecObj*)asettings);
    } catch (const Fail& e) {
        mi_set_error(e.what());
        rc = e.code;
    } catch (const std::bad_alloc&) {
        mi_set_error("host out of memory");
        rc = MI_ERR_DEVICE;
    }
    // the B index is consumed when it is not the A index (dalign/filter.c:2194-2202)
    if (btable != atable) mi_free_kmer_index(btable);
    g_stats.t_total = now_s() - t0;
    g_stats.graphs_captured = g_graphs.captured - gc0;
    g_stats.graphs_replayed = g_graphs.replayed - gr0;
    g_stats.launches_replayed = g_graphs.launches_replayed - gl0;
    return rc;
}

void Match_Filter(char* aname, HITS_DB* ablock, char* bname, HITS_DB* bblock, void* atable, int alen, void* btable,
                  int blen, int comp, Align_Spec* asettings)
{
    (void)alen;
    (void)blen;
    if (mi_match_filter(aname, ablock, bname, bblock, atable, btable, comp, asettings) != 0) {
        fprintf(stderr, "Match_Filter: %s\n", mi_last_error());
        exit(1);
    }
    if (VERBOSE) {
        const MiFilterStats& s = g_stats;
        printf("\n     %lld %d-mers, %lld seed hits, %lld confirmed hits (%.3f s)\n", s.nhits, g_kmer, s.naligned,
               s.nrecords, s.t_total);
        fflush(stdout);
    }
}

void mi_filter_stats(MiFilterStats* out) { *out = g_stats; }

}  // extern "C"

// ---- LAsort (dalign/LAsort.c) ------------------------------------------------------
// The record order of LAsort: (aread, bread, abpos), equal keys in input
// order (SORT_OVL, dalign/LAsort.c:60-86, under glibc's qsort, a stable merge
// sort).  The host walks the records once (their offsets depend on every
// earlier trace length) and builds the keys; the GPU sorts them with the
// stable radix sort and gathers the records into the output.

namespace mi {

// stage-2 keys in stage-1 order: out[j] = {ab[in[j].idx], in[j].idx}
__global__ void k_las_rekey(const RecKey* __restrict__ in, const unsigned long long* __restrict__ ab, long long n,
                            RecKey* __restrict__ out)
{
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const long long i = in[j].idx;
    out[j] = RecKey{ab[i], i};
}

__global__ void k_las_sizes(const RecKey* __restrict__ sorted, const long long* __restrict__ size, long long n,
                            u64* __restrict__ out)
{
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    out[j] = (u64)size[sorted[j].idx];
}

// one wave per output record: its bytes from the input body (records start at
// any byte offset: 1-byte trace values make odd lengths)
__global__ __launch_bounds__(256) void k_las_gather(const uint8_t* __restrict__ body, const long long* __restrict__ off,
                                                    const long long* __restrict__ size, const RecKey* __restrict__ sorted,
                                                    const u64* __restrict__ dpos, long long n, uint8_t* __restrict__ out)
{
    const long long j = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= n) return;
    const int l = threadIdx.x & 63;
    const long long i = sorted[j].idx;
    const uint8_t* src = body + off[i];
    uint8_t* dst = out + dpos[j];
    const long long m = size[i];
    for (long long b = l; b < m; b += 64) dst[b] = src[b];
}

}  // namespace mi

// with_comp: LAmerge's order (aread, bread, COMP, abpos; SORT_OVL,
// utils/LAmergeUtils.c:50-80), else LAsort's (aread, bread, abpos)
static int las_order(const void* body, long long size, long long novl, int tspace, bool with_comp, void* out,
                     long long* out_size)
{
    try {
        if (novl < 0 || size < 0 || (novl > 0 && (!body || !out)))
            throw Fail("mi_las_sort: bad arguments", MI_ERR_ARG);
        const int tb = tspace <= 125 ? 1 : 2;   // TRACE_XOVR (dalign/align.h)
        const uint8_t* b = (const uint8_t*)body;
        std::vector<long long> off(novl), sz(novl);
        std::vector<int> ar(novl), br(novl), ap(novl);
        std::vector<unsigned char> cp(novl);
        long long o = 0;
        int amin = INT32_MAX, amax = INT32_MIN, bmin = INT32_MAX, bmax = INT32_MIN, pmin = INT32_MAX, pmax = INT32_MIN;
        for (long long i = 0; i < novl; i++) {
            if (o + 40 > size) throw Fail("mi_las_sort: records run past the end of the file", MI_ERR_ARG);
            int tlen;
            memcpy(&tlen, b + o, 4);
            memcpy(&ap[i], b + o + 8, 4);
            memcpy(&ar[i], b + o + 28, 4);
            memcpy(&br[i], b + o + 32, 4);
            cp[i] = with_comp ? (b[o + 24] & 1) : 0;   // COMP flag, bit 0 of flags (lib/oflags.h)
            const long long n = 40 + (long long)tlen * tb;
            if (tlen < 0 || o + n > size) throw Fail("mi_las_sort: records run past the end of the file", MI_ERR_ARG);
            off[i] = o;
            sz[i] = n;
            o += n;
            amin = std::min(amin, ar[i]);
            amax = std::max(amax, ar[i]);
            bmin = std::min(bmin, br[i]);
            bmax = std::max(bmax, br[i]);
            pmin = std::min(pmin, ap[i]);
            pmax = std::max(pmax, ap[i]);
        }
        if (out_size) *out_size = o;
        if (novl == 0) return 0;
        if (novl == 1) {
            memcpy(out, body, o);
            return 0;
        }
        const int ra = bits_for((long long)amax - amin), rb = bits_for((long long)bmax - bmin),
                  pb = bits_for((long long)pmax - pmin) + (with_comp ? 1 : 0);   // COMP above abpos
        const bool one = ra + rb + pb <= 64;
        std::vector<RecKey> keys(novl);
        std::vector<unsigned long long> ab;
        for (long long i = 0; i < novl; i++) {
            const unsigned long long a = (unsigned long long)((long long)ar[i] - amin),
                                     c = (unsigned long long)((long long)br[i] - bmin),
                                     p = ((unsigned long long)cp[i] << (pb - 1)) |
                                         (unsigned long long)((long long)ap[i] - pmin);
            keys[i] = RecKey{one ? (a << (rb + pb)) | (c << pb) | p : p, i};
        }
        if (!one) {
            ab.resize(novl);
            for (long long i = 0; i < novl; i++)
                ab[i] = ((unsigned long long)((long long)ar[i] - amin) << rb) | (unsigned long long)((long long)br[i] - bmin);
        }
        hipStream_t st = fstream();
        DBuf<uint8_t> dbody, dout;
        DBuf<long long> doff, dsz;
        DBuf<RecKey> k1, k2;
        DBuf<unsigned long long> dab;
        DBuf<u64> s1, pos;
        dbody.ensure(o);
        dout.ensure(o);
        doff.ensure(novl);
        dsz.ensure(novl);
        k1.ensure(novl);
        k2.ensure(novl);
        s1.ensure(novl);
        pos.ensure(novl + 1);
        HC(hipMemcpyAsync(dbody.p, body, o, hipMemcpyHostToDevice, st));
        HC(hipMemcpyAsync(doff.p, off.data(), sizeof(long long) * novl, hipMemcpyHostToDevice, st));
        HC(hipMemcpyAsync(dsz.p, sz.data(), sizeof(long long) * novl, hipMemcpyHostToDevice, st));
        HC(hipMemcpyAsync(k1.p, keys.data(), sizeof(RecKey) * novl, hipMemcpyHostToDevice, st));
        RecKey* sorted = radix_sort(k1.p, k2.p, novl, 0, one ? ra + rb + pb : pb, RecKeyOf{});
        if (!one) {   // then (aread, bread), stable over the abpos order
            dab.ensure(novl);
            HC(hipMemcpyAsync(dab.p, ab.data(), sizeof(unsigned long long) * novl, hipMemcpyHostToDevice, st));
            RecKey* other = sorted == k1.p ? k2.p : k1.p;
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_las_rekey), dim3(nblk(novl)), dim3(256), 0, st, sorted, dab.p, novl, other);
            HC(hipGetLastError());
            sorted = radix_sort(other, sorted, novl, 0, ra + rb, RecKeyOf{});
        }
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_las_sizes), dim3(nblk(novl)), dim3(256), 0, st, sorted, dsz.p, novl, s1.p);
        HC(hipGetLastError());
        g_scan.run(s1.p, pos.p, novl);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_las_gather), dim3((unsigned)((novl + 3) / 4)), dim3(256), 0, st, dbody.p, doff.p,
                           dsz.p, sorted, pos.p, novl, dout.p);
        HC(hipGetLastError());
        HC(hipMemcpyAsync(out, dout.p, o, hipMemcpyDeviceToHost, st));
        HC(hipStreamSynchronize(st));
        return 0;
    } catch (const Fail& e) {
        mi_set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        mi_set_error("mi_las_sort: host out of memory");
        return MI_ERR_DEVICE;
    }
}

extern "C" int mi_las_sort(const void* body, long long size, long long novl, int tspace, void* out, long long* out_size)
{
    return las_order(body, size, novl, tspace, false, out, out_size);
}

extern "C" int mi_las_merge(const void* body, long long size, long long novl, int tspace, void* out, long long* out_size)
{
    return las_order(body, size, novl, tspace, true, out, out_size);
}
