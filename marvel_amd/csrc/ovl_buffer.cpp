// Overlap buffers of an Align_Spec and the .las writer (host side).
// Behaviour of dalign/align.c:5630-6056: records are buffered per thread,
// then Write_Overlap_Buffer sorts all of them stably by (aread, bread,
// COMP, abpos) and writes one .las file, or two split at lastRead, into
// the run's RRR_BBBBB/ directories.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <unistd.h>

#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/marvl_align.h"
#include "mi_common.h"
#include "records.h"

using mi::SpecObj;

namespace {

constexpr int kBufferOvls = 500000;   // initial records over all threads (dalign/align.c:5651)

std::string run_dir(int run_id, int block)
{
    char b[64];
    snprintf(b, sizeof b, "%03d_%05d", run_id, block);
    return b;
}

// "ROOT.<id>" -> (ROOT, id); no dot: (name, 0)
int split_block_name(const char* name, std::string& root)
{
    const char* dot = strrchr(name, '.');
    if (dot == nullptr) {
        root = name;
        return 0;
    }
    root.assign(name, dot - name);
    return atoi(dot + 1);
}

// Stable merge of the sorted ranges a[0, mid) and a[mid, end) in place (tmp:
// scratch of end entries), split by merge path over up to 8 threads.  On
// equal keys the first range goes first.
template <class T, class Less>
void parallel_merge(T* a, size_t mid, size_t end, T* tmp, Less less)
{
    const size_t n1 = mid, n2 = end - mid, n = end;
    const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, n / 65536));
    // co-rank: the number i of first-range entries among the first k outputs
    auto corank = [&](size_t k) {
        size_t lo = k > n2 ? k - n2 : 0, hi = std::min(k, n1);
        while (lo < hi) {
            const size_t i = (lo + hi) / 2, j = k - i;
            // a[i] goes before b[j - 1] unless b[j - 1] < a[i]
            if (j > 0 && i < n1 && !less(a[mid + j - 1], a[i])) lo = i + 1;
            else hi = i;
        }
        return lo;
    };
    std::vector<size_t> ci(nt + 1);
    for (int t = 0; t <= nt; t++) ci[t] = corank(n * t / nt);
    auto part = [&](int t) {
        const size_t k0 = n * t / nt, k1 = n * (t + 1) / nt;
        const size_t i0 = ci[t], i1 = ci[t + 1];
        std::merge(a + i0, a + i1, a + mid + (k0 - i0), a + mid + (k1 - i1), tmp + k0, less);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(part, t);
    part(0);
    for (auto& x : th) x.join();
    std::copy(tmp, tmp + n, a);
}

// bytes of one record on disk
inline size_t rec_bytes(const Overlap* v, int tbytes) { return 40 + (v->path.trace ? (size_t)v->path.tlen * tbytes : 0); }

// Grow-only scratch kept by the calling thread across calls (the output of
// a block pair is rebuilt every call; fresh multi-MB buffers would be paid
// again in page faults each time).
template <class T>
T* scratch(std::vector<T>& v, size_t n)
{
    if (v.size() < n) v.resize(n + n / 8);
    return v.data();
}

bool write_las(const std::string& path, const std::vector<const Overlap*>& recs, size_t lo, size_t hi, int tspace,
               int tbytes)
{
    const int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0666);
    if (fd < 0) {
        fprintf(stderr, "[ERROR] - Write_Overlap_Buffer: Cannot open file %s for writing\n", path.c_str());
        exit(1);
    }
    // up to 8 threads each serialise a contiguous range of records into their
    // slice of one buffer and write that slice at its file offset
    const size_t n = hi - lo;
    const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, n / 50000));
    std::vector<size_t> cut(nt + 1), start(nt + 1), part(nt, 0);
    for (int t = 0; t <= nt; t++) cut[t] = lo + n * t / nt;
    auto run = [&](auto&& fn) {
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(fn, t);
        fn(0);
        for (auto& x : th) x.join();
    };
    run([&](int t) {
        size_t b = 0;
        for (size_t j = cut[t]; j < cut[t + 1]; j++) b += rec_bytes(recs[j], tbytes);
        part[t] = b;
    });
    start[0] = 12;
    for (int t = 0; t < nt; t++) start[t + 1] = start[t] + part[t];
    const size_t bytes = start[nt];
    static thread_local std::vector<char> obuf;
    char* out = scratch(obuf, bytes);
    const int64_t nrec = (int64_t)n;
    memcpy(out, &nrec, 8);
    memcpy(out + 8, &tspace, 4);
    std::vector<char> fail(nt, 0);
    run([&](int t) {
        size_t o = start[t];
        for (size_t j = cut[t]; j < cut[t + 1]; j++) {
            const Overlap* v = recs[j];
            const int32_t rec[10] = {v->path.tlen,  v->path.diffs,      v->path.abpos, v->path.bbpos, v->path.aepos,
                                     v->path.bepos, (int32_t)v->flags, v->aread,      v->bread,      0};
            memcpy(out + o, rec, 40);
            o += 40;
            if (v->path.trace) {
                const size_t tb = (size_t)v->path.tlen * tbytes;
                memcpy(out + o, v->path.trace, tb);
                o += tb;
            }
        }
        const size_t b0 = t == 0 ? 0 : start[t];   // slice 0 also carries the header
        size_t done = 0, len = start[t + 1] - b0;
        while (done < len) {
            const ssize_t w = pwrite(fd, out + b0 + done, len - done, (off_t)(b0 + done));
            if (w <= 0) {
                fail[t] = 1;
                break;
            }
            done += (size_t)w;
        }
    });
    const bool ok = close(fd) == 0 && std::find(fail.begin(), fail.end(), 1) == fail.end();
    return ok;
}

}  // namespace

namespace {

// One background thread writes .las files: Write_Overlap_Buffer hands it a
// snapshot of the buffer headers and returns; the next write,
// Reset_Overlap_Buffer's swap, appends to buffers it is still reading,
// Free_Align_Spec, mi_flush_overlap_writes and process exit wait for it.
// It lives until the process ends (never destroyed, so exit cannot tear it
// down under a running job).
struct Writer {
    std::mutex m;
    std::condition_variable cv;
    std::function<bool()> job;
    bool busy = false, failed = false;
    std::vector<const Overlap*> reading;   // ovls arrays the running job reads
    Writer()
    {
        std::thread([this] {
            for (;;) {
                std::function<bool()> f;
                {
                    std::unique_lock<std::mutex> l(m);
                    cv.wait(l, [this] { return job != nullptr; });
                    f = std::move(job);
                    job = nullptr;
                }
                const bool ok = f();
                std::lock_guard<std::mutex> l(m);
                failed = failed || !ok;
                busy = false;
                reading.clear();
                cv.notify_all();
            }
        }).detach();
    }
    // waits until idle; true when every write so far succeeded
    bool wait()
    {
        std::unique_lock<std::mutex> l(m);
        cv.wait(l, [this] { return !busy; });
        const bool ok = !failed;
        failed = false;
        return ok;
    }
    void submit(std::function<bool()> f, std::vector<const Overlap*> arrays)
    {
        std::lock_guard<std::mutex> l(m);
        job = std::move(f);
        busy = true;
        reading = std::move(arrays);
        cv.notify_all();
    }
    bool reads(const Overlap* a)
    {
        std::lock_guard<std::mutex> l(m);
        return busy && std::find(reading.begin(), reading.end(), a) != reading.end();
    }
};

Writer& writer()
{
    static Writer* w = [] {
        Writer* x = new Writer;
        atexit([] { mi::wait_overlap_writes(); });
        return x;
    }();
    return *w;
}

// an append to a buffer the writer is reading first waits for the write
void before_append(const Overlap_IO_Buffer* b)
{
    Writer& w = writer();
    if (w.reads(b->ovls)) mi::wait_overlap_writes();
}

}  // namespace

namespace mi {

void wait_overlap_writes()
{
    if (!writer().wait()) {
        fprintf(stderr, "[ERROR] - Write_Overlap_Buffer: write failed\n");
        _exit(1);
    }
}

int mi_append_records(Overlap_IO_Buffer* b, const OutRec* recs, long long n, const uint8_t* tr, long long trbytes,
                      int tbytes)
{
    if (n <= 0) return 0;
    before_append(b);
    if (b->otop + n > b->omax) {
        const long long nmax = (long long)((b->otop + n) * 1.2) + 1000;
        if (nmax > INT32_MAX) return 1;
        Overlap* o = (Overlap*)realloc(b->ovls, sizeof(Overlap) * nmax);
        if (o == nullptr) return 1;
        b->ovls = o;
        b->omax = (int)nmax;
    }
    const bool with_trace = b->no_trace == 0;
    const uint64_t need = with_trace ? (uint64_t)trbytes : 0;
    if (with_trace && b->ttop + need > b->tmax) {
        const uint64_t nmax = (uint64_t)((b->ttop + need) * 1.2) + 1000;
        char* t = (char*)realloc(b->trace, nmax);
        if (t == nullptr) return 1;
        if (t != b->trace) {   // keep the buffered records' trace pointers valid
            for (int i = 0; i < b->otop; i++)
                if (b->ovls[i].path.trace != nullptr)
                    b->ovls[i].path.trace = t + ((char*)b->ovls[i].path.trace - (char*)b->trace);
        }
        b->trace = t;
        b->tmax = nmax;
    }
    char* tbase = (char*)b->trace + b->ttop;
    if (with_trace && need) memcpy(tbase, tr, need);
    Overlap* o = b->ovls + b->otop;
    for (long long i = 0; i < n; i++) {
        const OutRec& r = recs[i];
        o[i].path.diffs = r.diffs;
        o[i].path.abpos = r.abpos;
        o[i].path.bbpos = r.bbpos;
        o[i].path.aepos = r.aepos;
        o[i].path.bepos = r.bepos;
        o[i].flags = r.flags;
        o[i].aread = r.aread;
        o[i].bread = r.bread;
        if (with_trace) {
            o[i].path.tlen = r.tlen;
            o[i].path.trace = tbase + r.toff;
        } else {
            o[i].path.tlen = 0;
            o[i].path.trace = nullptr;
        }
    }
    (void)tbytes;
    b->otop += (int)n;
    b->ttop += need;
    return 0;
}

}  // namespace mi

extern "C" {

Overlap_IO_Buffer* CreateOverlapBuffer(int nthreads, int tbytes, int no_trace)
{
    Overlap_IO_Buffer* b = (Overlap_IO_Buffer*)calloc(1, sizeof(Overlap_IO_Buffer));
    if (b == nullptr) return nullptr;
    b->omax = kBufferOvls / (nthreads > 0 ? nthreads : 1) + 1;
    b->ovls = (Overlap*)malloc(sizeof(Overlap) * b->omax);
    b->no_trace = no_trace;
    if (b->ovls == nullptr) {
        free(b);
        return nullptr;
    }
    if (no_trace) return b;
    if (tbytes < 1 || tbytes > 2) {
        fprintf(stderr, "[ERROR] - Unsupported size of trace: %d!\n", tbytes);
        free(b->ovls);
        free(b);
        return nullptr;
    }
    b->tbytes = tbytes;
    b->tmax = (uint64_t)b->omax * 120 * tbytes;   // bytes
    b->trace = malloc(b->tmax);
    if (b->trace == nullptr) {
        free(b->ovls);
        free(b);
        return nullptr;
    }
    return b;
}

// Appends a record; its trace (tlen values of tbytes bytes) is copied.
// Trace pointers of buffered records are kept valid across growth.
int AddOverlapToBuffer(Overlap_IO_Buffer* b, Overlap* ovl, int tbytes)
{
    if (b == nullptr) return 1;
    before_append(b);
    if (b->otop == b->omax) {
        const int nmax = (int)(b->omax * 1.2) + 1000;
        Overlap* o = (Overlap*)realloc(b->ovls, sizeof(Overlap) * nmax);
        if (o == nullptr) return 1;
        b->ovls = o;
        b->omax = nmax;
    }
    const bool with_trace = ovl->path.trace != nullptr && b->no_trace == 0;
    const uint64_t tb = with_trace ? (uint64_t)tbytes * ovl->path.tlen : 0;
    if (with_trace && b->ttop + tb > b->tmax) {
        const uint64_t nmax = (uint64_t)((b->ttop + tb) * 1.2) + 1000;
        char* t = (char*)realloc(b->trace, nmax);
        if (t == nullptr) return 1;
        if (t != b->trace) {   // keep the buffered records' trace pointers valid
            for (int i = 0; i < b->otop; i++)
                if (b->ovls[i].path.trace != nullptr)
                    b->ovls[i].path.trace = t + ((char*)b->ovls[i].path.trace - (char*)b->trace);
        }
        b->trace = t;
        b->tmax = nmax;
    }
    Overlap* o = b->ovls + b->otop;
    o->aread = ovl->aread;
    o->bread = ovl->bread;
    o->flags = ovl->flags;
    o->path.abpos = ovl->path.abpos;
    o->path.aepos = ovl->path.aepos;
    o->path.bbpos = ovl->path.bbpos;
    o->path.bepos = ovl->path.bepos;
    o->path.diffs = ovl->path.diffs;
    if (!with_trace) {
        o->path.trace = nullptr;
        o->path.tlen = 0;
    } else {
        o->path.tlen = ovl->path.tlen;
        memcpy((char*)b->trace + b->ttop, ovl->path.trace, tb);
        o->path.trace = (char*)b->trace + b->ttop;
    }
    b->otop++;
    b->ttop += tb;
    return 0;
}

Overlap_IO_Buffer* OVL_IO_Buffer(Align_Spec* espec) { return ((SpecObj*)espec)->iobuf; }
int Num_Threads(Align_Spec* espec) { return ((SpecObj*)espec)->nthreads; }
int Only_Identity(Align_Spec* espec) { return ((SpecObj*)espec)->only_identity; }
int Symmetric(Align_Spec* espec) { return ((SpecObj*)espec)->symmetric; }

}  // extern "C"

namespace {

// Sorts the records of buffers buf[0 .. nthreads) and writes the .las
// file(s) of the block pair; false when a write failed.
bool write_pair(const Overlap_IO_Buffer* buf, int nthreads, int only_identity, int symmetric, int tspace, int RUN_ID,
                const std::string& ablock, const std::string& bblock, int lastRead)
{
    // sort keys of all buffered records, thread buffers in order (the stable
    // order keeps that order for equal keys)
    struct Key {
        int aread, bread, comp, abpos;
        const Overlap* o;
    };
    auto less = [](const Key& l, const Key& r) {
        if (l.aread != r.aread) return l.aread < r.aread;
        if (l.bread != r.bread) return l.bread < r.bread;
        if (l.comp != r.comp) return l.comp < r.comp;
        return l.abpos < r.abpos;
    };
    static thread_local std::vector<Key> keys, tmp;
    keys.clear();
    {
        size_t n = 0;
        for (int i = 0; i < nthreads; i++) n += buf[i].otop;
        keys.reserve(n);
        for (int i = 0; i < nthreads; i++)
            for (int j = 0; j < buf[i].otop; j++) {
                const Overlap* o = buf[i].ovls + j;
                if (!only_identity || o->aread == o->bread)
                    keys.push_back(Key{o->aread, o->bread, (int)COMP(o->flags), o->path.abpos, o});
            }
    }
    // the buffers usually hold a few already sorted runs (one per filter
    // pass): merge them stably; anything else gets a full stable sort
    std::vector<size_t> runs{0};
    for (size_t i = 1; i < keys.size(); i++)
        if (less(keys[i], keys[i - 1])) runs.push_back(i);
    if (runs.size() <= 16) {
        if (tmp.size() < keys.size()) tmp.resize(keys.size() + keys.size() / 8);
        for (size_t r = 1; r < runs.size(); r++) {
            const size_t end = (r + 1 < runs.size()) ? runs[r + 1] : keys.size();
            parallel_merge(keys.data(), runs[r], end, tmp.data(), less);
        }
    } else
        std::stable_sort(keys.begin(), keys.end(), less);
    static thread_local std::vector<const Overlap*> all;
    all.resize(keys.size());
    for (size_t i = 0; i < keys.size(); i++) all[i] = keys[i].o;
    std::string aroot, broot;
    const int aid = split_block_name(ablock.c_str(), aroot);
    const int bid = split_block_name(bblock.c_str(), broot);
    const int tbytes = buf[0].tbytes;
    auto name = [&](const std::string& r1, int i1, const std::string& r2, int i2) {
        std::string p;
        if (i1 > 0) p = run_dir(RUN_ID, i1) + "/" + r1 + "." + std::to_string(i1) + "." + r2;
        else p = r1 + "." + r2;
        if (i2 > 0) p += "." + std::to_string(i2);
        return p + ".las";
    };
    if (ablock == bblock || symmetric == 0)
        return write_las(name(aroot, aid, broot, bid), all, 0, all.size(), tspace, tbytes);
    std::string p1 = name(aroot, aid, broot, bid), p2 = name(broot, bid, aroot, aid);
    if (bid < aid) std::swap(p1, p2);
    size_t j = 0;
    while (j < all.size() && all[j]->aread <= lastRead) j++;
    return write_las(p1, all, 0, j, tspace, tbytes) && write_las(p2, all, j, all.size(), tspace, tbytes);
}

}  // namespace

extern "C" {

// The records go to the background writer (a snapshot of the buffer
// headers; the buffers themselves stay untouched until it is done): the
// call returns once the previous write has finished.
void Write_Overlap_Buffer(Align_Spec* spec, int RUN_ID, char* ablock, char* bblock, int lastRead)
{
    SpecObj* s = (SpecObj*)spec;
    mi::wait_overlap_writes();
    auto snap = std::make_shared<std::vector<Overlap_IO_Buffer>>(s->iobuf, s->iobuf + s->nthreads);
    std::vector<const Overlap*> arrays;
    for (const auto& b : *snap) arrays.push_back(b.ovls);
    const int nt = s->nthreads, oi = s->only_identity, sym = s->symmetric, ts = s->hs.tspace;
    const std::string a(ablock), b(bblock);
    writer().submit([=] { return write_pair(snap->data(), nt, oi, sym, ts, RUN_ID, a, b, lastRead); },
                    std::move(arrays));
    s->written = true;
}

// Empties the buffers.  After a write the writer may still be reading them:
// they trade places with the spare set (made here the first time), which
// the previous write finished with.
void Reset_Overlap_Buffer(Align_Spec* spec)
{
    SpecObj* s = (SpecObj*)spec;
    if (s->written) {
        if (s->spare == nullptr) {
            s->spare = (Overlap_IO_Buffer*)malloc(sizeof(Overlap_IO_Buffer) * s->nthreads);
            for (int i = 0; s->spare && i < s->nthreads; i++) {
                Overlap_IO_Buffer* ob = CreateOverlapBuffer(s->nthreads, s->iobuf[i].tbytes ? s->iobuf[i].tbytes : 1,
                                                            s->iobuf[i].no_trace);
                if (ob == nullptr) {
                    fprintf(stderr, "[ERROR] - Reset_Overlap_Buffer: out of memory\n");
                    exit(1);
                }
                s->spare[i] = *ob;
                free(ob);
            }
            if (s->spare == nullptr) {
                fprintf(stderr, "[ERROR] - Reset_Overlap_Buffer: out of memory\n");
                exit(1);
            }
        }
        for (int i = 0; i < s->nthreads; i++) std::swap(s->iobuf[i], s->spare[i]);
        s->written = false;
    }
    for (int i = 0; i < s->nthreads; i++) {
        s->iobuf[i].otop = 0;
        s->iobuf[i].ttop = 0;
    }
}

// Waits for the .las files handed to the background writer; 0, or 1 when
// one of them failed to write.
int mi_flush_overlap_writes(void) { return writer().wait() ? 0 : 1; }

}  // extern "C"
