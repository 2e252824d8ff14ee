// dalign/align.h drop-in entry points of libmarvl_align (host side).
// Local_Alignment forwards to the GPU batch path with one task; the record
// helpers are plain host code on the byte-compatible .las layout.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/marvl_align.h"
#include "mi_common.h"

namespace {

using mi::SpecObj;

struct WorkObj {
    std::vector<uint64_t> packed;
    std::vector<uint16_t> traces;
    std::vector<int> itrace;   // Compute_Trace_* output
    Path bpath;
};

[[noreturn]] void die(const char* what)
{
    fprintf(stderr, "libmarvl_align: %s (%s)\n", what, mi_last_error());
    exit(1);
}

// 2-bit pack `len` bases (values 0..3) into words starting at out[at]
void pack_into(std::vector<uint64_t>& out, long long at, const char* s, int len)
{
    for (int i = 0; i < len; i++) out[at + (i >> 5)] |= (uint64_t)(s[i] & 3) << (2 * (i & 31));
}

}  // namespace

extern "C" {

Align_Spec* New_Align_Spec(double ave_corr, int trace_space, float* freq, int nthreads, int symmetric,
                           int only_identity, int no_trace_points)
{
    SpecObj* s = new SpecObj;
    mi::spec_build(&s->hs, ave_corr, trace_space, freq);
    s->prm.ave_corr = ave_corr;
    s->prm.trace_space = trace_space;
    for (int i = 0; i < 4; i++) s->prm.freq[i] = freq[i];
    s->nthreads = nthreads > 0 ? nthreads : 1;
    s->symmetric = symmetric;
    s->only_identity = only_identity;
    s->no_trace = no_trace_points;
    // per-thread overlap buffers (dalign/align.c:260-276)
    s->iobuf = (Overlap_IO_Buffer*)malloc(sizeof(Overlap_IO_Buffer) * s->nthreads);
    for (int i = 0; i < s->nthreads; i++) {
        Overlap_IO_Buffer* ob = CreateOverlapBuffer(s->nthreads, trace_space <= TRACE_XOVR ? 1 : 2, no_trace_points);
        if (ob == nullptr) exit(1);
        s->iobuf[i] = *ob;
        free(ob);
    }
    return (Align_Spec*)s;
}

void Free_Align_Spec(Align_Spec* spec)
{
    SpecObj* s = (SpecObj*)spec;
    mi::wait_overlap_writes();
    for (int i = 0; i < s->nthreads; i++) {
        free(s->iobuf[i].ovls);
        free(s->iobuf[i].trace);
        if (s->spare) {
            free(s->spare[i].ovls);
            free(s->spare[i].trace);
        }
    }
    free(s->iobuf);
    free(s->spare);
    delete s;
}
int Trace_Spacing(Align_Spec* spec) { return ((SpecObj*)spec)->hs.tspace; }
double Average_Correlation(Align_Spec* spec) { return ((SpecObj*)spec)->hs.ave_corr; }
float* Base_Frequencies(Align_Spec* spec) { return ((SpecObj*)spec)->hs.freq; }

Work_Data* New_Work_Data(void) { return (Work_Data*)new WorkObj; }
void Free_Work_Data(Work_Data* work) { delete (WorkObj*)work; }

Path* Local_Alignment(Alignment* align, Work_Data* work, Align_Spec* spec, int low, int hgh, int anti, int lbord,
                      int hbord)
{
    SpecObj* s = (SpecObj*)spec;
    WorkObj* w = (WorkObj*)work;
    const int alen = align->alen, blen = align->blen;
    const long long aw = (alen + 31) / 32, bw = (blen + 31) / 32;
    const bool selfie = (align->aseq == align->bseq);
    // layout: [guard][A words][guard guard][B words][guard guard]
    const long long a0 = 1, b0 = selfie ? 1 : a0 + aw + 2;
    w->packed.assign((selfie ? a0 + aw : b0 + bw) + 2, 0);
    pack_into(w->packed, a0, align->aseq, alen);
    if (!selfie) pack_into(w->packed, b0, align->bseq, blen);
    MiLaTask t;
    t.aword = a0;
    t.bword = b0;
    t.alen = alen;
    t.blen = blen;
    t.comp = COMP(align->flags) ? 1 : 0;
    t.low = low;
    t.hgh = hgh;
    t.anti = anti;
    t.lbord = lbord;
    t.hbord = hbord;
    t.selfie = selfie ? 1 : 0;
    const int mx = alen > blen ? alen : blen;
    const long long cap = 4LL * (mx / s->hs.tspace + 2) + 16;
    w->traces.assign(cap, 0);
    MiLaResult r;
    int rc = mi_la_batch_host(&s->prm, w->packed.data(), (long long)w->packed.size(), &t, 1, &r, w->traces.data(), cap);
    if (rc != 0) die("Local_Alignment failed");
    if (r.status != 0) {
        mi::mi_set_error("wave reached a state the reference leaves undefined");
        die("Local_Alignment");
    }
    Path* ap = align->path;
    ap->abpos = r.abpos;
    ap->bbpos = r.bbpos;
    ap->aepos = r.aepos;
    ap->bepos = r.bepos;
    ap->diffs = r.diffs;
    ap->tlen = r.atlen;
    ap->trace = w->traces.data() + r.atrace;
    Path* bp = &w->bpath;
    bp->abpos = r.b_abpos;
    bp->bbpos = r.b_bbpos;
    bp->aepos = r.b_aepos;
    bp->bepos = r.b_bepos;
    bp->diffs = r.diffs;
    bp->tlen = r.btlen;
    bp->trace = w->traces.data() + r.btrace;
    return bp;
}

int Compute_Trace_PTS(Alignment* align, Work_Data* work, int trace_spacing, int mode, int debug)
{
    (void)debug;
    WorkObj* w = (WorkObj*)work;
    Path* p = align->path;
    const int alen = align->alen, blen = align->blen;
    const long long aw = (alen + 31) / 32, bw = (blen + 31) / 32;
    const bool selfie = (align->aseq == align->bseq);
    const long long a0 = 1, b0 = selfie ? 1 : a0 + aw + 2;
    w->packed.assign((selfie ? a0 + aw : b0 + bw) + 2, 0);
    pack_into(w->packed, a0, align->aseq, alen);
    if (!selfie) pack_into(w->packed, b0, align->bseq, blen);
    MiTraceTask t;
    t.aword = a0;
    t.bword = b0;
    t.alen = alen;
    t.blen = blen;
    t.abpos = p->abpos;
    t.bbpos = p->bbpos;
    t.aepos = p->aepos;
    t.bepos = p->bepos;
    t.tlen = p->tlen;
    t.points = 0;
    t.selfie = selfie ? 1 : 0;
    const long long cap = (long long)(p->aepos - p->abpos) + (p->bepos - p->bbpos) + 16;
    std::vector<int> out(cap);
    MiTraceResult r;
    const int rc = mi_trace_pts_host(trace_spacing, mode, w->packed.data(), (long long)w->packed.size(),
                                     (const uint16_t*)p->trace, p->tlen, &t, 1, &r, out.data(), cap);
    if (rc != 0) die("Compute_Trace_PTS failed");
    if (r.status != 0) {
        mi::mi_set_error("self comparison starts on diagonal 0");
        die("Compute_Trace_PTS");
    }
    w->itrace.assign(out.begin(), out.begin() + r.ntrace);
    if (w->itrace.empty()) w->itrace.push_back(0);
    p->trace = w->itrace.data();
    p->tlen = r.ntrace;
    p->diffs = r.diffs;
    return 0;
}

// ---- .las record helpers (dalign/align.c:3325-3449) -------------------------
// A record on disk is the Overlap struct without its leading trace pointer:
// 40 bytes = tlen, diffs, abpos, bbpos, aepos, bepos, flags, aread, bread, pad.

static const size_t kRecBytes = sizeof(Overlap) - sizeof(void*);

int Read_Overlap(FILE* input, Overlap* ovl)
{
    return fread(((char*)ovl) + sizeof(void*), kRecBytes, 1, input) != 1;
}

int Read_Trace(FILE* input, Overlap* ovl, int tbytes)
{
    if (tbytes > 0 && ovl->path.tlen > 0)
        if (fread(ovl->path.trace, (size_t)tbytes * ovl->path.tlen, 1, input) != 1) return 1;
    return 0;
}

void Write_Overlap(FILE* output, Overlap* ovl, int tbytes)
{
    fwrite(((char*)ovl) + sizeof(void*), kRecBytes, 1, output);
    if (ovl->path.trace != NULL) fwrite(ovl->path.trace, tbytes, ovl->path.tlen, output);
}

void Compress_TraceTo8(Overlap* ovl)
{
    uint16_t* t16 = (uint16_t*)ovl->path.trace;
    uint8_t* t8 = (uint8_t*)ovl->path.trace;
    for (int j = 0; j < ovl->path.tlen; j++) t8[j] = (uint8_t)t16[j];
}

void Decompress_TraceTo16(Overlap* ovl)
{
    uint16_t* t16 = (uint16_t*)ovl->path.trace;
    uint8_t* t8 = (uint8_t*)ovl->path.trace;
    for (int j = ovl->path.tlen - 1; j >= 0; j--) t16[j] = t8[j];
}

int Check_Trace_Points(Overlap* ovl, int tspace, int verbose, char* fname)
{
    if (((ovl->path.aepos - 1) / tspace - ovl->path.abpos / tspace) * 2 != ovl->path.tlen - 2) {
        if (verbose) fprintf(stderr, "  %s: Wrong number of trace points\n", fname);
        return 1;
    }
    int p = ovl->path.bbpos;
    if (tspace <= TRACE_XOVR) {
        uint8_t* t = (uint8_t*)ovl->path.trace;
        for (int i = 1; i < ovl->path.tlen; i += 2) p += t[i];
    } else {
        uint16_t* t = (uint16_t*)ovl->path.trace;
        for (int i = 1; i < ovl->path.tlen; i += 2) p += t[i];
    }
    if (p != ovl->path.bepos) {
        if (verbose) fprintf(stderr, "  %s: Trace point sum != aligned interval\n", fname);
        return 1;
    }
    return 0;
}

}  // extern "C"

// Print_Overlap (dalign/align.h:313): the record's reads, intervals and
// differences, then its trace as "diffs/B position" pairs, the B position
// running from bbpos, a line break before every pair whose value index is
// a multiple of 10.
template <typename T>
static void print_trace(FILE* out, const Overlap* ovl, const T* t, int indent)
{
    int p = ovl->path.bbpos + t[1];
    fprintf(out, "%*sTrace: %3d/%5d", indent, "", (int)t[0], p);
    for (int i = 3; i < ovl->path.tlen; i += 2) {
        if (i % 10 == 0) fprintf(out, "\n%*s", indent + 6, "");
        p += t[i];
        fprintf(out, " %3d/%5d", (int)t[i - 1], p);
    }
    fprintf(out, "\n");
}

extern "C" {

void Print_Overlap(FILE* output, Overlap* ovl, int tbytes, int indent)
{
    fprintf(output, "%*s%d vs. ", indent, "", ovl->aread);
    if (COMP(ovl->flags)) fprintf(output, "c(%d)\n", ovl->bread);
    else fprintf(output, "%d\n", ovl->bread);
    fprintf(output, "%*s  [%d,%d] vs [%d,%d] w. %d diffs\n", indent, "", ovl->path.abpos, ovl->path.aepos,
            ovl->path.bbpos, ovl->path.bepos, ovl->path.diffs);
    if (ovl->path.trace == NULL) return;
    if (tbytes == 1) print_trace(output, ovl, (const uint8_t*)ovl->path.trace, indent);
    else print_trace(output, ovl, (const uint16_t*)ovl->path.trace, indent);
}

// Complement_Seq (dalign/align.h:119): reverse complement of bases 0..3 in
// place (an involution).
void Complement_Seq(char* a, int n)
{
    for (int i = 0, j = n - 1; i <= j; i++, j--) {
        const char x = a[i], y = a[j];
        a[i] = (char)(3 - y);
        a[j] = (char)(3 - x);
    }
}

// Flip_Alignment (dalign/align.h:230-234): the same alignment with the roles
// of A and B exchanged.  Without COMP the intervals swap and (full) every
// indel position of an int trace changes sign.  With COMP the intervals are
// mirrored into the other read's complemented coordinates and (full) an
// indel position p < 0 (a gap in A at B position -p) becomes alen + 2 + p,
// a position p > 0 becomes p - (blen + 2), the trace read backwards.
void Flip_Alignment(Alignment* align, int full)
{
    Path* path = align->path;
    const int alen = align->alen, blen = align->blen;
    int* trace = (int*)path->trace;
    const int tlen = path->tlen;
    if (COMP(align->flags)) {
        const int ab = path->abpos, ae = path->aepos, bb = path->bbpos, be = path->bepos;
        path->abpos = blen - be;
        path->aepos = blen - bb;
        path->bbpos = alen - ae;
        path->bepos = alen - ab;
        if (full) {
            for (int i = 0; i < tlen; i++) trace[i] = trace[i] < 0 ? alen + 2 + trace[i] : trace[i] - (blen + 2);
            for (int i = 0, j = tlen - 1; i < j; i++, j--) std::swap(trace[i], trace[j]);
        }
    } else {
        std::swap(path->abpos, path->bbpos);
        std::swap(path->aepos, path->bepos);
        if (full)
            for (int i = 0; i < tlen; i++) trace[i] = -trace[i];
    }
    std::swap(align->aseq, align->bseq);
    align->alen = blen;
    align->blen = alen;
}

}  // extern "C"
