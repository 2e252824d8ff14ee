// Soft-mask tracks of daligner -m (host side).
//
// Loading follows read_DB's track part (dalign/daligner.c:373-431): each
// track comes from the MARVEL layout (.ROOT.NAME.a2 / .d2, zlib chunks;
// lib/tracks.c:19-170, lib/compression.c:64-100) or else the Dazzler one
// (.ROOT.<block>.NAME.anno / .data, else .ROOT.NAME.anno / .data read from
// the block's first read on; db/DB.c:967-1120).  Offsets are turned into
// point counts as read_DB does (byte offsets / 4), and several tracks are
// merged into their union (Merge_Size / Merge_Tracks,
// dalign/daligner.c:193-370).  complement_track restates the track part of
// complement_DB (dalign/daligner.c:519-560).
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/marvl_filter.h"
#include "mi_common.h"

using mi::mi_set_error;

namespace {

struct TrackErr {
    std::string msg;
};

// tracks this library allocated (freed by mi_release_tracks)
std::mutex g_mu;
std::unordered_set<HITS_TRACK*> g_owned;

HITS_TRACK* new_track(const char* name, std::vector<int64_t>& anno, std::vector<int32_t>& data)
{
    HITS_TRACK* t = (HITS_TRACK*)calloc(1, sizeof(HITS_TRACK));
    int64_t* a = (int64_t*)malloc(sizeof(int64_t) * anno.size());
    int32_t* d = (int32_t*)malloc(sizeof(int32_t) * (data.size() + 1));
    char* nm = strdup(name);
    if (!t || !a || !d || !nm) {
        free(t);
        free(a);
        free(d);
        free(nm);
        throw TrackErr{"host out of memory"};
    }
    memcpy(a, anno.data(), sizeof(int64_t) * anno.size());
    if (!data.empty()) memcpy(d, data.data(), sizeof(int32_t) * data.size());
    t->name = nm;
    t->size = 8;
    t->anno = a;
    t->data = d;
    std::lock_guard<std::mutex> g(g_mu);
    g_owned.insert(t);
    return t;
}

bool read_file(const std::string& path, std::vector<unsigned char>& buf)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseeko(f, 0, SEEK_END);
    const long long n = ftello(f);
    fseeko(f, 0, SEEK_SET);
    buf.resize(n);
    const bool ok = n == 0 || fread(buf.data(), n, 1, f) == 1;
    fclose(f);
    return ok;
}

// zlib chunks: u64 compressed length, then the stream of at most 8 MiB of input
bool unchunk(const unsigned char* in, long long len, unsigned char* out, long long olen)
{
    long long i = 0, o = 0;
    while (i < len) {
        uint64_t c;
        if (i + 8 > len) return false;
        memcpy(&c, in + i, 8);
        if ((long long)c > len - i - 8) return false;
        uLongf d = (uLongf)(olen - o);
        if (uncompress(out + o, &d, in + i + 8, (uLong)c) != Z_OK) return false;
        i += 8 + (long long)c;
        o += (long long)d;
    }
    return o == olen;
}

// "DIR/.ROOT": prefix of the DB's hidden files (db->path of mi_open_block is "DIR/ROOT")
std::string stem_of(const HITS_DB* db)
{
    if (!db->path) throw TrackErr{"block has no path (open it with mi_open_block)"};
    const std::string p(db->path);
    const size_t s = p.rfind('/');
    return s == std::string::npos ? "." + p : p.substr(0, s + 1) + "." + p.substr(s + 1);
}

// MARVEL layout: 64-byte header (u16 version, u16 size, u32 pad, u64 len,
// clen, cdlen, 4 reserved).  Returns false when the .a2 file is absent.
bool load_marvel(const HITS_DB* db, const std::string& stem, const char* name, std::vector<int64_t>& anno,
                 std::vector<int32_t>& pts)
{
    const std::string fa = stem + "." + name + ".a2";
    std::vector<unsigned char> a;
    if (!read_file(fa, a)) return false;
    if (a.size() < 64) throw TrackErr{fa + ": short header"};
    uint16_t size;
    uint64_t len, clen, cdlen;
    memcpy(&size, a.data() + 2, 2);
    memcpy(&len, a.data() + 8, 8);
    memcpy(&clen, a.data() + 16, 8);
    memcpy(&cdlen, a.data() + 24, 8);
    if (size != 8 || (db->part == 0 && len != (uint64_t)db->nreads) || len < (uint64_t)db->ufirst + db->nreads ||
        clen > a.size() - 64)
        throw TrackErr{fa + ": invalid header"};
    std::vector<int64_t> full(len + 1);
    if (!unchunk(a.data() + 64, (long long)clen, (unsigned char*)full.data(), 8 * (long long)(len + 1)))
        throw TrackErr{fa + ": bad compressed offsets"};
    const long long tot = full[len];
    if (tot < 0 || (tot & 3)) throw TrackErr{fa + ": bad offsets"};
    std::vector<unsigned char> all(tot), d;
    const std::string fd = stem + "." + name + ".d2";
    if (tot > 0 && (!read_file(fd, d) || cdlen > d.size() || !unchunk(d.data(), (long long)cdlen, all.data(), tot)))
        throw TrackErr{fd + ": missing or bad point data"};
    // the block's own slice (see the header for the reference's double trim)
    const int64_t lo = full[db->ufirst];
    anno.resize(db->nreads + 1);
    for (int i = 0; i <= db->nreads; i++) {
        const int64_t v = full[db->ufirst + i] - lo;
        if (v < 0 || (v & 3) || lo + v > tot) throw TrackErr{fa + ": bad offsets"};
        anno[i] = v / 4;
    }
    pts.resize(anno[db->nreads]);
    if (!pts.empty()) memcpy(pts.data(), all.data() + lo, 4 * pts.size());
    return true;
}

void load_dazzler(const HITS_DB* db, const std::string& stem, const char* name, std::vector<int64_t>& anno,
                  std::vector<int32_t>& pts)
{
    std::string fa, fd;
    FILE* af = nullptr;
    bool ispart = false;
    if (db->part > 0) {
        fa = stem + "." + std::to_string(db->part) + "." + name + ".anno";
        af = fopen(fa.c_str(), "rb");
        ispart = af != nullptr;
    }
    if (!af) {
        fa = stem + "." + name + ".anno";
        af = fopen(fa.c_str(), "rb");
    }
    if (!af) throw TrackErr{std::string("track ") + name + " does not exist"};
    fd = fa.substr(0, fa.size() - 5) + ".data";
    int tlen = 0, size = 0;
    const bool hdr = fread(&tlen, 4, 1, af) == 1 && fread(&size, 4, 1, af) == 1;
    if (!hdr || size != 8 || tlen != (ispart ? db->nreads : db->ureads)) {
        fclose(af);
        throw TrackErr{fa + ": junk annotation or not the size of the database"};
    }
    if (!ispart && db->part > 0) fseeko(af, 8LL * db->ufirst, SEEK_CUR);
    std::vector<int64_t> an(db->nreads + 1);
    const bool ok = fread(an.data(), 8, db->nreads + 1, af) == (size_t)(db->nreads + 1);
    fclose(af);
    if (!ok) throw TrackErr{fa + ": short annotation"};
    const int64_t off = an[0];
    anno.resize(db->nreads + 1);
    for (int i = 0; i <= db->nreads; i++) {
        const int64_t v = an[i] - off;
        if (v < 0 || (v & 3)) throw TrackErr{fa + ": bad offsets"};
        anno[i] = v / 4;
    }
    pts.resize(anno[db->nreads]);
    FILE* df = fopen(fd.c_str(), "rb");
    const bool dok = df && fseeko(df, off, SEEK_SET) == 0 && (pts.empty() || fread(pts.data(), 4 * pts.size(), 1, df) == 1);
    if (df) fclose(df);
    if (!dok) throw TrackErr{fd + ": missing or short point data"};
}

// per read an even number of non-decreasing points within [0, rlen]
void check_points(const HITS_DB* db, const char* name, const std::vector<int64_t>& anno, const std::vector<int32_t>& pts)
{
    for (int i = 0; i < db->nreads; i++) {
        if (anno[i + 1] < anno[i] || ((anno[i + 1] - anno[i]) & 1))
            throw TrackErr{std::string("track ") + name + ": odd number of points for a read"};
        int prev = 0;
        for (int64_t k = anno[i]; k < anno[i + 1]; k++) {
            if (pts[k] < prev || pts[k] > db->reads[i].rlen)
                throw TrackErr{std::string("track ") + name + ": points not increasing within the read"};
            prev = pts[k];
        }
    }
}

// Union of n tracks, read by read: a sweep over the points in increasing
// order, each toggling its track in or out; a point is kept where the count
// of tracks inside leaves or returns to 0 (Merge_Tracks' rule).  Points at
// equal positions may be taken in any order: the kept lists then differ only
// by empty gaps between intervals, which hold no k-mer.
void merge(const HITS_DB* db, const std::vector<std::vector<int64_t>>& an, const std::vector<std::vector<int32_t>>& pt,
           std::vector<int64_t>& anno, std::vector<int32_t>& pts)
{
    const int n = (int)an.size();
    anno.assign(db->nreads + 1, 0);
    pts.clear();
    std::vector<int64_t> cur(n);
    std::vector<char> in(n);
    for (int r = 0; r < db->nreads; r++) {
        anno[r] = (int64_t)pts.size();
        int level = 0;
        for (int t = 0; t < n; t++) {
            cur[t] = an[t][r];
            in[t] = 0;
        }
        for (;;) {
            int best = -1;
            for (int t = 0; t < n; t++)
                if (cur[t] < an[t][r + 1] && (best < 0 || pt[t][cur[t]] < pt[best][cur[best]])) best = t;
            if (best < 0) break;
            const int x = pt[best][cur[best]++];
            in[best] ^= 1;
            if (in[best]) {
                if (level++ == 0) pts.push_back(x);
            } else if (--level == 0) {
                pts.push_back(x);
            }
        }
    }
    anno[db->nreads] = (int64_t)pts.size();
}

}  // namespace

extern "C" {

int mi_load_mask_tracks(HITS_DB* db, const char* const* names, int n)
{
    if (!db || (n > 0 && !names) || n < 0) {
        mi_set_error("mi_load_mask_tracks: null block or names");
        return MI_ERR_ARG;
    }
    if (n == 0) return 0;
    try {
        const std::string stem = stem_of(db);
        std::vector<std::vector<int64_t>> an(n);
        std::vector<std::vector<int32_t>> pt(n);
        for (int t = 0; t < n; t++) {
            if (!names[t] || !names[t][0] || names[t][0] == '.')
                throw TrackErr{"track names must be non-empty and not begin with '.'"};
            for (int u = 0; u < t; u++)
                if (strcmp(names[u], names[t]) == 0) throw TrackErr{std::string("track ") + names[t] + " given twice"};
            if (!load_marvel(db, stem, names[t], an[t], pt[t])) load_dazzler(db, stem, names[t], an[t], pt[t]);
            check_points(db, names[t], an[t], pt[t]);
        }
        HITS_TRACK* t;
        if (n == 1) {
            t = new_track(names[0], an[0], pt[0]);
        } else {
            std::vector<int64_t> anno;
            std::vector<int32_t> pts;
            merge(db, an, pt, anno, pts);
            t = new_track("merge", anno, pts);
        }
        mi_release_tracks(db);
        db->tracks = t;
        return 0;
    } catch (const TrackErr& e) {
        mi_set_error(("mi_load_mask_tracks: " + e.msg).c_str());
        return MI_ERR_ARG;
    } catch (const std::bad_alloc&) {
        mi_set_error("mi_load_mask_tracks: host out of memory");
        return MI_ERR_ARG;
    }
}

void mi_release_tracks(HITS_DB* db)
{
    if (!db) return;
    HITS_TRACK* t = (HITS_TRACK*)db->tracks;
    HITS_TRACK* keep = nullptr;
    HITS_TRACK** tail = &keep;
    std::lock_guard<std::mutex> g(g_mu);
    while (t) {
        HITS_TRACK* nx = t->next;
        if (g_owned.erase(t)) {
            free(t->name);
            free(t->anno);
            free(t->data);
            free(t);
        } else {   // not ours: left in the list
            *tail = t;
            tail = &t->next;
            t->next = nullptr;
        }
        t = nx;
    }
    db->tracks = keep;
}

}  // extern "C"

namespace mi {

// complement of the first track of src (the mask) for the complemented block
void* complement_track(const HITS_DB* src)
{
    const HITS_TRACK* s = (const HITS_TRACK*)src->tracks;
    if (!s) return nullptr;
    try {
    const int64_t* sa = (const int64_t*)s->anno;
    const int32_t* sd = (const int32_t*)s->data;
    std::vector<int64_t> anno(sa, sa + src->nreads + 1);
    std::vector<int32_t> pts(sa[src->nreads] - sa[0]);
    for (int i = 0; i < src->nreads; i++) {
        const int rlen = src->reads[i].rlen;
        for (int64_t k = sa[i], j = sa[i + 1] - 1; k < sa[i + 1]; k++, j--) pts[k - sa[0]] = rlen - sd[j];
    }
    for (auto& v : anno) v -= sa[0];
    return new_track(s->name, anno, pts);
    } catch (...) {
        return nullptr;
    }
}

}  // namespace mi
