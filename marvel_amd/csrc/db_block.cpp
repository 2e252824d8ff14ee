// DB blocks for the daligner driver (host side): the stub / .idx / .bps
// layout of MARVEL's DB (db/DB.h:164-249) read into HITS_DB records the
// filter takes, plus the block complement of the driver.
//   stub  DIR/ROOT.db : "files = N", N file lines, "blocks = B", "size = S",
//                       then B+1 lines with the first read of each block
//   DIR/.ROOT.idx     : 88-byte DB header (ureads, freq[4], maxlen, totlen, ...)
//                       then 32-byte read records (rlen, -, boff, coff, flags)
//   DIR/.ROOT.bps     : ceil(rlen/4) bytes per read, base i in bits 6-2(i%4)
// In memory every read is one byte per base (0..3) with a 4 before and
// after it (db/DB.c:1349-1410).
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/marvl_filter.h"
#include "mi_common.h"

namespace mi {
// Complement of the mask track of block src (its first track) as a new
// library-owned HITS_TRACK for the complemented block (tracks.cpp); nullptr
// when src has no track or memory runs out.
void* complement_track(const HITS_DB* src);
}

namespace {

bool read_stub(const std::string& stub, int part, int* ufirst, int* ulast)
{
    FILE* f = fopen(stub.c_str(), "r");
    if (f == nullptr) return false;
    int nfiles = 0, nblocks = 0, last = 0;
    long long size = 0;
    char a[4096], b[4096];
    bool ok = fscanf(f, "files = %9d\n", &nfiles) == 1;
    for (int i = 0; ok && i < nfiles; i++) ok = fscanf(f, "  %9d %4095s %4095s\n", &last, a, b) == 3;
    if (!ok) {
        fclose(f);
        return false;
    }
    *ufirst = 0;
    *ulast = -1;   // whole DB
    if (fscanf(f, "blocks = %9d\n", &nblocks) != 1) {   // never split
        fclose(f);
        return part == 0;
    }
    if (fscanf(f, "size = %9lld\n", &size) != 1 || part > nblocks) {
        fclose(f);
        return false;
    }
    if (part > 0) {
        int v = 0;
        for (int i = 1; i <= part && ok; i++) ok = fscanf(f, " %9d\n", &v) == 1;
        *ufirst = v;
        ok = ok && fscanf(f, " %9d\n", &v) == 1;
        *ulast = v;
    }
    fclose(f);
    return ok;
}

}  // namespace

extern "C" {

int mi_open_block(const char* path, HITS_DB* db)
{
    memset(db, 0, sizeof *db);
    std::string p(path), dir = ".", root;
    const size_t sl = p.rfind('/');
    if (sl != std::string::npos) {
        dir = p.substr(0, sl);
        root = p.substr(sl + 1);
    } else
        root = p;
    if (root.size() > 4 && root.compare(root.size() - 4, 4, ".dam") == 0) root.resize(root.size() - 4);
    else if (root.size() > 3 && root.compare(root.size() - 3, 3, ".db") == 0) root.resize(root.size() - 3);
    int part = 0;
    const size_t dot = root.rfind('.');
    if (dot != std::string::npos && dot + 1 < root.size()) {
        char* end = nullptr;
        const long v = strtol(root.c_str() + dot + 1, &end, 10);
        if (*end == 0 && v > 0) {
            part = (int)v;
            root.resize(dot);
        }
    }
    int ufirst = 0, ulast = -1;
    // the stub is ROOT.db, else ROOT.dam (a DB of FA2dam; same index and bases files)
    std::string stub = dir + "/" + root + ".db";
    if (access(stub.c_str(), F_OK) != 0) stub = dir + "/" + root + ".dam";
    if (!read_stub(stub, part, &ufirst, &ulast)) {
        mi::mi_set_error("cannot read the DB stub");
        return MI_ERR_ARG;
    }
    FILE* idx = fopen((dir + "/." + root + ".idx").c_str(), "rb");
    if (idx == nullptr) {
        mi::mi_set_error("cannot open the .idx file");
        return MI_ERR_ARG;
    }
    unsigned char hdr[88];
    if (fread(hdr, 88, 1, idx) != 1) {
        fclose(idx);
        mi::mi_set_error("short .idx header");
        return MI_ERR_ARG;
    }
    memcpy(&db->ureads, hdr + 0, 4);
    memcpy(db->freq, hdr + 4, 16);
    memcpy(&db->maxlen, hdr + 20, 4);
    memcpy(&db->totlen, hdr + 24, 8);
    if (ulast < 0) ulast = db->ureads;
    const int n = ulast - ufirst;
    if (n < 0 || ufirst < 0 || ulast > db->ureads) {
        fclose(idx);
        mi::mi_set_error("block outside the DB");
        return MI_ERR_ARG;
    }
    HITS_READ* rd = (HITS_READ*)calloc((size_t)n + 2, sizeof(HITS_READ));
    std::vector<unsigned char> recs((size_t)n * 32 + 1);
    if (fseeko(idx, 88 + 32LL * ufirst, SEEK_SET) != 0 || (n > 0 && fread(recs.data(), 32, n, idx) != (size_t)n)) {
        fclose(idx);
        free(rd);
        mi::mi_set_error("short .idx read records");
        return MI_ERR_ARG;
    }
    fclose(idx);
    db->reads = rd + 1;   // reads[-1] exists, as in the reference's loader
    std::vector<long long> foff(n);
    long long tot = 0;
    int mx = 0;
    for (int i = 0; i < n; i++) {
        HITS_READ& r = db->reads[i];
        memcpy(&r.rlen, &recs[32 * (size_t)i], 4);
        memcpy(&foff[i], &recs[32 * (size_t)i + 8], 8);
        memcpy(&r.coff, &recs[32 * (size_t)i + 16], 8);
        memcpy(&r.flags, &recs[32 * (size_t)i + 24], 4);
        tot += r.rlen;
        mx = r.rlen > mx ? r.rlen : mx;
    }
    db->nreads = n;
    db->part = part;
    db->ufirst = ufirst;
    if (part > 0) {   // block statistics replace the DB's (db/DB.c:544-555)
        db->maxlen = mx;
        db->totlen = tot;
    }
    char* bases = (char*)malloc((size_t)tot + n + 2);
    bases[0] = 4;
    FILE* bps = fopen((dir + "/." + root + ".bps").c_str(), "rb");
    if (bps == nullptr) {
        free(bases);
        free(rd);
        db->reads = nullptr;
        mi::mi_set_error("cannot open the .bps file");
        return MI_ERR_ARG;
    }
    std::vector<unsigned char> pk;
    long long o = 1;
    for (int i = 0; i < n; i++) {
        const int len = db->reads[i].rlen, cl = (len + 3) / 4;
        pk.resize((size_t)cl + 1);
        if (fseeko(bps, foff[i], SEEK_SET) != 0 || (cl > 0 && fread(pk.data(), cl, 1, bps) != 1)) {
            fclose(bps);
            free(bases);
            free(rd);
            db->reads = nullptr;
            mi::mi_set_error("short .bps read");
            return MI_ERR_ARG;
        }
        char* s = bases + o;
        for (int j = 0; j < len; j++) s[j] = (char)((pk[j >> 2] >> (6 - 2 * (j & 3))) & 3);
        s[len] = 4;
        db->reads[i].boff = o;
        o += len + 1;
    }
    fclose(bps);
    db->bases = bases;
    db->loaded = 1;
    db->path = strdup((dir + "/" + root).c_str());
    return 0;
}

void mi_close_block(HITS_DB* db)
{
    mi_release_tracks(db);
    if (db->reads) free(db->reads - 1);
    free(db->bases);
    free(db->path);
    memset(db, 0, sizeof *db);
}

// Header of the reverse-complemented block (dalign/daligner.c:456-573):
// same reads, frequencies of A<->T and C<->G swapped.  The read records
// are shared with `src`; bases are not produced (the GPU index of the
// complement is derived from the source index).  A mask track of src is
// complemented for dst; like complement_DB, running out of memory there
// ends the process.
void mi_complement_header(const HITS_DB* src, HITS_DB* dst)
{
    *dst = *src;
    dst->freq[0] = src->freq[3];
    dst->freq[1] = src->freq[2];
    dst->freq[2] = src->freq[1];
    dst->freq[3] = src->freq[0];
    dst->bases = nullptr;
    dst->path = nullptr;
    dst->tracks = nullptr;
    if (src->tracks) {
        dst->tracks = mi::complement_track(src);
        if (!dst->tracks) {
            fprintf(stderr, "mi_complement_header: cannot allocate the complemented mask track\n");
            exit(1);
        }
    }
}

}  // extern "C"
