// `daligner` on the GPU: the command line and output files of MARVEL's
// dalign/daligner.c:615-1032 over libmarvl_align's Sort_Kmers /
// Match_Filter / Write_Overlap_Buffer.
//
//   daligner [-vbAITO] [-k<int(14)>] [-w<int(6)>] [-h<int(35)>] [-t<int>] [-M<int>]
//            [-e<double(.70)>] [-l<int(1000)>] [-s<int(100)>] [-H<int>]
//            [-r<int(1)>] [-j<int(4)>] [-m<track>]+ <subject:db> <target:db> ...
//
// -j keeps its reference meaning (the number of report chunks, a power of
// 2), which decides the pairs the reference skips at chunk ends; the GPU
// work itself is not split by it.  -m<track> (repeatable) soft-masks the
// intervals of the named tracks, Dazzler (.anno/.data) or MARVEL
// (.a2/.d2) layout; -b weighs k-mers by the base frequencies; -I also
// pairs every read with itself, -O keeps only those overlaps; a block may be
// of a .db or a .dam.
#include <strings.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/marvl_align.h"
#include "../../include/marvl_filter.h"

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Root(path, suffix): file name without directory and without `suffix`
// when it ends with it (any case, db/DB.c:104-140)
static std::string root_of(const char* path, const char* suffix)
{
    const char* s = strrchr(path, '/');
    std::string r = s ? s + 1 : path;
    const size_t n = strlen(suffix);
    if (r.size() > n && strcasecmp(r.c_str() + r.size() - n, suffix) == 0) r.resize(r.size() - n);
    return r;
}

// the driver's name of a block argument: without ".dam" for a .dam path, else without ".db"
static std::string root_of(const char* path)
{
    const size_t n = strlen(path);
    return root_of(path, n >= 4 && strcmp(path + n - 4, ".dam") == 0 ? ".dam" : ".db");
}

static void make_subdir(int run_id, int part)
{
    char d[64];
    snprintf(d, sizeof d, "%03d_%05d", run_id, part);
    struct stat s;
    if (stat(d, &s) != 0) mkdir(d, 0755);
}

static void usage()
{
    fprintf(stderr,
            "usage: daligner [-vbAITO] [-k<int(14)>] [-w<int(6)>] [-h<int(35)>] [-t<int>] [-M<int>]\n"
            "                [-e<double(.70)>] [-l<int(1000)>] [-s<int(100)>] [-H<int>] [-r<int(1)>] [-j<int(4)>]\n"
            "                [-m<track>]+\n"
            "                <subject:db> <target:db> ...\n");
}

static void print_stats(const char* what, double secs)
{
    MiFilterStats s;
    mi_filter_stats(&s);
    printf("{\"pass\": \"%s\", \"seconds\": %.4f, \"kmers_a\": %lld, \"kmers_b\": %lld, \"hits\": %lld, \"pairs\": %lld, "
           "\"jobs\": %lld, \"jobs_kept\": %lld, \"aligned\": %lld, \"records\": %lld, \"cells\": %lld, \"snake\": %lld, \"limit\": %d, "
           "\"edge_events\": %lld, \"edge_fixups\": %lld, \"retries\": %lld, \"t_filter\": %.4f, \"t_report\": %.4f, "
           "\"t_copy\": %.4f}\n",
           what, secs, s.nkmers_a, s.nkmers_b, s.nhits, s.npairs, s.njobs, s.njobs_kept, s.naligned, s.nrecords, s.cells, s.snake,
           s.limit, s.edge_events, s.edge_fixups, s.retries, s.t_filter, s.t_report, s.t_copy);
    fflush(stdout);
}

// The learned scorer's weights file (marvel_amd/data/scorer_weights.json
// layout: arrays "w1", "b1", "w2", "b2", "w3" and numbers "b3",
// "threshold"): the values of one key.
static bool json_floats(const std::string& js, const char* key, std::vector<float>& out)
{
    const std::string k = std::string("\"") + key + "\":";
    size_t p = js.find(k);
    if (p == std::string::npos) return false;
    p += k.size();
    while (p < js.size() && js[p] == ' ') p++;
    out.clear();
    const bool arr = p < js.size() && js[p] == '[';
    if (arr) p++;
    const char* s = js.c_str() + p;
    for (;;) {
        while (*s == ' ' || *s == ',') s++;
        if (*s == ']' || *s == 0) break;
        char* e = nullptr;
        const double v = strtod(s, &e);
        if (e == s) return false;
        out.push_back((float)v);
        s = e;
        if (!arr) break;
    }
    return !out.empty();
}

// MI_SCORER_WEIGHTS=<file>: gate candidate pairs with the learned scorer
// (not bit-exact with the reference; the default is the exact path)
static bool load_scorer_from_env()
{
    const char* f = getenv("MI_SCORER_WEIGHTS");
    if (f == nullptr || *f == 0) return true;
    FILE* fp = fopen(f, "rb");
    if (fp == nullptr) return false;
    std::string js;
    char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, fp)) > 0) js.append(buf, n);
    fclose(fp);
    std::vector<float> w1, b1, w2, b2, w3, b3, thr;
    if (!json_floats(js, "w1", w1) || !json_floats(js, "b1", b1) || !json_floats(js, "w2", w2) ||
        !json_floats(js, "b2", b2) || !json_floats(js, "w3", w3) || !json_floats(js, "b3", b3) ||
        !json_floats(js, "threshold", thr) || w1.size() != 16 * 64 || b1.size() != 64 || w2.size() != 64 * 64 ||
        b2.size() != 64 || w3.size() != 64)
        return false;
    return mi_scorer_load(w1.data(), b1.data(), w2.data(), b2.data(), w3.data(), b3[0], thr[0]) == 0 &&
           mi_scorer_mode(2) == 0;
}

int main(int argc, char** argv)
{
    int kmer = 14, binshift = 6, hitmin = 35, suppress = 0, tspace = 100, run_id = 1, no_trace = 0, l = 1000;
    double ave_err = .70;
    std::vector<const char*> masks;
    int only_identity = 0;
    VERBOSE = 0;
    IDENTITY = 0;
    SYMMETRIC = 1;
    HGAP_MIN = 0;
    NTHREADS = 4;
    NSHIFT = 2;
    MEM_PHYSICAL = (uint64_t)sysconf(_SC_PHYS_PAGES) * (uint64_t)sysconf(_SC_PAGESIZE);
    MEM_LIMIT = MEM_PHYSICAL;
    int c;
    while ((c = getopt(argc, argv, "vbOTAIk:w:h:t:M:e:l:s:H:D:m:r:j:")) != -1) {
        switch (c) {
        case 'v': VERBOSE += 1; break;
        case 'b': BIASED = 1; break;
        case 'I': IDENTITY = 1; break;
        case 'O':
            IDENTITY = 1;
            only_identity = 1;
            break;
        case 'T': no_trace = 1; break;
        case 'A': SYMMETRIC = 0; break;
        case 'k': kmer = atoi(optarg); break;
        case 'w': binshift = atoi(optarg); break;
        case 'h': hitmin = atoi(optarg); break;
        case 't': suppress = atoi(optarg); break;
        case 'H':
            // parsed into a local of the reference's main that shadows the
            // global the filter reads (dalign/daligner.c:641, :709-711): the
            // option is checked and has no effect on the output
            if (atoi(optarg) < 0) {
                fprintf(stderr, "invalid HGAP threshold of %d\n", atoi(optarg));
                return 1;
            }
            break;
        case 'e': ave_err = atof(optarg); break;
        case 'l': l = atoi(optarg); break;
        case 's': tspace = atoi(optarg); break;
        case 'r': run_id = atoi(optarg); break;
        case 'm': masks.push_back(optarg); break;
        case 'M': {
            const int t = atoi(optarg);
            if (t < 0) fprintf(stderr, "invalid memory limit of (%d)\n", t);
            MEM_LIMIT = (uint64_t)(long long)t * 0x40000000ULL;
            break;
        }
        case 'j': {
            const int t = atoi(optarg);
            int e = 0;
            while ((1 << (e + 1)) <= t) e++;
            if (t < 1 || (1 << e) != t) {
                fprintf(stderr, "invalid number of threads: %d. Must be a power of 2.\n", t);
                return 1;
            }
            NTHREADS = t;
            NSHIFT = e;
            break;
        }
        default:
            fprintf(stderr, "daligner: option -%c is not supported by this build\n", optopt ? optopt : c);
            usage();
            return 1;
        }
    }
    MINOVER = 2 * l;
    if (optind + 2 > argc) {
        usage();
        return 1;
    }
    if (kmer < 0 || binshift < 0 || hitmin < 0 || tspace <= 0 || run_id < 0) {
        fprintf(stderr, "invalid parameter\n");
        return 1;
    }
    if (Set_Filter_Params(kmer, binshift, suppress, hitmin)) {
        fprintf(stderr, "daligner: invalid k-mer length %d (%s)\n", kmer, mi_last_error());
        return 1;
    }
    if (!load_scorer_from_env()) {
        fprintf(stderr, "daligner: cannot load the scorer weights of MI_SCORER_WEIGHTS (%s)\n", mi_last_error());
        return 1;
    }
    const char* afile = argv[optind];
    HITS_DB A;
    if (mi_open_block(afile, &A)) {
        fprintf(stderr, "daligner: cannot open %s: %s\n", afile, mi_last_error());
        return 1;
    }
    // read_DB: the block's -m tracks, merged when there are several (dalign/daligner.c:373-431)
    if (!masks.empty() && mi_load_mask_tracks(&A, masks.data(), (int)masks.size())) {
        printf("[ERROR] - Unable to load track: %s\n", mi_last_error());
        return 1;
    }
    for (int i = 0; i < A.nreads; i++)
        if (A.reads[i].rlen < kmer) {
            fprintf(stderr, "[ERROR] - daligner: Block %s contains reads < %dbp long !  Run DBsplit.\n", afile, kmer);
            return 1;
        }
    std::string aroot = root_of(afile);
    // B blocks of another DB turn the symmetric output off (dalign/daligner.c:884-917)
    if (SYMMETRIC)
        for (int i = optind + 1; i < argc; i++) {
            if (strcmp(afile, argv[i]) == 0) continue;
            const std::string br = root_of(argv[i], ".dam");   // as dalign/daligner.c:893-896 takes it
            const size_t ad = aroot.rfind('.'), bd = br.rfind('.');
            const std::string ap = aroot.substr(0, ad == std::string::npos ? aroot.size() : ad + 1);
            const std::string bp = br.substr(0, bd == std::string::npos ? br.size() : bd + 1);
            if (ap != bp) {
                SYMMETRIC = 0;
                break;
            }
        }
    make_subdir(run_id, A.part);
    Align_Spec* spec = New_Align_Spec(ave_err, tspace, A.freq, NTHREADS, SYMMETRIC, only_identity, no_trace);
    char* aname = strdup(aroot.c_str());
    void* aindex = nullptr;
    for (int i = optind + 1; i < argc; i++) {
        const char* bfile = argv[i];
        const bool self = strcmp(afile, bfile) == 0;
        const double t0 = now_s();
        if (aindex == nullptr) aindex = Sort_Kmers(&A, nullptr);
        if (self) {
            Match_Filter(aname, &A, aname, &A, aindex, 0, aindex, 0, 0, spec);
            if (VERBOSE) print_stats("forward", now_s() - t0);
            HITS_DB cA;
            mi_complement_header(&A, &cA);
            void* cindex = mi_sort_kmers_complement(aindex, &cA, nullptr);
            if (cindex == nullptr) {
                fprintf(stderr, "daligner: %s\n", mi_last_error());
                return 1;
            }
            const double t1 = now_s();
            Match_Filter(aname, &A, aname, &cA, aindex, 0, cindex, 0, 1, spec);
            mi_release_tracks(&cA);
            if (VERBOSE) print_stats("complement", now_s() - t1);
            Write_Overlap_Buffer(spec, run_id, aname, aname, A.ufirst + A.nreads - 1);
            Reset_Overlap_Buffer(spec);
        } else {
            HITS_DB B;
            if (mi_open_block(bfile, &B)) {
                fprintf(stderr, "daligner: cannot open %s: %s\n", bfile, mi_last_error());
                return 1;
            }
            if (!masks.empty() && mi_load_mask_tracks(&B, masks.data(), (int)masks.size())) {
                printf("[ERROR] - Unable to load track: %s\n", mi_last_error());
                return 1;
            }
            for (int r = 0; r < B.nreads; r++)
                if (B.reads[r].rlen < kmer) {
                    fprintf(stderr, "[ERROR] - daligner: Block %s contains reads < %dbp long !  Run DBsplit.\n", bfile,
                            kmer);
                    return 1;
                }
            if (SYMMETRIC) make_subdir(run_id, B.part);
            char* bname = strdup(root_of(bfile).c_str());
            void* bindex = Sort_Kmers(&B, nullptr);
            HITS_DB cB;
            mi_complement_header(&B, &cB);
            void* cindex = mi_sort_kmers_complement(bindex, &cB, nullptr);
            if (cindex == nullptr) {
                fprintf(stderr, "daligner: %s\n", mi_last_error());
                return 1;
            }
            Match_Filter(aname, &A, bname, &B, aindex, 0, bindex, 0, 0, spec);
            if (VERBOSE) print_stats("forward", now_s() - t0);
            const double t1 = now_s();
            Match_Filter(aname, &A, bname, &cB, aindex, 0, cindex, 0, 1, spec);
            mi_release_tracks(&cB);
            if (VERBOSE) print_stats("complement", now_s() - t1);
            const int last = (B.part < A.part) ? B.ufirst + B.nreads - 1 : A.ufirst + A.nreads - 1;
            Write_Overlap_Buffer(spec, run_id, aname, bname, last);
            Reset_Overlap_Buffer(spec);
            free(bname);
            mi_close_block(&B);
        }
        if (VERBOSE) printf("{\"block_pair_seconds\": %.4f}\n", now_s() - t0);
    }
    mi_free_kmer_index(aindex);
    Free_Align_Spec(spec);
    free(aname);
    mi_close_block(&A);
    return 0;
}
