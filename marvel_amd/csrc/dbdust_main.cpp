// `DBdust` on the GPU: the command line and track files of db/DBdust.c
// (`DBdust [-b] [-w<int(64)>] [-t<double(2.)>] [-m<int(10)>] <path:db>`)
// over libmarvl_align's mi_dust.  Writes the Dazzler interval track
// .ROOT.dust.anno / .data beside the DB (ROOT = the argument without .db,
// so a block "DB.2" gets .DB.2.dust.*); for a whole DB whose track already
// covers some reads, only the new reads are scanned and appended.  The
// window is limited to 256.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../include/marvl_filter.h"
#include "las_cli.hpp"

static void usage()
{
    fprintf(stderr, "usage: [-b] [-w <int(64)>] [-t<double(2.)>] [-m<int(10)>] <path:db|dam>\n");
    fprintf(stderr, "options: -b ... biased, i.e. take into account the frequency of a given base\n");
    fprintf(stderr, "         -w ... scan window size\n");
    fprintf(stderr, "         -t ... thershold for being a low complexity interval\n");
    fprintf(stderr, "         -m ... intervals greater than -m bases are reported\n");
}

int main(int argc, char** argv)
{
    int window = 64, minlen = 9, biased = 0;
    double thresh = 2.;
    opterr = 0;
    int c;
    while ((c = getopt(argc, argv, "bw:t:m:")) != -1) {
        switch (c) {
        case 'b': biased = 1; break;
        case 'w':
            window = atoi(optarg);
            if (window <= 0) {
                fprintf(stderr, "[ERROR] Window must be positive (%d)\n", window);
                return 1;
            }
            break;
        case 't':
            thresh = atof(optarg);
            if (thresh <= 0.) {
                fprintf(stderr, "[ERROR] Threshold must be positive (%f)\n", thresh);
                return 1;
            }
            break;
        case 'm':
            minlen = (int)(atof(optarg) - 1);
            if (minlen <= 0) {
                fprintf(stderr, "[ERROR] Minimum hit size must be positive (%d)\n", minlen);
                return 1;
            }
            break;
        default:
            fprintf(stderr, "Unsupported option: %s\n", argv[optind - 1]);
            usage();
            return 1;
        }
    }
    if (optind + 1 > argc) {
        fprintf(stderr, "[ERROR] DB is required!\n");
        usage();
        return 1;
    }
    HITS_DB db;
    if (mi_open_block(argv[optind], &db)) {
        fprintf(stderr, "DBdust: cannot open %s: %s\n", argv[optind], mi_last_error());
        return 1;
    }
    const std::string stem = lascli::path_to(argv[optind]) + "/." + lascli::root_of(argv[optind], ".db");
    const std::string fa = stem + ".dust.anno", fd = stem + ".dust.data";
    // an existing whole-DB track: scan only the reads it does not cover yet
    int done = 0;
    long long base = 0;
    FILE* af = db.part > 0 ? nullptr : fopen(fa.c_str(), "r+");
    FILE* df = nullptr;
    if (af) {
        df = fopen(fd.c_str(), "r+");
        if (!df || fread(&done, 4, 1, af) != 1) {
            fprintf(stderr, "DBdust: cannot read %s\n", fa.c_str());
            return 1;
        }
        if (done >= db.nreads) return 0;
        fseeko(af, 0, SEEK_SET);
        const int size = 8;
        fwrite(&db.nreads, 4, 1, af);
        fwrite(&size, 4, 1, af);
        fseeko(af, 0, SEEK_END);
        fseeko(df, 0, SEEK_END);
        base = ftello(df);
    } else {
        af = fopen(fa.c_str(), "w");
        df = fopen(fd.c_str(), "w");
        if (!af || !df) {
            fprintf(stderr, "DBdust: cannot create %s\n", fa.c_str());
            return 1;
        }
        const int size = 8;
        const long long zero = 0;
        fwrite(&db.nreads, 4, 1, af);
        fwrite(&size, 4, 1, af);
        fwrite(&zero, 8, 1, af);
    }
    HITS_DB sub = db;   // the reads still to scan
    sub.reads = db.reads + done;
    sub.nreads = db.nreads - done;
    long long tot = 0;
    for (int i = 0; i < sub.nreads; i++) tot += sub.reads[i].rlen;
    std::vector<long long> anno(sub.nreads + 1);
    std::vector<int> pts(std::max(tot, 1LL));
    if (mi_dust(&sub, window, thresh, minlen, biased, anno.data(), pts.data())) {
        fprintf(stderr, "DBdust: %s\n", mi_last_error());
        return 1;
    }
    for (int i = 1; i <= sub.nreads; i++) {
        const long long v = base + anno[i];
        fwrite(&v, 8, 1, af);
    }
    if (anno[sub.nreads] > 0) fwrite(pts.data(), 1, anno[sub.nreads], df);
    const bool ok = fclose(af) == 0 && fclose(df) == 0;
    mi_close_block(&db);
    return ok ? 0 : 1;
}
