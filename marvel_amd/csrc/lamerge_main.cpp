// `LAmerge` with the GPU doing the ordering: the command line and input
// collection of utils/LAmerge.c + utils/LAmergeUtils.c,
//
//   LAmerge [-hksv] [-n <int(8)>] [-S <suffix>] [-f <file list>] <db> <out.las> [<in.las | dir> ...]
//
// inputs are the named files, the .las files of named directories (in
// readdir order, names not starting with '.'), the lines of the -f list, or,
// when none is given, DB.1<suffix>.las ... DB.<blocks><suffix>.las beside the
// DB.  One input is copied; several are read whole and ordered by
// mi_las_merge: the reference's output for sorted inputs whose equal keys lie
// in one file, and for -s.  The reference merges -n files per round through
// intermediate files; one GPU pass needs no rounds, so -n and -k (keep the
// intermediate files) are checked and change nothing.  -C (record checks) is
// not supported.
#include <dirent.h>
#include <getopt.h>
#include <sys/stat.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/marvl_filter.h"
#include "las_cli.hpp"

static void usage(const char* prog)
{
    fprintf(stderr,
            "usage: %s [-hksv] [-n <int(8)>] [-S <suffix>] [-f <file>] <db> <out.las> [<in.las | dir> ...]\n"
            "   -v       verbose\n"
            "   -k       keep intermediate merge files (no effect: one merge pass)\n"
            "   -s       sort the input files first\n"
            "   -n <int> files merged per round, in [2, 255] (no effect: one merge pass)\n"
            "   -S <s>   merge DB.<block><s>.las for every block when no input is named\n"
            "   -f <f>   read the input file names from f, one a line\n",
            prog);
}

// an input the reference accepts without record checks: non-empty, readable header
static bool usable(const std::string& f, int verbose, bool silent)
{
    if (!silent || verbose > 2) printf("Check file: %s\n", f.c_str());
    struct stat st;
    if (stat(f.c_str(), &st) != 0 || st.st_size == 0) return false;
    FILE* in = fopen(f.c_str(), "r");
    if (!in) return false;
    long long n;
    int ts;
    const bool ok = fread(&n, 8, 1, in) == 1 && fread(&ts, 4, 1, in) == 1 && n >= 0;
    fclose(in);
    return ok;
}

int main(int argc, char** argv)
{
    int verbose = 0, fway = 8;
    std::string suffix, list;
    static struct option longopts[] = {{"help", no_argument, 0, 'h'},         {"keep", no_argument, 0, 'k'},
                                       {"sort", no_argument, 0, 's'},         {"verbose", no_argument, 0, 'v'},
                                       {"nFiles", required_argument, 0, 'n'}, {"in", required_argument, 0, 'f'},
                                       {"suffix", required_argument, 0, 'S'}, {"check", required_argument, 0, 'C'},
                                       {0, 0, 0, 0}};
    int c, idx = 0;
    while ((c = getopt_long(argc, argv, "hksvn:S:f:C:", longopts, &idx)) != -1) {
        switch (c) {
        case 'v': verbose++; break;
        case 'k':
        case 's': break;   // -s: the merge order is a stable sort of every record already
        case 'n':
            fway = (int)strtol(optarg, nullptr, 10);
            if (fway < 2 || fway > 255) {
                fprintf(stderr, "Number of files to merge is not accpeted! Must be an integer in [2, 255]\n");
                return 1;
            }
            break;
        case 'S': suffix = optarg[0] == '\0' ? "" : (optarg[0] == '.' ? optarg : std::string(".") + optarg); break;
        case 'f': list = optarg; break;
        case 'C':
            fprintf(stderr, "[ERROR] - LAmerge: record checks (-C) are not supported by this build\n");
            return 1;
        default: usage(argv[0]); return 1;
        }
    }
    if (optind + 1 > argc) {
        fprintf(stderr, "At least a database is required!\n");
        usage(argv[0]);
        return 1;
    }
    std::string db = argv[optind++];
    if (db.size() < 3 || strcasecmp(db.c_str() + db.size() - 3, ".db") != 0) db += ".db";
    int nblocks = -1;
    {   // DB_Blocks (db/DB.c:1605-1658): "blocks = %d" after the file lines
        FILE* f = fopen(db.c_str(), "r");
        int nfiles;
        char line[30001];
        if (f && fscanf(f, "files = %d\n", &nfiles) == 1) {
            bool ok = true;
            for (int i = 0; i < nfiles && ok; i++) ok = fgets(line, sizeof line, f) != nullptr;
            if (ok && fscanf(f, "blocks = %d\n", &nblocks) != 1) nblocks = -1;
        }
        if (f) fclose(f);
    }
    if (nblocks < 0) {
        fprintf(stderr, "Cannot determine number of blocks of database file %s!\n", db.c_str());
        return 1;
    }
    const std::string dbname = lascli::root_of(db, ".db");
    if (optind == argc) {
        fprintf(stderr, "[ERROR] - LAmerge: The output file is required!\n");
        usage(argv[0]);
        return 1;
    }
    std::string out = argv[optind];
    if (out.size() > 4 && out.compare(out.size() - 4, 4, ".las") == 0) out.resize(out.size() - 4);
    {
        struct stat sb;
        if (stat(argv[optind], &sb) == 0) {
            if (S_ISDIR(sb.st_mode)) {
                fprintf(stderr, "[ERROR] - LAmerge: The output overlap file cannot be an existing directory: %s!\n",
                        argv[optind]);
                return 1;
            }
            if (S_ISREG(sb.st_mode))
                fprintf(stderr, "[WARNING] - LAmerge: The output overlap file %s will be overwritten!\n", argv[optind]);
        }
    }
    optind++;
    std::vector<std::string> in;
    auto add = [&](const std::string& f) {
        in.push_back(f);
        if (verbose) printf("Add input file: %s\n", f.c_str());
    };
    for (; optind < argc; optind++) {
        struct stat sb;
        if (stat(argv[optind], &sb) != 0) {
            fprintf(stderr, "[ERROR] - LAmerge: Cannot determine file type from argument: %s\n", argv[optind]);
            return 1;
        }
        if (S_ISDIR(sb.st_mode)) {
            DIR* d = opendir(argv[optind]);
            int failed = 0;
            if (!d) {
                fprintf(stderr, "[WARNING] - LAmerge: Could not open directory: %s\n", argv[optind]);
                continue;
            }
            while (struct dirent* e = readdir(d)) {
                const size_t n = strlen(e->d_name);
                if (e->d_name[0] == '.' || n < 4 || strcmp(e->d_name + n - 4, ".las") != 0) continue;
                const std::string f = std::string(argv[optind]) + "/" + e->d_name;
                if (usable(f, verbose, true)) add(f);
                else failed++;
            }
            closedir(d);
            if (failed) {
                fprintf(stderr, "[ERROR]: %d input overlap files are corrupt. Stop!\n", failed);
                return 1;
            }
        } else if (S_ISREG(sb.st_mode) || S_ISLNK(sb.st_mode)) {
            if (usable(argv[optind], verbose, true)) add(argv[optind]);
            else fprintf(stderr, "[WARNING] - LAmerge: skip file %s. File check fails!", argv[optind]);
        } else {
            fprintf(stderr, "[WARNING] - LAmerge: Cannot handle file type from argument: %s\n", argv[optind]);
        }
    }
    if (!list.empty()) {
        FILE* f = fopen(list.c_str(), "r");
        if (!f) {
            fprintf(stderr, "[WARNING] - LAmerge : Cannot open input file %s\n.\n", list.c_str());
        } else {
            char line[4096];
            int failed = 0;
            while (fgets(line, sizeof line, f)) {
                char* a = line;
                while (isspace((unsigned char)*a)) a++;
                char* z = a + strlen(a);
                while (z > a && isspace((unsigned char)z[-1])) *--z = '\0';
                if (usable(a, verbose, false)) add(a);
                else failed++;
            }
            fclose(f);
            if (failed) {
                fprintf(stderr, "[ERROR] - LAmerge : %d files do not pass check. Stop!\n", failed);
                return 1;
            }
        }
    }
    if (in.empty()) {   // DB.<block><suffix>.las for every block, all of them usable
        const std::string dir = lascli::path_to(db);
        bool failed = false;
        for (int i = 1; i <= nblocks; i++) {
            const std::string f = dir + "/" + dbname + "." + std::to_string(i) + suffix + ".las";
            if (!usable(f, verbose, true)) {
                fprintf(stderr, "[ERROR] - LAmerge : Overlap file %s did not pass check!\n", f.c_str());
                failed = true;
            }
        }
        if (failed) {
            fprintf(stderr, "[ERROR] - LAmerge : Some overlap files are not valid. Merging by suffix requires that all "
                            "Overlaps files [1, numBlocks] to be valid! Stop here!\n");
            return 1;
        }
        for (int i = 1; i <= nblocks; i++) add(dir + "/" + dbname + "." + std::to_string(i) + suffix + ".las");
    }
    if (in.empty()) {
        fprintf(stderr, "[WARNING] - LAmerge: No input files. Nothing to do!\n");
        return 0;
    }
    const std::string fout = out + ".las";
    if (in.size() == 1) {   // copied as it is
        FILE* a = fopen(in[0].c_str(), "rb");
        FILE* b = fopen(fout.c_str(), "wb");
        if (!a || !b) {
            fprintf(stderr, "Cannot open %s.\n", !a ? in[0].c_str() : fout.c_str());
            return 1;
        }
        char buf[1 << 16];
        size_t n;
        while ((n = fread(buf, 1, sizeof buf, a)) > 0)
            if (fwrite(buf, 1, n, b) != n) {
                fprintf(stderr, "Error writing destination file.\n");
                return 1;
            }
        fclose(a);
        return fclose(b) == 0 ? 0 : 1;
    }
    // read every input's records (the bytes of its header's record count) into one body
    long long total = 0;
    int tspace = 0;
    std::vector<char> body;
    for (size_t k = 0; k < in.size(); k++) {
        FILE* f = fopen(in[k].c_str(), "rb");
        long long n;
        int ts;
        if (!f || fread(&n, 8, 1, f) != 1 || fread(&ts, 4, 1, f) != 1) {
            fprintf(stderr, "[ERROR] - LAmerge: Cannot open file \"%s\" for reading\n", in[k].c_str());
            return 1;
        }
        if (k == 0) {
            tspace = ts;
        } else if (ts != tspace) {
            fprintf(stderr, "LAmerge: PT-point spacing conflict (%d vs %d)\n", tspace, ts);
            return 1;
        }
        const int tb = tspace <= 125 ? 1 : 2;
        for (long long j = 0; j < n; j++) {
            const size_t o = body.size();
            body.resize(o + 40);
            int tlen;
            if (fread(body.data() + o, 40, 1, f) != 1 || (memcpy(&tlen, body.data() + o, 4), tlen < 0)) {
                fprintf(stderr, "[ERROR] - LAmerge: %s holds fewer records than its header says\n", in[k].c_str());
                return 1;
            }
            const size_t t = (size_t)tlen * tb;
            body.resize(o + 40 + t);
            if (t > 0 && fread(body.data() + o + 40, t, 1, f) != 1) {
                fprintf(stderr, "[ERROR] - LAmerge: %s holds fewer records than its header says\n", in[k].c_str());
                return 1;
            }
        }
        fclose(f);
        total += n;
    }
    if (verbose) {
        printf("Merging %zu files totalling ", in.size());
        lascli::print_number(total, stdout);
        printf(" records\n");
    }
    std::vector<char> sorted(std::max<size_t>(body.size(), 1));
    long long osize = 0;
    if (mi_las_merge(body.data(), (long long)body.size(), total, tspace, sorted.data(), &osize) != 0) {
        fprintf(stderr, "[ERROR] - LAmerge: %s\n", mi_last_error());
        return 1;
    }
    FILE* g = fopen(fout.c_str(), "wb");
    if (!g) {
        fprintf(stderr, "[ERROR] - LAmerge: Cannot open file \"%s\" for writing\n", fout.c_str());
        return 1;
    }
    const bool ok = fwrite(&total, 8, 1, g) == 1 && fwrite(&tspace, 4, 1, g) == 1 &&
                    (osize == 0 || fwrite(sorted.data(), osize, 1, g) == 1);
    if (fclose(g) != 0 || !ok) {
        fprintf(stderr, "[ERROR] - LAmerge: cannot write %s\n", fout.c_str());
        return 1;
    }
    (void)fway;
    return usable(fout, verbose, false) ? 0 : 1;
}
