// `LAcat <source:las> > <target>.las`: the behaviour of dalign/LAcat.c --
// ROOT.1.las, ROOT.2.las, ... up to the first file that does not open,
// written to stdout under one header (the sum of the record counts, the
// first file's trace spacing; a file with another spacing is an error) with
// each file's records in order.  Host only: it is a copy.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "las_cli.hpp"

int main(int argc, char** argv)
{
    if (argc <= 1) {
        fprintf(stderr, "Usage: LAcat <source:las> > <target>.las\n");
        return 1;
    }
    const std::string pwd = lascli::path_to(argv[1]), root = lascli::root_of(argv[1], ".las");
    auto name = [&](int i) { return pwd + "/" + root + "." + std::to_string(i) + ".las"; };
    long long novl = 0;
    int tspace = 0;
    int nfiles = 0;
    for (int i = 1;; i++) {
        FILE* f = fopen(name(i).c_str(), "r");
        if (!f) break;
        long long n;
        int ts;
        if (fread(&n, 8, 1, f) != 1 || fread(&ts, 4, 1, f) != 1) {
            fprintf(stderr, "LAcat: System error, read failed!\n");
            return 1;
        }
        fclose(f);
        novl += n;
        if (i == 1) {
            tspace = ts;
        } else if (ts != tspace) {
            fprintf(stderr, "LAcat: PT-point spacing conflict (%d vs %d)\n", tspace, ts);
            return 1;
        }
        nfiles = i;
    }
    fwrite(&novl, 8, 1, stdout);
    fwrite(&tspace, 4, 1, stdout);
    const int tb = tspace <= 125 ? 1 : 2;
    std::vector<char> buf;
    for (int i = 1; i <= nfiles; i++) {
        FILE* f = fopen(name(i).c_str(), "r");
        if (!f) break;
        long long n;
        int ts;
        if (fread(&n, 8, 1, f) != 1 || fread(&ts, 4, 1, f) != 1) {
            fprintf(stderr, "LAcat: System error, read failed!\n");
            return 1;
        }
        // copy exactly n records (whatever follows them is not part of the file's records)
        for (long long j = 0; j < n; j++) {
            char hdr[40];
            if (fread(hdr, 40, 1, f) != 1) {
                fprintf(stderr, "LAcat: %s: fewer records than its header says\n", name(i).c_str());
                return 1;
            }
            int tlen;
            memcpy(&tlen, hdr, 4);
            const long long tsz = (long long)tlen * tb;
            if (tlen < 0) {
                fprintf(stderr, "LAcat: %s: negative trace length\n", name(i).c_str());
                return 1;
            }
            buf.resize(std::max(tsz, 1LL));
            if (tsz > 0 && fread(buf.data(), tsz, 1, f) != 1) {
                fprintf(stderr, "LAcat: %s: fewer records than its header says\n", name(i).c_str());
                return 1;
            }
            fwrite(hdr, 40, 1, stdout);
            if (tsz > 0) fwrite(buf.data(), tsz, 1, stdout);
        }
        fclose(f);
    }
    return fflush(stdout) == 0 ? 0 : 1;
}
