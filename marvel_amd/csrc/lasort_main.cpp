// `LAsort` on the GPU: the command line and files of dalign/LAsort.c
// (`LAsort [-v] <align:las> ...` writes ROOT.S.las beside each ROOT.las) with
// the sort and record gather in libmarvl_align's mi_las_sort.
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/marvl_filter.h"
#include "las_cli.hpp"

int main(int argc, char** argv)
{
    bool verbose = false;
    std::vector<std::string> files;
    for (int i = 1; i < argc; i++) {
        if (argv[i][0] == '-') {
            for (int k = 1; argv[i][k]; k++) {
                if (argv[i][k] != 'v') {
                    fprintf(stderr, "LAsort: -%c is an illegal option\n", argv[i][k]);
                    return 1;
                }
                verbose = true;
            }
        } else {
            files.push_back(argv[i]);
        }
    }
    if (files.empty()) {
        fprintf(stderr, "Usage: LAsort [-v] <align:las> ...\n");
        return 1;
    }
    std::vector<char> in, out;
    for (const std::string& a : files) {
        const std::string pwd = lascli::path_to(a), root = lascli::root_of(a, ".las");
        const std::string name = pwd + "/" + root + ".las";
        FILE* f = fopen(name.c_str(), "r");
        if (!f) {
            fprintf(stderr, "LAsort: Cannot open %s for 'r'\n", name.c_str());
            return 1;
        }
        struct stat info;
        stat(name.c_str(), &info);
        long long novl = 0;
        int tspace = 0;
        if (fread(&novl, 8, 1, f) != 1 || fread(&tspace, 4, 1, f) != 1) {
            fprintf(stderr, "LAsort: System error, read failed!\n");
            return 1;
        }
        const long long size = (long long)info.st_size - 12;
        if (verbose) {
            printf("  %s: ", root.c_str());
            lascli::print_number(novl, stdout);
            printf(" records ");
            lascli::print_number((long long)info.st_size - novl * 40, stdout);
            printf(" trace bytes\n");
            fflush(stdout);
        }
        in.resize(std::max(size, 1LL));
        if (size > 0 && fread(in.data(), size, 1, f) != 1) {
            fprintf(stderr, "LAsort: System error, read failed!\n");
            return 1;
        }
        fclose(f);
        out.resize(std::max(size, 1LL));
        long long osize = 0;
        if (mi_las_sort(in.data(), size, novl, tspace, out.data(), &osize) != 0) {
            fprintf(stderr, "LAsort: %s: %s\n", name.c_str(), mi_last_error());
            return 1;
        }
        const std::string oname = pwd + "/" + root + ".S.las";
        FILE* g = fopen(oname.c_str(), "w");
        if (!g) {
            fprintf(stderr, "LAsort: Cannot open %s for 'w'\n", oname.c_str());
            return 1;
        }
        const bool ok = fwrite(&novl, 8, 1, g) == 1 && fwrite(&tspace, 4, 1, g) == 1 &&
                        (osize == 0 || fwrite(out.data(), osize, 1, g) == 1);
        if (fclose(g) != 0 || !ok) {
            fprintf(stderr, "LAsort: cannot write %s\n", oname.c_str());
            return 1;
        }
    }
    return 0;
}
