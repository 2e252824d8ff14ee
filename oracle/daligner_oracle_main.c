/* Oracle CLI: `daligner_oracle` with the option set and output naming of
 * MARVEL dalign/daligner (dalign/daligner.c:615-1032,
 * dalign/align.c:5815-6043).  TEST INFRASTRUCTURE ONLY.
 *
 *   daligner_oracle [-vAT] [-k] [-w] [-h] [-t] [-M] [-e] [-l] [-s] [-H] [-r] [-j]
 *                   <A.block> <B.block> ...
 * -m (mask tracks), -b (biased), -I/-O (identity) and -D are rejected.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>
#include <time.h>

#include "dal_oracle.h"

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* Root(): file name without directory and without ".db" */
static void root_of(const char *path, char *out, size_t n)
{
    const char *s = strrchr(path, '/');
    snprintf(out, n, "%s", s ? s + 1 : path);
    size_t l = strlen(out);
    if (l > 3 && strcmp(out + l - 3, ".db") == 0) out[l - 3] = 0;
}

/* split "ROOT.<id>" into root and id (Write_Overlap_Buffer's parse) */
static int block_id(const char *name, char *root, size_t n)
{
    const char *dot = strrchr(name, '.');
    if (dot == NULL) {
        snprintf(root, n, "%s", name);
        return 0;
    }
    snprintf(root, n, "%.*s", (int)(dot - name), name);
    return atoi(dot + 1);
}

static void subdir(int run, int id, char *out, size_t n)
{
    if (id == 0) snprintf(out, n, ".");
    else snprintf(out, n, "%03d_%05d", run, id);
}

int main(int argc, char **argv)
{
    or_opts o;
    char **masks = NULL;
    int nmask = 0;
    int c, l = 1000;
    or_opts_default(&o);
    o.mem_limit = (uint64_t)sysconf(_SC_PHYS_PAGES) * (uint64_t)sysconf(_SC_PAGESIZE);
    while ((c = getopt(argc, argv, "vbOTAIk:w:h:t:M:e:l:s:H:D:m:r:j:")) != -1) {
        switch (c) {
        case 'v': o.verbose += 1; break;
        case 'b': o.biased = 1; break;
        case 'I': o.identity = 1; break;
        case 'O': o.identity = 1; o.only_identity = 1; break;
        case 'T': o.no_trace = 1; break;
        case 'A': o.symmetric = 0; break;
        case 'k': o.kmer = atoi(optarg); break;
        case 'w': o.binshift = atoi(optarg); break;
        case 'h': o.hitmin = atoi(optarg); break;
        case 't': o.suppress = atoi(optarg); break;
        case 'H':
            /* the reference parses -H into a local of main that shadows the
             * global HGAP_MIN the filter reads (dalign/daligner.c:641,
             * :709-711), so the option has no effect on the output */
            if (atoi(optarg) < 0) {
                fprintf(stderr, "invalid HGAP threshold of %d\n", atoi(optarg));
                return 1;
            }
            break;
        case 'e': o.ave_corr = atof(optarg); break;
        case 'l': l = atoi(optarg); break;
        case 's': o.tspace = atoi(optarg); break;
        case 'r': o.run_id = atoi(optarg); break;
        case 'm':
            masks = (char **)realloc(masks, sizeof(char *) * (nmask + 1));
            masks[nmask++] = optarg;
            break;
        case 'M': o.mem_limit = (uint64_t)atoi(optarg) * 0x40000000ULL; break;
        case 'j': {
            int t = atoi(optarg), e = 0;
            while ((1 << (e + 1)) <= t) e++;
            if (t < 1 || (1 << e) != t) {
                fprintf(stderr, "invalid number of threads: %d. Must be a power of 2.\n", t);
                return 1;
            }
            o.nthreads = t;
            o.nshift = e;
            break;
        }
        default:
            fprintf(stderr, "daligner_oracle: option -%c not supported\n", optopt ? optopt : c);
            return 1;
        }
    }
    o.minover = 2 * l;
    if (optind + 2 > argc) {
        fprintf(stderr, "usage: daligner_oracle [options] <A.block> <B.block> ...\n");
        return 1;
    }
    const char *afile = argv[optind];
    or_block A;
    char err[512];
    if (or_block_load(&A, afile)) {
        fprintf(stderr, "cannot open %s\n", afile);
        return 1;
    }
    if (nmask && or_block_load_masks(&A, masks, nmask, err, sizeof err)) {
        printf("[ERROR] - Unable to load track: %s\n", err);
        return 1;
    }
    for (int i = 0; i < A.nreads; i++)
        if (A.rlen[i] < o.kmer) {
            fprintf(stderr, "[ERROR] - daligner: Block %s contains reads < %dbp long !  Run DBsplit.\n", afile, o.kmer);
            return 1;
        }
    char aroot[4096], aname[4096], dir[64];
    root_of(afile, aroot, sizeof aroot);
    int aid = block_id(aroot, aname, sizeof aname);
    subdir(o.run_id, A.part, dir, sizeof dir);
    mkdir(dir, 0755);
    or_spec *sp = (or_spec *)malloc(sizeof(or_spec));
    or_spec_init(sp, o.ave_corr, o.tspace, A.freq);

    for (int bi = optind + 1; bi < argc; bi++) {
        const char *bfile = argv[bi];
        int self = strcmp(afile, bfile) == 0;
        or_block B;
        or_ovlset out;
        or_stats st;
        memset(&out, 0, sizeof out);
        memset(&st, 0, sizeof st);
        char broot[4096], bname[4096];
        root_of(bfile, broot, sizeof broot);
        int bid = block_id(broot, bname, sizeof bname);
        if (!self) {
            if (or_block_load(&B, bfile)) {
                fprintf(stderr, "cannot open %s\n", bfile);
                return 1;
            }
            if (nmask && or_block_load_masks(&B, masks, nmask, err, sizeof err)) {
                printf("[ERROR] - Unable to load track: %s\n", err);
                return 1;
            }
            if (strcmp(aname, bname) != 0) o.symmetric = 0;
            if (o.symmetric) {
                char d2[64];
                subdir(o.run_id, B.part, d2, sizeof d2);
                mkdir(d2, 0755);
            }
        }
        double t0 = now();
        or_daligner_pair(&o, sp, &A, self ? &A : &B, self, &out, &st);
        if (o.only_identity) {   /* -O: only a read's overlaps with itself are written (dalign/align.c:5786-5800) */
            int64_t k = 0;
            for (int64_t j = 0; j < out.n; j++)
                if (out.ovl[j].aread == out.ovl[j].bread) out.ovl[k++] = out.ovl[j];
            out.n = k;
        }
        or_ovlset_sort(&out);
        double t1 = now();
        char p1[9000], p2[9000], d1[64], d2[64];
        if (self || !o.symmetric) {
            subdir(o.run_id, aid, d1, sizeof d1);
            if (aid > 0 && bid > 0) snprintf(p1, sizeof p1, "%s/%s.%d.%s.%d.las", d1, aname, aid, bname, bid);
            else if (aid > 0) snprintf(p1, sizeof p1, "%s/%s.%d.%s.las", d1, aname, aid, bname);
            else if (bid > 0) snprintf(p1, sizeof p1, "%s.%s.%d.las", aname, bname, bid);
            else snprintf(p1, sizeof p1, "%s.%s.las", aname, bname);
            or_write_las(p1, &out, 0, out.n, o.tspace);
        } else {
            subdir(o.run_id, aid, d1, sizeof d1);
            subdir(o.run_id, bid, d2, sizeof d2);
            snprintf(p1, sizeof p1, "%s/%s.%d.%s.%d.las", d1, aname, aid, bname, bid);
            snprintf(p2, sizeof p2, "%s/%s.%d.%s.%d.las", d2, bname, bid, aname, aid);
            if (bid < aid) {
                char t[9000];
                memcpy(t, p1, sizeof t);
                memcpy(p1, p2, sizeof t);
                memcpy(p2, t, sizeof t);
            }
            int last = (B.part < A.part) ? B.ufirst + B.nreads - 1 : A.ufirst + A.nreads - 1;
            int64_t j = 0;
            while (j < out.n && out.ovl[j].aread <= last) j++;
            or_write_las(p1, &out, 0, j, o.tspace);
            or_write_las(p2, &out, j, out.n, o.tspace);
        }
        if (o.verbose)
            printf("{\"records\": %lld, \"seconds\": %.3f, \"kmers_a\": %lld, \"kmers_b\": %lld, \"hits\": %lld, "
                   "\"pairs\": %lld, \"seeds\": %lld, \"cells\": %lld, \"snake\": %lld, \"limit\": %d, "
                   "\"edge_events\": %lld}\n",
                   (long long)out.n, t1 - t0, (long long)st.nkmers_a, (long long)st.nkmers_b,
                   (long long)st.nhits, (long long)st.npairs, (long long)st.nseeds, (long long)st.cells,
                   (long long)st.snake, st.limit, (long long)st.edge_events);
        or_ovlset_free(&out);
        if (!self) or_block_free(&B);
    }
    free(sp);
    or_block_free(&A);
    return 0;
}
