/* trace_vectors: runs MARVEL's own Local_Alignment and Compute_Trace_PTS
 * (dalign/align.c, compiled from /root/reference by oracle/Makefile into
 * oracle/_ref/trace_vectors) over a la_vectors task file and writes, per
 * task and per mode (LOWERMOST, GREEDIEST, UPPERMOST), the difference count
 * and the trace.  TEST INFRASTRUCTURE ONLY.
 *
 * result file: ntasks x { int32 status (0: traced), int32 a_tlen;
 *                         3 x { int32 diffs, ntrace; int32 trace[ntrace] } }
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dalign/align.h"

static void rd(void *p, size_t n, FILE *f)
{
    if (fread(p, 1, n, f) != n) {
        fprintf(stderr, "trace_vectors: short task file\n");
        exit(1);
    }
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: trace_vectors <tasks.bin> <results.bin>\n");
        return 1;
    }
    FILE *in = fopen(argv[1], "rb"), *out = fopen(argv[2], "wb");
    if (in == NULL || out == NULL) return 1;
    int32_t n, ts;
    double corr;
    float freq[4];
    int64_t nb;
    rd(&n, 4, in);
    rd(&corr, 8, in);
    rd(&ts, 4, in);
    rd(freq, 16, in);
    rd(&nb, 8, in);
    char *seq = (char *)malloc(nb + 16);
    rd(seq, nb, in);
    Align_Spec *spec = New_Align_Spec(corr, ts, freq, 1, 1, 0, 0);
    Work_Data *work = New_Work_Data();
    uint16_t *keep = NULL;
    int kcap = 0;
    for (int i = 0; i < n; i++) {
        int64_t ao, bo;
        int32_t v[8];
        rd(&ao, 8, in);
        rd(&bo, 8, in);
        rd(v, 32, in);
        Path apath;
        Alignment al;
        al.path = &apath;
        al.flags = v[2] ? COMP_FLAG : 0;
        al.aseq = seq + ao;
        al.bseq = seq + bo;
        al.alen = v[0];
        al.blen = v[1];
        Local_Alignment(&al, work, spec, v[3], v[4], v[5], v[6], v[7]);
        Path saved = apath;
        if (saved.tlen > kcap) {
            kcap = saved.tlen * 2 + 64;
            keep = (uint16_t *)realloc(keep, sizeof(uint16_t) * kcap);
        }
        memcpy(keep, apath.trace, sizeof(uint16_t) * saved.tlen);
        int32_t head[2] = {0, saved.tlen};
        fwrite(head, 4, 2, out);
        for (int mode = -1; mode <= 1; mode++) {
            apath = saved;
            apath.trace = keep;
            Compute_Trace_PTS(&al, work, ts, mode, 0);
            int32_t r[2] = {apath.diffs, apath.tlen};
            fwrite(r, 4, 2, out);
            fwrite(apath.trace, 4, apath.tlen, out);
        }
    }
    fclose(out);
    fclose(in);
    return 0;
}
