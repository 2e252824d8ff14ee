/* la_vectors: runs MARVEL's own Local_Alignment (dalign/align.c, compiled from
 * /root/reference by oracle/Makefile into oracle/_ref/la_vectors) over a task
 * file and writes the golden results.  TEST INFRASTRUCTURE ONLY.
 *
 * task file (little endian):
 *   int32 ntasks; double ave_corr; int32 tspace; float freq[4];
 *   int64 nbytes; uint8 seq[nbytes]          (bases 0..3, value 4 around every read)
 *   ntasks x { int64 aoff, boff; int32 alen, blen, comp, low, hgh, anti, lbord, hbord; }
 * result file:
 *   ntasks x { int32 abpos, bbpos, aepos, bepos, diffs, atlen,
 *              b_abpos, b_bbpos, b_aepos, b_bepos, b_diffs, btlen;
 *              uint16 atrace[atlen]; uint16 btrace[btlen]; }
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dalign/align.h"

static void rd(void *p, size_t n, FILE *f)
{
    if (fread(p, 1, n, f) != n) {
        fprintf(stderr, "la_vectors: short task file\n");
        exit(1);
    }
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: la_vectors <tasks.bin> <results.bin>\n");
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
        Path *bp = Local_Alignment(&al, work, spec, v[3], v[4], v[5], v[6], v[7]);
        int32_t r[12] = {apath.abpos, apath.bbpos, apath.aepos, apath.bepos, apath.diffs, apath.tlen,
                         bp->abpos,   bp->bbpos,   bp->aepos,   bp->bepos,   bp->diffs,   bp->tlen};
        fwrite(r, 4, 12, out);
        fwrite(apath.trace, 2, apath.tlen, out);
        fwrite(bp->trace, 2, bp->tlen, out);
    }
    fclose(out);
    fclose(in);
    return 0;
}
