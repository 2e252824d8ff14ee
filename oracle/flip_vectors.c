/* Test harness of this repo (not a reference file): reads commands on stdin,
 * runs the reference's Complement_Seq, Flip_Alignment and Print_Overlap
 * (dalign/align.c, linked in by oracle/Makefile) on them and prints the
 * results, for the golden vectors of the library's versions.
 *
 *   C n b0 .. b(n-1)                                  -> Complement_Seq
 *   F comp full alen blen abpos aepos bbpos bepos tlen t0 ..  -> Flip_Alignment
 *   P tbytes indent aread bread flags abpos aepos bbpos bepos diffs tlen t0 ..
 *                                                     -> Print_Overlap
 * Every result ends with a line "%". */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dalign/align.h"

int main(void)
{
    char op[4];
    while (scanf("%3s", op) == 1) {
        if (op[0] == 'C') {
            int n;
            if (scanf("%d", &n) != 1) return 1;
            char *a = malloc(n + 1);
            for (int i = 0; i < n; i++) {
                int b;
                if (scanf("%d", &b) != 1) return 1;
                a[i] = (char)b;
            }
            Complement_Seq(a, n);
            for (int i = 0; i < n; i++) printf("%d%c", a[i], i + 1 < n ? ' ' : '\n');
            if (n == 0) printf("\n");
            free(a);
        } else if (op[0] == 'F') {
            int comp, full, alen, blen, tlen;
            Path path;
            Alignment al;
            if (scanf("%d %d %d %d %d %d %d %d %d", &comp, &full, &alen, &blen, &path.abpos, &path.aepos, &path.bbpos,
                      &path.bepos, &tlen) != 9)
                return 1;
            int *t = malloc(sizeof(int) * (tlen + 1));
            for (int i = 0; i < tlen; i++)
                if (scanf("%d", t + i) != 1) return 1;
            char sa[1], sb[1];
            path.trace = t;
            path.tlen = tlen;
            path.diffs = 0;
            al.path = &path;
            al.flags = comp ? COMP_FLAG : 0;
            al.aseq = sa;
            al.bseq = sb;
            al.alen = alen;
            al.blen = blen;
            Flip_Alignment(&al, full);
            printf("%d %d %d %d %d %d %d", al.alen, al.blen, path.abpos, path.aepos, path.bbpos, path.bepos,
                   al.aseq == sb && al.bseq == sa);
            for (int i = 0; i < tlen; i++) printf(" %d", t[i]);
            printf("\n");
            free(t);
        } else if (op[0] == 'P') {
            int tbytes, indent, tlen;
            Overlap o;
            memset(&o, 0, sizeof o);
            unsigned flags;
            if (scanf("%d %d %d %d %u %d %d %d %d %d %d", &tbytes, &indent, &o.aread, &o.bread, &flags, &o.path.abpos,
                      &o.path.aepos, &o.path.bbpos, &o.path.bepos, &o.path.diffs, &tlen) != 11)
                return 1;
            o.flags = flags;
            o.path.tlen = tlen;
            void *t = malloc(2 * (tlen + 1));
            for (int i = 0; i < tlen; i++) {
                int v;
                if (scanf("%d", &v) != 1) return 1;
                if (tbytes == 1) ((uint8 *)t)[i] = (uint8)v;
                else ((uint16 *)t)[i] = (uint16)v;
            }
            o.path.trace = tlen > 0 ? t : NULL;
            Print_Overlap(stdout, &o, tbytes, indent);
            free(t);
        } else
            return 1;
        printf("%%\n");
    }
    return 0;
}
