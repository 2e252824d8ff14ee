/* libmarvl_align: the k-mer match filter of MARVEL daligner on the GPU.
 *
 * Drop-in for dalign/filter.h: a C host that links this library in place of
 * filter.o / align.o keeps calling Set_Filter_Params, Sort_Kmers and
 * Match_Filter with the reference's arguments, and the records reach the
 * Align_Spec's overlap buffers exactly as report_thread would have put them
 * there (Write_Overlap_Buffer then writes byte-identical .las files).
 *
 * The block types below have the byte layout of db/DB.h:164-234 (HITS_READ
 * 32 bytes, HITS_DB 88 bytes on x86-64), so blocks opened by the
 * reference's Open_DB_Block / Read_All_Sequences can be passed unchanged.
 */
#ifndef MARVL_FILTER_H
#define MARVL_FILTER_H

#include <stdint.h>

#include "marvl_align.h"

#ifdef __cplusplus
extern "C" {
#endif

/* db/DB.h:164-171 */
typedef struct {
    int rlen;       /* length of the read */
    int64_t boff;   /* offset of the read's first base in HITS_DB.bases */
    int64_t coff;
    int flags;
} HITS_READ;

/* db/DB.h:179-185.  The filter reads the first record of HITS_DB.tracks as
 * the soft mask of daligner -m, as read_DB leaves it
 * (dalign/daligner.c:373-431): anno = int64[nreads+1] offsets counted in
 * points (the file's byte offsets divided by 4), data = int32 points, read
 * i's points at data[anno[i] .. anno[i+1]) in pairs (begin, end) of masked
 * intervals, non-decreasing and within [0, rlen]. */
typedef struct HITS_TRACK_ {
    struct HITS_TRACK_ *next;
    char *name;
    int size;
    void *anno;
    void *data;
} HITS_TRACK;

/* db/DB.h:213-234 */
typedef struct {
    int ureads;
    float freq[4];
    int maxlen;
    int64_t totlen;
    int nreads;
    int part;
    int ufirst;
    char *path;
    int loaded;
    void *bases;        /* bytes 0..3 per base, with a 4 before and after every read */
    HITS_READ *reads;   /* [0..nreads) */
    void *tracks;       /* HITS_TRACK list; the first one masks k-mers (-m) */
} HITS_DB;

/* Driver-owned parameters read by the filter (dalign/filter.h:54-73,
 * defined at dalign/daligner.c:60-71).  The library defines them; a host
 * that defines them itself takes precedence at link time. */
extern int BIASED;
extern int VERBOSE;
extern int MINOVER;
extern int HGAP_MIN;
extern int SYMMETRIC;
extern int IDENTITY;
extern uint64_t MEM_LIMIT;
extern uint64_t MEM_PHYSICAL;
extern int NTHREADS;   /* report chunks (-j): fixes which pairs are skipped, so it is part of the output */
extern int NSHIFT;     /* log2(NTHREADS) */

/* Set_Filter_Params: replaces dalign/filter.c:119-139 (dalign/filter.h:67).
 * Returns 1 for kmer <= 1, and for kmer > 31 (undefined in the reference).
 * A block whose keys do not fit 64 bits (2 kmer + bits of its read count
 * + bits of its longest read > 64) is an error of Sort_Kmers. */
int Set_Filter_Params(int kmer, int binshift, int suppress, int hitmin);

/* Sort_Kmers: replaces dalign/filter.c:698-903 (dalign/filter.h:69).
 * Builds the sorted k-mer index of `block` on the GPU and returns an opaque
 * handle (the block's 2-bit packed reads stay on the device with it);
 * *len = number of k-mers after -t suppression.  With block->tracks set,
 * only k-mers wholly outside the first track's intervals are indexed
 * (tuple_thread's masked branch, dalign/filter.c:416-470); a track whose
 * points are not increasing pairs within the reads is an error.  With
 * BIASED (-b) the k-mers are biased_tuple_thread's (dalign/filter.c:492-633),
 * weighted by the base frequencies of the first block indexed in the
 * process, as the reference computes them once.  The handle is released by Match_Filter when passed as its
 * btable with atable != btable (the reference frees its B index there,
 * dalign/filter.c:2194-2202), otherwise by mi_free_kmer_index. */
void *Sort_Kmers(HITS_DB *block, int *len);

/* Match_Filter: replaces dalign/filter.c:2056-2407 (dalign/filter.h:71-73).
 * Self comparison is decided like the reference: aname == bname as
 * pointers.  The records of both views are appended to the overlap buffer
 * of `asettings`.  Errors print a message and exit(1) like the reference's
 * batch mode; mi_match_filter is the returning variant. */
void Match_Filter(char *aname, HITS_DB *ablock, char *bname, HITS_DB *bblock, void *atable, int alen, void *btable,
                  int blen, int comp, Align_Spec *asettings);

/* ---- no reference counterpart ------------------------------------------- */

/* Statistics of the last Match_Filter / mi_match_filter call. */
typedef struct {
    long long nkmers_a, nkmers_b;   /* index sizes */
    long long nhits;                /* seeds after the -M cap */
    long long npairs;               /* read pairs with seeds */
    long long njobs;                /* pairs report_thread would scan */
    long long naligned;             /* Local_Alignment calls */
    long long nrecords;             /* records appended (both views) */
    long long cells, snake;         /* aligner work: wave widths, snake compare steps */
    long long edge_events;          /* band array accesses beyond [mindiag, maxdiag] */
    long long edge_fixups;          /* pairs re-run for the band cell shared across pairs */
    long long retries;              /* pairs re-run with wider windows / larger arenas */
    int limit;                      /* k-mer match cap from MEM_LIMIT */
    double t_filter, t_report, t_copy;   /* seconds (host clock): merge + seed sort, report launches, copy-out */
    double t_total;
    double t_report_kernel;         /* seconds in the report kernel (HIP events on the filter stream) */
    long long report_launches;
    double t_seed_sort;             /* seconds in the seed radix sort (HIP events) */
    long long seed_sort_passes;     /* 8-bit digit passes of that sort */
    long long njobs_kept;           /* pairs past the exact gate pre-check (the report kernel's jobs) */
    long long wave_steps;           /* aligner wave steps (one per difference level and pass) */
    long long cand_scored;          /* read pairs the scorer evaluated (gating mode) */
    long long cand_skipped;         /* read pairs it dropped before alignment */
    /* with MI_REDUND_STATS=1 in the environment: how often Handle_Redundancies
     * (dalign/filter.c:1403-1566) took each branch, final runs only --
     * 0-2: alignment j starts first and k reaches further, fused (A and B
     * views with COMP / without / A view only); 4: j contains k; 5: same
     * start, j longer; 6-8: k starts first and j reaches further, fused (as
     * 0-2); 3 unused */
    long long redund[9];
    /* radix sorts and scans recorded as HIP graphs by this pass, replayed
     * from earlier passes, and the kernel launches those replays stood for */
    long long graphs_captured, graphs_replayed, launches_replayed;
} MiFilterStats;

/* Match_Filter returning an error code (MI_ERR_*) instead of exiting. */
int mi_match_filter(char *aname, HITS_DB *ablock, char *bname, HITS_DB *bblock, void *atable, void *btable,
                    int comp, Align_Spec *asettings);
void mi_filter_stats(MiFilterStats *out);

/* ---- candidate scorer (no reference counterpart) ---------------------------
 * An MLP 16 -> 64 -> 64 -> 1 (ReLU; bf16 weights and activations, fp32
 * accumulation, gfx950 MFMA) over 16 features of each alignment candidate:
 * a read pair that passes the exact gate pre-check of report_thread's band
 * gate (dalign/filter.c:1740-1778), described by its seed count and its
 * strongest diagonal bins (seeds, apos span, mean / sd of the seed spacing,
 * read lengths, the overlap the bin's diagonal implies, short-period k-mers).
 * Weights are input-major: w1[16*64], w2[64*64], w3[64].  Mode 0 aligns every
 * pair (bit-exact with the reference); 1 does the same and records each
 * pair's features, the records its exact run wrote and the aligner work it
 * took; 2 drops the pairs whose logit is below the threshold before
 * alignment (every kept pair runs the exact report kernel). */
int mi_scorer_load(const float *w1, const float *b1, const float *w2, const float *b2, const float *w3, float b3,
                   float threshold);
int mi_scorer_mode(int mode);
long long mi_scorer_captured(void);
/* copies up to cap captured rows (16 floats, the record count and the work,
 * band cells + snake steps, of each) and clears the store */
long long mi_scorer_take(float *feat, int *nrec, unsigned long long *work, long long cap);
/* logits of n feature rows with the loaded weights, on the GPU */
int mi_mlp_score(const float *feat, long long n, float *logits);

/* Training on the GPU (bf16 MFMA forward and backward, fp32 Adam).  The
 * parameter vector holds 5313 floats: w1[16*64], b1[64], w2[64*64], b2[64],
 * w3[64], b3 (input-major).  Loss: binary cross-entropy on the logits with
 * per-row weights, averaged by the weights.
 * mi_scorer_grad: gradient of that loss over all n rows at `params` (and the
 * loss); mi_scorer_train: `epochs` passes of Adam (lr; beta1 0.9, beta2
 * 0.999) over shuffled minibatches of `batch` rows, params updated in place,
 * the weighted mean loss of each epoch in epoch_loss (optional). */
int mi_scorer_grad(const float *feat, const float *label, const float *weight, long long n, const float *params,
                   float *grad, float *loss);
int mi_scorer_train(const float *feat, const float *label, const float *weight, long long n, int epochs, int batch,
                    float lr, unsigned long long seed, float *params, float *epoch_loss);

/* Index of the reverse complement of the block indexed by `table`, built on
 * the GPU from the packed reads the handle holds (the complement the
 * driver makes with complement_DB, dalign/daligner.c:456-573, followed by
 * Sort_Kmers).  `cblock` describes the complemented block (read lengths,
 * ufirst, maxlen); its bases are not read. */
void *mi_sort_kmers_complement(void *table, HITS_DB *cblock, int *len);

/* Sort_Kmers in two steps, for callers that keep a block resident in HBM:
 * mi_block_device uploads and 2-bit packs the block's reads (no k-mers);
 * mi_sort_kmers_device builds a k-mer index of such a block without
 * touching host memory.  Both handles are released by mi_free_kmer_index;
 * only the second is a table for Match_Filter. */
void *mi_block_device(HITS_DB *block);
void *mi_sort_kmers_device(void *dblock, int *len);
/* Waits for all work of this process on the GPU. */
int mi_device_synchronize(void);
void mi_free_kmer_index(void *table);

/* Block loading for hosts without the reference's DB library: the
 * behaviour of Open_DB_Block + Read_All_Sequences (db/DB.c:360-585,
 * :1349-1410) for "DIR/ROOT.<block>" or "ROOT[.db|.dam]" (the stub
 * ROOT.db, else ROOT.dam); trimmed DBs are not supported.  Returns 0 or
 * MI_ERR_ARG. */
int mi_open_block(const char *path, HITS_DB *db);
void mi_close_block(HITS_DB *db);
/* Header of the block's reverse complement for mi_sort_kmers_complement
 * (read records shared with src, no bases, freq swapped as complement_DB
 * swaps it, dalign/daligner.c:456-573).  When src has a mask track, dst gets
 * its complement (each read's points reversed, x -> rlen - x,
 * dalign/daligner.c:519-560), released by mi_release_tracks(dst). */
void mi_complement_header(const HITS_DB *src, HITS_DB *dst);

/* The soft-mask tracks of daligner -m for a block opened by mi_open_block:
 * read_DB's track part (dalign/daligner.c:373-431).  Each name is loaded
 * from the MARVEL layout (.ROOT.NAME.a2/.d2, zlib chunks; lib/tracks.c:19-170)
 * or else the Dazzler one (.ROOT.<block>.NAME.anno/.data, then
 * .ROOT.NAME.anno/.data from the block's first read on; db/DB.c:967-1120);
 * n > 1 tracks are merged into their union (Merge_Tracks,
 * dalign/daligner.c:275-370).  db->tracks is set to the one resulting track.
 * The reference trims a MARVEL-layout track of a block whose first read is
 * beyond 1 twice, reading outside the block's slice; this loader takes the
 * block's own slice.  Returns 0 or MI_ERR_ARG (message in mi_last_error). */
int mi_load_mask_tracks(HITS_DB *db, const char *const *names, int n);
/* LAsort on the GPU: the sort and output loop of dalign/LAsort.c:183-235
 * (records ordered by (aread, bread, abpos), SORT_OVL at :60-86; equal keys
 * keep their input order, as under glibc's merge-sort qsort).  `body` is
 * the `size` bytes after a .las header holding `novl` records; tspace <= 125
 * means 1-byte trace values, else 2.  The sorted records go to `out` (at
 * least *out_size bytes: the records' bytes, without anything after the
 * last one).  Returns 0, MI_ERR_ARG when the records do not fit in `size`,
 * or MI_ERR_DEVICE. */
int mi_las_sort(const void *body, long long size, long long novl, int tspace, void *out, long long *out_size);
/* LAmerge on the GPU: the order of utils/LAmerge.c over the records of all
 * its inputs, concatenated in input order -- (aread, bread, COMP, abpos)
 * (SORT_OVL, utils/LAmergeUtils.c:50-80), equal keys in input order.  That
 * is the reference's output for sorted inputs whose equal keys lie in one
 * file (its heap merge, :97-141; block merges, where each file holds other
 * B reads) and for -s (stable qsort of the concatenation).  Arguments as
 * mi_las_sort. */
int mi_las_merge(const void *body, long long size, long long novl, int tspace, void *out, long long *out_size);

/* DBdust on the GPU: the SDUST scan of db/DBdust.c:236-470 over every read
 * of `db` (bases 0..3 at db->bases + reads[i].boff, as Read_All_Sequences
 * leaves them).  Unbiased scores: reads cut into segments of
 * MI_DUST_SEGMENT steps (default max(384, 4 window)), one lane each after
 * a warm-up of MI_DUST_WARMUP steps (default max(64, 2 window)); segments
 * that began in another state than their predecessor ended in are re-run;
 * MI_DUST_STATS prints the counts to stderr.  Biased scores and windows
 * below 3: one lane per read.  window (-w, 1..256), threshold (-t),
 * mlen (-m less one: intervals of at least mlen are kept), biased (-b, with
 * db->freq).  Writes anno[0..nreads] (byte offsets, as in the .dust.anno
 * file) and the points (pairs begin, end+1; at most the block's total read
 * length of them).  Returns 0, MI_ERR_ARG or MI_ERR_DEVICE. */
int mi_dust(const HITS_DB *db, int window, double threshold, int mlen, int biased, long long *anno, int *points);

/* Frees the tracks this library allocated for db (mi_load_mask_tracks,
 * mi_complement_header) and clears db->tracks; mi_close_block calls it. */
void mi_release_tracks(HITS_DB *db);

#ifdef __cplusplus
}
#endif
#endif
