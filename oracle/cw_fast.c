/* Test infrastructure only: a CPU restatement of ClustalW's k-tuple distances
 * (pairwise/FastPairwiseAlign.cpp, the -QUICKTREE path), the checker for the
 * device kernel cw_fast_pairs.  Nothing in the library links this file.
 *
 * Parity with the reference is unpinned beyond closed forms: the reference
 * ships no k-tuple (quick tree) output, so this restatement is checked only
 * against what the algorithm fixes by itself (identical sequences score 100,
 * sequences without a common k-tuple score 0, the DNA mapping's codes) and
 * otherwise follows the reference's source line by line as cited. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cw_oracle.h"

/* FastPairwiseAlign.cpp:147-168: protein codes 0 (A), 2 (C), 6 (G), 18 (T), 19 (U)
 * to DNA 0..3; every other code breaks a k-tuple */
static const int kAA2DNA[36] = {0, -1, 1, -1, -1, -1, 2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                                3, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};

typedef struct {
    int dna, ktup, window, wind_gap, signif, percent, max_aa, max_aln;
    int *zza, *zzb, *zzc, *zzd;        /* tptr / pl of sequence i and j, 1-based */
    int *displ, *slopes, *diag;        /* 1 .. 2 max_aln */
    int *acc[5];                       /* accum rows 0..4, 1 .. 2 max_aln */
    int curr, maxsofar, next;
} fast_t;

static int residue_of(const fast_t *f, int code)
{
    if (f->dna) return (code >= 0 && code < 36) ? kAA2DNA[code] : -1;
    return code;
}

static int max_res(const fast_t *f) { return f->dna ? 3 : f->max_aa; }

/* makePPtrs / makeNPtrs (:404-521): per k-tuple code the last position, and per
 * position the previous one with the same code.  seq is 0-based here;
 * the reference's (*seq)[p] is seq[p - 1]. */
static void make_ptrs(const fast_t *f, int *tptr, int *pl, const int *seq, int length)
{
    int pot[10], i, j, limit = 1;
    for (i = 1; i <= f->ktup; i++) {
        pot[i] = limit;
        limit *= max_res(f) + 1;
    }
    for (i = 1; i <= limit; i++) pl[i] = 0;
    for (i = 1; i <= length; i++) tptr[i] = 0;
    for (i = 1; i <= length - f->ktup + 1; i++) {
        int code = 0, bad = 0;
        for (j = 1; j <= f->ktup; j++) {
            const int r = residue_of(f, seq[i + j - 2]);
            if (r < 0 || r > max_res(f)) { bad = 1; break; }
            code += r * pot[j];
        }
        if (bad) continue;
        ++code;
        if (pl[code] != 0) tptr[i] = pl[code];
        pl[code] = i;
    }
}

/* desQuickSort (:584-659): ascending on a1 with a2 carried along, over
 * elements 1 .. n-1 (the reference's upper bound leaves element n in place) */
static void des_quick_sort(int *a1, int *a2, int n)
{
    int lst[50], ust[50], p = 1, i, j, piv, t;
    lst[1] = 1;
    ust[1] = n - 1;
    while (p > 0) {
        if (lst[p] >= ust[p]) {
            p--;
            continue;
        }
        i = lst[p] - 1;
        j = ust[p];
        piv = a1[j];
        while (i < j) {
            for (i = i + 1; a1[i] < piv; i++)
                ;
            for (j = j - 1; j > i; j--)
                if (a1[j] <= piv) break;
            if (i < j) {
                t = a1[i]; a1[i] = a1[j]; a1[j] = t;
                t = a2[i]; a2[i] = a2[j]; a2[j] = t;
            }
        }
        j = ust[p];
        t = a1[i]; a1[i] = a1[j]; a1[j] = t;
        t = a2[i]; a2[i] = a2[j]; a2[j] = t;
        if (i - lst[p] < ust[p] - i) {
            lst[p + 1] = lst[p];
            ust[p + 1] = i - 1;
            lst[p] = i + 1;
        } else {
            lst[p + 1] = i + 1;
            ust[p + 1] = ust[p];
            ust[p] = i - 1;
        }
        p = p + 1;
    }
}

/* putFrag (:523-559): the new fragment enters the list kept in descending
 * score order through accum[4] */
static void put_frag(fast_t *f, int fs, int v1, int v2, int flen)
{
    const int c = f->curr;
    f->acc[0][c] = fs;
    f->acc[1][c] = v1;
    f->acc[2][c] = v2;
    f->acc[3][c] = flen;
    if (!f->maxsofar) {
        f->maxsofar = 1;
        f->acc[4][c] = 0;
        return;
    }
    if (fs >= f->acc[0][f->maxsofar]) {
        f->acc[4][c] = f->maxsofar;
        f->maxsofar = c;
        return;
    }
    int end;
    f->next = f->maxsofar;
    for (;;) {
        end = f->next;
        f->next = f->acc[4][f->next];
        if (fs >= f->acc[0][f->next]) break;
    }
    f->acc[4][c] = f->next;
    f->acc[4][end] = c;
}

/* fragRelPos (:561-582) */
static int frag_rel_pos(const fast_t *f, int a1, int b1, int a2, int b2)
{
    if (a1 - b1 == a2 - b2) return a2 < a1;
    return a2 + f->ktup - 1 < a1 && b2 + f->ktup - 1 < b1;
}

/* pairAlign (:171-402) of sequence 1 (seq, l1; pointers zza/zzc) against
 * sequence 2 (l2; pointers zzb/zzd) */
static void pair_align(fast_t *f, const int *seq, int l1, int l2)
{
    int pot[10], i, j, l, m, limit = 1;
    for (i = 1; i <= f->ktup; i++) {
        pot[i] = limit;
        limit *= max_res(f) + 1;
    }
    const int tl1 = l1 + l2 - 1;
    for (i = 1; i <= tl1; i++) {
        f->slopes[i] = f->displ[i] = 0;
        f->diag[i] = i;
    }
    for (i = 1; i <= limit; i++)                 /* k-tuple matches per diagonal */
        for (int vn1 = f->zzc[i]; vn1; vn1 = f->zza[vn1])
            for (int vn2 = f->zzd[i]; vn2; vn2 = f->zzb[vn2]) ++f->displ[vn1 - vn2 + l2];
    des_quick_sort(f->displ, f->diag, tl1);
    j = tl1 - f->signif + 1;
    if (j < 1) j = 1;
    for (i = tl1; i >= j; i--)                   /* WINDOW around the top SIGNIF diagonals */
        if (f->displ[i] > 0) {
            const int pos = f->diag[i];
            l = 1 > pos - f->window ? 1 : pos - f->window;
            m = tl1 < pos + f->window ? tl1 : pos + f->window;
            for (; l <= m; l++) f->slopes[l] = 1;
        }
    for (i = 1; i <= tl1; i++) f->displ[i] = 0;
    f->curr = f->maxsofar = 0;
    for (i = 1; i <= l1 - f->ktup + 1; i++) {
        int encrypt = 0, bad = 0;
        for (j = 1; j <= f->ktup; j++) {
            const int r = residue_of(f, seq[i + j - 2]);
            if (r < 0 || r > max_res(f)) { bad = 1; break; }
            encrypt += r * pot[j];
        }
        if (bad) continue;
        ++encrypt;
        for (int vn2 = f->zzd[encrypt]; vn2; vn2 = f->zzb[vn2]) {
            const int osptr = i - vn2 + l2;
            if (f->slopes[osptr] != 1) continue;
            int flen = 0, fs = f->ktup;
            f->next = f->maxsofar;
            for (;;) {                           /* A-loop (:299-393) */
                if (!f->next) {
                    ++f->curr;
                    if (f->curr >= 2 * f->max_aln) return;    /* "(Partial alignment)" */
                    f->displ[osptr] = f->curr;
                    put_frag(f, fs, i, vn2, flen);
                    break;
                }
                const int nx = f->next;
                if (frag_rel_pos(f, i, vn2, f->acc[1][nx], f->acc[2][nx])) {
                    if (i - vn2 == f->acc[1][nx] - f->acc[2][nx]) {
                        if (i > f->acc[1][nx] + (f->ktup - 1)) fs = f->acc[0][nx] + f->ktup;
                        else fs = f->acc[0][nx] + (i - f->acc[1][nx]);
                    } else {
                        int subt1;
                        const int d = f->displ[osptr];
                        if (d == 0) subt1 = f->ktup;
                        else if (i > f->acc[1][d] + (f->ktup - 1)) subt1 = f->acc[0][d] + f->ktup;
                        else subt1 = f->acc[0][d] + (i - f->acc[1][d]);
                        const int subt2 = f->acc[0][nx] - f->wind_gap + f->ktup;
                        if (subt2 > subt1) {
                            flen = nx;
                            fs = subt2;
                        } else {
                            flen = d;
                            fs = subt1;
                        }
                        f->next = 0;
                        continue;
                    }
                    flen = nx;
                    f->next = 0;
                    continue;
                }
                f->next = f->acc[4][nx];
            }
        }
    }
}

/* FastPairwiseAlign::pairwiseAlign (:20-145): sequences i = iStart+1 .. iEnd
 * against j = max(jStart+2, i+1) .. jEnd (1-based); dist[k] for the k-th pair
 * in that order, pairs[2k], pairs[2k+1] the 0-based indices.  Returns the pair
 * count, or -1 on the reference's fatal range checks. */
long cwo_fast_pairwise(int nseq, const int *const *seqs, const int *lens, int dna, int ktup,
                       int window, int wind_gap, int signif, int percent, int max_aln, int iStart,
                       int iEnd, int jStart, int jEnd, int *pairs, double *dist)
{
    if (iStart < 0 || iEnd < iStart || jStart < 0 || jEnd < jStart || iEnd > nseq || jEnd > nseq)
        return -1;
    fast_t f;
    memset(&f, 0, sizeof f);
    f.dna = dna;
    f.ktup = ktup;
    f.window = window;
    f.wind_gap = wind_gap;
    f.signif = signif;
    f.percent = percent;
    f.max_aa = 24;
    f.max_aln = max_aln;
    int limit = 1, lmax = 0;
    for (int q = 0; q < ktup; q++) limit *= (dna ? 3 : 24) + 1;
    for (int s = 0; s < nseq; s++) lmax = lens[s] > lmax ? lens[s] : lmax;
    const size_t num = (size_t)2 * max_aln + 1;
    const size_t dnum = num > (size_t)2 * lmax + 2 ? num : (size_t)2 * lmax + 2;
    f.zza = (int *)calloc((size_t)lmax + 2, sizeof(int));
    f.zzb = (int *)calloc((size_t)lmax + 2, sizeof(int));
    f.zzc = (int *)calloc((size_t)limit + 2, sizeof(int));
    f.zzd = (int *)calloc((size_t)limit + 2, sizeof(int));
    f.displ = (int *)calloc(dnum, sizeof(int));
    f.slopes = (int *)calloc(dnum, sizeof(int));
    f.diag = (int *)calloc(dnum, sizeof(int));
    for (int r = 0; r < 5; r++) f.acc[r] = (int *)calloc(num + 1, sizeof(int));
    long k = 0;
    for (int i = iStart + 1; i <= iEnd; ++i) {
        make_ptrs(&f, f.zza, f.zzc, seqs[i - 1], lens[i - 1]);
        for (int j = (jStart + 2 > i + 1 ? jStart + 2 : i + 1); j <= jEnd; ++j) {
            make_ptrs(&f, f.zzb, f.zzd, seqs[j - 1], lens[j - 1]);
            pair_align(&f, seqs[i - 1], lens[i - 1], lens[j - 1]);
            double score;
            if (!f.maxsofar) {
                score = 0.0;
            } else {
                score = (double)f.acc[0][f.maxsofar];
                if (percent) {
                    const int dsr = lens[i - 1] < lens[j - 1] ? lens[i - 1] : lens[j - 1];
                    score = (score / (double)dsr) * 100.0;
                }
            }
            if (pairs) {
                pairs[2 * k] = i - 1;
                pairs[2 * k + 1] = j - 1;
            }
            if (dist) dist[k] = (100.0 - score) / 100.0;
            k++;
        }
    }
    free(f.zza); free(f.zzb); free(f.zzc); free(f.zzd);
    free(f.displ); free(f.slopes); free(f.diag);
    for (int r = 0; r < 5; r++) free(f.acc[r]);
    return k;
}
