/*
 * cw_oracle.c -- CPU restatement of ClustalW 2.0.11's full-DP pairwise path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * kernels and the CPU baseline of bench.py; the shipped library never links
 * it.  It restates, in plain C, the arithmetic of
 *   pairwise/FullPairwiseAlign.cpp           (forwardPass, reversePass, diff,
 *                                             add/del, tbgap/tegap, tracePath,
 *                                             the distance and the pair loop)
 *   substitutionMatrix/SubMatrix.cpp         (getPairwiseMatrix, getMatrix)
 * (paths relative to CPU/clustalw-2.0.11/src/).  The control flow of each
 * routine follows the reference line for line because the tie-breaking of
 * every comparison is part of the result; only storage differs (C arrays and
 * an explicit per-call context instead of class members and statics).
 *
 * Pinning: see tests/test_oracle.py -- NJ .dnd fixtures of the reference,
 * closed forms (matAvgScore, gap-open table), brute-force quadratic-space
 * Smith-Waterman / Needleman-Wunsch-Gotoh optima, edit-script re-scoring.
 */
#include "cw_oracle.h"
#include "../cuda-clustalw_amd/csrc/cw_matrices.inc"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXAA 24   /* UserParameters.cpp:132-133: strlen("ABCDEFGHIKLMNOPQRSTUVWXYZ-") - 2 */
#define GAPPOS1 30 /* UserParameters.cpp:38 */
#define GAPPOS2 31 /* UserParameters.cpp:39 */

/* ---------------------------------------------------------------------- */
/* SubMatrix::getMatrix (SubMatrix.cpp:689-829), negFlag=true, minimise=false */
static int get_matrix(const short *tab, int nletters, int mat[CWO_NUMRES][CWO_NUMRES],
                      int scale, int *mat_avg)
{
    int i, j, max_res, av3;
    for (i = 0; i < CWO_NUMRES; i++)
        for (j = 0; j < CWO_NUMRES; j++) mat[i][j] = 0;
    /* table is code-indexed; entries of codes outside the alphabet are 0 and
     * the reference never writes them either (xref == -1) */
    for (i = 0; i < 25; i++)
        for (j = 0; j < 25; j++) mat[i][j] = tab[i * 25 + j] * scale;
    max_res = nletters - 1;                       /* :713-733 count diag, --maxRes */
    av3 = 0;
    for (i = 0; i <= MAXAA; i++)                  /* :735-750 */
        for (j = 0; j < i; j++) av3 += mat[i][j];
    av3 = (int)(av3 / (((float)(max_res * max_res - max_res)) / 2));   /* :754 */
    *mat_avg = -av3;                              /* :755 */
    /* negFlag=true: no positive shift (:775-798); gap rows/cols (:799-813) */
    for (i = 0; i < GAPPOS1; i++) {
        mat[i][GAPPOS1] = mat[GAPPOS1][i] = 0;
        mat[i][GAPPOS2] = mat[GAPPOS2][i] = 0;
    }
    mat[GAPPOS1][GAPPOS1] = mat[GAPPOS2][GAPPOS2] = 0;
    mat[GAPPOS2][GAPPOS1] = mat[GAPPOS1][GAPPOS2] = 0;
    return max_res + 2;                           /* :826 */
}

/* SubMatrix::getPairwiseMatrix (SubMatrix.cpp:256-390), default matrices */
int cwo_pairwise_matrix(int which, float transition_weight, int mat[CWO_NUMRES][CWO_NUMRES],
                        int *mat_avg, int *int_scale, float *go_scale, float *ge_scale)
{
    int scale = 100;                              /* :274-278 (not OS_MAC) */
    int r;
    *go_scale = *ge_scale = 1.0f;
    if (which == CWO_MAT_IUB) {                   /* :288-318, "iub" */
        r = get_matrix(cw_iub_dna, CW_IUB_DNA_NLETTERS, mat, scale, mat_avg);
        mat[0][4] = (int)(transition_weight * mat[0][0]);
        mat[4][0] = (int)(transition_weight * mat[0][0]);
        mat[2][11] = (int)(transition_weight * mat[0][0]);
        mat[11][2] = (int)(transition_weight * mat[0][0]);
        mat[2][12] = (int)(transition_weight * mat[0][0]);
        mat[12][2] = (int)(transition_weight * mat[0][0]);
    } else {                                      /* :348-354, "gonnet" */
        scale /= 10;
        r = get_matrix(cw_gonnet250, CW_GONNET250_NLETTERS, mat, scale, mat_avg);
    }
    *int_scale = scale;
    return r;
}

/* FullPairwiseAlign.cpp:129-148 */
void cwo_gap_params(int dna, int mat_avg, int int_scale, float go_scale, float ge_scale,
                    float pw_go, float pw_ge, int n, int m, int *go, int *ge)
{
    int mn = n < m ? n : m;
    if (dna) {
        *go = (int)(2 * pw_go * int_scale * go_scale);
        *ge = (int)(pw_ge * int_scale * ge_scale);
    } else {
        if (mat_avg <= 0)
            *go = 2 * (int)((pw_go + log((double)mn)) * int_scale);
        else
            *go = (int)(2 * mat_avg * (pw_go + log((double)mn)) * go_scale);
        *ge = (int)(pw_ge * int_scale);
    }
}

/* FullPairwiseAlign.cpp:172-181 */
float cwo_distance(int count, int len1, int len2)
{
    float mm = (float)(100.0 * (float)count);     /* tracePath :276 */
    int mn = len1 < len2 ? len1 : len2;
    if (len1 == 0 || len2 == 0) mm = 0;
    else mm /= (float)mn;
    return ((float)100.0 - mm) / (float)100.0;
}

/* ---------------------------------------------------------------------- */
/* per-pair context: the member state of FullPairwiseAlign */
typedef struct {
    const int *s1, *s2;
    const int (*mat)[CWO_NUMRES];
    int go, ge;
    int *HH, *DD, *RR, *SS, *displ;
    int print_ptr, last_print;
    int max_score, se1, se2, sb1, sb2;
    int gap1, gap2;
    int leaf_count;
    int64_t cells_fwd, cells_rev, cells_mm, nodes;
} ctx_t;

#define SCORE(c, i, j) ((c)->mat[(c)->s1[i]][(c)->s2[j]])

/* FullPairwiseAlign::forwardPass (:280-341) */
static void forward_pass(ctx_t *c, int n, int m)
{
    int i, j, f, hh, p, t;
    int *HH = c->HH, *DD = c->DD;
    const int go = c->go, ge = c->ge;
    c->max_score = 0;
    c->se1 = c->se2 = 0;
    for (i = 0; i <= m; i++) { HH[i] = 0; DD[i] = -go; }
    for (i = 1; i <= n; i++) {
        hh = p = 0;
        f = -go;
        for (j = 1; j <= m; j++) {
            f -= ge;
            t = hh - go - ge;
            if (f < t) f = t;
            DD[j] -= ge;
            t = HH[j] - go - ge;
            if (DD[j] < t) DD[j] = t;
            hh = p + SCORE(c, i, j);
            if (hh < f) hh = f;
            if (hh < DD[j]) hh = DD[j];
            if (hh < 0) hh = 0;
            p = HH[j];
            HH[j] = hh;
            if (hh > c->max_score) { c->max_score = hh; c->se1 = i; c->se2 = j; }
        }
    }
    c->cells_fwd += (int64_t)n * m;
}

/* FullPairwiseAlign::reversePass (:343-416) */
static void reverse_pass(ctx_t *c)
{
    int i, j, f, hh, p, t, cost = 0;
    int *HH = c->HH, *DD = c->DD;
    const int go = c->go, ge = c->ge;
    c->sb1 = c->sb2 = 1;
    for (i = c->se2; i > 0; i--) { HH[i] = -1; DD[i] = -1; }
    for (i = c->se1; i > 0; i--) {
        hh = f = -1;
        p = (i == c->se1) ? 0 : -1;
        for (j = c->se2; j > 0; j--) {
            c->cells_rev++;
            f -= ge;
            t = hh - go - ge;
            if (f < t) f = t;
            DD[j] -= ge;
            t = HH[j] - go - ge;
            if (DD[j] < t) DD[j] = t;
            hh = p + SCORE(c, i, j);
            if (hh < f) hh = f;
            if (hh < DD[j]) hh = DD[j];
            p = HH[j];
            HH[j] = hh;
            if (hh > cost) {
                cost = hh;
                c->sb1 = i;
                c->sb2 = j;
                if (cost >= c->max_score) break;
            }
        }
        if (cost >= c->max_score) break;
    }
}

/* FullPairwiseAlign::add (:213-224) */
static void ed_add(ctx_t *c, int v)
{
    if (c->last_print < 0) {
        c->displ[c->print_ptr - 1] = v;
        c->displ[c->print_ptr++] = c->last_print;
    } else {
        c->last_print = c->displ[c->print_ptr++] = v;
    }
}

/* FullPairwiseAlign::del (:614-624) */
static void ed_del(ctx_t *c, int k)
{
    if (c->last_print < 0)
        c->last_print = c->displ[c->print_ptr - 1] -= k;
    else
        c->last_print = c->displ[c->print_ptr++] = -(k);
}

/* :638-660 */
static inline int tbgap(const ctx_t *c, int k, int tb) { return k <= 0 ? 0 : tb + c->ge * k; }
static inline int tegap(const ctx_t *c, int k, int te) { return k <= 0 ? 0 : te + c->ge * k; }

/* FullPairwiseAlign::diff (:418-612).  A and B are 0-based offsets into the
 * 1-based sequences, M rows of seq1 and N columns of seq2. */
static int diff(ctx_t *c, int A, int B, int M, int N, int tb, int te)
{
    int type, midi, midj, i, j, midh;
    int f, hh, e, s, t;
    int *HH = c->HH, *DD = c->DD, *RR = c->RR, *SS = c->SS;
    const int go = c->go, ge = c->ge;

    if (N <= 0) {
        if (M > 0) ed_del(c, M);
        return -tbgap(c, M, tb);
    }
    if (M <= 1) {
        if (M <= 0) {
            ed_add(c, N);
            return -tbgap(c, N, tb);
        }
        midh = -(tb + ge) - tegap(c, N, te);
        hh = -(te + ge) - tbgap(c, N, tb);
        if (hh > midh) midh = hh;
        midj = 0;
        for (j = 1; j <= N; j++) {
            hh = SCORE(c, 1 + A, j + B) - tegap(c, N - j, te) - tbgap(c, j - 1, tb);
            if (hh > midh) { midh = hh; midj = j; }
        }
        c->cells_mm += N;
        if (midj == 0) {
            ed_del(c, 1);
            ed_add(c, N);
        } else {
            int r1 = c->s1[A + 1], r2 = c->s2[B + midj];
            if (r1 != c->gap1 && r2 != c->gap2 && r1 == r2) c->leaf_count++;
            if (midj > 1) ed_add(c, midj - 1);
            c->displ[c->print_ptr++] = c->last_print = 0;
            if (midj < N) ed_add(c, N - midj);
        }
        return midh;
    }

    c->nodes++;
    c->cells_mm += (int64_t)M * N;
    midi = M / 2;
    HH[0] = 0;
    t = -tb;
    for (j = 1; j <= N; j++) {
        HH[j] = t = t - ge;
        DD[j] = t - go;
    }
    t = -tb;
    for (i = 1; i <= midi; i++) {
        s = HH[0];
        HH[0] = hh = t = t - ge;
        f = t - go;
        for (j = 1; j <= N; j++) {
            if ((hh = hh - go - ge) > (f = f - ge)) f = hh;
            if ((hh = HH[j] - go - ge) > (e = DD[j] - ge)) e = hh;
            hh = s + SCORE(c, i + A, j + B);
            if (f > hh) hh = f;
            if (e > hh) hh = e;
            s = HH[j];
            HH[j] = hh;
            DD[j] = e;
        }
    }
    DD[0] = HH[0];

    RR[N] = 0;
    t = -te;
    for (j = N - 1; j >= 0; j--) {
        RR[j] = t = t - ge;
        SS[j] = t - go;
    }
    t = -te;
    for (i = M - 1; i >= midi; i--) {
        s = RR[N];
        RR[N] = hh = t = t - ge;
        f = t - go;
        for (j = N - 1; j >= 0; j--) {
            if ((hh = hh - go - ge) > (f = f - ge)) f = hh;
            if ((hh = RR[j] - go - ge) > (e = SS[j] - ge)) e = hh;
            hh = s + SCORE(c, i + 1 + A, j + 1 + B);
            if (f > hh) hh = f;
            if (e > hh) hh = e;
            s = RR[j];
            RR[j] = hh;
            SS[j] = e;
        }
    }
    SS[N] = RR[N];

    midh = HH[0] + RR[0];
    midj = 0;
    type = 1;
    for (j = 0; j <= N; j++) {
        hh = HH[j] + RR[j];
        if (hh >= midh)
            if (hh > midh || (HH[j] != DD[j] && RR[j] == SS[j])) {
                midh = hh;
                midj = j;
            }
    }
    for (j = N; j >= 0; j--) {
        hh = DD[j] + SS[j] + go;
        if (hh > midh) {
            midh = hh;
            midj = j;
            type = 2;
        }
    }

    if (type == 1) {
        diff(c, A, B, midi, midj, tb, go);
        diff(c, A + midi, B + midj, M - midi, N - midj, go, te);
    } else {
        diff(c, A, B, midi - 1, midj, tb, 0);
        ed_del(c, 2);
        diff(c, A + midi + 1, B + midj, M - midi - 1, N - midj, 0, te);
    }
    return midh;
}

/* FullPairwiseAlign::tracePath (:231-278), returning the integer count */
static int trace_path(ctx_t *c, int tsb1, int tsb2)
{
    int i, k, i1 = tsb1, i2 = tsb2, count = 0, to_do = c->print_ptr - 1;
    for (i = 1; i <= to_do; ++i) {
        if (c->displ[i] == 0) {
            int r1 = c->s1[i1], r2 = c->s2[i2];
            if (r1 != c->gap1 && r2 != c->gap2 && r1 == r2) count++;
            ++i1;
            ++i2;
        } else if ((k = c->displ[i]) > 0) {
            i2 += k;
        } else {
            i1 -= k;
        }
    }
    return count;
}

void cwo_align_pair(const int *seq1, int n, const int *seq2, int m,
                    const int mat[CWO_NUMRES][CWO_NUMRES], int go, int ge,
                    int gap_pos1, int gap_pos2, int *displ, cwo_pair_result *res)
{
    ctx_t c;
    int mx = (n > m ? n : m) + 2;
    int *buf = (int *)malloc(sizeof(int) * (size_t)(4 * mx));
    int *own_displ = 0;
    memset(&c, 0, sizeof c);
    c.s1 = seq1; c.s2 = seq2; c.mat = (const int (*)[CWO_NUMRES])mat;
    c.go = go; c.ge = ge; c.gap1 = gap_pos1; c.gap2 = gap_pos2;
    c.HH = buf; c.DD = buf + mx; c.RR = buf + 2 * mx; c.SS = buf + 3 * mx;
    if (!displ) displ = own_displ = (int *)malloc(sizeof(int) * (size_t)(2 * (n + m) + 2));
    c.displ = displ;
    forward_pass(&c, n, m);
    reverse_pass(&c);
    c.last_print = 0;
    c.print_ptr = 1;
    res->mm_score = diff(&c, c.sb1 - 1, c.sb2 - 1, c.se1 - c.sb1 + 1, c.se2 - c.sb2 + 1, 0, 0);
    res->count = trace_path(&c, c.sb1, c.sb2);
    res->max_score = c.max_score; res->se1 = c.se1; res->se2 = c.se2;
    res->sb1 = c.sb1; res->sb2 = c.sb2;
    res->n_displ = c.print_ptr - 1;
    res->count_leaves = c.leaf_count;
    res->cells_fwd = c.cells_fwd; res->cells_rev = c.cells_rev; res->cells_mm = c.cells_mm;
    res->mm_nodes = c.nodes;
    free(buf);
    free(own_displ);
}

/* ---------------------------------------------------------------------- */
static int gapfree_len(const uint8_t *s, int n, int gap1, int gap2)
{
    int i, l = 0;
    for (i = 0; i < n; i++)
        if (s[i] != gap1 && s[i] != gap2) l++;
    return l;
}

static void one_pair(const uint8_t *codes, const int64_t *offs, const int32_t *lens,
                     int si, int sj, const int mat[CWO_NUMRES][CWO_NUMRES],
                     const int32_t *go_by_len, int ge, int gap1, int gap2,
                     int *b1, int *b2, int32_t *cnt, float *dist, cwo_pair_result *r)
{
    int n = lens[si], m = lens[sj], k, go;
    int len1, len2;
    cwo_pair_result rr;
    if (!r) r = &rr;
    memset(r, 0, sizeof *r);
    if (n == 0 || m == 0) {                       /* :113-118 */
        *cnt = -1;
        *dist = 1.0f;
        return;
    }
    len1 = gapfree_len(codes + offs[si], n, gap1, gap2);
    len2 = gapfree_len(codes + offs[sj], m, gap1, gap2);
    for (k = 0; k < n; k++) b1[k + 1] = codes[offs[si] + k];
    for (k = 0; k < m; k++) b2[k + 1] = codes[offs[sj] + k];
    b1[0] = b2[0] = 0;
    go = go_by_len[n < m ? n : m];
    cwo_align_pair(b1, n, b2, m, mat, go, ge, gap1, gap2, 0, r);
    *cnt = r->count;
    *dist = cwo_distance(r->count, len1, len2);
}

int64_t cwo_count_pairs(int nseq, int i_start, int i_end, int j_start, int j_end)
{
    int64_t np = 0;
    int si, sj0;
    for (si = i_start > 0 ? i_start : 0; si < nseq && si < i_end; si++) {
        sj0 = si + 1 > j_start + 1 ? si + 1 : j_start + 1;
        int hi = nseq < j_end ? nseq : j_end;
        if (hi > sj0) np += hi - sj0;
    }
    return np;
}

void cwo_pairwise_list(const uint8_t *codes, const int64_t *offs, const int32_t *lens,
                       const int32_t *pairs, int64_t npairs,
                       const int mat[CWO_NUMRES][CWO_NUMRES], const int32_t *go_by_len,
                       int ge, int gap_pos1, int gap_pos2, int nthreads,
                       int32_t *out_count, float *out_dist, cwo_pair_result *out_res)
{
    int maxl = 0;
    int64_t p;
    for (p = 0; p < npairs; p++) {
        int a = lens[pairs[2 * p]], b = lens[pairs[2 * p + 1]];
        if (a > maxl) maxl = a;
        if (b > maxl) maxl = b;
    }
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        int *b1 = (int *)malloc(sizeof(int) * (size_t)(maxl + 2));
        int *b2 = (int *)malloc(sizeof(int) * (size_t)(maxl + 2));
        int64_t q;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (q = 0; q < npairs; q++)
            one_pair(codes, offs, lens, pairs[2 * q], pairs[2 * q + 1], mat, go_by_len, ge,
                     gap_pos1, gap_pos2, b1, b2, &out_count[q], &out_dist[q],
                     out_res ? &out_res[q] : 0);
        free(b1);
        free(b2);
    }
}

int64_t cwo_pairwise_range(const uint8_t *codes, const int64_t *offs, const int32_t *lens,
                           int nseq, int i_start, int i_end, int j_start, int j_end,
                           const int mat[CWO_NUMRES][CWO_NUMRES], const int32_t *go_by_len,
                           int ge, int gap_pos1, int gap_pos2, int nthreads,
                           int32_t *out_pairs, int32_t *out_count, float *out_dist,
                           int64_t *cells)
{
    int64_t np = 0, p;
    int si, sj;
    cwo_pair_result *res;
    /* the loops of :97 and :110 */
    for (si = i_start > 0 ? i_start : 0; si < nseq && si < i_end; si++)
        for (sj = (si + 1 > j_start + 1 ? si + 1 : j_start + 1); sj < nseq && sj < j_end; sj++) {
            out_pairs[2 * np] = si;
            out_pairs[2 * np + 1] = sj;
            np++;
        }
    res = cells ? (cwo_pair_result *)calloc((size_t)(np ? np : 1), sizeof *res) : 0;
    cwo_pairwise_list(codes, offs, lens, out_pairs, np, mat, go_by_len, ge, gap_pos1, gap_pos2,
                      nthreads, out_count, out_dist, res);
    if (cells) {
        cells[0] = cells[1] = cells[2] = cells[3] = 0;
        for (p = 0; p < np; p++) {
            cells[0] += res[p].cells_fwd;
            cells[1] += res[p].cells_rev;
            cells[2] += res[p].cells_mm;
            cells[3] += res[p].mm_nodes;
        }
        free(res);
    }
    return np;
}
