/*
 * cw_msa.c -- CPU restatement of ClustalW 2.0.11's progressive profile
 * alignment.  TEST INFRASTRUCTURE ONLY: the checker for the profile DP and
 * the source of the reference's edit scripts in the parity tests.
 *
 * Restates (paths relative to CPU/clustalw-2.0.11/src/):
 *   substitutionMatrix/SubMatrix.cpp:403-677  getProfileAlignMatrix (gonnet
 *                                             series and iub, default names)
 *   substitutionMatrix/SubMatrix.cpp:689-829  getMatrix with negFlag=false
 *   substitutionMatrix/SubMatrix.cpp:172-246  the default cross references
 *   multipleAlign/ProfileBase.cpp:52-298      calcGapCoeff
 *   multipleAlign/ProfileBase.cpp:310-547     calcVPenalties, calcResidueSpecificPen,
 *                                             calcHydrophilicPen, localPenalty, percentId
 *   multipleAlign/ProfileWithSub.cpp:42-147   calcProfileWithSub
 *   multipleAlign/ProfileStandard.cpp:41-117  calcStandardProfile
 *   multipleAlign/MyersMillerProfileAlign.cpp:36-505   profileAlign
 *   multipleAlign/MyersMillerProfileAlign.cpp:521-811  progDiff
 *   multipleAlign/MyersMillerProfileAlign.cpp:818-934  addGGaps
 *   multipleAlign/MyersMillerProfileAlign.cpp:1002-1262 prfScore, progTracepath,
 *                                             progDel/progAdd/progAlign, the penalties
 *   multipleAlign/MSA.cpp:32-347              multiSeqAlign (iStart = 0)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cw_oracle.h"
#include "../cuda-clustalw_amd/csrc/cw_matrices.inc"

#define NUMRES 32
#define MAXAA 24
#define GAPPOS1 30
#define GAPPOS2 31
#define GAPCOL 32
#define LENCOL 33
#define PRF_W CWO_PRF_STRIDE          /* LENCOL + 2 ints per profile row */
#define ENDALN 127                    /* general/clustalw.h:113 */
#define EXTRA_END 2                   /* extraEndElemNum, general/clustalw.h:112 */
#define INT_SCALE_FACTOR 1000         /* general/clustalw.h:226 (OS_UNIX) */

static const char CODES[] = "ABCDEFGHIKLMNOPQRSTUVWXYZ-";  /* UserParameters.cpp:132 */
static const char AA_ORDER[] = "ABCDEFGHIKLMNPQRSTVWXYZ";   /* SubMatrix.cpp:179 */
static const char DNA_ORDER[] = "ABCDGHKMNRSTUVWXY";        /* SubMatrix.cpp:180 */

static int code_of(char c)
{
    for (int i = 0; CODES[i]; i++)
        if (CODES[i] == c) return i;
    return -1;
}

/* ---------------------------------------------------------------------- */
/* getMatrix (SubMatrix.cpp:689-829) on a code-indexed table; `order` is the
 * table's alphabet, whose letters are the codes the cross reference maps. */
static int get_matrix_flag(const short *tab, const char *order, int mat[NUMRES][NUMRES],
                           int neg_flag, int scale, int *mat_avg)
{
    int i, j, av3 = 0, mn;
    int nlet = (int)strlen(order);
    int max_res = nlet - 1;
    for (i = 0; i < NUMRES; i++)
        for (j = 0; j < NUMRES; j++) mat[i][j] = 0;
    for (i = 0; i < 25; i++)
        for (j = 0; j < 25; j++) mat[i][j] = tab[i * 25 + j] * scale;
    for (i = 0; i <= MAXAA; i++)
        for (j = 0; j < i; j++) av3 += mat[i][j];
    av3 = (int)(av3 / (((float)(max_res * max_res - max_res)) / 2));
    *mat_avg = -av3;
    mn = mat[0][0];
    for (i = 0; i <= MAXAA; i++)
        for (j = 1; j <= i; j++)
            if (mat[i][j] < mn) mn = mat[i][j];
    if (!neg_flag && mn < 0) {
        for (i = 0; i < nlet; i++) {
            int ti = code_of(order[i]);
            for (j = 0; j < nlet; j++) mat[ti][code_of(order[j])] -= mn;
        }
    }
    for (i = 0; i < GAPPOS1; i++) {
        mat[i][GAPPOS1] = mat[GAPPOS1][i] = 0;
        mat[i][GAPPOS2] = mat[GAPPOS2][i] = 0;
    }
    mat[GAPPOS1][GAPPOS1] = mat[GAPPOS2][GAPPOS2] = 0;
    mat[GAPPOS2][GAPPOS1] = mat[GAPPOS1][GAPPOS2] = 0;
    return max_res + 2;
}

int cwo_profile_matrix(int dna, double pcid, int min_len, float transition_weight,
                       int neg_matrix, int distance_tree, int mat[NUMRES][NUMRES],
                       float *scale_out, int *mat_avg)
{
    float scale;
    float int_scale = 100;
    int r;
    if (dna) {                                          /* :421-471, "iub" */
        scale = 1.0f;
        r = get_matrix_flag(cw_iub_dna, DNA_ORDER, mat, neg_matrix, (int)int_scale, mat_avg);
        /* the cross reference maps table letters 0,4,2,11,12 = A,G,C,T,U */
        int a = code_of('A'), g = code_of('G'), c = code_of('C'), t = code_of('T'),
            u = code_of('U');
        int tw = (int)(transition_weight * mat[0][0]);
        mat[a][g] = tw;
        mat[g][a] = tw;
        mat[c][t] = tw;
        mat[t][c] = tw;
        mat[c][u] = tw;
        mat[u][c] = tw;
    } else {                                            /* :561-604, "gonnet" */
        const short *tab;
        scale = 0.75f;
        scale /= 2.0;
        if (neg_matrix || !distance_tree) {
            tab = cw_gonnet250;
        } else if (pcid > 35.0) {
            tab = cw_gonnet80;
            scale /= 2.0;
        } else if (pcid > 25.0) {
            tab = (min_len < 100) ? cw_gonnet250 : cw_gonnet120;
        } else {
            tab = (min_len < 100) ? cw_gonnet350 : cw_gonnet160;
        }
        int_scale /= 10;
        r = get_matrix_flag(tab, AA_ORDER, mat, neg_matrix, (int)int_scale, mat_avg);
    }
    *scale_out = scale;
    return r;
}

/* the matrix of one profileAlign step (MyersMillerProfileAlign.cpp:249-258):
 * the caller's series when it gives one, else the default names */
static int step_matrix(const cwo_msa_params *pm, double pcid, int min_len, int mat[NUMRES][NUMRES],
                       float *scale, int *mat_avg)
{
    if (pm->profile_matrix) {
        *scale = 1.0f;
        *mat_avg = 0;
        memset(mat, 0, sizeof(int) * NUMRES * NUMRES);
        return pm->profile_matrix(pm->profile_matrix_user, pcid, min_len, &mat[0][0], scale, mat_avg);
    }
    return cwo_profile_matrix(pm->dna, pcid, min_len, pm->transition_weight, pm->neg_matrix,
                              pm->distance_tree, mat, scale, mat_avg);
}

/* ---------------------------------------------------------------------- */
/* Profile builders.  `rows` are the profile's sequences (0-based columns). */
typedef struct {
    int prf_len, first, last;     /* prfLength, firstSeq, lastSeq */
    int *profile;                 /* (prf_len + 2) x PRF_W */
    int nVarPen, nHydPen, nPrefPen;
    int gdist;
} prof_t;

#define P(pr, i, j) ((pr)->profile[(size_t)(i) * PRF_W + (j)])

static void prof_init(prof_t *p, int len, int first, int last)
{
    memset(p, 0, sizeof *p);
    p->prf_len = len;
    p->first = first;
    p->last = last;
    p->profile = (int *)calloc((size_t)(len + 2) * PRF_W, sizeof(int));
}

/* ProfileBase::percentId (:519-547) */
static float percent_id(const prof_t *p, const int *s1, const int *s2)
{
    int i, count = 0, total = 0;
    for (i = 0; i < p->prf_len; i++) {
        if (s1[i] >= 0 && s1[i] < MAXAA) {
            total++;
            if (s1[i] == s2[i]) count++;
        }
        if (s1[i] == -3 || s2[i] == -3) break;
    }
    if (total == 0) return 0;
    return (float)(100.0 * (float)count / (float)total);
}

/* ProfileBase::calcVPenalties (:310-353); vlut is the identity */
static void calc_v_penalties(const prof_t *p, int *const *aln, int *weight)
{
    const int vwindow = 5, vll = 50;
    for (int i = 0; i < p->prf_len; i++) {
        int t = 0;
        weight[i] = 0;
        for (int j = i - vwindow; j < i + vwindow; j++) {
            if (j >= 0 && j < p->prf_len) {
                int ix1 = aln[p->first][j], ix2 = aln[p->first + 1][j];
                if (ix1 < 0 || ix1 > MAXAA || ix2 < 0 || ix2 > MAXAA) continue;
                weight[i] += (ix1 == ix2) ? 1 : 0;
                t++;
            }
        }
        weight[i] += t;
        if (t > 0)
            weight[i] = (weight[i] * 100) / (2 * t);
        else
            weight[i] = 100;
        if (weight[i] < vll) weight[i] = vll;
    }
}

/* ProfileBase::calcResidueSpecificPen (:360-391), Pascarella table (:30-33) */
static void calc_residue_specific_pen(const prof_t *p, int *const *aln, int *weight)
{
    static const char pres[] = "ACDEFGHKILMNPQRSTVYW";
    static const int pprob[] = {87, 87, 104, 69, 80, 139, 100, 104, 68, 79,
                                71, 137, 126, 93, 128, 124, 111, 75, 100, 77};
    int numSeq = p->last - p->first;
    for (int i = 0; i < p->prf_len; i++) {
        weight[i] = 0;
        for (int k = p->first; k < p->last; k++) {
            for (int j = 0; j < 20; j++) {
                signed char ix = (signed char)aln[k][i];
                if (ix < 0 || ix > MAXAA) continue;
                if (CODES[(int)ix] == pres[j]) {
                    weight[i] += (180 - pprob[j]);
                    break;
                }
            }
        }
        weight[i] /= numSeq;
    }
}

/* ProfileBase::calcHydrophilicPen (:398-459) */
static void calc_hydrophilic_pen(const prof_t *p, int *const *aln, int *weight,
                                 const char *hyd_residues)
{
    int i, j, k, s, e;
    int nh = (int)strlen(hyd_residues);
    int *hyd = (int *)calloc((size_t)p->prf_len + 2, sizeof(int));
    float scale;
    for (i = 0; i < p->prf_len; i++) weight[i] = 0;
    for (k = p->first; k < p->last; k++) {
        for (i = 0; i < p->prf_len; i++) {
            hyd[i] = 0;
            for (j = 0; j < nh; j++) {
                int res = aln[k][i];
                if (res < 0 || res > MAXAA) continue;
                if (CODES[res] == hyd_residues[j]) {
                    hyd[i] = 1;
                    break;
                }
            }
        }
        i = 0;
        while (i < p->prf_len) {
            if (hyd[i] == 0) {
                i++;
            } else {
                s = i;
                while (hyd[i] != 0 && i < p->prf_len) i++;
                e = i;
                if (e - s > 3)
                    for (j = s; j < e; j++) weight[j] += 100;
            }
        }
    }
    scale = (float)(p->last - p->first);
    for (i = 0; i < p->prf_len; i++) weight[i] = (int)(weight[i] / scale);
    free(hyd);
}

/* ProfileBase::localPenalty (:468-512), protein only */
static int local_penalty(const prof_t *p, int penalty, int n, const int *resW,
                         const int *hydW, const int *vW)
{
    int h = 0;
    float gw = 1.0f;
    if (!p->nVarPen) gw *= vW[n] / 100.0;
    if (!p->nHydPen) {
        if (hydW[n] > 0) {
            gw *= 0.5;
            h = 1;
        }
    }
    if (!p->nPrefPen && !h) gw *= (resW[n] / 100.0);
    gw *= penalty;
    return (int)gw;
}

/* ProfileBase::calcGapCoeff (:52-298); `gaps` is the caller's shared vector */
static void calc_gap_coeff(prof_t *p, int *const *aln, int *gaps, const cwo_msa_params *pm,
                           int gapCoef, int lenCoef)
{
    int c, i, j, is, ie, pcid;
    int numSeq = p->last - p->first;
    int L = p->prf_len;
    int *gapPos, *vW = NULL, *resW = NULL, *hydW = NULL;
    float scale;
    const int dna = pm->dna;

    if (numSeq == 2)
        pcid = (int)percent_id(p, aln[p->first], aln[p->first + 1]);
    else
        pcid = 0;
    for (j = 0; j < L; j++) gaps[j] = 0;
    /* no structure penalties in this restatement (useStructPenalties false) */
    if (!pm->no_var_penalties && pcid > 60) {
        p->nHydPen = p->nPrefPen = 1;
        p->nVarPen = 0;
    } else {
        p->nVarPen = 1;
        p->nHydPen = pm->no_hyd_penalties;
        p->nPrefPen = pm->no_pref_penalties;
        p->gdist = pm->gap_dist;
    }
    for (i = p->first; i < p->last; i++) {
        is = 0;
        ie = L;
        if (!pm->use_end_gaps && !pm->end_gap_penalties) {
            for (j = 0; j < L; j++) {
                c = aln[i][j];
                if (c < 0 || c > MAXAA)
                    is++;
                else
                    break;
            }
            for (j = L - 1; j >= 0; j--) {
                c = aln[i][j];
                if (c < 0 || c > MAXAA)
                    ie--;
                else
                    break;
            }
        }
        for (j = is; j < ie; j++)
            if (aln[i][j] < 0 || aln[i][j] > MAXAA) gaps[j]++;
    }
    if (!dna && !p->nVarPen) {
        vW = (int *)calloc((size_t)L + 2, sizeof(int));
        calc_v_penalties(p, aln, vW);
    }
    if (!dna && !p->nPrefPen) {
        resW = (int *)calloc((size_t)L + 2, sizeof(int));
        calc_residue_specific_pen(p, aln, resW);
    }
    if (!dna && !p->nHydPen) {
        hydW = (int *)calloc((size_t)L + 2, sizeof(int));
        calc_hydrophilic_pen(p, aln, hydW, pm->hyd_residues);
    }
    gapPos = (int *)calloc((size_t)L + 2, sizeof(int));
    if (dna || p->gdist <= 0) {
        for (i = 0; i < L; i++) gapPos[i] = gaps[i];
    } else {
        const int gdist = p->gdist;
        i = 0;
        while (i < L) {
            if (gaps[i] <= 0) {
                gapPos[i] = gaps[i];
                i++;
            } else {
                for (j = -gdist + 1; j < 0; j++)
                    if (i + j >= 0 && i + j < L && (gaps[i + j] == 0 || gaps[i + j] < j))
                        gapPos[i + j] = j;
                while (gaps[i] > 0) {
                    if (i >= L) break;
                    gapPos[i] = gaps[i];
                    i++;
                }
                for (j = 0; j < gdist; j++) {
                    if (gaps[i + j] > 0) break;
                    if (i + j >= 0 && i + j < L && (gaps[i + j] == 0 || gaps[i + j] < -j))
                        gapPos[i + j] = -j - 1;
                }
                i += j;
            }
        }
    }
    for (j = 0; j < L; j++) {
        if (gapPos[j] <= 0) {
            if (!dna) {
                P(p, j + 1, GAPCOL) = local_penalty(p, gapCoef, j, resW, hydW, vW);
                P(p, j + 1, LENCOL) = lenCoef;
            } else {
                P(p, j + 1, GAPCOL) = gapCoef;
                P(p, j + 1, LENCOL) = lenCoef;
            }
            if (gapPos[j] < 0)
                P(p, j + 1, GAPCOL) = (int)((P(p, j + 1, GAPCOL) *
                                             (2.0 + 2.0 * (p->gdist + gapPos[j]) / p->gdist)));
        } else {
            scale = ((float)(numSeq - gaps[j]) / (float)numSeq) * 1.0f;   /* reducedGap */
            P(p, j + 1, GAPCOL) = (int)(scale * gapCoef);
            P(p, j + 1, LENCOL) = (int)(0.5 * lenCoef);
        }
        if (P(p, j + 1, GAPCOL) <= 0) P(p, j + 1, GAPCOL) = 1;
        if (P(p, j + 1, LENCOL) <= 0) P(p, j + 1, LENCOL) = 1;
    }
    if (pm->end_gap_penalties) {
        P(p, 0, GAPCOL) = gapCoef;
        P(p, 0, LENCOL) = lenCoef;
    } else {
        P(p, 0, GAPCOL) = 0;
        P(p, 0, LENCOL) = 0;
        P(p, L, GAPCOL) = 0;
        P(p, L, LENCOL) = 0;
    }
    free(gapPos);
    free(vW);
    free(resW);
    free(hydW);
}

/* ProfileWithSub::calcProfileWithSub (ProfileWithSub.cpp:42-147) */
static void calc_profile_with_sub(prof_t *p, int *const *aln, const int *gaps,
                                  int mat[NUMRES][NUMRES], const int *w)
{
    int L = p->prf_len;
    int numSeq = p->last - p->first;
    int sum2 = 0, aa, seq, res, col, pos, f;
    float scale;
    int *wt = (int *)calloc((size_t)(NUMRES + 2) * (L + 2), sizeof(int));
#define WT(a, c) wt[(size_t)(a) * (L + 2) + (c)]
    for (seq = p->first; seq < p->last; seq++) sum2 += w[seq];
    for (col = 0; col < L; col++) {
        for (aa = 0; aa <= MAXAA; aa++) {
            WT(aa, col) = 0;
            for (seq = p->first; seq < p->last; seq++)
                if (aa == aln[seq][col]) WT(aa, col) += w[seq];
        }
        WT(GAPPOS1, col) = 0;
        for (seq = p->first; seq < p->last; seq++)
            if (GAPPOS1 == aln[seq][col]) WT(GAPPOS1, col) += w[seq];
        WT(GAPPOS2, col) = 0;
        for (seq = p->first; seq < p->last; seq++)
            if (GAPPOS2 == aln[seq][col]) WT(GAPPOS2, col) += w[seq];
    }
    for (pos = 0; pos < L; pos++) {
        if (gaps[pos] == numSeq) {
            for (res = 0; res <= MAXAA; res++) P(p, pos + 1, res) = mat[res][GAPPOS1];
            P(p, pos + 1, GAPPOS1) = mat[GAPPOS1][GAPPOS1];
            P(p, pos + 1, GAPPOS2) = mat[GAPPOS2][GAPPOS1];
        } else {
            scale = (float)(numSeq - gaps[pos]) / (float)numSeq;
            for (res = 0; res <= MAXAA; res++) {
                f = 0;
                for (aa = 0; aa <= MAXAA; aa++) f += WT(aa, pos) * mat[aa][res];
                f += WT(GAPPOS1, pos) * mat[GAPPOS1][res];
                f += WT(GAPPOS2, pos) * mat[GAPPOS2][res];
                P(p, pos + 1, res) = (int)(((float)f / (float)sum2) * scale);
            }
            f = 0;
            for (aa = 0; aa <= MAXAA; aa++) f += WT(aa, pos) * mat[aa][GAPPOS1];
            f += WT(GAPPOS1, pos) * mat[GAPPOS1][GAPPOS1];
            f += WT(GAPPOS2, pos) * mat[GAPPOS2][GAPPOS1];
            P(p, pos + 1, GAPPOS1) = (int)(((float)f / (float)sum2) * scale);
            f = 0;
            for (aa = 0; aa <= MAXAA; aa++) f += WT(aa, pos) * mat[aa][GAPPOS2];
            f += WT(GAPPOS1, pos) * mat[GAPPOS1][GAPPOS2];
            f += WT(GAPPOS2, pos) * mat[GAPPOS2][GAPPOS2];
            P(p, pos + 1, GAPPOS2) = (int)(((float)f / (float)sum2) * scale);
        }
    }
#undef WT
    free(wt);
}

/* ProfileStandard::calcStandardProfile (ProfileStandard.cpp:41-117) */
static void calc_standard_profile(prof_t *p, int *const *aln, const int *w)
{
    for (int r = 0; r < p->prf_len; r++) {
        int sum1, sum2 = 0, i, d;
        for (i = p->first; i < p->last; i++) sum2 += w[i];
        if (sum2 == 0) {
            for (d = 0; d <= MAXAA; d++) P(p, r + 1, d) = 0;
            P(p, r + 1, GAPPOS1) = 0;
            P(p, r + 1, GAPPOS2) = 0;
        } else {
            for (d = 0; d <= MAXAA; d++) {
                sum1 = 0;
                for (i = p->first; i < p->last; i++)
                    if (d == aln[i][r]) sum1 += w[i];
                P(p, r + 1, d) = (int)(10 * (float)sum1 / (float)sum2);
            }
            sum1 = 0;
            for (i = p->first; i < p->last; i++)
                if (GAPPOS1 == aln[i][r]) sum1 += w[i];
            P(p, r + 1, GAPPOS1) = (int)(10 * (float)sum1 / (float)sum2);
            sum1 = 0;
            for (i = p->first; i < p->last; i++)
                if (GAPPOS2 == aln[i][r]) sum1 += w[i];
            P(p, r + 1, GAPPOS2) = (int)(10 * (float)sum1 / (float)sum2);
        }
    }
}

/* ---------------------------------------------------------------------- */
/* The profile DP: progDiff and its helpers (MyersMillerProfileAlign.cpp). */
typedef struct {
    const int *p1, *p2;           /* profile rows, PRF_W ints each */
    int L1, L2;                   /* prfLength1, prfLength2 */
    int end_gap_penalties;
    int *HH, *DD, *RR, *SS, *gS;
    int *displ;
    int printPtr, lastPrint;
    cwo_dp_stats st;
    int *trace, trace_cap, ntrace;     /* optional: one record per progDiff call */
} dp_t;

/* record: A B M N go1|go2<<1 type(0 leaf,1,2) midj midh */
static void trace_node(dp_t *c, int A, int B, int M, int N, int go1, int go2, int type, int midj,
                       int midh)
{
    if (!c->trace) return;
    if (c->ntrace < c->trace_cap) {
        int *r = c->trace + 8 * (size_t)c->ntrace;
        r[0] = A; r[1] = B; r[2] = M; r[3] = N; r[4] = (go1 ? 1 : 0) | (go2 ? 2 : 0);
        r[5] = type; r[6] = midj; r[7] = midh;
    }
    c->ntrace++;
}

#define P1(i, j) (c->p1[(size_t)(i) * PRF_W + (j)])
#define P2(i, j) (c->p2[(size_t)(i) * PRF_W + (j)])

/* prfScore (:1002-1016) */
static inline int prf_score(const dp_t *c, int n, int m)
{
    int score = 0;
    for (int ix = 0; ix <= MAXAA; ix++) score += P1(n, ix) * P2(m, ix);
    score += P1(n, GAPPOS1) * P2(m, GAPPOS1);
    score += P1(n, GAPPOS2) * P2(m, GAPPOS2);
    return score / 10;
}

/* openPenalty1 / extPenalty1 / openPenalty2 / extPenalty2 (:1126-1229) */
static inline int open1(const dp_t *c, int i, int j)
{
    if (!c->end_gap_penalties && (i == 0 || i == c->L1)) return 0;
    return P2(j, GAPCOL) + P1(i, GAPCOL);
}
static inline int ext1(const dp_t *c, int i, int j)
{
    if (!c->end_gap_penalties && (i == 0 || i == c->L1)) return 0;
    return P2(j, LENCOL);
}
static inline int open2(const dp_t *c, int i, int j)
{
    if (!c->end_gap_penalties && (j == 0 || j == c->L2)) return 0;
    return P1(i, GAPCOL) + P2(j, GAPCOL);
}
static inline int ext2(const dp_t *c, int i, int j)
{
    if (!c->end_gap_penalties && (j == 0 || j == c->L2)) return 0;
    return P1(i, LENCOL);
}

/* gapPenalty1 (:1166-1189) */
static int gap_penalty1(const dp_t *c, int i, int j, int k)
{
    int ix, h = 0;
    if (k <= 0) return 0;
    if (!c->end_gap_penalties && (i == 0 || i == c->L1)) return 0;
    int g = P2(j, GAPCOL) + P1(i, GAPCOL);
    for (ix = 0; ix < k && ix + j < c->L2; ix++) h += P2(ix + j, LENCOL);
    return g + h;
}

/* gapPenalty2 (:1239-1262) */
static int gap_penalty2(const dp_t *c, int i, int j, int k)
{
    int ix, h = 0;
    if (k <= 0) return 0;
    if (!c->end_gap_penalties && (j == 0 || j == c->L2)) return 0;
    int g = P1(i, GAPCOL) + P2(j, GAPCOL);
    for (ix = 0; ix < k && ix + i < c->L1; ix++) h += P1(ix + i, LENCOL);
    return g + h;
}

/* progDel / progAdd / progAlign (:1067-1117) */
static void prog_del(dp_t *c, int k)
{
    if (c->lastPrint < 0)
        c->lastPrint = c->displ[c->printPtr - 1] -= k;
    else
        c->lastPrint = c->displ[c->printPtr++] = -(k);
}
static void prog_add(dp_t *c, int k)
{
    if (c->lastPrint < 0) {
        c->displ[c->printPtr - 1] = k;
        c->displ[c->printPtr++] = c->lastPrint;
    } else {
        c->lastPrint = c->displ[c->printPtr++] = k;
    }
}
static void prog_align(dp_t *c) { c->displ[c->printPtr++] = c->lastPrint = 0; }

/* progDiff (:521-811) */
static int prog_diff(dp_t *c, int A, int B, int M, int N, int go1, int go2)
{
    int midi, midj, type, midh;
    int t, tl, g, h, i, j, hh, f, e, s;
    int *HH = c->HH, *DD = c->DD, *RR = c->RR, *SS = c->SS, *gS = c->gS;

    if (N <= 0) {
        if (M > 0) prog_del(c, M);
        trace_node(c, A, B, M, N, go1, go2, 0, 0, -gap_penalty1(c, A, B, M));
        return -gap_penalty1(c, A, B, M);
    }
    if (M <= 1) {
        if (M <= 0) {
            prog_add(c, N);
            trace_node(c, A, B, M, N, go1, go2, 0, 0, -gap_penalty2(c, A, B, N));
            return -gap_penalty2(c, A, B, N);
        }
        c->st.leaf_cells += N;
        if (go1 == 0)
            midh = -gap_penalty1(c, A + 1, B + 1, N);
        else
            midh = -gap_penalty2(c, A + 1, B, 1) - gap_penalty1(c, A + 1, B + 1, N);
        midj = 0;
        for (j = 1; j <= N; j++) {
            hh = -gap_penalty1(c, A, B + 1, j - 1) + prf_score(c, A + 1, B + j) -
                 gap_penalty1(c, A + 1, B + j + 1, N - j);
            if (hh > midh) {
                midh = hh;
                midj = j;
            }
        }
        trace_node(c, A, B, M, N, go1, go2, 0, midj, midh);
        if (midj == 0) {
            prog_add(c, N);
            prog_del(c, 1);
        } else {
            if (midj > 1) prog_add(c, midj - 1);
            prog_align(c);
            if (midj < N) prog_add(c, N - midj);
        }
        return midh;
    }
    c->st.nodes++;
    midi = M / 2;
    c->st.fwd_cells += (int64_t)midi * N;
    c->st.rev_cells += (int64_t)(M - midi) * N;

    /* forward phase: HH[j], DD[j] */
    HH[0] = 0;
    t = -open1(c, A, B + 1);
    tl = -ext1(c, A, B + 1);
    for (j = 1; j <= N; j++) {
        HH[j] = t = t + tl;
        DD[j] = t - open2(c, A + 1, B + j);
    }
    t = (go1 == 0) ? 0 : -open2(c, A + 1, B);
    tl = -ext2(c, A + 1, B);
    for (i = 1; i <= midi; i++) {
        s = HH[0];
        HH[0] = hh = t = t + tl;
        f = t - open1(c, A + i, B + 1);
        for (j = 1; j <= N; j++) {
            g = open1(c, A + i, B + j);
            h = ext1(c, A + i, B + j);
            if ((hh = hh - g - h) > (f = f - h)) f = hh;
            g = open2(c, A + i, B + j);
            h = ext2(c, A + i, B + j);
            if ((hh = HH[j] - g - h) > (e = DD[j] - h)) e = hh;
            hh = s + prf_score(c, A + i, B + j);
            if (f > hh) hh = f;
            if (e > hh) hh = e;
            s = HH[j];
            HH[j] = hh;
            DD[j] = e;
        }
    }
    DD[0] = HH[0];

    /* reverse phase: RR[j], SS[j], gS[j] */
    RR[N] = 0;
    tl = 0;
    for (j = N - 1; j >= 0; j--) {
        g = -open1(c, A + M, B + j + 1);
        tl -= ext1(c, A + M, B + j + 1);
        RR[j] = g + tl;
        SS[j] = RR[j] - open2(c, A + M, B + j);
        gS[j] = open2(c, A + M, B + j);
    }
    tl = 0;
    for (i = M - 1; i >= midi; i--) {
        s = RR[N];
        g = (go2 == 0) ? 0 : -open2(c, A + i + 1, B + N);
        tl -= ext2(c, A + i + 1, B + N);
        RR[N] = hh = g + tl;
        t = open1(c, A + i, B + N);
        f = RR[N] - t;
        for (j = N - 1; j >= 0; j--) {
            g = open1(c, A + i, B + j + 1);
            h = ext1(c, A + i, B + j + 1);
            if ((hh = hh - g - h) > (f = f - h - g + t)) f = hh;
            t = g;
            g = open2(c, A + i + 1, B + j);
            h = ext2(c, A + i + 1, B + j);
            hh = RR[j] - g - h;
            if (i == (M - 1)) {
                e = SS[j] - h;
            } else {
                e = SS[j] - h - g + open2(c, A + i + 2, B + j);
                gS[j] = g;
            }
            if (hh > e) e = hh;
            hh = s + prf_score(c, A + i + 1, B + j + 1);
            if (f > hh) hh = f;
            if (e > hh) hh = e;
            s = RR[j];
            RR[j] = hh;
            SS[j] = e;
        }
    }
    SS[N] = RR[N];
    gS[N] = open2(c, A + midi + 1, B + N);

    /* midpoint (:756-776) */
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
        hh = DD[j] + SS[j] + gS[j];
        if (hh > midh) {
            midh = hh;
            midj = j;
            type = 2;
        }
    }
    trace_node(c, A, B, M, N, go1, go2, type, midj, midh);
    if (type == 1) {
        prog_diff(c, A, B, midi, midj, go1, 1);
        prog_diff(c, A + midi, B + midj, M - midi, N - midj, 1, go2);
    } else {
        prog_diff(c, A, B, midi - 1, midj, go1, 0);
        prog_del(c, 2);
        prog_diff(c, A + midi + 1, B + midj, M - midi - 1, N - midj, 0, go2);
    }
    return midh;
}

int cwo_prog_diff_trace(const int *prf1, int len1, const int *prf2, int len2,
                        int end_gap_penalties, int *displ, int *n_displ, cwo_dp_stats *stats,
                        int *trace, int trace_cap, int *ntrace)
{
    dp_t c;
    int n = len1 + len2 + 3;
    memset(&c, 0, sizeof c);
    c.trace = trace;
    c.trace_cap = trace_cap;
    c.p1 = prf1;
    c.p2 = prf2;
    c.L1 = len1;
    c.L2 = len2;
    c.end_gap_penalties = end_gap_penalties;
    c.HH = (int *)calloc((size_t)n, sizeof(int));
    c.DD = (int *)calloc((size_t)n, sizeof(int));
    c.RR = (int *)calloc((size_t)n, sizeof(int));
    c.SS = (int *)calloc((size_t)n, sizeof(int));
    c.gS = (int *)calloc((size_t)n, sizeof(int));
    c.displ = (int *)calloc((size_t)n + 1, sizeof(int));
    c.lastPrint = 0;
    c.printPtr = 1;
    /* :428-429: the top call passes profile1's GAPCOL at 0 and at prfLength1 */
    int score = prog_diff(&c, 0, 0, len1, len2, prf1[GAPCOL],
                          prf1[(size_t)len1 * PRF_W + GAPCOL]);
    *n_displ = c.printPtr - 1;
    if (displ) memcpy(displ, c.displ + 1, sizeof(int) * (size_t)(c.printPtr - 1));
    if (stats) *stats = c.st;
    free(c.HH);
    free(c.DD);
    free(c.RR);
    free(c.SS);
    free(c.gS);
    free(c.displ);
    if (ntrace) *ntrace = c.ntrace;
    return score;
}

int cwo_prog_diff(const int *prf1, int len1, const int *prf2, int len2, int end_gap_penalties,
                  int *displ, int *n_displ, cwo_dp_stats *stats)
{
    return cwo_prog_diff_trace(prf1, len1, prf2, len2, end_gap_penalties, displ, n_displ, stats,
                               NULL, 0, NULL);
}

/* ---------------------------------------------------------------------- */
/* profileAlign (:36-505).  seqs[s] holds sequence s (0-based, no dummy
 * element) of length lens[s]; aligned group members are replaced by their
 * new rows (realloc'd).  group/aligned are 1-based like the reference's. */
static int profile_align(const cwo_msa_params *pm, int nseq, int **seqs, int *lens,
                         const double *sim, const int *weights, int *group,
                         const int *aligned, cwo_step_info *info)
{
    int i, j, nseqs1 = 0, nseqs2 = 0, count = 0, len, is, c;
    int prfLength1 = 0, prfLength2 = 0, len1, len2, minLen, matAvg = 0;
    int switchProfiles;
    double pcid = 0.0, logmin, logdiff;
    int mat[NUMRES][NUMRES];
    float scale;
    int gapcoef1, gapcoef2, lencoef1, lencoef2;
    const double *D = sim;
#define SIM(a, b) D[(size_t)(a) * (nseq + 1) + (b)]

    for (i = 0; i < nseq; i++)
        if (aligned[i + 1] == 0) group[i + 1] = 0;
    for (i = 0; i < nseq; i++) {
        if (group[i + 1] == 1)
            nseqs1++;
        else if (group[i + 1] == 2)
            nseqs2++;
    }
    if (nseqs1 == 0 || nseqs2 == 0) return 0;
    if (nseqs2 > nseqs1) {
        switchProfiles = 1;
        for (i = 0; i < nseq; i++) {
            if (group[i + 1] == 1)
                group[i + 1] = 2;
            else if (group[i + 1] == 2)
                group[i + 1] = 1;
        }
    } else {
        switchProfiles = 0;
    }
    (void)switchProfiles;   /* only selects structure masks, absent here */
    for (i = 0; i < nseq; i++)
        if (group[i + 1] == 1)
            for (j = 0; j < nseq; j++)
                if (group[j + 1] == 2) {
                    count++;
                    pcid += SIM(i + 1, j + 1);
                }
    pcid = pcid / (float)count;

    /* the two profiles' rows, group 1 first, padded with gapPos1 */
    int ntot = nseq;
    int **rows = (int **)calloc((size_t)ntot, sizeof(int *));
    int *w = (int *)calloc((size_t)ntot, sizeof(int));
    int *members = (int *)calloc((size_t)ntot, sizeof(int));
    for (i = 0; i < nseq; i++)
        if (group[i + 1] == 1 && lens[i] > prfLength1) prfLength1 = lens[i];
    nseqs1 = 0;
    for (i = 0; i < nseq; i++) {
        if (group[i + 1] == 1) {
            len = lens[i];
            rows[nseqs1] = (int *)calloc((size_t)prfLength1 + EXTRA_END, sizeof(int));
            memcpy(rows[nseqs1], seqs[i], sizeof(int) * (size_t)len);
            for (j = len; j < prfLength1; j++) rows[nseqs1][j] = GAPPOS1;
            w[nseqs1] = weights[i];
            members[nseqs1] = i;
            nseqs1++;
        }
    }
    for (i = 0; i < nseq; i++)
        if (group[i + 1] == 2 && lens[i] > prfLength2) prfLength2 = lens[i];
    nseqs2 = 0;
    for (i = 0; i < nseq; i++) {
        if (group[i + 1] == 2) {
            int k = nseqs1 + nseqs2;
            len = lens[i];
            rows[k] = (int *)calloc((size_t)prfLength2 + EXTRA_END, sizeof(int));
            memcpy(rows[k], seqs[i], sizeof(int) * (size_t)len);
            for (j = len; j < prfLength2; j++) rows[k][j] = GAPPOS1;
            rows[k][j] = ENDALN;
            w[k] = weights[i];
            members[k] = i;
            nseqs2++;
        }
    }
    len1 = 0;
    for (i = 0; i < nseqs1; i++) {
        is = 0;
        for (j = 0; j < prfLength1; j++) {
            c = rows[i][j];
            if (c != GAPPOS1 && c != GAPPOS2) is++;
        }
        len1 += is;
    }
    len1 = (int)(len1 / (float)nseqs1);
    len2 = 0;
    for (i = nseqs1; i < nseqs1 + nseqs2; i++) {
        is = 0;
        for (j = 0; j < prfLength2; j++) {
            c = rows[i][j];
            if (c != GAPPOS1 && c != GAPPOS2) is++;
        }
        len2 += is;
    }
    len2 = (int)(len2 / (float)nseqs2);
    minLen = len1 < len2 ? len1 : len2;
    int maxRes = step_matrix(pm, pcid, minLen, mat, &scale, &matAvg);
    if (maxRes == 0 || maxRes == -1) return -1;
    if (pm->dna) {
        gapcoef1 = gapcoef2 = (int)(100.0 * pm->gap_open * scale);
        lencoef1 = lencoef2 = (int)(100.0 * pm->gap_extend * scale);
    } else {
        if (len1 == 0 || len2 == 0) {
            logmin = 1.0;
            logdiff = 1.0;
        } else {
            logmin = 1.0 / log10((double)minLen);
            if (len2 < len1)
                logdiff = 1.0 + 0.5 * log10((double)((float)len2 / (float)len1));
            else if (len1 < len2)
                logdiff = 1.0 + 0.5 * log10((double)((float)len1 / (float)len2));
            else
                logdiff = 1.0;
            if (logdiff < 0.9) logdiff = 0.9;
        }
        if (pm->neg_matrix) {
            gapcoef1 = gapcoef2 = (int)(100.0 * (float)(pm->gap_open));
            lencoef1 = lencoef2 = (int)(100.0 * pm->gap_extend);
        } else {
            if (matAvg <= 0)
                gapcoef1 = gapcoef2 = (int)(100.0 * (float)(pm->gap_open + logmin));
            else
                gapcoef1 = gapcoef2 = (int)(scale * matAvg *
                                            (float)(pm->gap_open / (logdiff * logmin)));
            lencoef1 = lencoef2 = (int)(100.0 * pm->gap_extend);
        }
    }
    prof_t pw, ps;
    int *gaps = (int *)calloc((size_t)prfLength1 + prfLength2 + 3, sizeof(int));
    prof_init(&pw, prfLength1, 0, nseqs1);
    prof_init(&ps, prfLength2, nseqs1, nseqs1 + nseqs2);
    calc_gap_coeff(&pw, rows, gaps, pm, gapcoef1, lencoef1);
    calc_profile_with_sub(&pw, rows, gaps, mat, w);
    calc_gap_coeff(&ps, rows, gaps, pm, gapcoef2, lencoef2);
    calc_standard_profile(&ps, rows, w);

    int ndmax = prfLength1 + prfLength2 + 3;
    int *displ = (int *)calloc((size_t)ndmax + 1, sizeof(int));
    int nd = 0;
    cwo_dp_stats st;
    int score = cwo_prog_diff(pw.profile, prfLength1, ps.profile, prfLength2,
                              pm->end_gap_penalties, displ, &nd, &st);

    /* progTracepath (:1023-1062) */
    int *path1 = (int *)calloc((size_t)ndmax + 1, sizeof(int));
    int *path2 = (int *)calloc((size_t)ndmax + 1, sizeof(int));
    int pos = 0;
    for (i = 0; i < nd; ++i) {
        int k = displ[i];
        if (k == 0) {
            path1[pos] = 2;
            path2[pos] = 2;
            ++pos;
        } else if (k > 0) {
            for (j = 0; j <= k - 1; ++j) {
                path2[pos + j] = 2;
                path1[pos + j] = 1;
            }
            pos += k;
        } else {
            k = -k;
            for (j = 0; j <= k - 1; ++j) {
                path1[pos + j] = 2;
                path2[pos + j] = 1;
            }
            pos += k;
        }
    }
    int alen = pos;

    /* addGGaps (:818-934) and the write-back (:464-493) */
    for (int r = 0; r < nseqs1 + nseqs2; r++) {
        const int *path = (r < nseqs1) ? path1 : path2;
        int rowlen = (r < nseqs1) ? prfLength1 : prfLength2;   /* size - extraEndElemNum */
        int *ta = (int *)malloc(sizeof(int) * ((size_t)alen + 1));
        int ix = 0;
        for (i = 0; i < alen; i++) {
            if (path[i] == 2) {
                ta[i] = (ix < rowlen) ? rows[r][ix] : ENDALN;
                ix++;
            } else {
                ta[i] = GAPPOS1;
            }
        }
        int s = members[r];
        free(seqs[s]);
        seqs[s] = ta;
        lens[s] = alen;
    }
    if (info) {
        info->score = score;
        info->prf_len1 = prfLength1;
        info->prf_len2 = prfLength2;
        info->nseqs1 = nseqs1;
        info->nseqs2 = nseqs2;
        info->aln_len = alen;
        info->n_displ = nd;
        info->pcid = pcid;
        info->gapcoef = gapcoef1;
        info->lencoef = lencoef1;
        info->mat_avg = matAvg;
        info->stats = st;
    }
    for (i = 0; i < ntot; i++) free(rows[i]);
    free(rows);
    free(w);
    free(members);
    free(gaps);
    free(pw.profile);
    free(ps.profile);
    free(displ);
    free(path1);
    free(path2);
#undef SIM
    return score / 100;
}

/* ---------------------------------------------------------------------- */
/* The profiles of one step, for handing to another DP implementation. */
int cwo_step_profiles(const cwo_msa_params *pm, int nseq, int *const *seqs, const int *lens,
                      const double *sim, const int *weights, const int *group_in,
                      const int *aligned, int **prf1, int *len1, int **prf2, int *len2)
{
    /* a throw-away copy of the group rows; profile_align's set-up is run up
     * to the profiles and the DP result is discarded */
    int i;
    int *group = (int *)malloc(sizeof(int) * ((size_t)nseq + 1));
    memcpy(group, group_in, sizeof(int) * ((size_t)nseq + 1));
    for (i = 0; i < nseq; i++)
        if (aligned[i + 1] == 0) group[i + 1] = 0;
    int n1 = 0, n2 = 0;
    for (i = 0; i < nseq; i++) {
        if (group[i + 1] == 1) n1++;
        else if (group[i + 1] == 2) n2++;
    }
    if (n1 == 0 || n2 == 0) {
        free(group);
        return 0;
    }
    if (n2 > n1) {
        for (i = 0; i < nseq; i++) {
            if (group[i + 1] == 1) group[i + 1] = 2;
            else if (group[i + 1] == 2) group[i + 1] = 1;
        }
        int t = n1;          /* profile 1 takes the larger group, as profile_align recounts */
        n1 = n2;
        n2 = t;
    }
    double pcid = 0.0;
    int count = 0;
    for (i = 0; i < nseq; i++)
        if (group[i + 1] == 1)
            for (int j = 0; j < nseq; j++)
                if (group[j + 1] == 2) {
                    count++;
                    pcid += sim[(size_t)(i + 1) * (nseq + 1) + (j + 1)];
                }
    pcid = pcid / (float)count;
    int L1 = 0, L2 = 0, k = 0;
    int **rows = (int **)calloc((size_t)nseq, sizeof(int *));
    int *w = (int *)calloc((size_t)nseq, sizeof(int));
    for (i = 0; i < nseq; i++) if (group[i + 1] == 1 && lens[i] > L1) L1 = lens[i];
    for (i = 0; i < nseq; i++) if (group[i + 1] == 2 && lens[i] > L2) L2 = lens[i];
    for (int gsel = 1; gsel <= 2; gsel++) {
        int L = gsel == 1 ? L1 : L2;
        for (i = 0; i < nseq; i++) {
            if (group[i + 1] != gsel) continue;
            rows[k] = (int *)calloc((size_t)L + EXTRA_END, sizeof(int));
            memcpy(rows[k], seqs[i], sizeof(int) * (size_t)lens[i]);
            int j;
            for (j = lens[i]; j < L; j++) rows[k][j] = GAPPOS1;
            if (gsel == 2) rows[k][j] = ENDALN;
            w[k] = weights[i];
            k++;
        }
    }
    int l1 = 0, l2 = 0;
    for (i = 0; i < n1; i++) {
        int is = 0;
        for (int j = 0; j < L1; j++) if (rows[i][j] != GAPPOS1 && rows[i][j] != GAPPOS2) is++;
        l1 += is;
    }
    l1 = (int)(l1 / (float)n1);
    for (i = n1; i < n1 + n2; i++) {
        int is = 0;
        for (int j = 0; j < L2; j++) if (rows[i][j] != GAPPOS1 && rows[i][j] != GAPPOS2) is++;
        l2 += is;
    }
    l2 = (int)(l2 / (float)n2);
    int minLen = l1 < l2 ? l1 : l2, matAvg = 0, mat[NUMRES][NUMRES];
    float scale;
    if (step_matrix(pm, pcid, minLen, mat, &scale, &matAvg) <= 0) {
        for (i = 0; i < nseq; i++) free(rows[i]);
        free(rows);
        free(w);
        free(group);
        return -1;
    }
    int gapcoef, lencoef;
    if (pm->dna) {
        gapcoef = (int)(100.0 * pm->gap_open * scale);
        lencoef = (int)(100.0 * pm->gap_extend * scale);
    } else {
        double logmin, logdiff;
        if (l1 == 0 || l2 == 0) {
            logmin = 1.0;
            logdiff = 1.0;
        } else {
            logmin = 1.0 / log10((double)minLen);
            if (l2 < l1) logdiff = 1.0 + 0.5 * log10((double)((float)l2 / (float)l1));
            else if (l1 < l2) logdiff = 1.0 + 0.5 * log10((double)((float)l1 / (float)l2));
            else logdiff = 1.0;
            if (logdiff < 0.9) logdiff = 0.9;
        }
        if (pm->neg_matrix) {
            gapcoef = (int)(100.0 * (float)(pm->gap_open));
        } else if (matAvg <= 0) {
            gapcoef = (int)(100.0 * (float)(pm->gap_open + logmin));
        } else {
            gapcoef = (int)(scale * matAvg * (float)(pm->gap_open / (logdiff * logmin)));
        }
        lencoef = (int)(100.0 * pm->gap_extend);
    }
    prof_t pw, ps;
    int *gaps = (int *)calloc((size_t)L1 + L2 + 3, sizeof(int));
    prof_init(&pw, L1, 0, n1);
    prof_init(&ps, L2, n1, n1 + n2);
    calc_gap_coeff(&pw, rows, gaps, pm, gapcoef, lencoef);
    calc_profile_with_sub(&pw, rows, gaps, mat, w);
    calc_gap_coeff(&ps, rows, gaps, pm, gapcoef, lencoef);
    calc_standard_profile(&ps, rows, w);
    *prf1 = pw.profile;
    *prf2 = ps.profile;
    *len1 = L1;
    *len2 = L2;
    for (i = 0; i < nseq; i++) free(rows[i]);
    free(rows);
    free(w);
    free(gaps);
    free(group);
    return 1;
}

/* ---------------------------------------------------------------------- */
/* multiSeqAlign (MSA.cpp:32-347) with iStart = 0. */
int cwo_msa(const cwo_msa_params *pm, int nseq, int **seqs, int *lens, const double *sim,
            int *seq_weight, const int *steps, int nsteps, int *out_order,
            cwo_step_info *step_info, int *n_info)
{
    int i, j, ix = 0, set, entries, score, iseq, sum;
    int max;
    int *aligned = (int *)calloc((size_t)nseq + 1, sizeof(int));
    int *maxid = (int *)calloc((size_t)nseq + 1, sizeof(int));
    int *group = (int *)calloc((size_t)nseq + 1, sizeof(int));
    int *treeWeight = (int *)calloc((size_t)nseq + 1, sizeof(int));
    int *weights = (int *)calloc((size_t)nseq + 1, sizeof(int));
    int *stepbuf = (int *)malloc(sizeof(int) * ((size_t)nseq + 1));
    int ninfo = 0;
#define SIM(a, b) sim[(size_t)(a) * (nseq + 1) + (b)]
    memcpy(weights, seq_weight, sizeof(int) * (size_t)nseq);   /* addSeqWeight */
    for (i = 1; i <= nseq; i++) {
        maxid[i] = -1;
        for (j = 1; j <= nseq; j++)
            if (j != i && maxid[i] < SIM(i, j)) maxid[i] = (int)SIM(i, j);
    }
    for (set = 1; set <= nsteps; ++set) {
        const int *row = steps + (size_t)(set - 1) * (nseq + 1);
        entries = 0;
        for (i = 1; i <= nseq; i++) {
            if (row[i] != 0 && maxid[i] > pm->divergence_cutoff) {
                entries++;
                if (aligned[i] == 0) {
                    out_order[ix] = i;       /* outputOrder ALIGNED */
                    ++ix;
                    aligned[i] = 1;
                }
            }
        }
        if (entries > 0) {
            memcpy(stepbuf, row, sizeof(int) * ((size_t)nseq + 1));
            score = profile_align(pm, nseq, seqs, lens, sim, weights, stepbuf, aligned,
                                  step_info ? &step_info[ninfo] : NULL);
            if (step_info) step_info[ninfo].step = set;
            ninfo++;
        } else {
            score = 0;
        }
        if (score < 0) {
            ix = -1;
            goto out;
        }
    }
    for (i = 0; i < nseq; i++) treeWeight[i] = seq_weight[i];
    if (ix == 0) {
        max = -1;
        iseq = 0;
        for (i = 1; i <= nseq; i++)
            for (j = i + 1; j <= nseq; j++)
                if (max < SIM(i, j)) {
                    max = (int)SIM(i, j);
                    iseq = i;
                }
        aligned[iseq] = 1;
        out_order[ix] = iseq;
        ++ix;
    }
    while (ix < nseq) {
        for (i = 1; i <= nseq; i++) {
            if (aligned[i] == 0) {
                maxid[i] = -1;
                for (j = 1; j <= nseq; j++)
                    if (maxid[i] < SIM(i, j) && aligned[j] != 0) maxid[i] = (int)SIM(i, j);
            }
        }
        max = -1;
        iseq = 0;
        for (i = 1; i <= nseq; i++)
            if (aligned[i] == 0 && maxid[i] > max) {
                max = maxid[i];
                iseq = i;
            }
        if (!pm->no_weights) {
            for (j = 0; j < nseq; j++)
                if (aligned[j + 1] != 0) weights[j] = (int)(treeWeight[j] * SIM(j + 1, iseq));
            sum = 0;
            for (j = 0; j < nseq; j++)
                if (aligned[j + 1] != 0) sum += weights[j];
            if (sum == 0) {
                for (j = 0; j < nseq; j++) weights[j] = 1;
                sum = j;
            }
            for (j = 0; j < nseq; j++) {
                if (aligned[j + 1] != 0) {
                    weights[j] = (weights[j] * INT_SCALE_FACTOR) / sum;
                    if (weights[j] < 1) weights[j] = 1;
                }
            }
        }
        for (j = 1; j <= nseq; j++) {
            if (aligned[j] != 0)
                group[j] = 1;
            else if (iseq == j)
                group[j] = 2;
        }
        aligned[iseq] = 1;
        score = profile_align(pm, nseq, seqs, lens, sim, weights, group, aligned,
                              step_info ? &step_info[ninfo] : NULL);
        if (step_info) step_info[ninfo].step = -iseq;
        ninfo++;
        out_order[ix] = iseq;
        ++ix;
    }
out:
    if (n_info) *n_info = ninfo;
    free(aligned);
    free(maxid);
    free(group);
    free(treeWeight);
    free(weights);
    free(stepbuf);
#undef SIM
    return ix < 0 ? -1 : nseq;
}

void cwo_msa_params_default(int dna, cwo_msa_params *p)
{
    memset(p, 0, sizeof *p);
    p->dna = dna;
    p->gap_open = dna ? 15.0f : 10.0f;      /* DNAGapOpen / AAGapOpen */
    p->gap_extend = dna ? 6.66f : 0.2f;     /* DNAGapExtend / AAGapExtend */
    p->gap_dist = 4;
    p->no_var_penalties = 1;
    p->no_hyd_penalties = 0;
    p->no_pref_penalties = 0;
    p->use_end_gaps = 0;
    p->end_gap_penalties = 0;
    p->transition_weight = 0.5f;
    p->neg_matrix = 0;
    p->no_weights = 0;
    p->divergence_cutoff = 30;
    p->hyd_residues = "GPSNDQEKR";
    p->distance_tree = 1;
    p->profile_matrix = NULL;
    p->profile_matrix_user = NULL;
}
