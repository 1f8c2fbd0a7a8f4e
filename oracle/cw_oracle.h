/*
 * cw_oracle.h -- CPU restatement of ClustalW 2.0.11's full-DP pairwise
 * distance path, used ONLY as a test checker and as the CPU baseline of the
 * benchmark. Nothing in the shipped library links or calls this code.
 *
 * Every function cites the reference source it restates (paths relative to
 * CPU/clustalw-2.0.11/src/ of coldfunction/CUDA-clustalW).
 */
#ifndef CW_ORACLE_H
#define CW_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CWO_NUMRES 32

/* matrix ids for cwo_pairwise_matrix */
enum { CWO_MAT_GONNET250 = 0, CWO_MAT_IUB = 1 };

typedef struct {
    int max_score, se1, se2;     /* forwardPass result */
    int sb1, sb2;                /* reversePass result */
    int mm_score;                /* value returned by diff() */
    int count;                   /* identities counted by tracePath */
    int n_displ;                 /* printPtr - 1 */
    int count_leaves;            /* identities summed over diff's M==1 leaves */
    int64_t cells_fwd, cells_rev, cells_mm;
    int64_t mm_nodes;            /* internal (M>=2, N>=1) diff nodes */
} cwo_pair_result;

/* SubMatrix::getPairwiseMatrix + getMatrix (SubMatrix.cpp:256-390, 689-829).
 * Fills mat[32][32], returns maxRes (0 on failure); *mat_avg = matrixAvgScore,
 * *int_scale = PairScaleValues.intScale, *go_scale and *ge_scale the gap scales.
 * transition_weight is UserParameters' transitionWeight (DNA only). */
int cwo_pairwise_matrix(int which, float transition_weight, int mat[CWO_NUMRES][CWO_NUMRES],
                        int *mat_avg, int *int_scale, float *go_scale, float *ge_scale);

/* Gap parameters of FullPairwiseAlign::pairwiseAlign (FullPairwiseAlign.cpp:129-148)
 * for one pair whose gapped lengths are n and m. */
void cwo_gap_params(int dna, int mat_avg, int int_scale, float go_scale, float ge_scale,
                    float pw_go, float pw_ge, int n, int m, int *go, int *ge);

/* One pair: forwardPass, reversePass, diff, tracePath (FullPairwiseAlign.cpp:151-170).
 * seq1/seq2 are 1-based code arrays (element 0 unused) of gapped length n, m.
 * displ (optional) receives the edit script displ[1..n_displ]; it must hold
 * 2*(n+m)+2 ints. */
void cwo_align_pair(const int *seq1, int n, const int *seq2, int m,
                    const int mat[CWO_NUMRES][CWO_NUMRES], int go, int ge,
                    int gap_pos1, int gap_pos2, int *displ, cwo_pair_result *res);

/* Distance of FullPairwiseAlign.cpp:170-181 from an identity count and the two
 * gap-free lengths; returns the float that the reference stores. */
float cwo_distance(int count, int len1, int len2);

/* The pair loop of FullPairwiseAlign::pairwiseAlign (FullPairwiseAlign.cpp:97-183)
 * over packed uint8 codes (sequence s occupies codes[offs[s] .. offs[s]+lens[s])).
 * Pairs are visited in the reference's order; pair p gets out_count[p] (-1 when
 * n or m is 0), out_dist[p], and (si,sj) in out_pairs[2p..2p+1].  go_by_len[L]
 * is the gap-open for min(n,m)==L.  Returns the number of pairs.  nthreads>1
 * splits the pairs over OpenMP threads (results are identical).  cells, if
 * non-null, receives {fwd, rev, mm, nodes} summed over pairs. */
int64_t cwo_pairwise_range(const uint8_t *codes, const int64_t *offs, const int32_t *lens,
                           int nseq, int i_start, int i_end, int j_start, int j_end,
                           const int mat[CWO_NUMRES][CWO_NUMRES], const int32_t *go_by_len,
                           int ge, int gap_pos1, int gap_pos2, int nthreads,
                           int32_t *out_pairs, int32_t *out_count, float *out_dist,
                           int64_t *cells);

/* Number of pairs the range visits (same loop bounds as above). */
int64_t cwo_count_pairs(int nseq, int i_start, int i_end, int j_start, int j_end);

/* Explicit pair list variant: pairs[2p], pairs[2p+1] are 0-based (si, sj). */
void cwo_pairwise_list(const uint8_t *codes, const int64_t *offs, const int32_t *lens,
                       const int32_t *pairs, int64_t npairs,
                       const int mat[CWO_NUMRES][CWO_NUMRES], const int32_t *go_by_len,
                       int ge, int gap_pos1, int gap_pos2, int nthreads,
                       int32_t *out_count, float *out_dist, cwo_pair_result *out_res);


/* ---------------------------------------------------------------------- */
/* Progressive (profile) alignment: oracle/cw_msa.c and oracle/cw_guide.c.  */

#define CWO_PRF_STRIDE 35   /* ints per profile row: codes 0..31, GAPCOL 32, LENCOL 33, pad */

/* The UserParameters the progressive stage reads (UserParameters.cpp:33-175,
 * set*Params :501-523); cwo_msa_params_default fills the reference defaults. */
typedef struct {
    int dna;
    float gap_open, gap_extend;      /* gapOpen / gapExtend (multiple alignment) */
    int gap_dist;                    /* gapDist */
    int no_var_penalties, no_hyd_penalties, no_pref_penalties;
    int use_end_gaps, end_gap_penalties;
    float transition_weight;
    int neg_matrix;                  /* useNegMatrix */
    int no_weights;
    int divergence_cutoff;
    const char *hyd_residues;
    int distance_tree;               /* UserParameters::getDistanceTree (SubMatrix.cpp:490,534,565) */
    /* optional stand-in for SubMatrix::getProfileAlignMatrix with another
     * matrix series: fills mat, *scale, *mat_avg and returns maxRes */
    int (*profile_matrix)(void *user, double pcid, int min_len, int *mat, float *scale,
                          int *mat_avg);
    void *profile_matrix_user;
} cwo_msa_params;

typedef struct {
    int64_t fwd_cells, rev_cells, leaf_cells, nodes;
} cwo_dp_stats;

typedef struct {
    int step;                        /* guide-tree step (>0) or -seq of a delayed one */
    int score;                       /* progDiff's value (profileAlign returns score/100) */
    int prf_len1, prf_len2, nseqs1, nseqs2, aln_len, n_displ;
    double pcid;
    int gapcoef, lencoef, mat_avg;
    cwo_dp_stats stats;
} cwo_step_info;

void cwo_msa_params_default(int dna, cwo_msa_params *p);

/* SubMatrix::getProfileAlignMatrix (SubMatrix.cpp:403-677) for the default
 * matrix names (gonnet series / iub) and getMatrix(negFlag = neg_matrix). */
int cwo_profile_matrix(int dna, double pcid, int min_len, float transition_weight,
                       int neg_matrix, int distance_tree, int mat[CWO_NUMRES][CWO_NUMRES],
                       float *scale, int *mat_avg);

/* MyersMillerProfileAlign::progDiff on two profiles of (len+2) rows of
 * CWO_PRF_STRIDE ints; the top call's go1/go2 are prf1's GAPCOL at rows 0 and
 * len1 (MyersMillerProfileAlign.cpp:428-429).  displ receives the edit script
 * (printPtr-1 entries, at most len1+len2+2); returns progDiff's score. */
int cwo_prog_diff(const int *prf1, int len1, const int *prf2, int len2, int end_gap_penalties,
                  int *displ, int *n_displ, cwo_dp_stats *stats);

/* cwo_prog_diff that also records every progDiff call in depth-first order:
 * 8 ints each (A, B, M, N, go1 | go2 << 1, type (0 = base case, 1, 2), midj,
 * returned value); *ntrace gets the number of calls (records beyond
 * trace_cap are dropped). */
int cwo_prog_diff_trace(const int *prf1, int len1, const int *prf2, int len2,
                        int end_gap_penalties, int *displ, int *n_displ, cwo_dp_stats *stats,
                        int *trace, int trace_cap, int *ntrace);

/* The two profiles profileAlign would build for one step (malloc'd; free
 * with cwo_free).  Returns 0 when a group is empty, -1 when the profile
 * matrix fails, 1 otherwise. */
int cwo_step_profiles(const cwo_msa_params *pm, int nseq, int *const *seqs, const int *lens,
                      const double *sim, const int *weights, const int *group,
                      const int *aligned, int **prf1, int *len1, int **prf2, int *len2);

/* MSA::multiSeqAlign (MSA.cpp:32-347, iStart 0).  seqs[s] (malloc'd, lens[s]
 * codes) are replaced by their aligned rows; sim is the (n+1)^2 similarity
 * matrix, steps nsteps rows of n+1 ints (element 0 unused).  out_order gets
 * the 1-based output order.  step_info (optional, room for nsteps + n) gets
 * one record per profileAlign call.  Returns n, or -1. */
int cwo_msa(const cwo_msa_params *pm, int nseq, int **seqs, int *lens, const double *sim,
            int *seq_weight, const int *steps, int nsteps, int *out_order,
            cwo_step_info *step_info, int *n_info);

/* Tree::readTree + calcSeqWeights + calcSimilarities + createSets as
 * TreeInterface::getWeightsAndStepsUseOldGuideTreeNJ calls them
 * (TreeInterface.cpp:209-276), from the .dnd text.  names are the alignment's
 * sequence names.  weights[n], sim[(n+1)^2]; *steps (malloc'd) gets nsteps
 * rows of n+1 ints.  Returns 1 or a CWO_GUIDE_* error. */
enum { CWO_GUIDE_TOO_FEW = -1, CWO_GUIDE_FORMAT = -2, CWO_GUIDE_MISMATCH = -3 };
int cwo_guide_from_dnd(const char *dnd, int n, const char **names, int *weights,
                       double *sim, int **steps, int *nsteps);

/* FastPairwiseAlign::pairwiseAlign (pairwise/FastPairwiseAlign.cpp:20-145),
 * the k-tuple distances: sequences i = iStart+1 .. iEnd against
 * j = max(jStart+2, i+1) .. jEnd (1-based), seqs 0-based code arrays.
 * max_aln is Alignment::getMaxAlnLength (2 x the longest sequence).  pairs
 * (2 per pair, 0-based) and dist may be NULL.  Returns the pair count, or -1. */
long cwo_fast_pairwise(int nseq, const int *const *seqs, const int *lens, int dna, int ktup,
                       int window, int wind_gap, int signif, int percent, int max_aln, int iStart,
                       int iEnd, int jStart, int jEnd, int *pairs, double *dist);

#ifdef __cplusplus
}
#endif
#endif
