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

#ifdef __cplusplus
}
#endif
#endif
