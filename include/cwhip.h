/*
 * cwhip.h -- C ABI of the MI355X (gfx950) implementation of ClustalW 2.0.11's
 * full-DP pairwise distance stage.
 *
 * The ABI carries plain pointers and sizes only.  A context owns the device
 * buffers (sequences resident in HBM, per-wave scratch, results) and is used
 * from one host thread at a time.  Functions return 0 on success and a
 * negative CWHIP_E* code on failure; cwhip_last_error() describes the last
 * failure of a context.
 *
 * Reference interfaces replaced (paths in coldfunction/CUDA-clustalW under
 * CPU/clustalw-2.0.11/src/):
 *   cwhip_pairwise_range  <- FullPairwiseAlign::pairwiseAlign(Alignment*, DistMatrix*,
 *                            int iStart, int iEnd, int jStart, int jEnd)
 *                            (pairwise/FullPairwiseAlign.cpp:34-211,
 *                             declared pairwise/PairwiseAlignBase.h:19-20)
 *   cwhip_pair_params_default <- SubMatrix::getPairwiseMatrix + the gap
 *                            parameters of FullPairwiseAlign.cpp:129-148
 *                            (substitutionMatrix/SubMatrix.cpp:256-390, 689-829)
 *   cwhip_distance        <- the float distance of FullPairwiseAlign.cpp:170-181
 *   cwhip_fast_pairwise_range <- FastPairwiseAlign::pairwiseAlign (k-tuple
 *                            distances, pairwise/FastPairwiseAlign.cpp:20-145)
 * The GPU variant of the reference exposes the same operation as the free
 * function cudaFullPairwiseAlign(Alignment*, DistMatrix*, int,int,int,int)
 * (GPU/cuda_clustalw_final/src/cuda/cudaFullPairwiseAlign.h:20).
 */
#ifndef CWHIP_H
#define CWHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CWHIP_NUMRES 32

enum {
    CWHIP_OK = 0,
    CWHIP_E_ARG = -1,       /* bad argument (null pointer, bad range, size) */
    CWHIP_E_HIP = -2,       /* HIP runtime error */
    CWHIP_E_NOSEQ = -3,     /* no sequences loaded */
    CWHIP_E_TOOLONG = -4,   /* a sequence exceeds CWHIP_MAX_LEN */
    CWHIP_E_KERNEL = -5,    /* a kernel reported an internal inconsistency */
    CWHIP_E_NOMEM = -6,     /* device allocation failed, or a job larger than the HBM budget */
    CWHIP_E_TIMEOUT = -7,   /* cwhip_server_wait: the job did not finish in time */
    CWHIP_E_MATRIX = -8     /* the profile matrix of a step could not be set up (the caller's
                               cwhip_profile_matrix_fn returned <= 0) */
};

/* Longest sequence the pairwise kernels accept: a DP cell's (row, column) is
 * packed as row << 16 | column in one signed 32-bit key.  Both sequences of a
 * pair sit in a wave's LDS, so beyond ~16000 residues a workgroup runs fewer
 * than four waves (two at 32767). */
#define CWHIP_MAX_LEN 32767
#define CWHIP_MAX_PROFILE_LEN 64000  /* longest profile (columns) cwhip_profile_dp accepts */

/* kernel choice for cwhip_pairwise*; both give identical results */
enum {
    CWHIP_KERNEL_PLAIN = 0,   /* one wavefront per pair, launched over pair blocks */
    CWHIP_KERNEL_BATCHED = 1  /* persistent kernel over a length-binned device work queue,
                                 shapes by tuning class, stages replayed as a hipGraph */
};

/* Scoring of one pairwise run.  matrix is NUMRES x NUMRES, row-major, indexed
 * by residue code; gap_open_by_len[L] is the scaled gap-open penalty of a pair
 * whose shorter gapped length is L (0 <= L <= max_len); gap_extend is the
 * scaled extension penalty; gap_pos1/gap_pos2 are the two gap codes. */
typedef struct {
    int32_t matrix[CWHIP_NUMRES * CWHIP_NUMRES];
    int32_t gap_extend;
    int32_t gap_pos1;
    int32_t gap_pos2;
    int32_t max_len;
    const int32_t *gap_open_by_len;
} cwhip_pair_params;

typedef struct cwhip_ctx cwhip_ctx;

/* Create a context on HIP device `device`. */
int cwhip_create(int device, cwhip_ctx **out);
void cwhip_destroy(cwhip_ctx *ctx);
const char *cwhip_last_error(const cwhip_ctx *ctx);

/* Scoring set up the way the reference does for the pairwise stage
 * (SubMatrix.cpp:256-390; FullPairwiseAlign.cpp:129-148): dna=0 -> Gonnet 250
 * x10, gap open int(2*matAvg*(pw_gap_open + ln L)) in double, extend
 * int(pw_gap_extend*10); dna=1 -> IUB x100 with transition weight, open
 * int(2*pw_gap_open*100), extend int(pw_gap_extend*100).  gap_open_buf must
 * hold max_len+1 ints; p->gap_open_by_len is pointed at it. */
int cwhip_pair_params_default(int dna, float pw_gap_open, float pw_gap_extend,
                              float transition_weight, int max_len,
                              cwhip_pair_params *p, int32_t *gap_open_buf);

/* Upload sequences: sequence s is codes[offsets[s] .. offsets[s]+lens[s]) (codes
 * as Sequence::encodeSequence produces them, without the leading dummy).  The
 * sequences stay resident on the device until the next call or destroy. */
int cwhip_load_sequences(cwhip_ctx *ctx, const uint8_t *codes, const int64_t *offsets,
                         const int32_t *lens, int nseq);

/* Number of pairs the reference's loops visit for a range (FullPairwiseAlign.cpp:97,110). */
int64_t cwhip_range_npairs(int nseq, int i_start, int i_end, int j_start, int j_end);

/* The pairwise stage over the reference's range semantics: pairs (si, sj) with
 * si in [max(0,i_start), min(N,i_end)), sj in [max(si+1, j_start+1), min(N,j_end)),
 * in that visiting order.  out_pairs (2 ints per pair, may be NULL), out_count
 * (identities; -1 when either sequence is empty, may be NULL) and out_dist
 * (the float distance the reference stores; 1.0 for empty sequences) receive
 * one entry per pair.  Returns CWHIP_E_ARG for the reference's fatal range
 * check (i_start<0, i_end<i_start, j_start<0, j_end<j_start). */
int cwhip_pairwise_range(cwhip_ctx *ctx, const cwhip_pair_params *p, int i_start, int i_end,
                         int j_start, int j_end, int kernel, int32_t *out_pairs,
                         int32_t *out_count, float *out_dist);

/* Same for an explicit list of 0-based (si, sj) pairs. */
int cwhip_pairwise_pairs(cwhip_ctx *ctx, const cwhip_pair_params *p, const int32_t *pairs,
                         int64_t npairs, int kernel, int32_t *out_count, float *out_dist);

/* ---- k-tuple distances (the -QUICKTREE path) ------------------------------
 * FastPairwiseAlign::pairwiseAlign(Alignment*, DistMatrix*, int iStart,
 * int iEnd, int jStart, int jEnd) (pairwise/FastPairwiseAlign.cpp:20-145,
 * declared pairwise/FastPairwiseAlign.h, the same PairwiseAlignBase.h:19-20
 * interface).  Parameters as UserParameters keeps them for the k-tuple stage
 * (ktup, window, windowGap, signif, percent; general/UserParameters.cpp:86-94,
 * 501-532); max_aln is Alignment::getMaxAlnLength (0: twice the longest loaded
 * sequence, what the reference has after reading the sequences). */
typedef struct {
    int32_t dna;
    int32_t ktup;        /* protein 1..2, DNA 1..4 */
    int32_t window, wind_gap, signif;
    int32_t percent;     /* score as percent of the shorter length */
    int32_t max_aln;
} cwhip_fast_params;

/* The reference defaults: protein ktup 1, window 5, window gap 3, 5 top
 * diagonals; DNA 2, 4, 5, 4; percent on. */
int cwhip_fast_params_default(int dna, cwhip_fast_params *fp);

/* k-tuple distances of the loaded sequences over the reference's range:
 * sequences i = i_start+1 .. i_end against j = max(j_start+2, i+1) .. j_end
 * (1-based, :64-80), in that order; *out_npairs receives the count, out_pairs
 * (2 ints per pair, 0-based, may be NULL) and out_dist (the double distance
 * (100 - score) / 100 the reference stores with DistMatrix::SetAt) one entry
 * per pair.  Returns CWHIP_E_ARG for the reference's fatal range checks and for
 * a range beyond the loaded sequences. */
int cwhip_fast_pairwise_range(cwhip_ctx *ctx, const cwhip_fast_params *fp, int i_start, int i_end,
                              int j_start, int j_end, int32_t *out_pairs, double *out_dist,
                              int64_t *out_npairs);

/* Kernel time (ms) of the last cwhip_fast_pairwise_range. */
float cwhip_fast_last_ms(const cwhip_ctx *ctx);

/* Distance of FullPairwiseAlign.cpp:170-181 from an identity count and the
 * two gap-free lengths (IEEE single arithmetic, as the reference). */
float cwhip_distance(int32_t count, int32_t len1, int32_t len2);

/* ---- device-resident timing interface (inputs already in HBM) ---------- */

/* Stage a pair list and the scoring on the device without running it. */
int cwhip_stage_pairs(cwhip_ctx *ctx, const cwhip_pair_params *p, const int32_t *pairs,
                      int64_t npairs);
/* Run the staged pairs `reps` times with `kernel`; *ms receives the device time of
 * the last repetition measured with HIP events on the context's stream, and
 * *kernel_ms the time of the main DP kernel alone. */
int cwhip_run_staged(cwhip_ctx *ctx, int kernel, int reps, float *ms, float *kernel_ms);
/* Copy the identity counts of the last run (staged order) to the host. */
int cwhip_fetch_counts(cwhip_ctx *ctx, int32_t *out_count, int64_t npairs);
/* Per-pair diagnostics of the last run, 8 ints per pair:
 * [maxScore, se1, se2, sb1, sb2, status, Myers-Miller internal nodes,
 *  Myers-Miller cell updates of the reference's diff()].  The reference's
 * forward pass updates n*m cells and its reverse pass (se1-sb1)*se2 +
 * (se2-sb2+1) cells (0 when maxScore is 0).  The last field is an unsigned
 * 32-bit count: pairs of two sequences near CWHIP_MAX_LEN reach ~3.4e9. */
int cwhip_fetch_info(cwhip_ctx *ctx, int32_t *out_info, int64_t npairs);

/* Per-wave tuning knobs (0 = keep the built-in default): rows per lane limit,
 * waves per workgroup, persistent workgroups per CU. */
int cwhip_set_tuning(cwhip_ctx *ctx, int rmax, int waves_per_group, int groups_per_cu);

/* Schedule of the profile alignments' Myers-Miller levels (no counterpart in
 * the reference, which runs progDiff on one thread; for callers and tests that
 * compare schedules, all of which give the same alignment): about how many
 * rows each pipelined stripe of a large root's half sweeps takes when those
 * run over several CUs (0: the root stays in its pair's workgroup), how many
 * levels sweep their halves one wave per workgroup before the rest run in the
 * pair's workgroup (1..6; 1: none), the most profile pairs a batch may hold
 * for either (0: none), and how many sleeps such a stripe waits for the one
 * above it before it gives up and the call fails with CWHIP_E_KERNEL (0: the
 * built-in 2^20, about a second).  -1 keeps a value. */
int cwhip_set_msa_tuning(cwhip_ctx *ctx, int wide_rows, int wide_levels, int wide_pairs, int wait_limit);

/* Shapes of the batched kernel by tuning class (the longer sequence of a pair:
 * <=128, <=256, <=512, <=1024, <=2048, longer): stripe height in rows per lane
 * (a stripe is 64*rmax rows) and the widest Myers-Miller node that runs on the
 * column path (lanes own 1, 2, 4 or 8 columns; 32..256).  The plain kernel
 * always uses the built-in defaults.  No counterpart in the reference (its
 * pairwise loop has no tunable shape); results do not depend on the table. */
#define CWHIP_TUNE_CLASSES 6
typedef struct {
    int32_t rmax[CWHIP_TUNE_CLASSES];
    int32_t colcap[CWHIP_TUNE_CLASSES];
} cwhip_tune_table;
int cwhip_get_tune_table(cwhip_ctx *ctx, cwhip_tune_table *t);
int cwhip_set_tune_table(cwhip_ctx *ctx, const cwhip_tune_table *t);
/* Autotuner: for every class present in `pairs` (the caller's loaded
 * sequences), times the batched kernel on a strided sample of that class's
 * pairs (four per resident wave) under the class's current shape and each
 * candidate, best of three runs, and moves to a candidate only when it is at
 * least 3% faster than the best so far.  The table goes to
 * *out (may be NULL), becomes this context's table and the default of contexts
 * created later on the same device.  *tune_ms (may be NULL) receives the device
 * time spent.  A pair list staged before the call is staged again afterwards. */
int cwhip_autotune(cwhip_ctx *ctx, const cwhip_pair_params *p, const int32_t *pairs,
                   int64_t npairs, cwhip_tune_table *out, float *tune_ms);
/* A batched launch that repeats the previous one (same staged job, tables and
 * shape) is captured once (queue reset, binning, persistent kernel) and replayed
 * as one hipGraph; a launch seen for the first time runs directly.  0 launches
 * the stages one by one always. */
int cwhip_set_graph(cwhip_ctx *ctx, int enable);


/* ---- the pairwise stage without device allocation ----------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The same batched kernels as cwhip_pairwise_pairs with
 * CWHIP_KERNEL_BATCHED (queue reset, length binning, the persistent kernel)
 * run on the caller's stream over caller memory; their work space (queue
 * counter, bins, the binned order and each wave's DP rows) is one file-scope
 * device array of CWHIP_NOALLOC_SCRATCH_BYTES that every GPU the library runs
 * on holds until the process ends.  The call allocates nothing, does not
 * synchronise and returns once the work is enqueued.
 *
 * Every pointer is device memory of the current device (hipSetDevice): the
 * sequences (codes[offsets[s] .. + lens[s]], s < nseq), npairs pairs of 0-based
 * indices, the scoring (matrix NUMRES x NUMRES, gap_open_by_len[0..max_len]),
 * out_count (npairs identities, -1 when either sequence is empty, as
 * cwhip_pairwise_pairs) and out_status (one int: 0 when every pair ran; bit 4
 * when a pair names a sequence outside 0..nseq-1 or longer than max_len; such a
 * pair gets count -2 and is not aligned).  The distance of a pair is
 * cwhip_distance(count, gap-free len1, gap-free len2).
 *
 * Limits: npairs <= CWHIP_NOALLOC_MAX_PAIRS, max_len <= CWHIP_MAX_LEN, and at
 * least one workgroup's DP rows must fit beside the pair order (4 waves x 8 rows
 * of max_len + 2 ints; the number of resident waves is cut to what fits).
 * Returns CWHIP_E_ARG (null pointer, negative size), CWHIP_E_TOOLONG
 * (max_len), CWHIP_E_NOMEM (npairs, or rows that do not fit), CWHIP_E_HIP.
 * All calls on one GPU share its array: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller. */
#define CWHIP_NOALLOC_SCRATCH_BYTES (128ll << 20)
#define CWHIP_NOALLOC_MAX_PAIRS (1ll << 22)
typedef struct {
    const int32_t *matrix;            /* device, NUMRES x NUMRES */
    const int32_t *gap_open_by_len;   /* device, max_len + 1 entries */
    int32_t gap_extend;
    int32_t gap_pos1;
    int32_t gap_pos2;
    int32_t max_len;
} cwhip_pair_dev_params;

int cwhip_pairwise_noalloc(const cwhip_pair_dev_params *p, const uint8_t *codes,
                           const int64_t *offsets, const int32_t *lens, int nseq,
                           const int32_t *pairs, int64_t npairs, int32_t *out_count,
                           int32_t *out_status, void *stream);

/* ------------------------------------------------------------------------ */
/* Progressive (profile) alignment                                          */
/* ------------------------------------------------------------------------ */

#define CWHIP_PRF_W 35   /* ints per profile row: codes 0..31, GAPCOL 32, LENCOL 33, pad */

/* One profile pair for cwhip_profile_dp.  prf1/prf2 hold len+2 rows of
 * CWHIP_PRF_W ints laid out as MyersMillerProfileAlign's profile1/profile2
 * (row 0 and rows 1..len; column 32 = GAPCOL, 33 = LENCOL).  displ receives
 * the edit script displ[1..printPtr-1] of the reference (0 = aligned column,
 * k > 0 = k columns of profile 2 against gaps, k < 0 = -k columns of profile 1
 * against gaps); it needs room for len1 + len2 + 2 entries. */
typedef struct {
    const int32_t *prf1;
    const int32_t *prf2;
    int32_t len1, len2;
    int32_t *displ;      /* out */
    int32_t n_displ;     /* out: printPtr - 1 */
    int32_t score;       /* out: the value progDiff returns (profileAlign returns score/100) */
} cwhip_prf_pair;

/* MyersMillerProfileAlign::progDiff + progTracepath's input for a batch of
 * independent profile pairs (multipleAlign/MyersMillerProfileAlign.cpp:521-811,
 * 1002-1262; the top call of :428-429 with go1/go2 = prf1 GAPCOL at rows 0 and
 * len1).  end_gap_penalties is UserParameters::endGapPenalties. */
int cwhip_profile_dp(cwhip_ctx *ctx, cwhip_prf_pair *pairs, int npairs, int end_gap_penalties);

/* Optional replacement of the built-in matrix series (gonnet for protein, iub
 * for DNA) of a profileAlign step: fills matrix (NUMRES x NUMRES, row-major),
 * *scale and *mat_avg the way SubMatrix::getProfileAlignMatrix(matrix, pcid,
 * minLen, scaleValues, matAvg) does (substitutionMatrix/SubMatrix.cpp:403-677)
 * and returns its maxRes.  A return of 0 or less fails the step, as the
 * `return -1` of MyersMillerProfileAlign.cpp:258-263 does.  The library calls
 * it from one thread at a time (host worker threads, serialised). */
typedef int (*cwhip_profile_matrix_fn)(void *user, double pcid, int32_t min_len, int32_t *matrix,
                                       float *scale, int32_t *mat_avg);

/* The UserParameters the progressive stage reads (general/UserParameters.cpp:
 * 33-175, setDNAParams/setProtParams :501-523). */
typedef struct {
    int32_t dna;
    float gap_open, gap_extend;       /* gapOpen / gapExtend of the multiple alignment */
    int32_t gap_dist;                 /* gapDist */
    int32_t no_var_penalties, no_hyd_penalties, no_pref_penalties;
    int32_t use_end_gaps, end_gap_penalties;
    float transition_weight;
    int32_t neg_matrix;               /* useNegMatrix */
    int32_t no_weights;
    int32_t divergence_cutoff;
    const char *hyd_residues;         /* hydResidues, e.g. "GPSNDQEKR" */
    int32_t distance_tree;            /* UserParameters::getDistanceTree: 0 after a guide tree
                                         without branch lengths (tree/Tree.cpp:564) or the
                                         profile-to-profile path (multipleAlign/MSA.cpp:572);
                                         then every step takes the series' fixed matrix
                                         (gon250, SubMatrix.cpp:565) */
    cwhip_profile_matrix_fn profile_matrix;   /* NULL: the built-in series */
    void *profile_matrix_user;                /* passed to profile_matrix */
} cwhip_msa_params;

/* The reference defaults for protein (dna = 0) or DNA (dna = 1); distance_tree 1,
 * no profile_matrix callback. */
int cwhip_msa_params_default(int dna, cwhip_msa_params *p);

/* MSA::multiSeqAlign with iStart = 0 (multipleAlign/MSA.cpp:32-347) driving
 * MyersMillerProfileAlign::profileAlign (MyersMillerProfileAlign.cpp:36-505)
 * for every step; the profile DPs run on the device, guide-tree steps on
 * disjoint subtrees in one batch.  Inputs: the nseq sequences as codes
 * (codes[offsets[s] .. + lens[s]]), the (nseq+1)^2 row-major similarity matrix
 * of Tree::calcSimilarities (1-based, row/column 0 unused), the sequence
 * weights of Tree::calcSeqWeights, and nsteps guide-tree steps of nseq+1 ints
 * each (element 0 unused; 1 / 2 = group, 0 = not in the step) as
 * Tree::createSets makes them.  out_order[k] (1-based) is the k-th sequence of
 * the output order (outputOrder ALIGNED); *aln_len is the alignment length.
 * The aligned rows stay in the context: fetch them with cwhip_msa_rows.
 * Returns CWHIP_E_MATRIX when a step's profile matrix fails (MSA.cpp:147-152
 * returns -1 then). */
int cwhip_msa(cwhip_ctx *ctx, const cwhip_msa_params *p, int nseq, const uint8_t *codes,
              const int64_t *offsets, const int32_t *lens, const double *sim,
              const int32_t *seq_weight, const int32_t *steps, int nsteps, int32_t *out_order,
              int32_t *aln_len);

/* MyersMillerProfileAlign::profileAlign (multipleAlign/MyersMillerProfileAlign.cpp:
 * 36-505) for one step: the two profiles of group (1-based, nseq+1 entries,
 * 1 / 2 / 0) are built from the sequences (codes[offsets[s] .. + lens[s]])
 * with the weights (nseq) and the similarity matrix ((nseq+1)^2, 1-based), the
 * profile DP runs on the device, and every member's new row (info[1] codes)
 * is written to rows_out + s*stride (stride >= info[2] + info[3] + 2).  As in
 * the reference, group is modified (members with aligned[s] == 0 dropped, the
 * groups swapped when group 2 is larger).  info: [0] profileAlign's return
 * value (score / 100; 0 when a group is empty and nothing was aligned; -1,
 * with nothing aligned, when the profile matrix fails, :258-263),
 * [1] the new alignment length, [2] prfLength1, [3] prfLength2. */
int cwhip_profile_align(cwhip_ctx *ctx, const cwhip_msa_params *p, int nseq, const uint8_t *codes,
                        const int64_t *offsets, const int32_t *lens, const double *sim,
                        const int32_t *weights, int32_t *group, const int32_t *aligned,
                        uint8_t *rows_out, int64_t stride, int32_t *info);

/* The two profiles profileAlign builds for one step (MyersMillerProfileAlign.cpp:
 * 80-300: ProfileBase::calcGapCoeff, ProfileBase.cpp:52-549; profile 1 by
 * ProfileWithSub::calcProfileWithSub, ProfileWithSub.cpp:42-147; profile 2 by
 * ProfileStandard::calcStandardProfile, ProfileStandard.cpp:41-117), built on the
 * device as cwhip_msa builds them.  Arguments as cwhip_profile_align (group is not
 * modified here); prf1/prf2 receive (len+2) x CWHIP_PRF_W ints (cap1/cap2 in ints);
 * len12 receives len1 and len2, or -1 and -1 when a group is empty. */
int cwhip_step_profiles(cwhip_ctx *ctx, const cwhip_msa_params *p, int nseq, const uint8_t *codes,
                        const int64_t *offsets, const int32_t *lens, const double *sim,
                        const int32_t *weights, const int32_t *group, const int32_t *aligned,
                        int32_t *prf1, int64_t cap1, int32_t *prf2, int64_t cap2, int32_t *len12);

/* Copies the rows of the last cwhip_msa (sequence s at out + s*stride,
 * aln_len codes each; gaps as code 30 or 31). */
int cwhip_msa_rows(cwhip_ctx *ctx, uint8_t *out, int64_t stride);

/* Statistics of the last cwhip_msa: [0] profileAlign calls, [1] device batches,
 * [2] Myers-Miller cell updates (forward + reverse halves of every node),
 * [3] host milliseconds, [4] device milliseconds (kernels), [5] prfScore
 * entries computed (sum of len1 * len2 over the profile pairs). */
int cwhip_msa_stats(cwhip_ctx *ctx, double *out6);


/* ------------------------------------------------------------------------ */
/* Guide tree and output text (host code; no device needed)                 */
/* ------------------------------------------------------------------------ */

/* NJTree::generateTree + ClusterTreeOutput::printPhylipTree
 * (tree/NJTree.cpp:161-751, tree/ClusterTreeOutput.cpp:45-220; two sequences:
 * tree/UnRootedClusterTree.cpp:225-233): the .dnd text of the neighbour-joining
 * guide tree of nseq >= 2 sequences.  dist is the (nseq+1)^2 row-major
 * DistMatrix (1-based, as FullPairwiseAlign fills it); names[0..nseq-1].  The
 * text (NUL-terminated) goes to out when cap > its length; *len is set
 * either way, so a first call with out = NULL sizes the buffer. */
int cwhip_nj_tree(int nseq, const double *dist, const char *const *names, char *out,
                  int64_t cap, int64_t *len);

/* The same tree, built on the context's device (tree/NJTree.cpp:161-751 as
 * above, every double computed by the reference's expression in its order, so
 * the text is byte-identical to cwhip_nj_tree's):
 *   cwhip_nj_tree_from_run   over the DistMatrix of the context's last
 *       finished pairwise run, which must have covered every pair of the
 *       loaded sequences (cwhip_pairwise_range over 0..nseq x 0..nseq): the
 *       distances are formed and joined in device memory and never cross to
 *       the host (only the tree does); nseq and names as loaded;
 *   cwhip_nj_tree_on_device  over a caller's (nseq+1)^2 DistMatrix, uploaded.
 * The kernel takes 4 .. CWHIP_NJ_DEVICE_MAX sequences; other sizes (and a
 * matrix that is not symmetric) run the host code of cwhip_nj_tree.  Sizing
 * as cwhip_nj_tree.  CWHIP_E_ARG when the context has no such run. */
#define CWHIP_NJ_DEVICE_MAX 3000
int cwhip_nj_tree_from_run(cwhip_ctx *ctx, const char *const *names, char *out, int64_t cap, int64_t *len);
int cwhip_nj_tree_on_device(cwhip_ctx *ctx, int nseq, const double *dist, const char *const *names,
                            char *out, int64_t cap, int64_t *len);

/* Tree::readTree, calcSeqWeights, calcSimilarities and createSets as
 * TreeInterface::getWeightsAndStepsUseOldGuideTreeNJ calls them
 * (tree/TreeInterface.cpp:209-276, tree/Tree.cpp): from .dnd text to the
 * sequence weights (nseq), the similarity matrix ((nseq+1)^2, row-major,
 * 1-based) and the alignment steps (room for (nseq-1) x (nseq+1) ints; element
 * 0 of each step unused).  *nsteps gets the number of steps; *distance_tree
 * (may be NULL) gets UserParameters::distanceTree as readTree leaves it: 1, or
 * 0 when a leaf has no branch length (tree/Tree.cpp:145, :561-565), for
 * cwhip_msa_params.distance_tree. */
int cwhip_guide_from_tree(int nseq, const char *dnd, const char *const *names, int32_t *weights,
                          double *sim, int32_t *steps, int32_t *nsteps, int32_t *distance_tree);

/* AlignmentOutput::clustalOut (alignment/AlignmentOutput.cpp:1573-2116): the
 * CLUSTAL text of an alignment.  rows: sequence s (0-based) at rows + s*stride,
 * lens[s] codes; order: 1-based output order; dna selects the conservation
 * line's rules.  Sizing as cwhip_nj_tree. */
int cwhip_clustal_text(int nseq, const char *const *names, const uint8_t *rows, int64_t stride,
                       const int32_t *lens, const int32_t *order, int dna, char *out, int64_t cap,
                       int64_t *len);


/* ------------------------------------------------------------------------ */
/* Serving: jobs of different sizes arriving over time                      */
/* ------------------------------------------------------------------------ */
/* No counterpart in the reference, which runs one alignment per process; the
 * jobs run the entry points above (the reference operations they replace are
 * listed there).  A server owns `slots` device contexts, each with its own
 * stream and worker thread.  Jobs start in arrival order (none overtakes an
 * older one) as soon as a slot is free and the job's device reservation fits
 * in hbm_budget beside what the other slots hold: a running slot its job's
 * reservation, an idle slot the buffers it keeps from earlier jobs (released
 * when a job needs the room).  A running job's context cannot allocate beyond
 * its reservation; a job that needs more is run again with twice the
 * reservation.  So the device bytes the server holds never exceed hbm_budget.
 * Admitted jobs run concurrently and their kernels overlap on the device.
 * Results do not depend on the mix. */

typedef struct cwhip_server cwhip_server;

enum {
    CWHIP_JOB_DISTANCES = 0,  /* the DistMatrix of all pairs (FullPairwiseAlign over 0..n) */
    CWHIP_JOB_ALIGNMENT = 1   /* distances + NJ tree (.dnd) + progressive alignment (.aln) */
};

int cwhip_server_create(int device, int slots, int64_t hbm_budget, cwhip_server **out);
/* stops the workers after their running jobs; pending jobs are dropped */
void cwhip_server_destroy(cwhip_server *srv);
/* copies the inputs; names are needed for CWHIP_JOB_ALIGNMENT.  Returns
 * CWHIP_E_NOMEM if the job alone exceeds the budget. */
int cwhip_server_submit(cwhip_server *srv, int kind, int dna, int nseq, const uint8_t *codes,
                        const int64_t *offsets, const int32_t *lens, const char *const *names,
                        int64_t *job);
/* waits up to timeout_s seconds (negative: no limit); CWHIP_OK when the job is done */
int cwhip_server_wait(cwhip_server *srv, int64_t job, double timeout_s);
/* nseq*(nseq-1)/2 float distances in the pair order of the range 0..nseq */
int cwhip_server_distances(cwhip_server *srv, int64_t job, float *dist);
/* which = 0: the .dnd text, 1: the .aln text (sizing as cwhip_nj_tree); 2: for a job that
 * failed, the error text of the context that ran it */
int cwhip_server_text(cwhip_server *srv, int64_t job, int which, char *out, int64_t cap,
                      int64_t *len);
/* [0] submitted, [1] started, [2] finished (seconds since the server was created),
 * [3] conventional pairwise cells + Myers-Miller cell updates of the job */
int cwhip_server_times(cwhip_server *srv, int64_t job, double *t4);
int cwhip_server_release(cwhip_server *srv, int64_t job);
/* [0] jobs finished, [1] cells, [2] peak device bytes held by the slots'
 * contexts (counted at every allocation and release), [3] most jobs running at
 * once, [4] budget (bytes), [5] slots */
int cwhip_server_stats(cwhip_server *srv, double *out6);

#ifdef __cplusplus
}
#endif
#endif
