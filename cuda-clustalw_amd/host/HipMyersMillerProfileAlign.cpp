/*
 * HipMyersMillerProfileAlign.cpp -- see HipMyersMillerProfileAlign.h.
 *
 * Host side of MyersMillerProfileAlign::profileAlign
 * (multipleAlign/MyersMillerProfileAlign.cpp:36-505 of the reference): packs the
 * Alignment's 1-based rows into uint8 codes, passes the sequence weights
 * (Alignment::getSeqWeight), the similarity matrix the caller reads pcid from
 * (:124-137), the group and aligned vectors and the multiple-alignment
 * UserParameters to cwhip_profile_align, and writes the new rows back with
 * the maximum lengths updated as :206 and :446-493 do.  The profile matrix of
 * each step comes from the reference's own SubMatrix, so every matrix series
 * and the distance-tree rule behave as in the CPU path.
 */
#include "HipMyersMillerProfileAlign.h"

#include <cstdlib>
#include <iostream>
#include <vector>

#include "cwhip.h"

namespace clustalw
{

// SubMatrix::getProfileAlignMatrix of the reference (SubMatrix.cpp:403-677),
// called by the library for every step: whatever matrix series the user chose
// (gonnet, blosum, pam, id, a user matrix or series; iub, clustalw or a user
// DNA matrix) and the distance-tree rule come out as the CPU path has them.
static int refProfileAlignMatrix(void*, double pcid, int32_t minLen, int32_t* out, float* scale,
                                 int32_t* matAvg)
{
    int matrix[NUMRES][NUMRES];
    PrfScaleValues scaleVals;
    scaleVals.scale = 1.0;                 // MyersMillerProfileAlign.cpp:249-251
    scaleVals.intScale = 100.0;
    int avg = 0;
    const int maxRes = subMatrix->getProfileAlignMatrix(matrix, pcid, minLen, scaleVals, avg);
    for (int i = 0; i < NUMRES; ++i)
        for (int j = 0; j < NUMRES; ++j) out[i * NUMRES + j] = matrix[i][j];
    *scale = scaleVals.scale;
    *matAvg = avg;
    return maxRes;
}

HipMyersMillerProfileAlign::HipMyersMillerProfileAlign() : ctx(0), simOwner(0), simSize(0)
{
    if (cwhip_create(0, &ctx) != CWHIP_OK)
    {
        std::cerr << "Could not open the GPU for the profile alignment.\n"
                  << "Need to terminate program.\n";
        exit(1);
    }
}

HipMyersMillerProfileAlign::~HipMyersMillerProfileAlign()
{
    cwhip_destroy(ctx);
}

int HipMyersMillerProfileAlign::profileAlign(Alignment* alnPtr, DistMatrix* distMat,
                                             vector<int>* group, int* aligned)
{
    const int numSeq = alnPtr->getNumSeqs();
    if (numSeq <= 0) return 0;
    cwhip_msa_params p;
    cwhip_msa_params_default(userParameters->getDNAFlag() ? 1 : 0, &p);
    p.gap_open = userParameters->getGapOpen();
    p.gap_extend = userParameters->getGapExtend();
    p.gap_dist = userParameters->getGapDist();
    p.no_var_penalties = userParameters->getNoVarPenalties() ? 1 : 0;
    p.no_hyd_penalties = userParameters->getNoHydPenalties() ? 1 : 0;
    p.no_pref_penalties = userParameters->getNoPrefPenalties() ? 1 : 0;
    p.use_end_gaps = userParameters->getUseEndGaps() ? 1 : 0;
    p.end_gap_penalties = userParameters->getEndGapPenalties() ? 1 : 0;
    p.transition_weight = userParameters->getTransitionWeight();
    p.neg_matrix = userParameters->getUseNegMatrix() ? 1 : 0;
    p.no_weights = userParameters->getNoWeights() ? 1 : 0;
    p.divergence_cutoff = userParameters->getDivergenceCutoff();
    std::string hyd = userParameters->getHydResidues();
    p.hyd_residues = hyd.c_str();
    p.distance_tree = userParameters->getDistanceTree() ? 1 : 0;
    p.profile_matrix = &refProfileAlignMatrix;
    p.profile_matrix_user = 0;

    // the rows and weights of every sequence (codes are < 256 by construction)
    std::vector<int64_t> offs(numSeq);
    std::vector<int32_t> lens(numSeq), weights(numSeq);
    std::vector<uint8_t> codes;
    int64_t total = 0;
    for (int s = 0; s < numSeq; ++s)
    {
        const std::vector<int>* seq = alnPtr->getSequence(s + 1);
        const int n = alnPtr->getSeqLength(s + 1);
        offs[s] = (int64_t)codes.size();
        lens[s] = n;
        for (int k = 1; k <= n; ++k) codes.push_back((uint8_t)(*seq)[k]);
        weights[s] = alnPtr->getSeqWeight(s);
        total += n;
    }
    if (simOwner != distMat || simSize != numSeq + 1)
    {
        simOwner = distMat;
        simSize = numSeq + 1;
        sim.assign((size_t)simSize * simSize, 0.0);
        for (int i = 1; i <= numSeq; ++i)
            for (int j = 1; j <= numSeq; ++j) sim[(size_t)i * simSize + j] = (*distMat)(i, j);
    }
    std::vector<int32_t> g(numSeq + 1), al(numSeq + 1);
    for (int i = 0; i <= numSeq; ++i)
    {
        g[i] = (*group)[i];
        al[i] = aligned[i];
    }
    const int64_t stride = total + 2;
    std::vector<uint8_t> rows((size_t)numSeq * stride);
    int32_t info[4];
    if (cwhip_profile_align(ctx, &p, numSeq, codes.data(), offs.data(), lens.data(), sim.data(),
                            weights.data(), g.data(), al.data(), rows.data(), stride, info) != CWHIP_OK)
    {
        std::cerr << "An exception has occured in the MyersMillerProfileAlign class.\n"
                  << cwhip_last_error(ctx) << "\n";
        exit(1);
    }
    for (int i = 0; i <= numSeq; ++i) (*group)[i] = g[i];
    const int alen = info[1];
    if (alen == 0) return info[0];
    alnPtr->setMaxAlnLength(info[2] + info[3] + 2);
    SeqArray* newSequences = alnPtr->getSeqArrayForRealloc();
    for (int j = 0; j < numSeq; ++j)
    {
        if ((*group)[j + 1] != 1 && (*group)[j + 1] != 2) continue;
        (*newSequences)[j + 1].clear();
        (*newSequences)[j + 1].resize(alen + 1);
        for (int i = 0; i < alen; ++i)
            (*newSequences)[j + 1][i + 1] = rows[(size_t)j * stride + i];
    }
    alnPtr->calculateMaxLengths();
    return info[0];
}

}
