/*
 * HipFullPairwiseAlign.cpp -- see HipFullPairwiseAlign.h.
 *
 * Mirrors FullPairwiseAlign::pairwiseAlign (pairwise/FullPairwiseAlign.cpp:34-211
 * of the reference) step by step on the host side: size and range checks
 * (:47-58), the empty-alignment return (:63-66), the substitution matrix from
 * SubMatrix::getPairwiseMatrix (:81-86), the gap penalties of :129-148 (here
 * tabulated by min(n,m) once, evaluated in double exactly as the reference
 * does per pair), the pair loops (:97, :110) and the two SetAt calls per pair
 * (:182-183).  The DP itself runs in the gfx950 kernels behind cwhip.h.
 */
#include "HipFullPairwiseAlign.h"

#include <math.h>

#include <cstdlib>
#include <iostream>
#include <vector>

#include "cwhip.h"

namespace clustalw
{

HipFullPairwiseAlign::HipFullPairwiseAlign() : ctx(0)
{
    if (cwhip_create(0, &ctx) != CWHIP_OK)
    {
        std::cerr << "Could not open the GPU for the pairwise alignment.\n"
                  << "Need to terminate program.\n";
        exit(1);
    }
}

HipFullPairwiseAlign::~HipFullPairwiseAlign()
{
    cwhip_destroy(ctx);
}

static void die(cwhip_ctx *ctx, const char *what)
{
    std::cerr << "An exception has occured in the FullPairwiseAlign class.\n"
              << what << ": " << cwhip_last_error(ctx) << "\n";
    exit(1);
}

void HipFullPairwiseAlign::pairwiseAlign(Alignment *alignPtr, DistMatrix *distMat, int iStart,
                                         int iEnd, int jStart, int jEnd)
{
    if (distMat->getSize() != alignPtr->getNumSeqs() + 1)
    {
        std::cout << "The distance matrix is not the right size!\n"
                  << "Need to terminate program.\n";
        exit(1);
    }
    if ((iStart < 0) || (iEnd < iStart) || (jStart < 0) || (jEnd < jStart))
    {
        std::cout << "The range for pairwise Alignment is incorrect.\n"
                  << "Need to terminate program.\n";
        exit(1);
    }
    const int numSeqs = alignPtr->getNumSeqs();
    if (numSeqs == 0)
    {
        return;
    }

    // scoring, as FullPairwiseAlign.cpp:79-94 and :129-148
    cwhip_pair_params p;
    int matrix[NUMRES][NUMRES];
    int matAvgScore = 0;
    PairScaleValues scaleValues;
    int maxRes = subMatrix->getPairwiseMatrix(matrix, scaleValues, matAvgScore);
    if (maxRes == 0)
    {
        std::cout << "Could not get the substitution matrix\n";
        return;
    }
    for (int i = 0; i < NUMRES; ++i)
        for (int j = 0; j < NUMRES; ++j)
            p.matrix[i * CWHIP_NUMRES + j] = matrix[i][j];
    const bool dna = userParameters->getDNAFlag();
    const float pwGapOpen = userParameters->getPWGapOpen();
    const float pwGapExtend = userParameters->getPWGapExtend();
    const int intScale = scaleValues.intScale;
    const float gapOpenScale = scaleValues.gapOpenScale;
    const float gapExtendScale = scaleValues.gapExtendScale;

    int maxLen = 0;
    std::vector<int64_t> offs(numSeqs);
    std::vector<int32_t> lens(numSeqs);
    std::vector<uint8_t> codes;
    for (int s = 0; s < numSeqs; ++s)
    {
        const std::vector<int> *seq = alignPtr->getSequence(s + 1);
        const int n = alignPtr->getSeqLength(s + 1);
        offs[s] = (int64_t)codes.size();
        lens[s] = n;
        for (int k = 1; k <= n; ++k)
            codes.push_back((uint8_t)(*seq)[k]);
        if (n > maxLen) maxLen = n;
    }
    std::vector<int32_t> gapOpenByLen(maxLen + 1, 0);
    for (int L = 1; L <= maxLen; ++L)
    {
        if (dna)
            gapOpenByLen[L] = static_cast<int>(2 * pwGapOpen * intScale * gapOpenScale);
        else if (matAvgScore <= 0)
            gapOpenByLen[L] = 2 * static_cast<int>((pwGapOpen + log(static_cast<double>(L))) * intScale);
        else
            gapOpenByLen[L] = static_cast<int>(2 * matAvgScore * (pwGapOpen +
                              log(static_cast<double>(L))) * gapOpenScale);
    }
    p.gap_extend = dna ? static_cast<int>(pwGapExtend * intScale * gapExtendScale)
                       : static_cast<int>(pwGapExtend * intScale);
    p.gap_pos1 = userParameters->getGapPos1();
    p.gap_pos2 = userParameters->getGapPos2();
    p.max_len = maxLen;
    p.gap_open_by_len = gapOpenByLen.data();

    if (cwhip_load_sequences(ctx, codes.data(), offs.data(), lens.data(), numSeqs) != CWHIP_OK)
        die(ctx, "cwhip_load_sequences");
    const int64_t np = cwhip_range_npairs(numSeqs, iStart, iEnd, jStart, jEnd);
    std::vector<int32_t> pairs(2 * (np > 0 ? np : 1));
    std::vector<int32_t> count(np > 0 ? np : 1);
    std::vector<float> dist(np > 0 ? np : 1);
    if (cwhip_pairwise_range(ctx, &p, iStart, iEnd, jStart, jEnd, CWHIP_KERNEL_BATCHED,
                             pairs.data(), count.data(), dist.data()) != CWHIP_OK)
        die(ctx, "cwhip_pairwise_range");

    for (int64_t k = 0; k < np; ++k)
    {
        const int si = pairs[2 * k], sj = pairs[2 * k + 1];
        const double score = dist[k];
        distMat->SetAt(si + 1, sj + 1, score);
        distMat->SetAt(sj + 1, si + 1, score);
        if (userParameters->getDisplayInfo() && count[k] >= 0)
        {
            // the percent identity the reference prints (:185-189)
            int len1 = 0, len2 = 0;
            for (int i = 0; i < lens[si]; ++i)
                if (codes[offs[si] + i] != p.gap_pos1 && codes[offs[si] + i] != p.gap_pos2) ++len1;
            for (int i = 0; i < lens[sj]; ++i)
                if (codes[offs[sj] + i] != p.gap_pos1 && codes[offs[sj] + i] != p.gap_pos2) ++len2;
            float mmScore = (float)(100.0 * (float)count[k]);
            if (len1 == 0 || len2 == 0) mmScore = 0;
            else mmScore /= (float)(len1 < len2 ? len1 : len2);
            utilityObject->info("Sequences (%d:%d) Aligned. Score:  %d", si + 1, sj + 1,
                                (int)mmScore);
        }
    }
    if (userParameters->getDisplayInfo())
    {
        distMat->printArray();
    }
}

}
