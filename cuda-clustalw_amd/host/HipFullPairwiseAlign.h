/*
 * HipFullPairwiseAlign.h -- drop-in replacement of clustalw::FullPairwiseAlign
 * (pairwise/FullPairwiseAlign.h:14-20 in CPU/clustalw-2.0.11/src of
 * coldfunction/CUDA-clustalW) that runs the full-DP pairwise stage on an
 * MI355X through the C ABI of libcwhip (include/cwhip.h).
 *
 * The class implements the reference's PairwiseAlignBase interface
 * (pairwise/PairwiseAlignBase.h:19-20) with the same range semantics, the
 * same scoring set-up (the reference's own SubMatrix and UserParameters), the
 * same fatal checks and messages, and writes the same float distances into the
 * DistMatrix.  It is meant to be compiled inside the reference tree, next to
 * FullPairwiseAlign.cpp; see INTEGRATION.md for the two-line hook in Clustal.cpp.
 */
#ifndef HIPFULLPAIRWISEALIGN_H
#define HIPFULLPAIRWISEALIGN_H

#include "PairwiseAlignBase.h"

struct cwhip_ctx;

namespace clustalw
{

class HipFullPairwiseAlign : public PairwiseAlignBase
{
    public:
        HipFullPairwiseAlign();
        virtual ~HipFullPairwiseAlign();
        virtual void pairwiseAlign(Alignment *alignPtr, DistMatrix *distMat, int iStart,
                                   int iEnd, int jStart, int jEnd);
    private:
        cwhip_ctx *ctx;
};

}
#endif
