/*
 * HipMyersMillerProfileAlign.h -- drop-in replacement of
 * clustalw::MyersMillerProfileAlign (multipleAlign/MyersMillerProfileAlign.h:16-21
 * in CPU/clustalw-2.0.11/src of coldfunction/CUDA-clustalW) that runs the
 * profile-profile Myers-Miller DP on an MI355X through the C ABI of libcwhip
 * (include/cwhip.h: cwhip_profile_align).
 *
 * It implements the reference's ProfileAlignAlgorithm interface
 * (multipleAlign/ProfileAlignAlgorithm.h:18-19): same group semantics (members
 * not yet aligned dropped, groups swapped when group 2 is larger), same return
 * value (score / 100, 0 for an empty group), and it rewrites the Alignment's
 * rows and maximum lengths as the reference does (:446-493).  The scoring
 * parameters come from the reference's UserParameters.  It is meant to be
 * compiled inside the reference tree; MSA::multiSeqAlign (MSA.cpp:59) then
 * creates it in place of MyersMillerProfileAlign.
 */
#ifndef HIPMYERSMILLERPROFILEALIGN_H
#define HIPMYERSMILLERPROFILEALIGN_H

#include <vector>

#include "ProfileAlignAlgorithm.h"

struct cwhip_ctx;

namespace clustalw
{

class HipMyersMillerProfileAlign : public ProfileAlignAlgorithm
{
    public:
        HipMyersMillerProfileAlign();
        virtual ~HipMyersMillerProfileAlign();
        virtual int profileAlign(Alignment* alnPtr, DistMatrix* distMat, vector<int>* group,
                                 int* aligned);
    private:
        cwhip_ctx *ctx;
        // dense copy of the similarity matrix, refreshed when another matrix comes in
        const DistMatrix *simOwner;
        int simSize;
        std::vector<double> sim;
};

}
#endif
