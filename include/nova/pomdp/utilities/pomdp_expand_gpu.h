// nova-mi355x: random belief expansion driven from the device
// (replaces include/nova/pomdp/utilities/pomdp_expand_gpu.h:47-48).
// Bnew must be null on entry; it is allocated with new[numDesiredBeliefPoints*n]. Row 0 is
// belief 0 of the set; each further row is a step of a random trajectory from it.
#ifndef NOVA_POMDP_EXPAND_GPU_H
#define NOVA_POMDP_EXPAND_GPU_H

#include <nova/pomdp/pomdp.h>

namespace nova {

extern "C" int pomdp_expand_random_gpu(const POMDP *pomdp, unsigned int numThreads,
        unsigned int numDesiredBeliefPoints, unsigned int &maxNonZeroValues, float *&Bnew);

} // namespace nova

#endif
