// nova-mi355x: host-side POMDP allocation and belief utilities
// (replaces include/nova/pomdp/utilities/pomdp_model_cpu.h:47-77).
#ifndef NOVA_POMDP_MODEL_CPU_H
#define NOVA_POMDP_MODEL_CPU_H

#include <nova/pomdp/pomdp.h>

namespace nova {

// Allocates S (= -1), T, O, R (= 0), Z (= 0, as the reference does), B (= 0);
// rejects n < ns, zero sizes, gamma outside [0,1], horizon 0 (pomdp_model_cpu.cpp:34-86).
extern "C" int pomdp_initialize_cpu(POMDP *pomdp, unsigned int n, unsigned int ns, unsigned int m,
        unsigned int z, unsigned int r, unsigned int rz, float gamma, unsigned int horizon);

// bp = tau(b, a, o): bp(s') ∝ O[a,s',o] * sum_s T(s,a,s') b(s). bp must be null on entry and
// is allocated with new[n]. Returns NOVA_WARNING_INVALID_BELIEF when the normaliser is
// below 1e-9 in magnitude (pomdp_model_cpu.cpp:89-137).
extern "C" int pomdp_belief_update_cpu(const POMDP *pomdp, const float *b,
        unsigned int a, unsigned int o, float *&bp);

// Appends numBeliefsToAdd dense beliefs (rows of Bnew, n each) to the sparse Z/B set, growing
// rz to the largest support (pomdp_model_cpu.cpp:157-227).
extern "C" int pomdp_add_new_raw_beliefs_cpu(POMDP *pomdp, unsigned int numBeliefsToAdd,
        float *Bnew);

extern "C" int pomdp_uninitialize_cpu(POMDP *pomdp);

} // namespace nova

#endif
