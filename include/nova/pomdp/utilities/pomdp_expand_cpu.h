// nova-mi355x: belief-set expansion on the host
// (replaces include/nova/pomdp/utilities/pomdp_expand_cpu.h:43-61).
#ifndef NOVA_POMDP_EXPAND_CPU_H
#define NOVA_POMDP_EXPAND_CPU_H

#include <nova/pomdp/pomdp.h>
#include <nova/pomdp/policies/pomdp_alpha_vectors.h>

namespace nova {

// Random trajectories from random existing beliefs (pomdp_expand_cpu.cpp:82-163).
// The host RNG is libc rand(), seeded with srand(time) as in the reference.
extern "C" int pomdp_expand_random_cpu(POMDP *pomdp, unsigned int numBeliefsToAdd);
// For each belief, the successor tau(b,a,o) farthest (L1) from the set (pomdp_expand_cpu.cpp:166-266).
extern "C" int pomdp_expand_distinct_beliefs_cpu(POMDP *pomdp);
// PEMA: appends the successor with the largest expected approximation error
// (pomdp_expand_cpu.cpp:269-466). By default it reproduces nova's alpha choice, whose argmax is
// defective (:391-402); nova_amd_pema_mode(1) (nova/amd/nova_amd.h) corrects it.
extern "C" int pomdp_expand_pema_cpu(POMDP *pomdp, const POMDPAlphaVectors *policy);

} // namespace nova

#endif
