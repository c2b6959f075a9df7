// nova-mi355x: host-side MDP allocation (replaces include/nova/mdp/utilities/mdp_model_cpu.h:46-54).
#ifndef NOVA_MDP_MODEL_CPU_H
#define NOVA_MDP_MODEL_CPU_H

#include <nova/mdp/mdp.h>

namespace nova {

// Allocates S (= -1), T (= 0), R (= 0) and goals; rejects n < ns, s0 >= n, ng >= n,
// gamma outside [0,1], horizon 0, epsilon < 0, or non-null arrays (mdp_model_cpu.cpp:35-82).
extern "C" int mdp_initialize_cpu(MDP *mdp, unsigned int n, unsigned int ns, unsigned int m,
        float gamma, unsigned int horizon, float epsilon, unsigned int s0, unsigned int ng);

extern "C" int mdp_uninitialize_cpu(MDP *mdp);

} // namespace nova

#endif
