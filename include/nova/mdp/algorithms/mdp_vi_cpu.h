// nova-mi355x: MDP value iteration on the host
// (replaces include/nova/mdp/algorithms/mdp_vi_cpu.h:42-93).
// x86-64 sizeof 40: VInitial@0 currentHorizon@8 V@16 VPrime@24 pi@32.
#ifndef NOVA_MDP_VI_CPU_H
#define NOVA_MDP_VI_CPU_H

#include <nova/mdp/mdp.h>
#include <nova/mdp/policies/mdp_value_function.h>

namespace nova {

typedef struct NovaMDPVICPU {
    float *VInitial;
    unsigned int currentHorizon;
    float *V;
    float *VPrime;
    unsigned int *pi;
} MDPVICPU;

// Allocates V, VPrime, pi; copies VInitial (or zeros) into V and VPrime (mdp_vi_cpu.cpp:66-98).
extern "C" int mdp_vi_initialize_cpu(const MDP *mdp, MDPVICPU *vi);
// initialize + `horizon` sweeps + get_policy + uninitialize (mdp_vi_cpu.cpp:101-156).
extern "C" int mdp_vi_execute_cpu(const MDP *mdp, MDPVICPU *vi, MDPValueFunction *policy);
extern "C" int mdp_vi_uninitialize_cpu(const MDP *mdp, MDPVICPU *vi);
// One Bellman sweep: even currentHorizon reads V and writes VPrime, odd the reverse
// (mdp_vi_cpu.cpp:184-195). Does not validate its arguments.
extern "C" int mdp_vi_update_cpu(const MDP *mdp, MDPVICPU *vi);
// V by parity of currentHorizon, pi from the last sweep (mdp_vi_cpu.cpp:198-223).
extern "C" int mdp_vi_get_policy_cpu(const MDP *mdp, MDPVICPU *vi, MDPValueFunction *policy);

} // namespace nova

#endif
