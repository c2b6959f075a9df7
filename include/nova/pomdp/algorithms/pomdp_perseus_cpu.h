// nova-mi355x: Perseus on the host (replaces include/nova/pomdp/algorithms/pomdp_perseus_cpu.h:47-107).
// x86-64 sizeof 64: GammaInitial@0 currentHorizon@8 rGamma@12 rGammaPrime@16 rTilde@20
// BTilde@24 Gamma@32 GammaPrime@40 pi@48 piPrime@56.
#ifndef NOVA_POMDP_PERSEUS_CPU_H
#define NOVA_POMDP_PERSEUS_CPU_H

#include <nova/pomdp/pomdp.h>
#include <nova/pomdp/policies/pomdp_alpha_vectors.h>

namespace nova {

typedef struct NovaPOMDPPerseusCPU {
    float *GammaInitial;
    unsigned int currentHorizon;
    unsigned int rGamma;
    unsigned int rGammaPrime;
    unsigned int rTilde;
    unsigned int *BTilde;
    float *Gamma;
    float *GammaPrime;
    unsigned int *pi;
    unsigned int *piPrime;
} POMDPPerseusCPU;

extern "C" int pomdp_perseus_initialize_cpu(const POMDP *pomdp, POMDPPerseusCPU *per);
extern "C" int pomdp_perseus_execute_cpu(const POMDP *pomdp, POMDPPerseusCPU *per, POMDPAlphaVectors *policy);
extern "C" int pomdp_perseus_uninitialize_cpu(const POMDP *pomdp, POMDPPerseusCPU *per);
// One randomized backup; returns NOVA_CONVERGED when every belief has improved
// (pomdp_perseus_cpu.cpp:312-429).
extern "C" int pomdp_perseus_update_cpu(const POMDP *pomdp, POMDPPerseusCPU *per);
extern "C" int pomdp_perseus_get_policy_cpu(const POMDP *pomdp, POMDPPerseusCPU *per, POMDPAlphaVectors *policy);

} // namespace nova

#endif
