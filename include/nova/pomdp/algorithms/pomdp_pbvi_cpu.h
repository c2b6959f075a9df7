// nova-mi355x: point-based value iteration on the host
// (replaces include/nova/pomdp/algorithms/pomdp_pbvi_cpu.h:42-93).
// x86-64 sizeof 40: GammaInitial@0 currentHorizon@8 Gamma@16 GammaPrime@24 pi@32.
#ifndef NOVA_POMDP_PBVI_CPU_H
#define NOVA_POMDP_PBVI_CPU_H

#include <nova/pomdp/pomdp.h>
#include <nova/pomdp/policies/pomdp_alpha_vectors.h>

namespace nova {

typedef struct NovaPOMDPPBVICPU {
    float *GammaInitial;
    unsigned int currentHorizon;
    float *Gamma;
    float *GammaPrime;
    unsigned int *pi;
} POMDPPBVICPU;

extern "C" int pomdp_pbvi_initialize_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi);
extern "C" int pomdp_pbvi_execute_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi, POMDPAlphaVectors *policy);
extern "C" int pomdp_pbvi_uninitialize_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi);
extern "C" int pomdp_pbvi_update_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi);
extern "C" int pomdp_pbvi_get_policy_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi, POMDPAlphaVectors *policy);

} // namespace nova

#endif
