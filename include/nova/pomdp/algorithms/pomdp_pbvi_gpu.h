// nova-mi355x: point-based value iteration on one MI355X
// (replaces include/nova/pomdp/algorithms/pomdp_pbvi_gpu.h:44-100).
// x86-64 sizeof 48: GammaInitial@0 numThreads@8 currentHorizon@12 d_Gamma@16
// d_GammaPrime@24 d_pi@32 d_alphaBA@40. The r*m*n alphaBA scratch of the reference is not
// needed by this implementation; d_alphaBA instead points at the library's device workspace:
// the model products O*T (m*z*n*ns floats), a state-major copy of Gamma (n*r floats) and the
// per-iteration argmax record (per belief, action and observation). It is allocated by
// pomdp_pbvi_initialize_gpu and released by pomdp_pbvi_uninitialize_gpu.
#ifndef NOVA_POMDP_PBVI_GPU_H
#define NOVA_POMDP_PBVI_GPU_H

#include <nova/pomdp/pomdp.h>
#include <nova/pomdp/policies/pomdp_alpha_vectors.h>

namespace nova {

typedef struct NovaPOMDPPBVIGPU {
    float *GammaInitial;
    unsigned int numThreads;
    unsigned int currentHorizon;
    float *d_Gamma;
    float *d_GammaPrime;
    unsigned int *d_pi;
    float *d_alphaBA;
} POMDPPBVIGPU;

extern "C" int pomdp_pbvi_initialize_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi);
extern "C" int pomdp_pbvi_execute_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, POMDPAlphaVectors *policy);
extern "C" int pomdp_pbvi_uninitialize_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi);
extern "C" int pomdp_pbvi_update_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi);
extern "C" int pomdp_pbvi_get_policy_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, POMDPAlphaVectors *policy);

} // namespace nova

#endif
