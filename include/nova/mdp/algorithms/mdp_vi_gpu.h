// nova-mi355x: MDP value iteration on one MI355X
// (replaces include/nova/mdp/algorithms/mdp_vi_gpu.h:43-95).
// x86-64 sizeof 40: VInitial@0 numThreads@8 currentHorizon@12 d_V@16 d_VPrime@24 d_pi@32.
// numThreads is kept for compatibility (it must be a multiple of 32, as in
// mdp_vi_gpu.cu:175-178) but is only a hint: the wave64 tiling is chosen by the library.
#ifndef NOVA_MDP_VI_GPU_H
#define NOVA_MDP_VI_GPU_H

#include <nova/mdp/mdp.h>
#include <nova/mdp/policies/mdp_value_function.h>

namespace nova {

typedef struct NovaMDPVIGPU {
    float *VInitial;
    unsigned int numThreads;
    unsigned int currentHorizon;
    float *d_V;
    float *d_VPrime;
    unsigned int *d_pi;
} MDPVIGPU;

// hipMalloc d_V, d_VPrime, d_pi; uploads VInitial (zeros when null), pi = 0
// (mdp_vi_gpu.cu:90-160). Buffers left by an earlier initialize are released first.
extern "C" int mdp_vi_initialize_gpu(const MDP *mdp, MDPVIGPU *vi);
// Validates, initializes, runs `horizon` sweeps as one captured HIP graph, downloads the
// policy, uninitializes (mdp_vi_gpu.cu:163-221).
extern "C" int mdp_vi_execute_gpu(const MDP *mdp, MDPVIGPU *vi, MDPValueFunction *policy);
extern "C" int mdp_vi_uninitialize_gpu(const MDP *mdp, MDPVIGPU *vi);
// One sweep on the device, synchronous at return (mdp_vi_gpu.cu:269-307).
extern "C" int mdp_vi_update_gpu(const MDP *mdp, MDPVIGPU *vi);
// Downloads V (by parity) and pi into a new MDPValueFunction (mdp_vi_gpu.cu:310-346).
extern "C" int mdp_vi_get_policy_gpu(const MDP *mdp, MDPVIGPU *vi, MDPValueFunction *policy);

} // namespace nova

#endif
