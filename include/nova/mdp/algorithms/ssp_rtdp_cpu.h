// nova-mi355x: SSP RTDP (replaces include/nova/mdp/algorithms/ssp_rtdp_cpu.h:44-106), on the
// host as in the reference. Same struct layout and entry points; R holds positive costs and
// goals are absorbing.
#ifndef NOVA_SSP_RTDP_CPU_H
#define NOVA_SSP_RTDP_CPU_H

#include <nova/mdp/mdp.h>
#include <nova/mdp/policies/mdp_value_function.h>

namespace nova {

typedef struct NovaSSPRTDPCPU {
    float *VInitial;
    unsigned int trials;
    unsigned int currentTrial;
    unsigned int currentHorizon;
    float *V;
    unsigned int *pi;
} SSPRTDPCPU;

extern "C" int ssp_rtdp_initialize_cpu(const MDP *mdp, SSPRTDPCPU *rtdp);
extern "C" int ssp_rtdp_execute_cpu(const MDP *mdp, SSPRTDPCPU *rtdp, MDPValueFunction *policy);
extern "C" int ssp_rtdp_uninitialize_cpu(const MDP *mdp, SSPRTDPCPU *rtdp);
extern "C" int ssp_rtdp_update_cpu(const MDP *mdp, SSPRTDPCPU *rtdp);
extern "C" int ssp_rtdp_get_policy_cpu(const MDP *mdp, SSPRTDPCPU *rtdp, MDPValueFunction *policy);

} // namespace nova

#endif
