// nova-mi355x: SSP LAO* (replaces include/nova/mdp/algorithms/ssp_lao_star_cpu.h:42-97), on the
// host as in the reference. Same struct layout and entry points; R holds positive costs and
// goals are absorbing.
#ifndef NOVA_SSP_LAO_STAR_CPU_H
#define NOVA_SSP_LAO_STAR_CPU_H

#include <nova/mdp/mdp.h>
#include <nova/mdp/policies/mdp_value_function.h>

namespace nova {

typedef struct NovaSSPLAOStarCPU {
    float *VInitial;
    unsigned int currentHorizon;
    float *V;
    unsigned int *pi;
} SSPLAOStarCPU;

extern "C" int ssp_lao_star_initialize_cpu(const MDP *mdp, SSPLAOStarCPU *lao);
extern "C" int ssp_lao_star_execute_cpu(const MDP *mdp, SSPLAOStarCPU *lao, MDPValueFunction *policy);
extern "C" int ssp_lao_star_uninitialize_cpu(const MDP *mdp, SSPLAOStarCPU *lao);
extern "C" int ssp_lao_star_update_cpu(const MDP *mdp, SSPLAOStarCPU *lao);
extern "C" int ssp_lao_star_get_policy_cpu(const MDP *mdp, SSPLAOStarCPU *lao, MDPValueFunction *policy);

} // namespace nova

#endif
