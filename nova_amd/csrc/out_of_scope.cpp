// nova-mi355x: exported entry points outside the Bellman-backup path this library provides.
// The reference's ctypes layer binds every symbol at import (e.g. python/nova/nova_ssp_rtdp.py:59-74),
// so they must exist; each reports that it is not provided and returns NOVA_ERROR_INVALID_DATA.
#include <nova/mdp/algorithms/ssp_lao_star_cpu.h>
#include <nova/mdp/algorithms/ssp_rtdp_cpu.h>
#include <nova/pomdp/utilities/pomdp_sigma_cpu.h>

#include "nova_common.h"

#define NOVA_NOT_PROVIDED(fn)                                                    \
    do {                                                                         \
        NOVA_ERR(fn, "Not provided by this library (outside the Bellman-backup path)."); \
        return NOVA_ERROR_INVALID_DATA;                                          \
    } while (0)

namespace nova {

int ssp_lao_star_initialize_cpu(const MDP *, SSPLAOStarCPU *) { NOVA_NOT_PROVIDED("ssp_lao_star_initialize_cpu"); }
int ssp_lao_star_execute_cpu(const MDP *, SSPLAOStarCPU *, MDPValueFunction *) { NOVA_NOT_PROVIDED("ssp_lao_star_execute_cpu"); }
int ssp_lao_star_uninitialize_cpu(const MDP *, SSPLAOStarCPU *) { NOVA_NOT_PROVIDED("ssp_lao_star_uninitialize_cpu"); }
int ssp_lao_star_update_cpu(const MDP *, SSPLAOStarCPU *) { NOVA_NOT_PROVIDED("ssp_lao_star_update_cpu"); }
int ssp_lao_star_get_policy_cpu(const MDP *, SSPLAOStarCPU *, MDPValueFunction *) { NOVA_NOT_PROVIDED("ssp_lao_star_get_policy_cpu"); }

int ssp_rtdp_initialize_cpu(const MDP *, SSPRTDPCPU *) { NOVA_NOT_PROVIDED("ssp_rtdp_initialize_cpu"); }
int ssp_rtdp_execute_cpu(const MDP *, SSPRTDPCPU *, MDPValueFunction *) { NOVA_NOT_PROVIDED("ssp_rtdp_execute_cpu"); }
int ssp_rtdp_uninitialize_cpu(const MDP *, SSPRTDPCPU *) { NOVA_NOT_PROVIDED("ssp_rtdp_uninitialize_cpu"); }
int ssp_rtdp_update_cpu(const MDP *, SSPRTDPCPU *) { NOVA_NOT_PROVIDED("ssp_rtdp_update_cpu"); }
int ssp_rtdp_get_policy_cpu(const MDP *, SSPRTDPCPU *, MDPValueFunction *) { NOVA_NOT_PROVIDED("ssp_rtdp_get_policy_cpu"); }

int pomdp_sigma_cpu(POMDP *, unsigned int, float &) { NOVA_NOT_PROVIDED("pomdp_sigma_cpu"); }

} // namespace nova
