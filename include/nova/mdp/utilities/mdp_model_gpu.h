// nova-mi355x: HBM residency of the MDP arrays
// (replaces include/nova/mdp/utilities/mdp_model_gpu.h:38-101).
// Each helper hipMallocs the d_* field and copies the host array into it; the aggregate
// functions return the SUM of the helpers' codes (mdp_model_gpu.cu:41-50): nonzero = failure.
#ifndef NOVA_MDP_MODEL_GPU_H
#define NOVA_MDP_MODEL_GPU_H

#include <nova/mdp/mdp.h>

namespace nova {

extern "C" int mdp_initialize_gpu(MDP *mdp);
extern "C" int mdp_uninitialize_gpu(MDP *mdp);
extern "C" int mdp_initialize_successors_gpu(MDP *mdp);
extern "C" int mdp_uninitialize_successors_gpu(MDP *mdp);
extern "C" int mdp_initialize_state_transitions_gpu(MDP *mdp);
extern "C" int mdp_uninitialize_state_transitions_gpu(MDP *mdp);
extern "C" int mdp_initialize_rewards_gpu(MDP *mdp);
extern "C" int mdp_uninitialize_rewards_gpu(MDP *mdp);
extern "C" int mdp_initialize_goals_gpu(MDP *mdp);
extern "C" int mdp_uninitialize_goals_gpu(MDP *mdp);

} // namespace nova

#endif
