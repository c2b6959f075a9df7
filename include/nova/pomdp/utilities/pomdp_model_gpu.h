// nova-mi355x: HBM residency of the POMDP arrays
// (replaces include/nova/pomdp/utilities/pomdp_model_gpu.h:38-129).
// Aggregates return the sum of the helpers' codes (pomdp_model_gpu.cu:34-76).
#ifndef NOVA_POMDP_MODEL_GPU_H
#define NOVA_POMDP_MODEL_GPU_H

#include <nova/pomdp/pomdp.h>

namespace nova {

extern "C" int pomdp_initialize_gpu(POMDP *pomdp);
extern "C" int pomdp_uninitialize_gpu(POMDP *pomdp);
extern "C" int pomdp_initialize_successors_gpu(POMDP *pomdp);
extern "C" int pomdp_uninitialize_successors_gpu(POMDP *pomdp);
extern "C" int pomdp_initialize_state_transitions_gpu(POMDP *pomdp);
extern "C" int pomdp_uninitialize_state_transitions_gpu(POMDP *pomdp);
extern "C" int pomdp_initialize_observation_transitions_gpu(POMDP *pomdp);
extern "C" int pomdp_uninitialize_observation_transitions_gpu(POMDP *pomdp);
extern "C" int pomdp_initialize_rewards_gpu(POMDP *pomdp);
extern "C" int pomdp_uninitialize_rewards_gpu(POMDP *pomdp);
extern "C" int pomdp_initialize_nonzero_beliefs_gpu(POMDP *pomdp);
extern "C" int pomdp_uninitialize_nonzero_beliefs_gpu(POMDP *pomdp);
extern "C" int pomdp_initialize_belief_points_gpu(POMDP *pomdp);
extern "C" int pomdp_uninitialize_belief_points_gpu(POMDP *pomdp);

} // namespace nova

#endif
