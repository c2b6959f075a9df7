// nova-mi355x: compile-time checks of the ABI structure layouts.
//
// Callers built against nova's headers (the C++ core, the Python ctypes layer) pass these
// structures by pointer, so every size and field offset must equal the reference's x86-64
// layout: include/nova/mdp/mdp.h:53-75, include/nova/pomdp/pomdp.h:61-85,
// include/nova/mdp/algorithms/mdp_vi_cpu.h and mdp_vi_gpu.h, include/nova/pomdp/algorithms/
// pomdp_pbvi_cpu.h, pomdp_pbvi_gpu.h and pomdp_perseus_cpu.h,
// include/nova/mdp/policies/mdp_value_function.h:48-55 and
// include/nova/pomdp/policies/pomdp_alpha_vectors.h:39-45 (sizes and offsets measured with
// g++ 11 offsetof against those headers). A mismatch stops the build here.
#include <cstddef>

#include <nova/mdp/algorithms/mdp_vi_cpu.h>
#include <nova/mdp/algorithms/mdp_vi_gpu.h>
#include <nova/mdp/algorithms/ssp_lao_star_cpu.h>
#include <nova/mdp/algorithms/ssp_rtdp_cpu.h>
#include <nova/mdp/mdp.h>
#include <nova/mdp/policies/mdp_value_function.h>
#include <nova/pomdp/algorithms/pomdp_pbvi_cpu.h>
#include <nova/pomdp/algorithms/pomdp_pbvi_gpu.h>
#include <nova/pomdp/algorithms/pomdp_perseus_cpu.h>
#include <nova/pomdp/policies/pomdp_alpha_vectors.h>
#include <nova/pomdp/pomdp.h>

namespace {

using namespace nova;

#define NOVA_AMD_AT(T, f, off) static_assert(offsetof(T, f) == (off), #T "::" #f " offset")

static_assert(sizeof(MDP) == 96, "MDP size");
NOVA_AMD_AT(MDP, n, 0);
NOVA_AMD_AT(MDP, ns, 4);
NOVA_AMD_AT(MDP, m, 8);
NOVA_AMD_AT(MDP, gamma, 12);
NOVA_AMD_AT(MDP, horizon, 16);
NOVA_AMD_AT(MDP, epsilon, 20);
NOVA_AMD_AT(MDP, s0, 24);
NOVA_AMD_AT(MDP, ng, 28);
NOVA_AMD_AT(MDP, goals, 32);
NOVA_AMD_AT(MDP, S, 40);
NOVA_AMD_AT(MDP, T, 48);
NOVA_AMD_AT(MDP, R, 56);
NOVA_AMD_AT(MDP, d_goals, 64);
NOVA_AMD_AT(MDP, d_S, 72);
NOVA_AMD_AT(MDP, d_T, 80);
NOVA_AMD_AT(MDP, d_R, 88);

static_assert(sizeof(POMDP) == 128, "POMDP size");
NOVA_AMD_AT(POMDP, n, 0);
NOVA_AMD_AT(POMDP, ns, 4);
NOVA_AMD_AT(POMDP, m, 8);
NOVA_AMD_AT(POMDP, z, 12);
NOVA_AMD_AT(POMDP, r, 16);
NOVA_AMD_AT(POMDP, rz, 20);
NOVA_AMD_AT(POMDP, gamma, 24);
NOVA_AMD_AT(POMDP, horizon, 28);
NOVA_AMD_AT(POMDP, S, 32);
NOVA_AMD_AT(POMDP, T, 40);
NOVA_AMD_AT(POMDP, O, 48);
NOVA_AMD_AT(POMDP, R, 56);
NOVA_AMD_AT(POMDP, Z, 64);
NOVA_AMD_AT(POMDP, B, 72);
NOVA_AMD_AT(POMDP, d_S, 80);
NOVA_AMD_AT(POMDP, d_T, 88);
NOVA_AMD_AT(POMDP, d_O, 96);
NOVA_AMD_AT(POMDP, d_R, 104);
NOVA_AMD_AT(POMDP, d_Z, 112);
NOVA_AMD_AT(POMDP, d_B, 120);

static_assert(sizeof(MDPVICPU) == 40, "MDPVICPU size");
NOVA_AMD_AT(MDPVICPU, VInitial, 0);
NOVA_AMD_AT(MDPVICPU, currentHorizon, 8);
NOVA_AMD_AT(MDPVICPU, V, 16);
NOVA_AMD_AT(MDPVICPU, VPrime, 24);
NOVA_AMD_AT(MDPVICPU, pi, 32);

static_assert(sizeof(MDPVIGPU) == 40, "MDPVIGPU size");
NOVA_AMD_AT(MDPVIGPU, VInitial, 0);
NOVA_AMD_AT(MDPVIGPU, numThreads, 8);
NOVA_AMD_AT(MDPVIGPU, currentHorizon, 12);
NOVA_AMD_AT(MDPVIGPU, d_V, 16);
NOVA_AMD_AT(MDPVIGPU, d_VPrime, 24);
NOVA_AMD_AT(MDPVIGPU, d_pi, 32);

static_assert(sizeof(POMDPPBVICPU) == 40, "POMDPPBVICPU size");
NOVA_AMD_AT(POMDPPBVICPU, GammaInitial, 0);
NOVA_AMD_AT(POMDPPBVICPU, currentHorizon, 8);
NOVA_AMD_AT(POMDPPBVICPU, Gamma, 16);
NOVA_AMD_AT(POMDPPBVICPU, GammaPrime, 24);
NOVA_AMD_AT(POMDPPBVICPU, pi, 32);

static_assert(sizeof(POMDPPBVIGPU) == 48, "POMDPPBVIGPU size");
NOVA_AMD_AT(POMDPPBVIGPU, GammaInitial, 0);
NOVA_AMD_AT(POMDPPBVIGPU, numThreads, 8);
NOVA_AMD_AT(POMDPPBVIGPU, currentHorizon, 12);
NOVA_AMD_AT(POMDPPBVIGPU, d_Gamma, 16);
NOVA_AMD_AT(POMDPPBVIGPU, d_GammaPrime, 24);
NOVA_AMD_AT(POMDPPBVIGPU, d_pi, 32);
NOVA_AMD_AT(POMDPPBVIGPU, d_alphaBA, 40);

static_assert(sizeof(POMDPPerseusCPU) == 64, "POMDPPerseusCPU size");
NOVA_AMD_AT(POMDPPerseusCPU, GammaInitial, 0);
NOVA_AMD_AT(POMDPPerseusCPU, currentHorizon, 8);
NOVA_AMD_AT(POMDPPerseusCPU, rGamma, 12);
NOVA_AMD_AT(POMDPPerseusCPU, rGammaPrime, 16);
NOVA_AMD_AT(POMDPPerseusCPU, rTilde, 20);
NOVA_AMD_AT(POMDPPerseusCPU, BTilde, 24);
NOVA_AMD_AT(POMDPPerseusCPU, Gamma, 32);
NOVA_AMD_AT(POMDPPerseusCPU, GammaPrime, 40);
NOVA_AMD_AT(POMDPPerseusCPU, pi, 48);
NOVA_AMD_AT(POMDPPerseusCPU, piPrime, 56);

static_assert(sizeof(MDPValueFunction) == 40, "MDPValueFunction size");
NOVA_AMD_AT(MDPValueFunction, n, 0);
NOVA_AMD_AT(MDPValueFunction, m, 4);
NOVA_AMD_AT(MDPValueFunction, r, 8);
NOVA_AMD_AT(MDPValueFunction, S, 16);
NOVA_AMD_AT(MDPValueFunction, V, 24);
NOVA_AMD_AT(MDPValueFunction, pi, 32);

static_assert(sizeof(POMDPAlphaVectors) == 32, "POMDPAlphaVectors size");
NOVA_AMD_AT(POMDPAlphaVectors, n, 0);
NOVA_AMD_AT(POMDPAlphaVectors, m, 4);
NOVA_AMD_AT(POMDPAlphaVectors, r, 8);
NOVA_AMD_AT(POMDPAlphaVectors, Gamma, 16);
NOVA_AMD_AT(POMDPAlphaVectors, pi, 24);

// The SSP solver structs (include/nova/mdp/algorithms/ssp_lao_star_cpu.h, ssp_rtdp_cpu.h)
static_assert(sizeof(SSPLAOStarCPU) == 32, "SSPLAOStarCPU size");
NOVA_AMD_AT(SSPLAOStarCPU, VInitial, 0);
NOVA_AMD_AT(SSPLAOStarCPU, currentHorizon, 8);
NOVA_AMD_AT(SSPLAOStarCPU, V, 16);
NOVA_AMD_AT(SSPLAOStarCPU, pi, 24);

static_assert(sizeof(SSPRTDPCPU) == 40, "SSPRTDPCPU size");
NOVA_AMD_AT(SSPRTDPCPU, VInitial, 0);
NOVA_AMD_AT(SSPRTDPCPU, trials, 8);
NOVA_AMD_AT(SSPRTDPCPU, currentTrial, 12);
NOVA_AMD_AT(SSPRTDPCPU, currentHorizon, 16);
NOVA_AMD_AT(SSPRTDPCPU, V, 24);
NOVA_AMD_AT(SSPRTDPCPU, pi, 32);

#undef NOVA_AMD_AT

} // namespace
