// nova-mi355x: entry points added beside nova's ABI. None of them has a counterpart in the
// reference; they do not change any of its 75 signatures.
#ifndef NOVA_AMD_H
#define NOVA_AMD_H

#include <nova/mdp/mdp.h>
#include <nova/mdp/algorithms/mdp_vi_gpu.h>
#include <nova/pomdp/pomdp.h>
#include <nova/pomdp/algorithms/pomdp_pbvi_gpu.h>

namespace nova {

// Runs `sweeps` Bellman sweeps on the device continuing from vi's state (the V/VPrime
// ping-pong by parity of currentHorizon, exactly as `sweeps` calls of mdp_vi_update_gpu), as
// one captured HIP graph on the library stream, then synchronises. When d_V0 is non-null it
// is first copied (device to device, n floats) into d_V and d_VPrime and currentHorizon is
// reset to 0. elapsedMs (nullable) receives the device time of the sweeps alone, measured
// with HIP events on the stream the sweeps run on.
extern "C" int mdp_vi_run_gpu(const MDP *mdp, MDPVIGPU *vi, const float *d_V0,
        unsigned int sweeps, float *elapsedMs);

// ---- PBVI (include/nova/pomdp/algorithms/pomdp_pbvi_gpu.h)

// Selects the arithmetic of the GPU PBVI backup for later calls in this process and returns
// the previous mode: 0 = fast (default; FMA accumulation, results within float rounding of
// nova's), 1 = exact (every product and sum rounded in nova's order: Gamma and pi equal
// nova's CPU results bit for bit). Other values only query.
extern "C" int nova_amd_pbvi_mode(int mode);

// Runs `iterations` PBVI iterations on the device continuing from pbvi's state (the
// Gamma/GammaPrime ping-pong by parity, as repeated pomdp_pbvi_update_gpu) as one captured
// HIP graph, then synchronises. elapsedMs (nullable) gets their device time from HIP events
// on the library stream. pbvi must have been initialized with pomdp_pbvi_initialize_gpu.
extern "C" int pomdp_pbvi_run_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi,
        unsigned int iterations, float *elapsedMs);

// Copies the argmax record of the last iteration to the host: jstar[(i*m + a)*z + o] = the
// alpha chosen for belief i, action a, observation o (r*m*z ints, nullable) and actions[i]
// = the action chosen for belief i, or -1 when none beat -1e10 (r ints, nullable).
extern "C" int pomdp_pbvi_argmax_gpu(const POMDP *pomdp, const POMDPPBVIGPU *pbvi,
        int *jstar, unsigned int *actions);

// ---- belief expansion (include/nova/pomdp/utilities/pomdp_expand_gpu.h)

// pomdp_expand_random_gpu with libc srand(seed) instead of srand(time): the same random
// stream as a seeded host run of the reference's loop, for reproducible expansions.
extern "C" int pomdp_expand_random_seeded_gpu(const POMDP *pomdp, unsigned int seed,
        unsigned int numDesiredBeliefPoints, unsigned int *maxNonZeroValues, float **Bnew);

// delete[] of a float array the library allocated for the caller (Bnew of the expansions,
// bp of pomdp_belief_update_cpu), for callers that cannot call delete[] themselves.
extern "C" int nova_amd_free_floats(float *p);

// Library build facts for callers and tests: writes up to len-1 characters of a
// NUL-terminated description (target architecture, kernel variants) into buf.
extern "C" int nova_amd_build_info(char *buf, unsigned int len);

} // namespace nova

#endif
