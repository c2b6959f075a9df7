// nova-mi355x: entry points added beside nova's ABI. None of them has a counterpart in the
// reference; they do not change any of its 75 signatures.
#ifndef NOVA_AMD_H
#define NOVA_AMD_H

#include <nova/mdp/mdp.h>
#include <nova/mdp/algorithms/mdp_vi_gpu.h>
#include <nova/pomdp/pomdp.h>
#include <nova/pomdp/algorithms/pomdp_pbvi_gpu.h>
#include <nova/pomdp/algorithms/pomdp_perseus_cpu.h>

namespace nova {

// Runs `sweeps` Bellman sweeps on the device continuing from vi's state (the V/VPrime
// ping-pong by parity of currentHorizon, exactly as `sweeps` calls of mdp_vi_update_gpu), as
// one captured HIP graph on the library stream, then synchronises. When d_V0 is non-null it
// is first copied (device to device, n floats) into d_V and d_VPrime and currentHorizon is
// reset to 0. elapsedMs (nullable) receives the device time of the sweeps alone, measured
// with HIP events on the stream the sweeps run on.
extern "C" int mdp_vi_run_gpu(const MDP *mdp, MDPVIGPU *vi, const float *d_V0,
        unsigned int sweeps, float *elapsedMs);

// ---- differentiable value iteration

// Which sweep kernel a solver runs: 0 = one thread per state, 1 = the wave-tiled sweep,
// 2 = the successor-sorted tiles, 3 = the successor-chunked passes (V chunks in LDS, products
// through HBM) that mdp_vi_initialize_gpu lays out for models of at least 2^24 (state, action,
// successor) slots with ns <= 255 (NOVA_AMD_VI_KERNEL=tiled|sorted|chunked overrides the size
// rule where the shape allows). All of them give the same V and pi bit for bit.
extern "C" int mdp_vi_sweep_kernel_gpu(const MDP *mdp, const MDPVIGPU *vi, unsigned int *kernel);

// Batched value iteration (no counterpart in the reference): solves count MDPs (each resident
// on the device, each with its own horizon, gamma and solver struct, VInitial or zeros) exactly
// as mdp_vi_execute_gpu would solve each one (mdp_vi_cpu.cpp:36-63 per sweep; V and pi bit for
// bit), with every sweep of every instance captured into one HIP graph: one launch per sweep
// for all instances, over a grid of every instance's state tiles. m <= 1024 per instance.
// policies (nullable) is an array of count empty policy structs that receive the results;
// the solver structs keep V/V'/pi on the device as after an initialize and the sweeps
// (currentHorizon = horizon). elapsedMs (nullable) gets the device time of the graph.
extern "C" int mdp_vi_batch_execute_gpu(const MDP *const *mdps, MDPVIGPU *const *vis,
        unsigned int count, MDPValueFunction *policies, float *elapsedMs);

// Solves V* = max_a (R + gamma T.V*) on the device to an FP64 residual max|V' - V| <= tol.
// Starts from vi->d_V (float; vi initialized or its d_* fields set by the caller) and runs
// FP64 sweeps until the residual is <= tol or has stopped falling (the FP64 rounding floor).
// Between blocks of 4 sweeps the iterate is moved along the slow mode of the error when the
// last step shows it: by MacQueen's midpoint shift along the constant vector when the step is
// of one sign and nearly uniform (rows of T summing to 1), else by the remaining steps of a
// mode that shrinks by a steady ratio, once two checks agree on it. The fixed point is that
// of plain sweeps; the solve stops on the residual alone. flags bit 1: plain sweeps, without
// the moves, first in float (on the state's sweep layout if it has one) while their residual
// still falls by 0.1% per 8 sweeps, then in FP64, checked every 8. flags bit 0: start the
// FP64 sweeps from d_V64[0..n) (a warm start). d_V64 is caller device memory of 2n doubles;
// on return d_V64[0..n) holds V* and vi->d_pi the greedy policy of the last sweep.
// sweepsDone / residual (nullable) report the sweeps run and the last FP64 residual. Only
// device arrays of mdp (d_S, d_T, d_R) are used.
extern "C" int mdp_vi_solve_gpu(const MDP *mdp, MDPVIGPU *vi, double *d_V64, double tol,
        unsigned int maxSweeps, unsigned int flags, unsigned int *sweepsDone, double *residual);

// The adjoint of that solve at policy d_pi (n) and values d_V64 (n doubles): solves
// (I - gamma P_pi)^T lambda = rho for d_rho (n floats) by FP64 fixed-point iterations (every
// 8, when the last two steps agree on one slowly shrinking mode, its remaining steps are
// added at once) until
// max|lambda_{k+1} - lambda_k| <= tol * max|lambda - rho| (or maxIters), writes lambda to
// d_lambda (n floats) and, when non-null, dL/dR into d_dR (n*m floats) and dL/dT into d_dT
// (n*m*ns floats). residual64 gets max|lambda - rho - gamma P_pi^T lambda| in FP64.
extern "C" int mdp_vi_adjoint_gpu(const MDP *mdp, const unsigned int *d_pi,
        const double *d_V64, const float *d_rho, float *d_lambda, float *d_dR, float *d_dT,
        float tol, unsigned int maxIters, unsigned int *itersDone, double *residual64);

// ---- PBVI (include/nova/pomdp/algorithms/pomdp_pbvi_gpu.h)

// Selects the arithmetic of the GPU PBVI backup for solvers initialized afterwards in this
// process and returns the previous mode: 0 = fast (default; FMA accumulation, results within
// float rounding of nova's), 1 = exact (every product and sum rounded in nova's order: Gamma
// and pi equal nova's CPU results bit for bit), 2 = dense beliefs scored on the matrix cores
// in bf16 with fp32 accumulation (any support length; n <= 36864; host S/T required),
// 3 = the same with fp32 matrix-core arithmetic. Other values only query.
extern "C" int nova_amd_pbvi_mode(int mode);

// Runs `iterations` PBVI iterations on the device continuing from pbvi's state (the
// Gamma/GammaPrime ping-pong by parity, as repeated pomdp_pbvi_update_gpu) as one captured
// HIP graph, then synchronises. elapsedMs (nullable) gets their device time from HIP events
// on the library stream. pbvi must have been initialized with pomdp_pbvi_initialize_gpu.
extern "C" int pomdp_pbvi_run_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi,
        unsigned int iterations, float *elapsedMs);

// Batched PBVI: solves count POMDPs (each resident on the device, each solver struct with
// GammaInitial set) exactly as pomdp_pbvi_execute_gpu would solve each one, but with every
// iteration of every instance captured into one HIP graph of grouped launches (one launch per
// stage for all instances). policies (nullable) is an array of count empty policy structs that
// receive the results. elapsedMs (nullable) gets the device time of the graph.
extern "C" int pomdp_pbvi_batch_execute_gpu(const POMDP *const *pomdps,
        POMDPPBVIGPU *const *pbvis, unsigned int count, POMDPAlphaVectors *policies,
        float *elapsedMs);

// One PBVI iteration with HIP events between its four kernels (transpose, score, action,
// materialize) on the library stream; stageMs[4] receives their device times in ms.
// Advances currentHorizon like pomdp_pbvi_update_gpu.
extern "C" int pomdp_pbvi_profile_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, float *stageMs);

// The scoring kernel an initialized GPU PBVI solver uses: 1, 2, 4 = alphas per lane on the
// vector ALUs, 8, 16, 32 = the fp32 matrix-core kernel with 4 waves, 8 waves, or 8 waves and a
// deeper prefetch per workgroup, 64 = its form with 16 observation rows (z <= 16), 128 = its
// form that sweeps every alpha tile in one workgroup (rz * ns <= 512), 256, 512 = the 4- and
// 8-wave kernels with operands requested three trips ahead (Gamma^T below 4 GB), 1024, 2048 =
// the same for the 16-row and the all-tiles kernels, 4096 = the 16-row kernel sweeping every
// alpha tile in one workgroup (z <= 16, rz * ns <= 512), 0 = the dense-belief matrix-core GEMM.
extern "C" int pomdp_pbvi_scoring_variant_gpu(const POMDPPBVIGPU *pbvi, unsigned int *variant);

// ---- PBVI argmax replay (gradients of L = sum_b max_alpha alpha . b with respect to R)

// Runs `iterations` iterations like pomdp_pbvi_run_gpu and records each one's choices:
// d_recJ[(k*r + i)*z + o] = alpha index chosen for alpha i, observation o at iteration k, and
// d_recA[k*r + i] = action chosen (-1: none; then d_recJ holds -1). Device arrays of
// iterations*r*z and iterations*r ints.
extern "C" int pomdp_pbvi_run_recorded_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi,
        unsigned int iterations, int *d_recJ, int *d_recA);

// d_values[i] = max_j Gamma[j] . b_i over belief i's sparse support (first maximiser from
// -1e10) and d_arg[i] = that j, for the rGamma alphas of d_Gamma (rGamma x n, device).
extern "C" int pomdp_pbvi_belief_values_gpu(const POMDP *pomdp, const float *d_Gamma,
        unsigned int rGamma, float *d_values, int *d_arg);

// dL/dR (n*m floats, device) of L = sum_i max_j Gamma_H[j] . b_i by replaying the records of
// `iterations` iterations from GammaInitial (which does not depend on R), d_finalArg being the
// maximisers of the final Gamma (pomdp_pbvi_belief_values_gpu). Uses the host S/T and the
// device O/Z/B of pomdp.
extern "C" int pomdp_pbvi_replay_grad_gpu(const POMDP *pomdp, unsigned int iterations,
        const int *d_recJ, const int *d_recA, const int *d_finalArg, float *d_dR);

// Copies the argmax record of the last iteration to the host: jstar[(i*m + a)*z + o] = the
// alpha chosen for belief i, action a, observation o (r*m*z ints, nullable) and actions[i]
// = the action chosen for belief i, or -1 when none beat -1e10 (r ints, nullable).
extern "C" int pomdp_pbvi_argmax_gpu(const POMDP *pomdp, const POMDPPBVIGPU *pbvi,
        int *jstar, unsigned int *actions);

// ---- Perseus (include/nova/pomdp/algorithms/pomdp_perseus_cpu.h)

// pomdp_perseus_execute_cpu with its backups and value updates on the GPU: same arguments,
// same randomized loop (beliefs drawn with libc rand() by nova's formula), same result
// (bit for bit in PBVI exact mode under the same srand seed; within float rounding otherwise).
// The POMDP must be resident on the device; per is left uninitialized on return.
extern "C" int pomdp_perseus_execute_gpu(const POMDP *pomdp, POMDPPerseusCPU *per,
        POMDPAlphaVectors *policy);

// ---- belief expansion (include/nova/pomdp/utilities/pomdp_expand_gpu.h)

// pomdp_expand_random_gpu with libc srand(seed) instead of srand(time): the same random
// stream as a seeded host run of the reference's loop, for reproducible expansions.
extern "C" int pomdp_expand_random_seeded_gpu(const POMDP *pomdp, unsigned int seed,
        unsigned int numDesiredBeliefPoints, unsigned int *maxNonZeroValues, float **Bnew);

// pomdp_expand_distinct_beliefs_cpu with the scoring of every (belief, action, observation)
// against every belief on the GPU; the appended beliefs equal the host function's bit for bit
// (src/pomdp/utilities/pomdp_expand_cpu.cpp:166-266). Needs rz*ns <= 1024.
extern "C" int pomdp_expand_distinct_beliefs_gpu(POMDP *pomdp);

// ---- stochastic shortest paths on the GPU (no counterpart in the reference, which runs LAO*
//      and RTDP on the CPU only)

// Jacobi min-cost Bellman backups of the k states listed in d_states (device array) from d_V:
// for each listed s, d_Vout[s] = min over a of R[s,a] + sum_i T[s,a,i] * f(d_V[S[s,a,i]]) and
// d_pi[s] = the first minimising action, with f = |.| when absV (the backup of LAO*,
// src/mdp/algorithms/ssp_lao_star_cpu.cpp:39-68, whose sign of V marks visitation) and f = the
// identity otherwise (RTDP's, src/mdp/algorithms/ssp_rtdp_cpu.cpp:39-74); each equals nova's
// backup of s from the same V bit for bit. Other entries of d_Vout and d_pi are left as they
// are. d_Vout must not be d_V. The MDP must be resident on the device. Synchronous.
extern "C" int nova_amd_ssp_backup_gpu(const MDP *mdp, unsigned int k,
        const unsigned int *d_states, int absV, const float *d_V, float *d_Vout,
        unsigned int *d_pi);

// SSP value iteration: Jacobi sweeps of that backup over every state from d_V0, with the goal
// states (mdp->d_goals) held at V = 0, pi = 0, until the largest change max_s |V'(s) - V(s)| of
// a sweep is at most epsilon or maxSweeps sweeps are done. d_V receives the last V and d_pi
// its greedy actions; *sweepsDone and *residual (when not null) the sweeps run and the last
// change. The MDP must be resident on the device with its goals. Synchronous.
extern "C" int nova_amd_ssp_sweeps_gpu(const MDP *mdp, int absV, const float *d_V0,
        unsigned int maxSweeps, float epsilon, float *d_V, unsigned int *d_pi,
        unsigned int *sweepsDone, float *residual);

// ---- executing alpha-vector policies over many beliefs (no counterpart in the reference,
//      which evaluates one belief per call)

// pomdp_alpha_vectors_value_and_action (src/pomdp/policies/pomdp_alpha_vectors.cpp:52-79) for
// k dense beliefs at once. Device arrays: d_Gamma (r x n alphas, row-major), d_pi (r actions),
// d_B (k x n beliefs, row-major); outputs d_V (k values), d_action (k actions) and, when not
// null, d_alpha (k indices of the chosen alpha, -1 when no alpha beats -1e10, in which case the
// value is -1e10 and the action 0, as the reference leaves them). precision: 0 = exact (each
// dot product summed over the states in order with separately rounded products and sums:
// equal to the reference bit for bit), 1 = fp32 on the matrix cores, 2 = bf16 operands with
// fp32 accumulation on the matrix cores (the -1e10 sentinel itself rounds to -9.9992e9 there).
// r <= 4,194,240. Synchronous; calls are serialised.
extern "C" int nova_amd_alpha_vectors_value_and_action_gpu(unsigned int n, unsigned int r,
        const float *d_Gamma, const unsigned int *d_pi, unsigned int k, const float *d_B,
        int precision, float *d_V, unsigned int *d_action, int *d_alpha);

// No counterpart in the reference. The exact mode (precision 0) of
// nova_amd_alpha_vectors_value_and_action_gpu for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc): it calls no
// hipMalloc, hipFree or allocating library routine. Its per-(belief, 64-alpha tile) partials
// live in a static 16 MiB device array that each GPU the library runs on holds until the
// process ends, so ceil(r / 64) * k <= 2,097,152 (larger calls return NOVA_ERROR_INVALID_DATA).
// Same arguments and results as the exact mode of that entry point, bit for bit. Calls on one
// GPU share the array; calls that overlap on one GPU race (the library serialises the calls of
// one process). Synchronous.
extern "C" int nova_amd_alpha_vectors_value_and_action_noalloc_gpu(unsigned int n,
        unsigned int r, const float *d_Gamma, const unsigned int *d_pi, unsigned int k,
        const float *d_B, float *d_V, unsigned int *d_action, int *d_alpha);

// pomdp_belief_update_cpu (src/pomdp/utilities/pomdp_model_cpu.cpp:89-137) for k (belief,
// action, observation) triples at once, bit for bit: d_B (k x n beliefs), d_action and
// d_observation (k each) on the device; d_Bp (k x n) receives the updated beliefs and
// d_status (k ints) NOVA_SUCCESS, NOVA_WARNING_INVALID_BELIEF (the observation is impossible:
// the row keeps the unnormalised values, as the reference returns them) or
// NOVA_ERROR_INVALID_DATA (action or observation out of range). The POMDP must be resident
// on the device and keep its host S and T, from which predecessor lists are built on first
// use and kept until another model is passed. Synchronous; calls are serialised.
extern "C" int nova_amd_belief_update_batch_gpu(const POMDP *pomdp, unsigned int k,
        const float *d_B, const unsigned int *d_action, const unsigned int *d_observation,
        float *d_Bp, int *d_status);

// Alpha choice inside pomdp_expand_pema_cpu (no counterpart in the reference): 0 (default)
// reproduces nova's loop (pomdp_expand_cpu.cpp:391-402), which never accumulates the dot
// product and so keeps alpha 0 unless a single product lies below -1e10; 1 takes the first
// maximiser of alpha . b over the policy's alphas. Applies to later calls in the process.
extern "C" int nova_amd_pema_mode(int mode);

// pomdp_expand_pema_cpu (pomdp_expand_cpu.cpp:269-466) with the work of every (belief,
// action, observation) on the GPU: the same appended belief bit for bit, in either alpha
// choice. The POMDP must be resident on the device and keep its host arrays; its host Z and B
// grow by one belief (call pomdp_initialize_gpu again to refresh the device copy).
extern "C" int pomdp_expand_pema_gpu(POMDP *pomdp, const POMDPAlphaVectors *policy);

// pomdp_sigma_cpu (src/pomdp/utilities/pomdp_sigma_cpu.cpp:50-118) on the device copy of the
// belief set (no counterpart in the reference): each belief keeps its numDesiredNonZeroValues
// largest entries, renormalised by their sum, and *sigma receives min(1, the smallest kept
// mass). The new set (rz = numDesiredNonZeroValues) replaces d_Z/d_B and, when the POMDP keeps
// host arrays, Z/B as pomdp_sigma_cpu replaces them. Entries of equal value are ordered by their
// position in the row (nova's std::sort leaves their order to the library's algorithm), so for
// rows of distinct values the result equals nova's bit for bit. rz <= 8,192. Synchronous.
extern "C" int pomdp_sigma_gpu(POMDP *pomdp, unsigned int numDesiredNonZeroValues, float *sigma);

// delete[] of a float array the library allocated for the caller (Bnew of the expansions,
// bp of pomdp_belief_update_cpu), for callers that cannot call delete[] themselves.
extern "C" int nova_amd_free_floats(float *p);

// Library build facts for callers and tests: writes up to len-1 characters of a
// NUL-terminated description (target architecture, kernel variants) into buf.
extern "C" int nova_amd_build_info(char *buf, unsigned int len);

} // namespace nova

#endif
