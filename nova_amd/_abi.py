"""ctypes view of libnova.so: the ABI structures and the argtypes of every entry point.

Field order and types are interface facts shared with nova's own ctypes mirrors
(python/nova/nova_mdp.py:47-62, nova_pomdp.py:50-69, nova_mdp_vi.py:50-54 and 79-84,
nova_pomdp_pbvi.py:50-54 and 79-85, nova_pomdp_perseus.py:50-59,
nova_mdp_value_function.py:45-50, nova_pomdp_alpha_vectors.py:45-49). As in the reference no
restype is set: every entry point returns a C int status (0 = success).

The library is nova_amd/lib/libnova.so, built by ``python -m nova_amd.build``. Loading fails
with an ImportError when it is missing: there is no fallback implementation.
"""
import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# NOVA_AMD_LIB names another build of the library (the host-sanitizer build of the tests)
LIB_PATH = os.environ.get("NOVA_AMD_LIB") or os.path.join(HERE, "lib", "libnova.so")

c_float_p = ct.POINTER(ct.c_float)
c_int_p = ct.POINTER(ct.c_int)
c_uint_p = ct.POINTER(ct.c_uint)


class NovaMDP(ct.Structure):
    _fields_ = [("n", ct.c_uint), ("ns", ct.c_uint), ("m", ct.c_uint),
                ("gamma", ct.c_float), ("horizon", ct.c_uint), ("epsilon", ct.c_float),
                ("s0", ct.c_uint), ("ng", ct.c_uint), ("goals", c_uint_p),
                ("S", c_int_p), ("T", c_float_p), ("R", c_float_p),
                ("d_goals", c_uint_p), ("d_S", c_int_p), ("d_T", c_float_p), ("d_R", c_float_p)]


class NovaPOMDP(ct.Structure):
    _fields_ = [("n", ct.c_uint), ("ns", ct.c_uint), ("m", ct.c_uint), ("z", ct.c_uint),
                ("r", ct.c_uint), ("rz", ct.c_uint), ("gamma", ct.c_float),
                ("horizon", ct.c_uint),
                ("S", c_int_p), ("T", c_float_p), ("O", c_float_p), ("R", c_float_p),
                ("Z", c_int_p), ("B", c_float_p),
                ("d_S", c_int_p), ("d_T", c_float_p), ("d_O", c_float_p), ("d_R", c_float_p),
                ("d_Z", c_int_p), ("d_B", c_float_p)]


class NovaMDPVICPU(ct.Structure):
    _fields_ = [("VInitial", c_float_p), ("currentHorizon", ct.c_uint),
                ("V", c_float_p), ("VPrime", c_float_p), ("pi", c_uint_p)]


class NovaMDPVIGPU(ct.Structure):
    _fields_ = [("VInitial", c_float_p), ("numThreads", ct.c_uint),
                ("currentHorizon", ct.c_uint),
                ("d_V", c_float_p), ("d_VPrime", c_float_p), ("d_pi", c_uint_p)]


class NovaPOMDPPBVICPU(ct.Structure):
    _fields_ = [("GammaInitial", c_float_p), ("currentHorizon", ct.c_uint),
                ("Gamma", c_float_p), ("GammaPrime", c_float_p), ("pi", c_uint_p)]


class NovaPOMDPPBVIGPU(ct.Structure):
    _fields_ = [("GammaInitial", c_float_p), ("numThreads", ct.c_uint),
                ("currentHorizon", ct.c_uint),
                ("d_Gamma", c_float_p), ("d_GammaPrime", c_float_p), ("d_pi", c_uint_p),
                ("d_alphaBA", c_float_p)]


class NovaPOMDPPerseusCPU(ct.Structure):
    _fields_ = [("GammaInitial", c_float_p), ("currentHorizon", ct.c_uint),
                ("rGamma", ct.c_uint), ("rGammaPrime", ct.c_uint), ("rTilde", ct.c_uint),
                ("BTilde", c_uint_p), ("Gamma", c_float_p), ("GammaPrime", c_float_p),
                ("pi", c_uint_p), ("piPrime", c_uint_p)]


class NovaMDPValueFunction(ct.Structure):
    _fields_ = [("n", ct.c_uint), ("m", ct.c_uint), ("r", ct.c_uint),
                ("S", c_uint_p), ("V", c_float_p), ("pi", c_uint_p)]


class NovaPOMDPAlphaVectors(ct.Structure):
    _fields_ = [("n", ct.c_uint), ("m", ct.c_uint), ("r", ct.c_uint),
                ("Gamma", c_float_p), ("pi", c_uint_p)]


class NovaSSPLAOStarCPU(ct.Structure):
    _fields_ = [("VInitial", c_float_p), ("currentHorizon", ct.c_uint),
                ("V", c_float_p), ("pi", c_uint_p)]


class NovaSSPRTDPCPU(ct.Structure):
    _fields_ = [("VInitial", c_float_p), ("trials", ct.c_uint), ("currentTrial", ct.c_uint),
                ("currentHorizon", ct.c_uint), ("V", c_float_p), ("pi", c_uint_p)]


# Expected x86-64 layouts (sizeof, {field: offset}); measured from the reference headers.
LAYOUTS = {
    NovaMDP: (96, dict(n=0, ns=4, m=8, gamma=12, horizon=16, epsilon=20, s0=24, ng=28,
                       goals=32, S=40, T=48, R=56, d_goals=64, d_S=72, d_T=80, d_R=88)),
    NovaPOMDP: (128, dict(n=0, ns=4, m=8, z=12, r=16, rz=20, gamma=24, horizon=28, S=32,
                          T=40, O=48, R=56, Z=64, B=72, d_S=80, d_T=88, d_O=96, d_R=104,
                          d_Z=112, d_B=120)),
    NovaMDPVICPU: (40, dict(VInitial=0, currentHorizon=8, V=16, VPrime=24, pi=32)),
    NovaMDPVIGPU: (40, dict(VInitial=0, numThreads=8, currentHorizon=12, d_V=16,
                            d_VPrime=24, d_pi=32)),
    NovaPOMDPPBVICPU: (40, dict(GammaInitial=0, currentHorizon=8, Gamma=16, GammaPrime=24,
                                pi=32)),
    NovaPOMDPPBVIGPU: (48, dict(GammaInitial=0, numThreads=8, currentHorizon=12, d_Gamma=16,
                                d_GammaPrime=24, d_pi=32, d_alphaBA=40)),
    NovaPOMDPPerseusCPU: (64, dict(GammaInitial=0, currentHorizon=8, rGamma=12,
                                   rGammaPrime=16, rTilde=20, BTilde=24, Gamma=32,
                                   GammaPrime=40, pi=48, piPrime=56)),
    NovaMDPValueFunction: (40, dict(n=0, m=4, r=8, S=16, V=24, pi=32)),
    NovaPOMDPAlphaVectors: (32, dict(n=0, m=4, r=8, Gamma=16, pi=24)),
}

_P = ct.POINTER
U = ct.c_uint
_MDP, _POMDP = _P(NovaMDP), _P(NovaPOMDP)
_VIC, _VIG = _P(NovaMDPVICPU), _P(NovaMDPVIGPU)
_PBC, _PBG = _P(NovaPOMDPPBVICPU), _P(NovaPOMDPPBVIGPU)
_PER = _P(NovaPOMDPPerseusCPU)
_MVF, _PAV = _P(NovaMDPValueFunction), _P(NovaPOMDPAlphaVectors)
_LAO, _RTDP = _P(NovaSSPLAOStarCPU), _P(NovaSSPRTDPCPU)

# The 75 entry points nova's ctypes layer binds at import (python/nova/nova_*.py).
ABI_SIGNATURES = {
    # python/nova/nova_mdp.py:67-89
    "mdp_initialize_cpu": [_MDP, U, U, U, ct.c_float, U, ct.c_float, U, U],
    "mdp_uninitialize_cpu": [_MDP],
    "mdp_initialize_gpu": [_MDP], "mdp_uninitialize_gpu": [_MDP],
    "mdp_initialize_successors_gpu": [_MDP], "mdp_uninitialize_successors_gpu": [_MDP],
    "mdp_initialize_state_transitions_gpu": [_MDP],
    "mdp_uninitialize_state_transitions_gpu": [_MDP],
    "mdp_initialize_rewards_gpu": [_MDP], "mdp_uninitialize_rewards_gpu": [_MDP],
    "mdp_initialize_goals_gpu": [_MDP], "mdp_uninitialize_goals_gpu": [_MDP],
    # python/nova/nova_mdp_vi.py:58-103
    "mdp_vi_initialize_cpu": [_MDP, _VIC], "mdp_vi_execute_cpu": [_MDP, _VIC, _MVF],
    "mdp_vi_uninitialize_cpu": [_MDP, _VIC], "mdp_vi_update_cpu": [_MDP, _VIC],
    "mdp_vi_get_policy_cpu": [_MDP, _VIC, _MVF],
    "mdp_vi_initialize_gpu": [_MDP, _VIG], "mdp_vi_execute_gpu": [_MDP, _VIG, _MVF],
    "mdp_vi_uninitialize_gpu": [_MDP, _VIG], "mdp_vi_update_gpu": [_MDP, _VIG],
    "mdp_vi_get_policy_gpu": [_MDP, _VIG, _MVF],
    # python/nova/nova_mdp_value_function.py:55-59
    "mdp_value_function_initialize": [_MVF, U, U, U],
    "mdp_value_function_uninitialize": [_MVF],
    # python/nova/nova_ssp_lao_star.py:57-72 and nova_ssp_rtdp.py:59-74
    "ssp_lao_star_initialize_cpu": [_MDP, _LAO], "ssp_lao_star_execute_cpu": [_MDP, _LAO, _MVF],
    "ssp_lao_star_uninitialize_cpu": [_MDP, _LAO], "ssp_lao_star_update_cpu": [_MDP, _LAO],
    "ssp_lao_star_get_policy_cpu": [_MDP, _LAO, _MVF],
    "ssp_rtdp_initialize_cpu": [_MDP, _RTDP], "ssp_rtdp_execute_cpu": [_MDP, _RTDP, _MVF],
    "ssp_rtdp_uninitialize_cpu": [_MDP, _RTDP], "ssp_rtdp_update_cpu": [_MDP, _RTDP],
    "ssp_rtdp_get_policy_cpu": [_MDP, _RTDP, _MVF],
    # python/nova/nova_pomdp.py:74-128
    "pomdp_initialize_cpu": [_POMDP, U, U, U, U, U, U, ct.c_float, U],
    "pomdp_uninitialize_cpu": [_POMDP],
    "pomdp_belief_update_cpu": [_POMDP, c_float_p, U, U, _P(c_float_p)],
    "pomdp_add_new_raw_beliefs_cpu": [_POMDP, U, c_float_p],
    "pomdp_expand_random_cpu": [_POMDP, U],
    "pomdp_expand_distinct_beliefs_cpu": [_POMDP],
    "pomdp_expand_pema_cpu": [_POMDP, _PAV],
    "pomdp_sigma_cpu": [_POMDP, U, _P(ct.c_float)],
    "pomdp_initialize_gpu": [_POMDP], "pomdp_uninitialize_gpu": [_POMDP],
    "pomdp_initialize_successors_gpu": [_POMDP], "pomdp_uninitialize_successors_gpu": [_POMDP],
    "pomdp_initialize_state_transitions_gpu": [_POMDP],
    "pomdp_uninitialize_state_transitions_gpu": [_POMDP],
    "pomdp_initialize_observation_transitions_gpu": [_POMDP],
    "pomdp_uninitialize_observation_transitions_gpu": [_POMDP],
    "pomdp_initialize_rewards_gpu": [_POMDP], "pomdp_uninitialize_rewards_gpu": [_POMDP],
    "pomdp_initialize_nonzero_beliefs_gpu": [_POMDP],
    "pomdp_uninitialize_nonzero_beliefs_gpu": [_POMDP],
    "pomdp_initialize_belief_points_gpu": [_POMDP],
    "pomdp_uninitialize_belief_points_gpu": [_POMDP],
    "pomdp_expand_random_gpu": [_POMDP, U, U, _P(ct.c_uint), _P(c_float_p)],
    # python/nova/nova_pomdp_pbvi.py:58-104
    "pomdp_pbvi_initialize_cpu": [_POMDP, _PBC], "pomdp_pbvi_execute_cpu": [_POMDP, _PBC, _PAV],
    "pomdp_pbvi_uninitialize_cpu": [_POMDP, _PBC], "pomdp_pbvi_update_cpu": [_POMDP, _PBC],
    "pomdp_pbvi_get_policy_cpu": [_POMDP, _PBC, _PAV],
    "pomdp_pbvi_initialize_gpu": [_POMDP, _PBG], "pomdp_pbvi_execute_gpu": [_POMDP, _PBG, _PAV],
    "pomdp_pbvi_uninitialize_gpu": [_POMDP, _PBG], "pomdp_pbvi_update_gpu": [_POMDP, _PBG],
    "pomdp_pbvi_get_policy_gpu": [_POMDP, _PBG, _PAV],
    # python/nova/nova_pomdp_perseus.py:63-78
    "pomdp_perseus_initialize_cpu": [_POMDP, _PER],
    "pomdp_perseus_execute_cpu": [_POMDP, _PER, _PAV],
    "pomdp_perseus_uninitialize_cpu": [_POMDP, _PER], "pomdp_perseus_update_cpu": [_POMDP, _PER],
    "pomdp_perseus_get_policy_cpu": [_POMDP, _PER, _PAV],
    # python/nova/nova_pomdp_alpha_vectors.py:54-62
    "pomdp_alpha_vectors_initialize": [_PAV, U, U, U],
    "pomdp_alpha_vectors_value_and_action": [_PAV, c_float_p, _P(ct.c_float), _P(ct.c_uint)],
    "pomdp_alpha_vectors_uninitialize": [_PAV],
}

# Entry points added beside nova's ABI (include/nova/amd/*.h).
EXTRA_SIGNATURES = {
    "mdp_vi_run_gpu": [_MDP, _VIG, ct.c_void_p, U, _P(ct.c_float)],
    "nova_amd_build_info": [ct.c_char_p, U],
    "nova_amd_pbvi_mode": [ct.c_int],
    "pomdp_pbvi_run_gpu": [_POMDP, _PBG, U, _P(ct.c_float)],
    "pomdp_pbvi_argmax_gpu": [_POMDP, _PBG, c_int_p, c_uint_p],
    "pomdp_pbvi_profile_gpu": [_POMDP, _PBG, _P(ct.c_float)],
    "pomdp_pbvi_run_recorded_gpu": [_POMDP, _PBG, U, ct.c_void_p, ct.c_void_p],
    "pomdp_pbvi_belief_values_gpu": [_POMDP, ct.c_void_p, U, ct.c_void_p, ct.c_void_p],
    "pomdp_pbvi_replay_grad_gpu": [_POMDP, U, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                   ct.c_void_p],
    "pomdp_pbvi_batch_execute_gpu": [ct.c_void_p, ct.c_void_p, U, ct.c_void_p, _P(ct.c_float)],
    "pomdp_expand_random_seeded_gpu": [_POMDP, U, U, _P(ct.c_uint), _P(c_float_p)],
    "nova_amd_free_floats": [c_float_p],
    "pomdp_expand_distinct_beliefs_gpu": [_POMDP],
    "pomdp_pbvi_scoring_variant_gpu": [_PBG, _P(ct.c_uint)],
    "pomdp_perseus_execute_gpu": [_POMDP, _PER, _PAV],
    "nova_amd_alpha_vectors_value_and_action_gpu": [U, U, ct.c_void_p, ct.c_void_p, U,
                                                    ct.c_void_p, ct.c_int, ct.c_void_p,
                                                    ct.c_void_p, ct.c_void_p],
    "nova_amd_alpha_vectors_value_and_action_noalloc_gpu": [U, U, ct.c_void_p, ct.c_void_p, U,
                                                            ct.c_void_p, ct.c_void_p,
                                                            ct.c_void_p, ct.c_void_p],
    "nova_amd_ssp_backup_gpu": [_MDP, U, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p,
                                ct.c_void_p],
    "nova_amd_ssp_sweeps_gpu": [_MDP, ct.c_int, ct.c_void_p, U, ct.c_float, ct.c_void_p,
                                ct.c_void_p, _P(ct.c_uint), _P(ct.c_float)],
    "nova_amd_belief_update_batch_gpu": [_POMDP, U, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                         ct.c_void_p, ct.c_void_p],
    "pomdp_sigma_gpu": [_POMDP, U, _P(ct.c_float)],
    "nova_amd_pema_mode": [ct.c_int],
    "pomdp_expand_pema_gpu": [_POMDP, _PAV],
    "mdp_vi_sweep_kernel_gpu": [_MDP, _VIG, _P(ct.c_uint)],
    "mdp_vi_batch_execute_gpu": [ct.c_void_p, ct.c_void_p, U, ct.c_void_p, _P(ct.c_float)],
    "mdp_vi_solve_gpu": [_MDP, _VIG, ct.c_void_p, ct.c_double, U, U, _P(ct.c_uint),
                         _P(ct.c_double)],
    "mdp_vi_adjoint_gpu": [_MDP, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                           ct.c_void_p, ct.c_void_p, ct.c_float, U, _P(ct.c_uint),
                           _P(ct.c_double)],
}

_lib = None


def load():
    """Returns the loaded libnova.so (argtypes set). Raises ImportError when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    # PyTorch ships its own HIP runtime (torch/lib/libamdhip64.so, SONAME libamdhip64.so.7).
    # When it is loaded first, libnova.so's dependency on libamdhip64.so.7 binds to that same
    # runtime; loading libnova.so first would bring in /opt/rocm's copy and torch would then
    # start a second HIP runtime in the process, which finds no GPU. So torch goes first.
    if not os.environ.get("NOVA_AMD_NO_TORCH"):
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    if not os.path.exists(LIB_PATH):
        raise ImportError("nova_amd: %s is missing; build it with `python -m nova_amd.build`"
                          % LIB_PATH)
    lib = ct.CDLL(LIB_PATH)
    for name, argtypes in list(ABI_SIGNATURES.items()) + list(EXTRA_SIGNATURES.items()):
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.argtypes = argtypes
            fn.restype = ct.c_int
    _lib = lib
    return lib


def missing_symbols(lib=None):
    """Names of the reference's 75 entry points (plus the added ones) the library lacks."""
    lib = lib or load()
    return [n for n in list(ABI_SIGNATURES) + list(EXTRA_SIGNATURES) if not hasattr(lib, n)]
