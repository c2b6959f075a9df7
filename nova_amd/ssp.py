"""Stochastic-shortest-path solvers (mirror python/nova/ssp_lao_star.py:35-120 and
python/nova/ssp_rtdp.py:35-125 of the reference). R holds positive costs; the MDP's goals are
absorbing and s0 is the start. Both solvers run on the host, as in the reference, and return
an MDPValueFunction over the states they kept: LAO* the final greedy solution graph from s0,
RTDP the states its trials expanded.

The heuristic VInitial defaults to zeros (admissible for positive costs)."""
import ctypes as ct

import numpy as np

from . import _abi
from ._arrays import float_array, view
from .mdp_value_function import MDPValueFunction


class _SSPSolver:
    _prefix = ""

    def _setup(self, mdpObject, VInitial):
        self._lib = _abi.load()
        self.mdp = mdpObject
        self.mdpPtr = ct.pointer(self.mdp)
        v0 = np.zeros(self.mdp.n, np.float32) if VInitial is None else VInitial
        self._v0 = float_array(v0)
        self.VInitial = ct.cast(self._v0, _abi.c_float_p)
        if getattr(self._lib, self._prefix + "initialize_cpu")(self.mdpPtr, self) != 0:
            raise Exception("Failed to initialize the %s (CPU) algorithm." % self._name)

    def __del__(self):
        try:
            getattr(self._lib, self._prefix + "uninitialize_cpu")(self.mdpPtr, self)
        except Exception:
            pass

    def solve(self):
        policy = MDPValueFunction()
        if getattr(self._lib, self._prefix + "execute_cpu")(self.mdpPtr, self, policy) != 0:
            raise Exception("Failed to execute the CPU %s solver." % self._name)
        return policy

    def update(self):
        """One update step; returns the library's code (LAO*: 9 = NOVA_CONVERGED)."""
        rc = getattr(self._lib, self._prefix + "update_cpu")(self.mdpPtr, self)
        if rc not in (0, 9):
            raise Exception("%s update failed with code %d." % (self._name, rc))
        return rc

    def get_policy(self):
        policy = MDPValueFunction()
        if getattr(self._lib, self._prefix + "get_policy_cpu")(self.mdpPtr, self, policy) != 0:
            raise Exception("Failed to get the policy.")
        return policy

    def V_array(self):
        return view(self.V, self.mdp.n, np.float32)

    def pi_array(self):
        return view(self.pi, self.mdp.n, np.uint32)


class SSPLAOStarCPU(_SSPSolver, _abi.NovaSSPLAOStarCPU):
    """LAO* (src/mdp/algorithms/ssp_lao_star_cpu.cpp): expand the greedy partial solution from
    s0 and back it up in reverse depth-first order until no unexpanded state remains on it and
    a check pass changes no action with a residual below the MDP's epsilon, or the MDP's
    horizon of passes is used up."""
    _prefix = "ssp_lao_star_"
    _name = "LAO*"

    def __init__(self, mdpObject, VInitial=None):
        _abi.NovaSSPLAOStarCPU.__init__(self)
        self._setup(mdpObject, VInitial)


class SSPRTDPCPU(_SSPSolver, _abi.NovaSSPRTDPCPU):
    """RTDP (src/mdp/algorithms/ssp_rtdp_cpu.cpp): `trials` greedy trials from s0, each at
    most the MDP's horizon long, backing up every state it passes. Successors are drawn with
    the C library's rand(); seed it with srand() for a reproducible run."""
    _prefix = "ssp_rtdp_"
    _name = "RTDP"

    def __init__(self, mdpObject, VInitial=None, trials=1):
        _abi.NovaSSPRTDPCPU.__init__(self)
        self.trials = int(trials)
        self._setup(mdpObject, VInitial)
