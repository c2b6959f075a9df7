"""Value iteration solvers (mirror python/nova/mdp_vi.py:38-203 of the reference).

The initial values default to Rmin / (1 - gamma), or 0 when gamma == 1 (mdp_vi.py:57-64).
"""
import ctypes as ct

import numpy as np

from . import _abi
from .mdp_value_function import MDPValueFunction
from ._arrays import float_array, view


def _default_v0(mdp):
    if mdp.gamma < 1.0:
        return np.full(mdp.n, mdp.Rmin / (1.0 - mdp.gamma), dtype=np.float32)
    return np.zeros(mdp.n, dtype=np.float32)


class MDPValueIterationCPU(_abi.NovaMDPVICPU):
    def __init__(self, mdpObject, Vinitial=None):
        super().__init__()
        self._lib = _abi.load()
        self.mdp = mdpObject
        self.mdpPtr = ct.pointer(self.mdp)
        self._v0 = float_array(_default_v0(self.mdp) if Vinitial is None else Vinitial)
        self.VInitial = ct.cast(self._v0, _abi.c_float_p)
        if self._lib.mdp_vi_initialize_cpu(self.mdpPtr, self) != 0:
            raise Exception("Failed to initialize the value iteration (CPU) algorithm.")

    def __del__(self):
        try:
            self._lib.mdp_vi_uninitialize_cpu(self.mdpPtr, self)
        except Exception:
            pass

    def solve(self):
        policy = MDPValueFunction()
        if self._lib.mdp_vi_execute_cpu(self.mdpPtr, self, policy) != 0:
            raise Exception("Failed to execute the CPU MDP solver.")
        return policy

    def update(self):
        if self._lib.mdp_vi_update_cpu(self.mdpPtr, self) != 0:
            raise Exception("Failed to run a Bellman update on the CPU.")

    def get_policy(self):
        policy = MDPValueFunction()
        if self._lib.mdp_vi_get_policy_cpu(self.mdpPtr, self, policy) != 0:
            raise Exception("Failed to get the policy.")
        return policy


class MDPValueIterationGPU(_abi.NovaMDPVIGPU):
    def __init__(self, mdpObject, numThreads=1024, Vinitial=None):
        super().__init__()
        self._lib = _abi.load()
        self.mdp = mdpObject
        self.mdpPtr = ct.pointer(self.mdp)
        self.numThreads = numThreads
        self._v0 = float_array(_default_v0(self.mdp) if Vinitial is None else Vinitial)
        self.VInitial = ct.cast(self._v0, _abi.c_float_p)
        if self._lib.mdp_vi_initialize_gpu(self.mdpPtr, self) != 0:
            raise Exception("Failed to initialize the value iteration (GPU) algorithm.")

    def __del__(self):
        try:
            self._lib.mdp_vi_uninitialize_gpu(self.mdpPtr, self)
        except Exception:
            pass

    def solve(self):
        policy = MDPValueFunction()
        if self._lib.mdp_vi_execute_gpu(self.mdpPtr, self, policy) != 0:
            raise Exception("Failed to execute the GPU MDP solver.")
        return policy

    def update(self):
        if self._lib.mdp_vi_update_gpu(self.mdpPtr, self) != 0:
            raise Exception("Failed to run a Bellman update on the GPU.")

    def get_policy(self):
        policy = MDPValueFunction()
        if self._lib.mdp_vi_get_policy_gpu(self.mdpPtr, self, policy) != 0:
            raise Exception("Failed to get the policy.")
        return policy

    def run(self, sweeps, d_V0=None):
        """Runs `sweeps` sweeps on the device (one graph); returns their device time in ms.
        d_V0: optional device pointer (int) of n floats to restart from."""
        ms = ct.c_float(0.0)
        rc = self._lib.mdp_vi_run_gpu(self.mdpPtr, self, d_V0, sweeps, ct.byref(ms))
        if rc != 0:
            raise Exception("mdp_vi_run_gpu failed with code %d" % rc)
        return ms.value

    def sweep_kernel(self):
        """0: one thread per state, 1: wave-tiled sweep, 2: successor-sorted tiles,
        3: successor chunks (V chunks in LDS, products through HBM)."""
        k = ct.c_uint(0)
        if self._lib.mdp_vi_sweep_kernel_gpu(self.mdpPtr, self, ct.byref(k)) != 0:
            raise Exception("mdp_vi_sweep_kernel_gpu failed")
        return k.value

    def device_values(self):
        """(V, pi) read back from the device without building a policy."""
        pol = self.get_policy()
        return pol.V_array(), pol.pi_array()


def batch_execute(solvers, policies=True):
    """Runs mdp_vi_batch_execute_gpu over MDPValueIterationGPU solvers (their MDPs resident on
    the device): every instance solved as its plain solve would be, all sweeps in one graph.
    Returns (list of (V, pi) or None, device ms)."""
    lib = _abi.load()
    count = len(solvers)
    MP = ct.POINTER(_abi.NovaMDP) * count
    VP = ct.POINTER(_abi.NovaMDPVIGPU) * count
    marr = MP(*[ct.cast(s.mdpPtr, ct.POINTER(_abi.NovaMDP)) for s in solvers])
    varr = VP(*[ct.cast(ct.pointer(s), ct.POINTER(_abi.NovaMDPVIGPU)) for s in solvers])
    pols = (_abi.NovaMDPValueFunction * count)() if policies else None
    ms = ct.c_float(0.0)
    rc = lib.mdp_vi_batch_execute_gpu(ct.cast(marr, ct.c_void_p), ct.cast(varr, ct.c_void_p),
                                      count, ct.cast(pols, ct.c_void_p) if pols else None,
                                      ct.byref(ms))
    if rc != 0:
        raise Exception("mdp_vi_batch_execute_gpu failed with code %d" % rc)
    out = None
    if pols is not None:
        out = []
        for k in range(count):
            p = pols[k]
            V = np.ctypeslib.as_array(p.V, (p.n,)).copy()
            pi = np.ctypeslib.as_array(p.pi, (p.n,)).copy()
            lib.mdp_value_function_uninitialize(ct.byref(p))
            out.append((V, pi))
    return out, ms.value
