"""PBVI solvers (mirror python/nova/pomdp_pbvi.py:38-214 of the reference).

GammaInitial defaults to Rmin / (1 - gamma) in every entry, or 0 when gamma == 1
(pomdp_pbvi.py:57-67).
"""
import ctypes as ct

import numpy as np

from . import _abi
from ._arrays import float_array
from .pomdp_alpha_vectors import POMDPAlphaVectors

FAST, EXACT = 0, 1


def _default_gamma0(p):
    v = p.Rmin / (1.0 - p.gamma) if p.gamma < 1.0 else 0.0
    return np.full(p.r * p.n, v, dtype=np.float32)


def set_mode(mode):
    """GPU backup arithmetic: FAST (default) or EXACT (nova's rounding, bit for bit).
    Returns the previous mode."""
    return _abi.load().nova_amd_pbvi_mode(int(mode))


class POMDPPBVICPU(_abi.NovaPOMDPPBVICPU):
    def __init__(self, pomdpObject, GammaInitial=None):
        super().__init__()
        self._lib = _abi.load()
        self.pomdp = pomdpObject
        self.pomdpPtr = ct.pointer(self.pomdp)
        g0 = _default_gamma0(self.pomdp) if GammaInitial is None else GammaInitial
        self._g0 = float_array(g0)
        self.GammaInitial = ct.cast(self._g0, _abi.c_float_p)
        if self._lib.pomdp_pbvi_initialize_cpu(self.pomdpPtr, self) != 0:
            raise Exception("Failed to initialize the PBVI (CPU) algorithm.")

    def __del__(self):
        try:
            self._lib.pomdp_pbvi_uninitialize_cpu(self.pomdpPtr, self)
        except Exception:
            pass

    def solve(self):
        policy = POMDPAlphaVectors()
        if self._lib.pomdp_pbvi_execute_cpu(self.pomdpPtr, self, policy) != 0:
            raise Exception("Failed to execute the CPU PBVI solver.")
        return policy

    def update(self):
        if self._lib.pomdp_pbvi_update_cpu(self.pomdpPtr, self) != 0:
            raise Exception("PBVI (CPU) update failed.")

    def get_policy(self):
        policy = POMDPAlphaVectors()
        if self._lib.pomdp_pbvi_get_policy_cpu(self.pomdpPtr, self, policy) != 0:
            raise Exception("Failed to get the policy.")
        return policy


class POMDPPBVIGPU(_abi.NovaPOMDPPBVIGPU):
    def __init__(self, pomdpObject, numThreads=1024, GammaInitial=None):
        super().__init__()
        self._lib = _abi.load()
        self.pomdp = pomdpObject
        self.pomdpPtr = ct.pointer(self.pomdp)
        self.numThreads = numThreads
        g0 = _default_gamma0(self.pomdp) if GammaInitial is None else GammaInitial
        self._g0 = float_array(g0)
        self.GammaInitial = ct.cast(self._g0, _abi.c_float_p)
        if self._lib.pomdp_pbvi_initialize_gpu(self.pomdpPtr, self) != 0:
            raise Exception("Failed to initialize the PBVI (GPU) algorithm.")

    def __del__(self):
        try:
            self._lib.pomdp_pbvi_uninitialize_gpu(self.pomdpPtr, self)
        except Exception:
            pass

    def solve(self):
        policy = POMDPAlphaVectors()
        if self._lib.pomdp_pbvi_execute_gpu(self.pomdpPtr, self, policy) != 0:
            raise Exception("Failed to execute the GPU PBVI solver.")
        return policy

    def update(self):
        if self._lib.pomdp_pbvi_update_gpu(self.pomdpPtr, self) != 0:
            raise Exception("PBVI (GPU) update failed.")

    def run(self, iterations):
        """Runs iterations on the device as one graph; returns their device time in ms."""
        ms = ct.c_float(0.0)
        rc = self._lib.pomdp_pbvi_run_gpu(self.pomdpPtr, self, iterations, ct.byref(ms))
        if rc != 0:
            raise Exception("pomdp_pbvi_run_gpu failed with code %d" % rc)
        return ms.value

    def scoring_variant(self):
        """1, 2, 4: alphas per lane on the vector ALUs; 8, 16, 32: fp32 matrix cores (4 waves,
        8 waves, 8 waves with a deeper prefetch); 64: fp32 matrix cores with 16 observation
        rows (z <= 16); 128: fp32 matrix cores, every alpha tile in one workgroup (rz * ns <=
        512); 256, 512: the 4- and 8-wave matrix-core kernels with operands requested three
        trips ahead; 1024, 2048: the same for the 16-row and the all-tiles kernels; 4096: the
        16-row kernel sweeping every alpha tile in one workgroup; 0: dense GEMM."""
        v = ct.c_uint(0)
        if self._lib.pomdp_pbvi_scoring_variant_gpu(self, ct.byref(v)) != 0:
            raise Exception("pomdp_pbvi_scoring_variant_gpu failed")
        return v.value

    def profile(self):
        """One iteration; returns the device ms of its (transpose, score, action, materialize)."""
        ms = (ct.c_float * 4)()
        if self._lib.pomdp_pbvi_profile_gpu(self.pomdpPtr, self, ms) != 0:
            raise Exception("pomdp_pbvi_profile_gpu failed")
        return list(ms)

    def get_policy(self):
        policy = POMDPAlphaVectors()
        if self._lib.pomdp_pbvi_get_policy_gpu(self.pomdpPtr, self, policy) != 0:
            raise Exception("Failed to get the policy.")
        return policy

    def argmax_record(self):
        """(jstar[r,m,z], actions[r]) of the last iteration."""
        p = self.pomdp
        js = np.empty((p.r, p.m, p.z), np.int32)
        act = np.empty(p.r, np.int32)
        rc = self._lib.pomdp_pbvi_argmax_gpu(self.pomdpPtr, self,
                                             js.ctypes.data_as(_abi.c_int_p),
                                             act.ctypes.data_as(_abi.c_uint_p))
        if rc != 0:
            raise Exception("pomdp_pbvi_argmax_gpu failed")
        return js, act


def batch_execute(solvers, policies=True):
    """Runs pomdp_pbvi_batch_execute_gpu over POMDPPBVIGPU solvers (their POMDPs resident on
    the device). Returns (policies or None, device ms)."""
    lib = _abi.load()
    count = len(solvers)
    PP = ct.POINTER(_abi.NovaPOMDP) * count
    SP = ct.POINTER(_abi.NovaPOMDPPBVIGPU) * count
    parr = PP(*[ct.cast(ct.pointer(s.pomdp), ct.POINTER(_abi.NovaPOMDP)) for s in solvers])
    sarr = SP(*[ct.cast(ct.pointer(s), ct.POINTER(_abi.NovaPOMDPPBVIGPU)) for s in solvers])
    pols = (_abi.NovaPOMDPAlphaVectors * count)() if policies else None
    ms = ct.c_float(0.0)
    rc = lib.pomdp_pbvi_batch_execute_gpu(ct.cast(parr, ct.c_void_p), ct.cast(sarr, ct.c_void_p),
                                          count, ct.cast(pols, ct.c_void_p) if pols else None,
                                          ct.byref(ms))
    if rc != 0:
        raise Exception("pomdp_pbvi_batch_execute_gpu failed with code %d" % rc)
    out = None
    if pols is not None:
        out = []
        for k in range(count):
            p = pols[k]
            G = np.ctypeslib.as_array(p.Gamma, (p.r * p.n,)).reshape(p.r, p.n).copy()
            pi = np.ctypeslib.as_array(p.pi, (p.r,)).copy()
            lib.pomdp_alpha_vectors_uninitialize(ct.byref(p))
            out.append((G, pi))
    return out, ms.value
