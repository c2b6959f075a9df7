"""Perseus solver (mirrors python/nova/pomdp_perseus.py of the reference)."""
import ctypes as ct

from . import _abi
from ._arrays import float_array
from .pomdp_alpha_vectors import POMDPAlphaVectors
from .pomdp_pbvi import _default_gamma0


class POMDPPerseusCPU(_abi.NovaPOMDPPerseusCPU):
    def __init__(self, pomdpObject, GammaInitial=None):
        super().__init__()
        self._lib = _abi.load()
        self.pomdp = pomdpObject
        self.pomdpPtr = ct.pointer(self.pomdp)
        g0 = _default_gamma0(self.pomdp) if GammaInitial is None else GammaInitial
        self._g0 = float_array(g0)
        self.GammaInitial = ct.cast(self._g0, _abi.c_float_p)
        if self._lib.pomdp_perseus_initialize_cpu(self.pomdpPtr, self) != 0:
            raise Exception("Failed to initialize the Perseus (CPU) algorithm.")

    def __del__(self):
        try:
            self._lib.pomdp_perseus_uninitialize_cpu(self.pomdpPtr, self)
        except Exception:
            pass

    def solve(self):
        policy = POMDPAlphaVectors()
        if self._lib.pomdp_perseus_execute_cpu(self.pomdpPtr, self, policy) != 0:
            raise Exception("Failed to execute the CPU Perseus solver.")
        return policy

    def solve_gpu(self):
        """pomdp_perseus_execute_gpu: the same randomized loop with its backups and V(b)
        updates on the device (the POMDP must be resident)."""
        policy = POMDPAlphaVectors()
        if self._lib.pomdp_perseus_execute_gpu(self.pomdpPtr, self, policy) != 0:
            raise Exception("Failed to execute the GPU Perseus solver.")
        return policy
