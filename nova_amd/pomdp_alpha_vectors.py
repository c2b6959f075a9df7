"""Alpha-vector policy (mirrors python/nova/pomdp_alpha_vectors.py:34-98)."""
import ctypes as ct

import numpy as np

from . import _abi
from ._arrays import view


class POMDPAlphaVectors(_abi.NovaPOMDPAlphaVectors):
    def __init__(self):
        super().__init__()
        self._lib = _abi.load()

    def __del__(self):
        try:
            self._lib.pomdp_alpha_vectors_uninitialize(self)
        except Exception:
            pass

    def value_and_action(self, b):
        """(V(b), action) for a dense belief b (pomdp_alpha_vectors_value_and_action)."""
        b = np.ascontiguousarray(b, np.float32)
        Vb = ct.c_float()
        a = ct.c_uint()
        rc = self._lib.pomdp_alpha_vectors_value_and_action(
            self, b.ctypes.data_as(_abi.c_float_p), ct.byref(Vb), ct.byref(a))
        if rc != 0:
            raise Exception("Failed to compute the value and action.")
        return Vb.value, a.value

    def Gamma_array(self):
        return view(self.Gamma, self.r * self.n, np.float32).reshape(self.r, self.n)

    def pi_array(self):
        return view(self.pi, self.r, np.uint32)
