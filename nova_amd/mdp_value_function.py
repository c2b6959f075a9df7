"""MDP value-function policy (mirrors python/nova/mdp_value_function.py:34-98)."""
import numpy as np

from . import _abi
from ._arrays import view


class MDPValueFunction(_abi.NovaMDPValueFunction):
    """V and pi of a solved MDP; the arrays are library-allocated and freed on deletion."""

    def __init__(self):
        super().__init__()
        self._lib = _abi.load()

    def __del__(self):
        try:
            self._lib.mdp_value_function_uninitialize(self)
        except Exception:
            pass

    def V_array(self):
        count = self.r if self.r > 0 else self.n
        return view(self.V, count, np.float32)

    def pi_array(self):
        count = self.r if self.r > 0 else self.n
        return view(self.pi, count, np.uint32)
