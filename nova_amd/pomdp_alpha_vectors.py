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

    PRECISIONS = {"exact": 0, "fp32": 1, "bf16": 2}

    def value_and_action_batch(self, B, precision="exact", return_alpha=False):
        """V(b) and the action for every row of B (k x n dense beliefs) on the GPU through
        nova_amd_alpha_vectors_value_and_action_gpu. precision: "exact" (equal to
        value_and_action bit for bit), "fp32" or "bf16" (matrix cores). B may be a numpy array
        or a torch tensor; returns torch tensors on the GPU (V, actions[, alpha indices])."""
        import torch
        if precision not in self.PRECISIONS:
            raise ValueError("precision must be one of %s" % sorted(self.PRECISIONS))
        dev = torch.device("cuda")
        Bt = torch.as_tensor(B, dtype=torch.float32).to(dev).contiguous()
        k = Bt.shape[0]
        if Bt.dim() != 2 or Bt.shape[1] != self.n:
            raise ValueError("B must be k x n")
        G = torch.from_numpy(self.Gamma_array()).to(dev)
        pi = torch.from_numpy(self.pi_array().astype(np.int32)).to(dev)
        V = torch.empty(k, dtype=torch.float32, device=dev)
        act = torch.empty(k, dtype=torch.int32, device=dev)
        j = torch.empty(k, dtype=torch.int32, device=dev) if return_alpha else None
        torch.cuda.synchronize()
        rc = self._lib.nova_amd_alpha_vectors_value_and_action_gpu(
            self.n, self.r, G.data_ptr(), pi.data_ptr(), k, Bt.data_ptr(),
            self.PRECISIONS[precision], V.data_ptr(), act.data_ptr(),
            j.data_ptr() if j is not None else None)
        if rc != 0:
            raise Exception("nova_amd_alpha_vectors_value_and_action_gpu failed with code %d"
                            % rc)
        return (V, act, j) if return_alpha else (V, act)

    def Gamma_array(self):
        return view(self.Gamma, self.r * self.n, np.float32).reshape(self.r, self.n)

    def pi_array(self):
        return view(self.pi, self.r, np.uint32)
