"""The POMDP model object (mirrors python/nova/pomdp.py:40-402 of the reference)."""
import ctypes as ct

import numpy as np

from . import _abi
from ._arrays import fill, view


class POMDP(_abi.NovaPOMDP):
    """A POMDP whose arrays live in library-owned memory (new[] in pomdp_initialize_cpu)."""

    def __init__(self):
        super().__init__()
        self.gamma = 0.9
        self.horizon = 1
        self.Rmin = None
        self.Rmax = None
        self.cpuIsInitialized = False
        self.gpuIsInitialized = False
        self._lib = _abi.load()

    def __del__(self):
        try:
            self.uninitialize_gpu()
            self.uninitialize()
        except Exception:
            pass

    def initialize(self, n, ns, m, z, r, rz, gamma, horizon):
        if self.cpuIsInitialized:
            return
        if self._lib.pomdp_initialize_cpu(self, n, ns, m, z, r, rz, gamma, horizon) != 0:
            raise Exception("Failed to initialize the POMDP.")
        self.cpuIsInitialized = True

    def uninitialize(self):
        if not self.cpuIsInitialized:
            return
        if self._lib.pomdp_uninitialize_cpu(self) != 0:
            raise Exception("Failed to uninitialize the POMDP.")
        self.cpuIsInitialized = False

    def initialize_gpu(self):
        if self.gpuIsInitialized:
            return
        if self._lib.pomdp_initialize_gpu(self) != 0:
            raise Exception("Failed to initialize the GPU variables for the POMDP.")
        self.gpuIsInitialized = True

    def _refresh_beliefs_gpu(self):
        """Re-uploads Z and B after the host belief set changed (r or rz may have grown), so
        that device kernels never read the old arrays with the new sizes."""
        if not self.gpuIsInitialized or self.r == 0:
            return
        if (self._lib.pomdp_initialize_nonzero_beliefs_gpu(self) != 0 or
                self._lib.pomdp_initialize_belief_points_gpu(self) != 0):
            raise Exception("Failed to refresh the device copy of the belief set.")

    def uninitialize_gpu(self):
        if not self.gpuIsInitialized:
            return
        if self._lib.pomdp_uninitialize_gpu(self) != 0:
            raise Exception("Failed to uninitialize the GPU variables for the POMDP.")
        self.gpuIsInitialized = False

    @classmethod
    def from_arrays(cls, S, T, O, R, Z, B, gamma, horizon):
        """S[n,m,ns] int32 (-1 ends a list), T[n,m,ns], O[m,n,z], R[n,m], Z[r,rz] int32,
        B[r,rz]."""
        S = np.asarray(S)
        n, m, ns = S.shape
        z = np.asarray(O).shape[2]
        r, rz = np.asarray(Z).shape
        p = cls()
        p.initialize(n, ns, m, z, r, rz, float(gamma), int(horizon))
        fill(p.S, S, np.int32)
        fill(p.T, T, np.float32)
        fill(p.O, O, np.float32)
        fill(p.R, R, np.float32)
        fill(p.Z, Z, np.int32)
        fill(p.B, B, np.float32)
        R32 = np.asarray(R, np.float32)
        p.Rmin = float(R32.min())
        p.Rmax = float(R32.max())
        return p

    def load(self, filename, filetype="cassandra", scalarize=lambda x: x[0]):
        """Loads a POMDP file (python/nova/pomdp.py:149-200): 'cassandra' (.pomdp text) or
        'raw' (nova's CSV layout); scalarize picks one reward from a raw file's k rewards.
        Replaces this object's model; the initial belief of the file is belief row 0."""
        from .file_loader import FileLoader
        if filetype not in ("cassandra", "raw"):
            raise ValueError("Invalid file type '%s'." % filetype)
        f = FileLoader()
        if filetype == "cassandra":
            f.load_cassandra(filename)
        else:
            f.load_raw_pomdp(filename, scalarize)
        if f.Z is None:
            raise ValueError("'%s' gives no initial belief" % filename)
        self.uninitialize_gpu()
        self.uninitialize()
        self.initialize(f.n, f.ns, f.m, f.z, f.r, f.rz, f.gamma, f.horizon)
        fill(self.S, f.S, np.int32)
        fill(self.T, f.T, np.float32)
        fill(self.O, f.O, np.float32)
        fill(self.R, f.R, np.float32)
        fill(self.Z, f.Z, np.int32)
        fill(self.B, f.B, np.float32)
        self.Rmin, self.Rmax = f.Rmin, f.Rmax
        return self

    def arrays(self):
        """dict(S, T, O, R, Z, B, gamma, horizon) as numpy copies in nova's layouts."""
        n, m, ns, z, r, rz = self.n, self.m, self.ns, self.z, self.r, self.rz
        return dict(S=view(self.S, n * m * ns, np.int32).reshape(n, m, ns),
                    T=view(self.T, n * m * ns, np.float32).reshape(n, m, ns),
                    O=view(self.O, m * n * z, np.float32).reshape(m, n, z),
                    R=view(self.R, n * m, np.float32).reshape(n, m),
                    Z=view(self.Z, r * rz, np.int32).reshape(r, rz),
                    B=view(self.B, r * rz, np.float32).reshape(r, rz),
                    gamma=np.float32(self.gamma), horizon=int(self.horizon))

    @classmethod
    def from_dict(cls, P):
        return cls.from_arrays(P["S"], P["T"], P["O"], P["R"], P["Z"], P["B"], P["gamma"],
                               P["horizon"])

    def beliefs(self):
        """(Z, B) as numpy copies shaped [r, rz]."""
        return (view(self.Z, self.r * self.rz, np.int32).reshape(self.r, self.rz),
                view(self.B, self.r * self.rz, np.float32).reshape(self.r, self.rz))

    def expand(self, method="random", numBeliefsToAdd=1000, pemaPolicy=None, gpu=False,
               seed=None, pemaAlgorithm=None):
        """Expands the belief set (python/nova/pomdp.py:202-229). With gpu=True the random
        trajectories (pomdp_expand_random_gpu, or its seeded variant when seed is given) or the
        distinct-belief scoring run on the device; the POMDP must then be resident. When the
        POMDP is resident, the device copy of Z/B is refreshed after the host set grows."""
        if gpu and method == "random":
            nnz = ct.c_uint(0)
            Bnew = _abi.c_float_p()
            if seed is None:
                rc = self._lib.pomdp_expand_random_gpu(self, 1024, numBeliefsToAdd,
                                                       ct.byref(nnz), ct.byref(Bnew))
            else:
                rc = self._lib.pomdp_expand_random_seeded_gpu(self, seed, numBeliefsToAdd,
                                                              ct.byref(nnz), ct.byref(Bnew))
            if rc == 0:
                rc = self._lib.pomdp_add_new_raw_beliefs_cpu(self, numBeliefsToAdd, Bnew)
            if Bnew:
                self._lib.nova_amd_free_floats(Bnew)
        elif gpu and method == "distinct_beliefs":
            rc = self._lib.pomdp_expand_distinct_beliefs_gpu(self)
        elif gpu and method == "pema":
            if pemaPolicy is None:
                pemaPolicy = pemaAlgorithm.solve()
            rc = self._lib.pomdp_expand_pema_gpu(self, ct.byref(pemaPolicy))
        elif gpu:
            raise Exception("Expansion '%s' has no GPU variant." % method)
        elif seed is not None:
            raise Exception("A seed applies to the GPU random expansion only (the host one "
                            "seeds with srand(time) as nova does).")
        elif method == "random":
            rc = self._lib.pomdp_expand_random_cpu(self, numBeliefsToAdd)
        elif method == "distinct_beliefs":
            rc = self._lib.pomdp_expand_distinct_beliefs_cpu(self)
        elif method == "pema":
            if pemaPolicy is None:
                pemaPolicy = pemaAlgorithm.solve()      # python/nova/pomdp.py:226-229
            rc = self._lib.pomdp_expand_pema_cpu(self, ct.byref(pemaPolicy))
        else:
            raise Exception("Failed to expand. Method '%s' is not defined." % method)
        if rc != 0:
            raise Exception("Belief expansion '%s' failed with code %d." % (method, rc))
        self._refresh_beliefs_gpu()

    def sigma_approximate(self, numDesiredNonZeroValues, gpu=False):
        """Keeps each belief's numDesiredNonZeroValues largest entries, renormalised
        (pomdp_sigma_cpu, or pomdp_sigma_gpu on the device copy with gpu=True); returns sigma,
        the smallest kept mass. Both the host and the device copies of Z/B end up updated."""
        sig = ct.c_float()
        if gpu:
            rc = self._lib.pomdp_sigma_gpu(self, int(numDesiredNonZeroValues), ct.byref(sig))
            if rc != 0:
                raise Exception("pomdp_sigma_gpu failed with code %d" % rc)
            return sig.value
        rc = self._lib.pomdp_sigma_cpu(self, int(numDesiredNonZeroValues), ct.byref(sig))
        if rc != 0:
            raise Exception("pomdp_sigma_cpu failed with code %d" % rc)
        self._refresh_beliefs_gpu()
        return sig.value

    def belief_update_batch(self, B, actions, observations):
        """tau(b, a, o) for every row of B (k x n) with its action and observation, on the GPU
        through nova_amd_belief_update_batch_gpu (bit for bit pomdp_belief_update_cpu). The
        POMDP must be resident (initialize_gpu). Returns torch tensors on the GPU: (B', status)
        with status 0, 10 (impossible observation: the row is left unnormalised) or 1."""
        import torch
        dev = torch.device("cuda")
        Bt = torch.as_tensor(B, dtype=torch.float32).to(dev).contiguous()
        k = Bt.shape[0]
        if Bt.dim() != 2 or Bt.shape[1] != self.n:
            raise ValueError("B must be k x n")
        a = torch.as_tensor(np.asarray(actions, np.int32)).to(dev).contiguous()
        o = torch.as_tensor(np.asarray(observations, np.int32)).to(dev).contiguous()
        if a.numel() != k or o.numel() != k:
            raise ValueError("one action and one observation per belief")
        Bp = torch.empty_like(Bt)
        status = torch.empty(k, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        rc = self._lib.nova_amd_belief_update_batch_gpu(self, k, Bt.data_ptr(), a.data_ptr(),
                                                       o.data_ptr(), Bp.data_ptr(),
                                                       status.data_ptr())
        if rc != 0:
            raise Exception("nova_amd_belief_update_batch_gpu failed with code %d" % rc)
        return Bp, status

    def belief_update(self, b, a, o):
        """tau(b, a, o) through pomdp_belief_update_cpu; returns (code, b')."""
        b = np.ascontiguousarray(b, np.float32)
        bp = _abi.c_float_p()
        rc = self._lib.pomdp_belief_update_cpu(self, b.ctypes.data_as(_abi.c_float_p), a, o,
                                               ct.byref(bp))
        out = view(bp, self.n, np.float32) if bp else None
        if bp:
            # the library allocated bp with new[]; nova's ABI has no matching free, so keep the
            # convention of its Python layer and let it be reclaimed by the process.
            pass
        return rc, out
