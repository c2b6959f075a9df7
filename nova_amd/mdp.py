"""The MDP model object (mirrors python/nova/mdp.py:40-305 of the reference)."""
import ctypes as ct

import numpy as np

from . import _abi
from ._arrays import fill, view


class MDP(_abi.NovaMDP):
    """An MDP whose arrays live in library-owned memory (new[] in mdp_initialize_cpu)."""

    def __init__(self):
        super().__init__()
        self.gamma = 0.9
        self.horizon = 1
        self.epsilon = 0.01
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

    def initialize(self, n, ns, m, gamma, horizon, epsilon=0.01, s0=0, ng=0):
        """Allocates S (= -1), T, R (= 0) and goals through mdp_initialize_cpu."""
        if self.cpuIsInitialized:
            return
        if self._lib.mdp_initialize_cpu(self, n, ns, m, gamma, horizon, epsilon, s0, ng) != 0:
            raise Exception("Failed to initialize the MDP.")
        self.cpuIsInitialized = True

    def uninitialize(self):
        if not self.cpuIsInitialized:
            return
        if self._lib.mdp_uninitialize_cpu(self) != 0:
            raise Exception("Failed to uninitialize the MDP.")
        self.cpuIsInitialized = False

    def initialize_gpu(self):
        """Uploads S, T, R (and goals) into HBM (python/nova/mdp.py:119-136)."""
        if self.gpuIsInitialized:
            return
        rc = self._lib.mdp_initialize_successors_gpu(self)
        rc += self._lib.mdp_initialize_state_transitions_gpu(self)
        rc += self._lib.mdp_initialize_rewards_gpu(self)
        if self.ng > 0:
            rc += self._lib.mdp_initialize_goals_gpu(self)
        if rc != 0:
            raise Exception("Failed to initialize the GPU variables for the MDP.")
        self.gpuIsInitialized = True

    def uninitialize_gpu(self):
        if not self.gpuIsInitialized:
            return
        rc = self._lib.mdp_uninitialize_successors_gpu(self)
        rc += self._lib.mdp_uninitialize_state_transitions_gpu(self)
        rc += self._lib.mdp_uninitialize_rewards_gpu(self)
        if self.ng > 0:
            rc += self._lib.mdp_uninitialize_goals_gpu(self)
        if rc != 0:
            raise Exception("Failed to uninitialize the GPU variables for the MDP.")
        self.gpuIsInitialized = False

    @classmethod
    def from_arrays(cls, S, T, R, gamma, horizon, epsilon=0.01, s0=0, goals=None):
        """Builds an MDP from S[n,m,ns] (int32, -1 ends a list), T[n,m,ns], R[n,m]."""
        S = np.asarray(S)
        n, m, ns = S.shape
        goals = np.zeros(0, np.uint32) if goals is None else np.asarray(goals, np.uint32)
        mdp = cls()
        mdp.initialize(n, ns, m, gamma, horizon, epsilon, s0, goals.size)
        fill(mdp.S, S, np.int32)
        fill(mdp.T, T, np.float32)
        fill(mdp.R, R, np.float32)
        if goals.size:
            fill(mdp.goals, goals, np.uint32)
        R32 = np.asarray(R, np.float32)
        mdp.Rmin = float(R32.min())
        mdp.Rmax = float(R32.max())
        return mdp

    def load(self, filename, filetype="cassandra", scalarize=lambda x: x[0]):
        """Loads an MDP file (python/nova/mdp.py:157-200): 'cassandra' (.pomdp text without
        observations) or 'raw' (nova's CSV layout with goals). Replaces this object's model."""
        from .file_loader import FileLoader
        if filetype not in ("cassandra", "raw"):
            raise ValueError("Invalid file type '%s'." % filetype)
        f = FileLoader()
        if filetype == "cassandra":
            f.load_cassandra(filename)
        else:
            f.load_raw_mdp(filename, scalarize)
        goals = np.zeros(0, np.uint32) if f.goals is None else np.asarray(f.goals, np.uint32)
        self.uninitialize_gpu()
        self.uninitialize()
        self.initialize(f.n, f.ns, f.m, f.gamma, f.horizon, f.epsilon, f.s0, goals.size)
        fill(self.S, f.S, np.int32)
        fill(self.T, f.T, np.float32)
        fill(self.R, f.R, np.float32)
        if goals.size:
            fill(self.goals, goals, np.uint32)
        self.Rmin, self.Rmax = f.Rmin, f.Rmax
        return self

    def arrays(self):
        """(S, T, R) as numpy copies shaped [n,m,ns], [n,m,ns], [n,m]."""
        n, m, ns = self.n, self.m, self.ns
        return (view(self.S, n * m * ns, np.int32).reshape(n, m, ns),
                view(self.T, n * m * ns, np.float32).reshape(n, m, ns),
                view(self.R, n * m, np.float32).reshape(n, m))

    def pointer(self):
        return ct.pointer(self)
