"""nova_amd: nova's Bellman-backup path (MDP value iteration, POMDP PBVI/Perseus backups and
belief expansion) for one AMD Instinct MI355X, behind nova's unchanged C ABI.

The Python layer mirrors nova's own (python/nova/*.py): the same module and class names
(MDP, POMDP, MDPValueIterationCPU/GPU, POMDPPBVICPU/GPU, POMDPPerseusCPU, MDPValueFunction,
POMDPAlphaVectors) over ctypes bindings of nova_amd/lib/libnova.so.
"""
from ._abi import LIB_PATH, load  # noqa: F401

__all__ = ["LIB_PATH", "load"]
