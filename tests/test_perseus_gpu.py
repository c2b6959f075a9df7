"""Perseus with its backups on the GPU (pomdp_perseus_execute_gpu) against nova's own CPU
Perseus (oracle/_ref), both drawing beliefs from this process's libc rand() after the same
srand(seed): bit for bit in exact mode, within float rounding of the values in fast mode."""
import ctypes as ct

import numpy as np
import pytest

from nova_amd import _abi, synth
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_pbvi import EXACT, FAST, set_mode

pytestmark = pytest.mark.gpu
libc = ct.CDLL("libc.so.6")


def run_ref(reflib, P, seed):
    pr, keep = reflib._pomdp(P)
    g0 = np.ascontiguousarray(P["Gamma0"], np.float32)
    per = _abi.NovaPOMDPPerseusCPU()
    per.GammaInitial = ct.cast(g0.ctypes.data, _abi.c_float_p)
    pol = _abi.NovaPOMDPAlphaVectors()
    libc.srand(seed)
    assert reflib.lib.pomdp_perseus_execute_cpu(ct.byref(pr), ct.byref(per), ct.byref(pol)) == 0
    n = P["S"].shape[0]
    G = np.ctypeslib.as_array(pol.Gamma, (pol.r * n,)).reshape(pol.r, n).copy()
    pi = np.ctypeslib.as_array(pol.pi, (pol.r,)).copy()
    reflib.lib.pomdp_alpha_vectors_uninitialize(ct.byref(pol))
    return G, pi


def run_gpu(lib, P, seed, mode):
    set_mode(mode)
    try:
        p = POMDP.from_dict(P)
        p.initialize_gpu()
        g0 = np.ascontiguousarray(P["Gamma0"], np.float32)
        per = _abi.NovaPOMDPPerseusCPU()
        per.GammaInitial = ct.cast(g0.ctypes.data, _abi.c_float_p)
        pol = _abi.NovaPOMDPAlphaVectors()
        libc.srand(seed)
        assert lib.pomdp_perseus_execute_gpu(ct.byref(p), ct.byref(per), ct.byref(pol)) == 0
        n = P["S"].shape[0]
        G = np.ctypeslib.as_array(pol.Gamma, (pol.r * n,)).reshape(pol.r, n).copy()
        pi = np.ctypeslib.as_array(pol.pi, (pol.r,)).copy()
        lib.pomdp_alpha_vectors_uninitialize(ct.byref(pol))
        return G, pi
    finally:
        set_mode(FAST)


@pytest.mark.parametrize("params", [dict(n=25, m=3, z=3, ns=4, r=20, rz=5, seed=43, horizon=40),
                                    dict(n=80, m=4, z=6, ns=5, r=60, rz=8, seed=44, horizon=30),
                                    dict(n=300, m=5, z=8, ns=6, r=200, rz=12, seed=45,
                                         horizon=25)])
def test_perseus_gpu_exact_bitwise(reflib, lib, params):
    P = synth.pomdp_arrays(gamma=0.9, **params)
    G, pi = run_ref(reflib, P, 99)
    Gg, pig = run_gpu(lib, P, 99, EXACT)
    print("perseus exact mode n=%d: %d alphas, bitwise %s" % (params["n"], G.shape[0],
          Gg.shape == G.shape and np.array_equal(Gg.view(np.uint32), G.view(np.uint32))))
    assert Gg.shape == G.shape
    assert np.array_equal(Gg.view(np.uint32), G.view(np.uint32)) and np.array_equal(pig, pi)


def test_perseus_gpu_fast_values(reflib, lib):
    P = synth.pomdp_arrays(n=80, m=4, z=6, ns=5, r=60, rz=8, seed=45, gamma=0.9, horizon=30)
    G, _ = run_ref(reflib, P, 7)
    Gg, _ = run_gpu(lib, P, 7, FAST)
    Z, B = P["Z"], P["B"]

    def vals(Gm):
        return np.array([np.max(Gm[:, Z[i]] @ B[i].astype(np.float64)) for i in range(Z.shape[0])])
    # different rounding can change which beliefs are drawn later, so compare the values the two
    # alpha sets give the belief set
    v, vg = vals(G), vals(Gg)
    err = np.max(np.abs(v - vg)) / max(1.0, np.max(np.abs(v)))
    print("perseus fast mode: max |V_gpu(b) - V_nova(b)| / ||V|| = %.3g" % err)
    assert err <= 1e-4
