"""The library's own host entry points (the *_cpu half of nova's ABI) against nova's CPU
sources compiled where they lie (oracle/_ref): bit for bit, on seeded inputs. CPU only."""
import ctypes as ct

import numpy as np
import pytest

from conftest import bits
from nova_amd import _abi, synth
from nova_amd.mdp import MDP
from nova_amd.mdp_vi import MDPValueIterationCPU
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_pbvi import POMDPPBVICPU
from nova_amd.pomdp_perseus import POMDPPerseusCPU

libc = ct.CDLL("libc.so.6")


def test_vi_cpu_bitwise(reflib):
    w = synth.mdp_arrays(n=2000, m=5, ns=9, seed=41, gamma=0.98)
    S = w["S"].copy()
    S[10, 2, 3] = -1
    S[99, 0, 0] = -1
    mdp = MDP.from_arrays(S, w["T"], w["R"], gamma=w["gamma"], horizon=17)
    v0 = np.linspace(-2, 2, 2000).astype(np.float32)
    pol = MDPValueIterationCPU(mdp, Vinitial=v0).solve()
    Vr, pir = reflib.vi_execute(S, w["T"], w["R"], w["gamma"], 17, V0=v0)
    assert np.array_equal(bits(pol.V_array()), bits(Vr))
    assert np.array_equal(pol.pi_array(), pir)


def test_pbvi_cpu_bitwise(reflib):
    P = synth.pomdp_arrays(n=40, m=3, z=4, ns=5, r=30, rz=6, seed=42, gamma=0.93, horizon=7)
    P["Z"] = P["Z"].copy()
    P["Z"][2, 4] = -1
    p = POMDP.from_dict(P)
    pol = POMDPPBVICPU(p, GammaInitial=P["Gamma0"]).solve()
    Gr, pir = reflib.pbvi_execute(P)
    assert np.array_equal(bits(pol.Gamma_array()), bits(Gr))
    assert np.array_equal(pol.pi_array(), pir)


def test_perseus_cpu_bitwise_seeded(reflib):
    """Both libraries draw beliefs with the process's libc rand(); seeding it identically
    before each run makes the randomized Perseus comparable bit for bit."""
    P = synth.pomdp_arrays(n=80, m=4, z=6, ns=5, r=60, rz=8, seed=44, gamma=0.9, horizon=30)
    p = POMDP.from_dict(P)
    libc.srand(1234)
    pol = POMDPPerseusCPU(p, GammaInitial=P["Gamma0"]).solve()
    G, pi = pol.Gamma_array(), pol.pi_array()

    pr, keep = reflib._pomdp(P)
    g0 = np.ascontiguousarray(P["Gamma0"], np.float32)
    per = _abi.NovaPOMDPPerseusCPU()
    per.GammaInitial = ct.cast(g0.ctypes.data, _abi.c_float_p)
    ref_pol = _abi.NovaPOMDPAlphaVectors()
    libc.srand(1234)
    assert reflib.lib.pomdp_perseus_execute_cpu(ct.byref(pr), ct.byref(per), ct.byref(ref_pol)) == 0
    rr = ref_pol.r
    Gr = np.ctypeslib.as_array(ref_pol.Gamma, (rr * P["S"].shape[0],)).reshape(rr, -1).copy()
    pir = np.ctypeslib.as_array(ref_pol.pi, (rr,)).copy()
    reflib.lib.pomdp_alpha_vectors_uninitialize(ct.byref(ref_pol))
    assert G.shape == Gr.shape
    assert np.array_equal(bits(G), bits(Gr)) and np.array_equal(pi, pir)


def test_belief_update_and_distinct_bitwise(reflib):
    P = synth.pomdp_arrays(n=20, m=3, z=3, ns=4, r=6, rz=5, seed=44)
    p = POMDP.from_dict(P)
    rng = np.random.default_rng(2)
    b = rng.dirichlet(np.ones(20)).astype(np.float32)
    for a in range(3):
        for o in range(3):
            rc, bp = p.belief_update(b, a, o)
            rc2, bp2 = reflib.belief_update(P, b, a, o)
            assert rc == rc2 and np.array_equal(bits(bp), bits(bp2))
    p.expand("distinct_beliefs")
    Z, B = p.beliefs()
    Z2, B2 = reflib.expand_distinct(P)
    assert np.array_equal(Z, Z2) and np.array_equal(bits(B), bits(B2))


def test_random_expansion_invariants():
    """pomdp_expand_random_cpu is seeded with time (as in the reference), so only invariants
    are checked: r grows by exactly k, every new belief is a distribution reached by one tau
    step from the set or from an earlier new belief."""
    P = synth.pomdp_arrays(n=12, m=2, z=3, ns=3, r=4, rz=4, seed=45, horizon=5)
    p = POMDP.from_dict(P)
    p.expand("random", numBeliefsToAdd=15)
    assert p.r == 4 + 15
    Z, B = p.beliefs()
    dense = np.zeros((p.r, 12))
    for i in range(p.r):
        k = Z[i] >= 0
        dense[i, Z[i][k]] = B[i][k]
    for i in range(4, p.r):
        assert abs(dense[i].sum() - 1) < 1e-5
        found = False
        for j in range(i):
            for a in range(2):
                for o in range(3):
                    rc, bp = p.belief_update(dense[j].astype(np.float32), a, o)
                    if rc == 0 and np.allclose(bp, dense[i], atol=1e-6):
                        found = True
                        break
                if found:
                    break
            if found:
                break
        assert found, "belief %d is not one tau step from an earlier belief" % i
