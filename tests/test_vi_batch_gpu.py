"""Batched value iteration (mdp_vi_batch_execute_gpu): many MDPs of uneven size, every sweep
of every instance in one HIP graph. Each instance must equal the float oracle's run
(mdp_vi_cpu.cpp:36-63 per sweep) and its plain GPU solve bit for bit, whatever its size,
horizon (odd, even, 1), gamma, initial values and ragged successor lists."""
import numpy as np
import pytest

from conftest import bits
from nova_amd import synth
from nova_amd.mdp import MDP
from nova_amd.mdp_vi import MDPValueIterationGPU, batch_execute

pytestmark = pytest.mark.gpu


def uneven(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(5, 3000))
        m = int(rng.integers(1, 21))
        ns = int(rng.integers(1, min(n, 17) + 1))
        w = synth.mdp_arrays(n=n, m=m, ns=ns, seed=seed * 100 + k,
                             gamma=float(rng.choice([0.9, 0.95, 0.99, 1.0])))
        S = w["S"].copy()
        cut = rng.random((n, m)) < 0.25
        pos = rng.integers(0, ns, size=(n, m))
        for s, a in zip(*np.nonzero(cut)):
            S[s, a, pos[s, a]] = -1
        w["S"] = S
        w["horizon"] = int(rng.choice([1, 2, 7, 20, 33]))
        w["V0"] = rng.uniform(-3, 3, n).astype(np.float32)
        out.append(w)
    return out


def solvers(batch):
    out = []
    for w in batch:
        mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=w["horizon"])
        mdp.initialize_gpu()
        out.append(MDPValueIterationGPU(mdp, Vinitial=w.get("V0")))
    return out


def test_vi_batch_uneven_vs_oracle_and_plain(oracle):
    batch = uneven(14, 3)
    res, ms = batch_execute(solvers(batch))
    plain = [s.solve() for s in solvers(batch)]
    for k, (w, (V, pi), pol) in enumerate(zip(batch, res, plain)):
        Vo, pio = oracle.vi_run(w["S"], w["T"], w["R"], w["gamma"], w["horizon"], V0=w["V0"])
        assert np.array_equal(bits(V), bits(Vo)) and np.array_equal(pi, pio), k
        assert np.array_equal(bits(V), bits(pol.V_array())), k
        assert np.array_equal(pi, pol.pi_array()), k
    print("vi batch: 14 uneven MDPs equal the oracle and the plain solves (graph %.2f ms)" % ms,
          flush=True)


def test_vi_batch_256_equals_plain():
    batch = synth.mdp_batch_instances()
    res, ms = batch_execute(solvers(batch))
    differ = []
    for k, (s, (V, pi)) in enumerate(zip(solvers(batch), res)):
        pol = s.solve()
        if not (np.array_equal(bits(V), bits(pol.V_array())) and
                np.array_equal(pi, pol.pi_array())):
            differ.append(k)
    assert not differ, differ
    print("vi batch: 256 MDPs equal their plain solves (graph %.1f ms)" % ms, flush=True)


def test_vi_batch_rejects_wide_action_sets():
    w = synth.mdp_arrays(n=4, m=1025, ns=2, seed=1)
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=0.9, horizon=2)
    mdp.initialize_gpu()
    with pytest.raises(Exception):
        batch_execute([MDPValueIterationGPU(mdp)])
