"""Stochastic-shortest-path Bellman backups on the GPU (nova_amd_ssp_backup_gpu,
nova_amd_ssp_sweeps_gpu), through the C ABI, against the oracle.

* Sweeps: every Jacobi sweep (goals held at 0) equals the oracle's float sweep bit for bit,
  sweep after sweep, with f = |V| (LAO*'s backup, src/mdp/algorithms/ssp_lao_star_cpu.cpp:39-68)
  and f = V (RTDP's, ssp_rtdp_cpu.cpp:39-74). The oracle's backup is itself pinned bit for bit
  against both of nova's functions (tests/test_ssp.py).
* Convergence (models whose every state reaches a goal, so that V* is finite; in
  another_grid_world_ssp some states cannot): sweeps from V = 0 to a 1e-6
  change reach the FP64 optimal values (the oracle's orc_ssp_vi_f64) within 1e-4 relative,
  and the greedy actions are optimal (each chosen action's FP64 Q value within 1e-4 relative
  of the minimum).
* Listed states: the backups of a list with repeats equal the oracle's bit for bit; entries not
  listed, and listed indices past n, are left untouched.
Cases: nova's SSP grid worlds (tests/grid_world/domains/*_ssp.raw) and seeded random SSPs
with and without dead ends."""
import ctypes as ct
import os

import numpy as np
import pytest

from conftest import bits
from nova_amd import _abi
from nova_amd.mdp import MDP

pytestmark = pytest.mark.gpu

RAW = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "raw")
CASES = ["grid_world_ssp", "another_grid_world_ssp", "random", "random_dead_ends"]


def random_ssp(n, m, ns, seed, n_goals=3, n_dead=0):
    rng = np.random.default_rng(seed)
    S = np.stack([rng.choice(n, ns, replace=False) for _ in range(n * m)]).reshape(n, m, ns)
    T = rng.exponential(size=(n, m, ns))
    T = (T / T.sum(axis=2, keepdims=True)).astype(np.float32)
    R = rng.uniform(1, 2, (n, m)).astype(np.float32)
    S = S.astype(np.int32)
    goals = rng.choice(np.arange(1, n), n_goals, replace=False).astype(np.uint32)
    for g in goals:
        S[g, :, :] = -1
        S[g, :, 0] = g
        T[g, :, :] = 0
        T[g, :, 0] = 1
        R[g, :] = 0
    others = np.setdiff1d(np.arange(1, n), goals)
    for d in rng.choice(others, n_dead, replace=False):
        S[d, :, :] = -1
        S[d, :, 0] = d
        T[d, :, :] = 0
        T[d, :, 0] = 1
        R[d, :] = 5
    # a few ragged rows
    cut = rng.random((n, m)) < 0.2
    for s, a in zip(*np.nonzero(cut)):
        if S[s, a, 1] >= 0 and ns > 2:
            S[s, a, 2:] = -1
            T[s, a, 2:] = 0
            T[s, a, :2] /= T[s, a, :2].sum()
    return S, T, R, goals


def model(case):
    if case.startswith("random"):
        S, T, R, goals = random_ssp(500, 4, 5, 31, n_dead=6 if "dead" in case else 0)
    else:
        mdp = MDP().load(os.path.join(RAW, case + ".raw"), filetype="raw")
        S, T, R = mdp.arrays()
        goals = np.ctypeslib.as_array(mdp.goals, (mdp.ng,)).copy()
    mdp = MDP.from_arrays(S, T, R, gamma=1.0, horizon=1000, goals=goals)
    mdp.initialize_gpu()
    return mdp, S, T, R, goals


def sweeps(mdp, V0, count, eps, absV):
    import torch
    lib = _abi.load()
    n = V0.size
    d0 = torch.from_numpy(np.ascontiguousarray(V0, np.float32)).cuda()
    dV = torch.empty(n, dtype=torch.float32, device="cuda")
    dpi = torch.empty(n, dtype=torch.int32, device="cuda")
    done, res = ct.c_uint(0), ct.c_float(0)
    rc = lib.nova_amd_ssp_sweeps_gpu(mdp, int(absV), d0.data_ptr(), count, eps, dV.data_ptr(),
                                     dpi.data_ptr(), ct.byref(done), ct.byref(res))
    assert rc == 0
    return dV.cpu().numpy(), dpi.cpu().numpy().astype(np.uint32), done.value, res.value


@pytest.mark.parametrize("absV", [0, 1])
@pytest.mark.parametrize("case", CASES)
def test_ssp_sweeps_bitwise_vs_oracle(oracle, case, absV):
    mdp, S, T, R, goals = model(case)
    n = S.shape[0]
    rng = np.random.default_rng(n + absV)
    V = rng.uniform(-5, 20, n).astype(np.float32) if absV else np.zeros(n, np.float32)
    for k in range(12):
        Vg, pig, done, _ = sweeps(mdp, V, 1, 0.0, absV)
        Vo, pio = oracle.ssp_jacobi(S, T, R, V, goals=goals, use_abs=bool(absV))
        assert done == 1
        assert np.array_equal(bits(Vg), bits(Vo)), (case, k)
        assert np.array_equal(pig, pio), (case, k)
        V = Vg
    print("ssp %s (n=%d, |V|=%d): 12 sweeps bit-identical to the oracle" % (case, n, absV),
          flush=True)


@pytest.mark.parametrize("case", ["grid_world_ssp", "random"])
def test_ssp_sweeps_converge_to_fp64_optimum(oracle, case):
    mdp, S, T, R, goals = model(case)
    n, m, ns = S.shape
    V, pi, done, res = sweeps(mdp, np.zeros(n, np.float32), 100000, 1e-6, 0)
    Vs, pis, it, r64 = oracle.ssp_vi_f64(S, T, R, goals)
    assert res <= 1e-6
    scale = max(1.0, float(np.abs(Vs).max()))
    err = np.max(np.abs(V - Vs)) / scale
    # FP64 Q values of the greedy actions against the FP64 minimum
    Q = R.astype(np.float64).copy()
    for i in range(ns):
        sp = S[:, :, i]
        ok = sp >= 0
        Q += np.where(ok, T[:, :, i] * Vs[np.where(ok, sp, 0)], 0.0)
    qbest = Q.min(axis=1)
    qpi = Q[np.arange(n), pi]
    nongoal = np.setdiff1d(np.arange(n), goals)
    gap = np.max((qpi - qbest)[nongoal] / np.maximum(1.0, np.abs(qbest[nongoal])))
    print("ssp %s: %d sweeps to a %.2g change; |V - V*| / ||V*|| = %.3g; greedy gap %.3g"
          % (case, done, res, err, gap), flush=True)
    assert err <= 1e-4
    assert gap <= 1e-4


@pytest.mark.parametrize("absV", [0, 1])
def test_ssp_backup_listed_states(oracle, absV):
    import torch
    mdp, S, T, R, goals = model("random_dead_ends")
    n = S.shape[0]
    rng = np.random.default_rng(7 + absV)
    V = rng.uniform(-5, 20, n).astype(np.float32)
    states = rng.integers(0, n, 300).astype(np.uint32)
    states[:10] = states[10:20]                       # repeats
    lib = _abi.load()
    dV = torch.from_numpy(V).cuda()
    dVo = torch.full((n,), 123.0, dtype=torch.float32, device="cuda")
    dpi = torch.full((n,), 77, dtype=torch.int32, device="cuda")
    with_oob = np.concatenate([states, np.array([n, n + 5], np.uint32)])   # skipped
    dst = torch.from_numpy(with_oob.view(np.int32)).cuda()
    assert lib.nova_amd_ssp_backup_gpu(mdp, len(with_oob), dst.data_ptr(), absV, dV.data_ptr(),
                                       dVo.data_ptr(), dpi.data_ptr()) == 0
    Vg, pig = dVo.cpu().numpy(), dpi.cpu().numpy().astype(np.uint32)
    Vo, pio = oracle.ssp_jacobi(S, T, R, V, use_abs=bool(absV), states=states,
                                Vout=np.full(n, 123.0, np.float32),
                                pi=np.full(n, 77, np.uint32))
    assert np.array_equal(bits(Vg), bits(Vo)) and np.array_equal(pig, pio)
    untouched = np.setdiff1d(np.arange(n), states)
    assert np.all(Vg[untouched] == 123.0) and np.all(pig[untouched] == 77)
