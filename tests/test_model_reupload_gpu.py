"""Model arrays uploaded again after a solver was set up, through the C ABI.

The reference's GPU kernels read the device model arrays on every call, so a re-upload between
calls takes effect. This library derives data from the model when a solver or cache is set up
(PBVI's O*T products, pair lists and dense operands; the VI successor-sorted layout; the
predecessor lists of the batched belief update). These tests re-upload T, whose new device copy
may land at the old address, and check that every path uses the new model:
  * PBVI (fast sparse, exact sparse, dense bf16, dense fp32): one update equals a fresh
    solver's on the new model, bit for bit;
  * VI with the sorted layout forced: the run equals the oracle on the new model;
  * batched belief updates: equal the oracle's float twin on the new model, bit for bit.
They also check that a solver keeps the arithmetic mode it was initialized with, and that a
solver refuses to run once the POMDP's sizes changed (its buffers were allocated for the old
ones)."""
import ctypes as ct

import numpy as np
import pytest

from conftest import bits
from nova_amd import _abi, synth
from nova_amd.mdp import MDP
from nova_amd.mdp_vi import MDPValueIterationGPU
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_pbvi import EXACT, FAST, POMDPPBVIGPU, set_mode

pytestmark = pytest.mark.gpu


def new_transitions(T, seed):
    """Other row-stochastic values on the same successor lists."""
    rng = np.random.default_rng(seed)
    T2 = np.where(T > 0, rng.uniform(0.1, 1.0, T.shape), 0.0)
    return (T2 / np.maximum(T2.sum(axis=2, keepdims=True), 1e-30)).astype(np.float32)


def reupload_T(p, T2, upload):
    np.ctypeslib.as_array(p.T, (T2.size,))[:] = T2.reshape(-1)
    assert upload(ct.byref(p)) == 0


def pomdp_instance(dense, seed):
    if dense:
        P = synth.pomdp_arrays(n=60, m=3, z=4, ns=3, r=150, rz=4, seed=seed, gamma=0.95,
                               horizon=2)
        rng = np.random.default_rng(seed)
        P["Z"] = np.tile(np.arange(60, dtype=np.int32), (150, 1))
        P["B"] = rng.dirichlet(np.ones(60), size=150).astype(np.float32)
        return P
    return synth.pomdp_arrays(n=80, m=3, z=4, ns=4, r=40, rz=6, seed=seed, gamma=0.95,
                              horizon=2)


def one_update(P, Gt):
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    s = POMDPPBVIGPU(p, GammaInitial=Gt)
    s.update()
    js, act = s.argmax_record()
    return js.copy(), act.copy(), s.get_policy().Gamma_array().copy()


@pytest.mark.parametrize("mode", [FAST, EXACT, 2, 3])
def test_pbvi_update_after_transitions_reupload(mode):
    P = pomdp_instance(dense=mode in (2, 3), seed=40 + mode)
    rng = np.random.default_rng(mode)
    Gt = (P["Gamma0"] + rng.uniform(-1, 1, P["Gamma0"].shape)).astype(np.float32)
    T2 = new_transitions(P["T"], seed=mode)
    lib = _abi.load()
    set_mode(mode)
    try:
        p = POMDP.from_dict(P)
        p.initialize_gpu()
        s = POMDPPBVIGPU(p, GammaInitial=Gt)
        reupload_T(p, T2, lib.pomdp_initialize_state_transitions_gpu)
        s.update()
        js, act = s.argmax_record()
        G = s.get_policy().Gamma_array().copy()
        P2 = dict(P)
        P2["T"] = T2
        jf, af, Gf = one_update(P2, Gt)
    finally:
        set_mode(FAST)
    assert np.array_equal(js, jf) and np.array_equal(act, af)
    assert np.array_equal(bits(G), bits(Gf))
    print("pbvi mode %d: update after a T re-upload equals a fresh solver" % mode, flush=True)


def test_pbvi_keeps_its_initialize_mode():
    """A solver initialized in fast mode with supports longer than exact mode's LDS staging
    (rz * ns > 512) keeps running in fast mode after the mode is switched to exact."""
    P = synth.pomdp_arrays(n=400, m=2, z=3, ns=8, r=30, rz=80, seed=3, gamma=0.9, horizon=2)
    Gt = P["Gamma0"].astype(np.float32)
    set_mode(FAST)
    try:
        p = POMDP.from_dict(P)
        p.initialize_gpu()
        s = POMDPPBVIGPU(p, GammaInitial=Gt)
        set_mode(EXACT)
        s.update()
        G = s.get_policy().Gamma_array().copy()
        set_mode(FAST)
        _, _, Gf = one_update(P, Gt)
    finally:
        set_mode(FAST)
    assert np.array_equal(bits(G), bits(Gf))


def test_pbvi_refuses_changed_sizes():
    """After the belief set grows (r changes), update and run return an error instead of
    running on buffers sized for the old r."""
    P = synth.pomdp_arrays(n=50, m=2, z=3, ns=3, r=20, rz=5, seed=8, gamma=0.9, horizon=2)
    set_mode(FAST)
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    s = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"].astype(np.float32))
    s.update()
    p.expand("random", numBeliefsToAdd=5)
    assert p.r > 20
    lib = _abi.load()
    assert lib.pomdp_pbvi_update_gpu(ct.byref(p), ct.byref(s)) != 0
    assert lib.pomdp_pbvi_run_gpu(ct.byref(p), ct.byref(s), 1, None) != 0


def test_vi_sorted_layout_after_transitions_reupload(oracle, monkeypatch):
    monkeypatch.setenv("NOVA_AMD_VI_KERNEL", "sorted")
    w = synth.mdp_arrays(n=600, m=4, ns=10, seed=12, gamma=0.95)
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=15)
    mdp.initialize_gpu()
    v0 = np.random.default_rng(1).uniform(-5, 5, 600).astype(np.float32)
    vi = MDPValueIterationGPU(mdp, numThreads=512, Vinitial=v0)
    assert vi.sweep_kernel() == 2
    T2 = new_transitions(w["T"], seed=12)
    reupload_T(mdp, T2, _abi.load().mdp_initialize_state_transitions_gpu)
    assert vi.sweep_kernel() != 2          # the layout holds the old T: not used
    vi.run(15)                             # sweeps without initializing again
    pol = vi.get_policy()
    Vo, pio = oracle.vi_run(w["S"], T2, w["R"], w["gamma"], 15, V0=v0)
    assert np.array_equal(bits(pol.V_array()), bits(Vo)) and np.array_equal(pol.pi_array(), pio)


def test_belief_updates_after_transitions_reupload(oracle):
    P = synth.pomdp_arrays(n=90, m=3, z=4, ns=5, r=4, rz=3, seed=21)
    rng = np.random.default_rng(21)
    k = 50
    B = rng.dirichlet(np.ones(90), size=k).astype(np.float32)
    a = rng.integers(0, 3, k).astype(np.uint32)
    o = rng.integers(0, 4, k).astype(np.uint32)
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    p.belief_update_batch(B, a, o)                    # builds the predecessor lists
    T2 = new_transitions(P["T"], seed=21)
    reupload_T(p, T2, _abi.load().pomdp_initialize_state_transitions_gpu)
    Bp, status = p.belief_update_batch(B, a, o)
    Bp, status = Bp.cpu().numpy(), status.cpu().numpy()
    P2 = dict(P)
    P2["T"] = T2
    for i in range(k):
        rc, bp = oracle.belief_update(P2, B[i], int(a[i]), int(o[i]))
        assert status[i] == rc and np.array_equal(bits(Bp[i]), bits(bp)), i
