"""Dense-belief PBVI scoring on the matrix cores (modes 2 = bf16 operands, 3 = fp32 operands,
both fp32 accumulation) against the float oracle from the same Gamma (one backup): with fp32
operands the chosen alphas must match except at near-ties (conftest.tie_rel with K = n) and
the beliefs' values agree within 1e-4 relative; bf16 operands are reported as a disagreement
rate and a value error, and must keep the beliefs' values within 1e-2 relative."""
import numpy as np
import pytest

from conftest import check_choices
from nova_amd import synth
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_pbvi import FAST, POMDPPBVIGPU, set_mode

pytestmark = pytest.mark.gpu


def dense_pomdp(n, m, z, ns, r, seed):
    P = synth.pomdp_arrays(n=n, m=m, z=z, ns=ns, r=r, rz=4, seed=seed, gamma=0.95, horizon=1)
    rng = np.random.default_rng(seed)
    P["Z"] = np.tile(np.arange(n, dtype=np.int32), (r, 1))
    P["B"] = rng.dirichlet(np.ones(n), size=r).astype(np.float32)
    return P


def sparse_values(P, G):
    return (G.astype(np.float64) @ P["B"].astype(np.float64).T).max(axis=0)


@pytest.mark.parametrize("mode,label", [(3, "fp32"), (2, "bf16")])
def test_dense_step_vs_oracle(oracle, mode, label):
    P = dense_pomdp(n=200, m=3, z=5, ns=4, r=300, seed=71)
    rng = np.random.default_rng(1)
    Gt = (P["Gamma0"] + rng.uniform(-1, 1, P["Gamma0"].shape)).astype(np.float32)
    ref = oracle.pbvi_backup(P, Gt)
    set_mode(mode)
    try:
        p = POMDP.from_dict(P)
        p.initialize_gpu()
        s = POMDPPBVIGPU(p, GammaInitial=Gt)
        s.update()
        js, act = s.argmax_record()
        G1 = s.get_policy().Gamma_array()
    finally:
        set_mode(FAST)
    jo = ref["jstar"]
    bad = js != jo
    scale = np.maximum(1.0, np.abs(ref["aval"]))[:, :, None]
    rate = bad.mean()
    v, vo = sparse_values(P, G1), sparse_values(P, ref["GammaPrime"])
    verr = np.max(np.abs(v - vo)) / max(1.0, np.max(np.abs(vo)))
    print("dense %s: alpha-choice disagreement %.4f, action disagreement %.4f, value error %.3g"
          % (label, rate, np.mean(act != ref["pi"].astype(np.int32)), verr))
    if mode == 3:
        # the score sums K = n products per (belief, action, observation) in the GEMM
        check_choices(js, jo, ref["gap"], scale, P["S"].shape[0], "dense fp32 alpha choices")
        assert verr < 1e-4
    else:
        # bf16 operands carry about 3 significant digits: reported, and held to 1e-2; the
        # measured value error at this shape is about 3e-4
        assert verr < 1e-2


def test_dense_bf16_blocked_equals_legacy(monkeypatch):
    """Dense bf16 PBVI through the chunked scaled-pair path and the k-blocked GEMM equals the
    row-major kernel that scales Gamma in registers, bit for bit: the next Gamma and every
    (belief, action, observation) argmax of one backup (ragged r; the 12 pairs in chunks of
    5, 5 and 2)."""
    P = dense_pomdp(n=150, m=3, z=4, ns=4, r=270, seed=5)
    rng = np.random.default_rng(2)
    Gt = (P["Gamma0"] + rng.uniform(-1, 1, P["Gamma0"].shape)).astype(np.float32)
    out = {}
    monkeypatch.setenv("NOVA_AMD_DENSE_CHUNK_PAIRS", "5")
    set_mode(2)
    try:
        for legacy in (False, True):
            if legacy:
                monkeypatch.setenv("NOVA_AMD_DENSE_KERNEL", "legacy")
            else:
                monkeypatch.delenv("NOVA_AMD_DENSE_KERNEL", raising=False)
            p = POMDP.from_dict(P)
            p.initialize_gpu()
            s = POMDPPBVIGPU(p, GammaInitial=Gt)
            s.update()
            js, act = s.argmax_record()
            out[legacy] = (js.copy(), act.copy(), s.get_policy().Gamma_array().copy())
            del s, p
    finally:
        set_mode(FAST)
        monkeypatch.delenv("NOVA_AMD_DENSE_KERNEL", raising=False)
    (j0, a0, G0), (j1, a1, G1) = out[False], out[True]
    assert np.array_equal(j0, j1) and np.array_equal(a0, a1)
    assert np.array_equal(G0.view(np.uint32), G1.view(np.uint32))
    print("dense bf16 blocked == legacy: r=270 n=150 m=3 z=4", flush=True)


@pytest.mark.parametrize("mode", [2, 3])
def test_dense_reinitialize_then_run_like_fresh(mode):
    """A repeated pomdp_pbvi_initialize_gpu on the same solver (its buffers, and the dense
    workspace, reallocated) followed by a graph run gives the Gamma and pi of a fresh solver:
    the run must not replay a graph captured with the released addresses."""
    import ctypes as ct
    from nova_amd import _abi
    P = dense_pomdp(n=90, m=2, z=3, ns=3, r=200, seed=9)
    Gt = P["Gamma0"].astype(np.float32)
    set_mode(mode)
    try:
        p = POMDP.from_dict(P)
        p.initialize_gpu()
        s = POMDPPBVIGPU(p, GammaInitial=Gt)
        s.run(3)                                   # captures and caches the run graph
        lib = _abi.load()
        assert lib.pomdp_pbvi_initialize_gpu(ct.byref(p), ct.byref(s)) == 0
        s.run(3)
        G1, pi1 = s.get_policy().Gamma_array().copy(), s.get_policy().pi_array().copy()
        f = POMDPPBVIGPU(p, GammaInitial=Gt)
        f.run(3)
        G2, pi2 = f.get_policy().Gamma_array().copy(), f.get_policy().pi_array().copy()
    finally:
        set_mode(FAST)
    assert np.array_equal(G1.view(np.uint32), G2.view(np.uint32)) and np.array_equal(pi1, pi2)
    print("dense mode %d: re-initialized solver equals a fresh one" % mode, flush=True)
