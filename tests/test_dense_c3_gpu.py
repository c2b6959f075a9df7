"""Dense-belief PBVI at its full size (n = 20,000, m = 16, z = 32, ns = 8, r = 4,096 beliefs
with full support, gamma = 0.95): one backup from a perturbed Gamma with fp32 operands on the
matrix cores (mode 3), checked on sampled beliefs against an FP64 evaluation written from the
mathematics: for belief b and pair (a, o), u_ao(s') = gamma * O[a,s',o] * sum_s b(s) T[s,a,s']
(one sparse pass), the scores are Gamma u_ao (one GEMV per pair), the alpha choice is the first
maximiser, and the action value is b.R_a + sum_o max_j score. Alpha and action choices must
match except at near-ties (conftest.tie_rel with K = n), and the value of the new alpha at b
within 1e-4 relative. The bf16-operand mode (mode 2) runs the same backup and its disagreement
with fp32 is reported beside it."""
import numpy as np
import pytest

from conftest import check_choices
from nova_amd import synth
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_pbvi import FAST, POMDPPBVIGPU, set_mode

pytestmark = pytest.mark.gpu


def _fp64_backup(P, G64, i):
    """FP64 scores of belief i: (best alpha per (a, o), best-minus-runner-up gap, best score,
    action values)."""
    S, T, O, R = P["S"], P["T"], P["O"], P["R"]
    n, m, ns = S.shape
    z = O.shape[2]
    gamma = float(P["gamma"])
    b = P["B"][i].astype(np.float64)
    jstar = np.empty((m, z), np.int64)
    gap = np.empty((m, z))
    best = np.empty((m, z))
    aval = np.empty(m)
    for a in range(m):
        w = np.bincount(S[:, a, :].ravel(), weights=(b[:, None] * T[:, a, :]).ravel(),
                        minlength=n)
        U = gamma * w[:, None] * O[a].astype(np.float64)          # n x z
        sc = G64 @ U                                              # r x z
        j = np.argmax(sc, axis=0)
        top = sc[j, np.arange(z)]
        sc2 = sc.copy()
        sc2[j, np.arange(z)] = -np.inf
        jstar[a], best[a], gap[a] = j, top, top - sc2.max(axis=0)
        aval[a] = b @ R[:, a].astype(np.float64) + top.sum()
    return jstar, gap, best, aval


def test_dense_full_size_fp32_vs_fp64():
    base = synth.workload("pbvi_n20000", rz=4)
    P = synth.dense_beliefs(base, seed=8)
    del base
    r, n = P["Gamma0"].shape
    m, z = P["R"].shape[1], P["O"].shape[2]
    rng = np.random.default_rng(9)
    G = (P["Gamma0"] + rng.uniform(-1, 1, P["Gamma0"].shape)).astype(np.float32)
    out = {}
    for mode, label in ((3, "fp32"), (2, "bf16")):
        set_mode(mode)
        try:
            p = POMDP.from_dict(P)
            p.initialize_gpu()
            s = POMDPPBVIGPU(p, GammaInitial=G)
            s.update()
            js, act = s.argmax_record()
            out[label] = (js.copy(), act.copy(), s.get_policy().Gamma_array().copy())
            del s, p
        finally:
            set_mode(FAST)
        print("dense %s backup at n=%d, r=%d done" % (label, n, r), flush=True)
    js, act, G1 = out["fp32"]
    G64 = G.astype(np.float64)
    for i in (0, 1, r // 2, r - 1):
        jo, gap, best, aval = _fp64_backup(P, G64, i)
        check_choices(js[i], jo, gap, np.abs(best), n,
                      "dense fp32 belief %d alpha choices" % i)
        ao = int(np.argmax(aval))
        if act[i] != ao:
            check_choices(np.array([act[i]]), np.array([ao]), np.array([aval[ao] - aval[act[i]]]),
                          aval[ao], n, "dense fp32 belief %d action" % i)
        v = float(G1[i].astype(np.float64) @ P["B"][i].astype(np.float64))
        err = abs(v - aval[ao]) / max(1.0, abs(aval[ao]))
        print("dense fp32 belief %d: value %.6f vs FP64 %.6f, relative error %.3g"
              % (i, v, aval[ao], err), flush=True)
        assert err <= 1e-4
    jb, ab, _ = out["bf16"]
    print("dense bf16 vs fp32 at full size: %.4f of alpha choices and %.4f of actions differ"
          % (float(np.mean(jb != js)), float(np.mean(ab != act))), flush=True)
