"""PEMA expansion and the sigma-approximation (src/pomdp/utilities/pomdp_expand_cpu.cpp:269-466,
src/pomdp/utilities/pomdp_sigma_cpu.cpp:50-118), host entry points of the drop-in ABI. CPU only.

* PEMA with nova's alpha choice: the appended beliefs equal nova's build bit for bit over
  several calls; with either alpha choice, the appended belief is the one an FP64 restatement
  of the algorithm (written from the mathematics with numpy) selects.
* sigma: Z, B and sigma equal nova's build bit for bit, ties between equal values included
  (both sort with libstdc++'s std::sort), on inputs nova's code handles (r >= n, k <= rz).
"""
import ctypes as ct
import os

import numpy as np
import pytest

from conftest import bits
from nova_amd import _abi, synth
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_alpha_vectors import POMDPAlphaVectors

DOMAINS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "domains")


def policy(G, pi, m):
    p = POMDPAlphaVectors()
    r, n = G.shape
    assert _abi.load().pomdp_alpha_vectors_initialize(ct.byref(p), n, m, r) == 0
    np.ctypeslib.as_array(p.Gamma, (r * n,))[:] = np.asarray(G, np.float32).reshape(-1)
    np.ctypeslib.as_array(p.pi, (r,))[:] = pi
    return p


def instance(seed, n=30, m=3, z=4, ns=4, r=10, rz=5):
    P = synth.pomdp_arrays(n=n, m=m, z=z, ns=ns, r=r, rz=rz, seed=seed, gamma=0.9, horizon=5)
    rng = np.random.default_rng(seed)
    G = rng.uniform(-5, 5, (6, n)).astype(np.float32)
    pi = rng.integers(0, m, 6).astype(np.uint32)
    return P, G, pi


def product_pema(P, G, pi, steps, mode=0):
    lib = _abi.load()
    assert lib.nova_amd_pema_mode(mode) == 0
    try:
        p = POMDP.from_dict(P)
        pol = policy(G, pi, P["S"].shape[1])
        for _ in range(steps):
            p.expand("pema", pemaPolicy=pol)
        return p.beliefs()
    finally:
        lib.nova_amd_pema_mode(0)


def pema_fp64(P, G, mode):
    """The belief PEMA appends, restated in FP64: tau(b, a, o) = normalised O[a,.,o] * (b T_a),
    Pr(o | b, a) its normaliser, c the first L1-closest set belief, alpha = alpha 0 (mode 0) or
    the first maximiser of alpha . c (mode 1), eps = sum_s (lim_s - alpha_s)(b'_s - c_s) with
    lim_s = (Rmax if b'_s >= b_s else Rmin) / (1 - gamma); first maxima throughout."""
    S, T, O, R, Z, B = P["S"], P["T"], P["O"], P["R"], P["Z"], P["B"]
    n, m, ns = S.shape
    z = O.shape[2]
    r = Z.shape[0]
    g = float(P["gamma"])
    D = np.zeros((r, n))
    for i in range(r):
        for k in range(Z.shape[1]):
            if Z[i, k] < 0:
                break
            D[i, Z[i, k]] = B[i, k]
    Td = np.zeros((m, n, n))
    for s in range(n):
        for a in range(m):
            for l in range(ns):
                if S[s, a, l] < 0:
                    break
                Td[a, s, S[s, a, l]] += T[s, a, l]
    Rmax, Rmin = float(R.max()), float(R.min())
    G64 = np.asarray(G, np.float64)

    def tau(b, a, o):
        v = (b @ Td[a]) * O[a, :, o]
        c = v.sum()
        return (v / c, c) if abs(c) >= 1e-9 else (None, c)

    best_i, bound, choice = None, -1e10, None
    for i in range(r):
        b = D[i]
        aStar, aStarV = 0, -1e10
        oStar = [0] * m
        for a in range(m):
            oV, aV = -1e10, 0.0
            for o in range(z):
                bp, pr = tau(b, a, o)
                if bp is None:
                    continue
                c = D[np.argmin(np.abs(D - bp).sum(axis=1))]
                al = G64[0] if mode == 0 else G64[int(np.argmax(G64 @ c))]
                lim = np.where(bp >= b, Rmax, Rmin) / (1.0 - g)
                v = pr * float(((lim - al) * (bp - c)).sum())
                if v > oV:
                    oV, oStar[a] = v, o
                aV += v
            if aV > aStarV:
                aStar, aStarV = a, aV
        if aStarV > bound:
            bound, best_i, choice = aStarV, i, (aStar, oStar[aStar])
    bp, _ = tau(D[best_i], *choice)
    return D[best_i] if bp is None else bp


def appended_dense(Zb, Bb, n, row):
    d = np.zeros(n)
    for k in range(Zb.shape[1]):
        if Zb[row, k] >= 0:
            d[Zb[row, k]] = Bb[row, k]
    return d


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pema_bitwise_vs_reference(reflib, seed):
    P, G, pi = instance(seed)
    Z, B = product_pema(P, G, pi, steps=3)
    Zr, Br = reflib.expand_pema(P, G, pi, steps=3)
    assert Z.shape == Zr.shape and np.array_equal(Z, Zr)
    assert np.array_equal(bits(B), bits(Br))


def test_pema_tiger_bitwise_vs_reference(reflib, oracle):
    """tests/tiger/tiger.py:56-66 expands tiger_95 with four PEMA calls against a solved
    policy; here the policy is the oracle's PBVI solution on the file's initial belief."""
    p = POMDP().load(os.path.join(DOMAINS, "tiger_95.pomdp"))
    P = p.arrays()
    P["Gamma0"] = np.full((1, 2), float(P["R"].min()) / (1 - float(P["gamma"])), np.float32)
    G, pi = oracle.pbvi_run(P, horizon=30)
    Z, B = product_pema(P, G, pi, steps=4)
    Zr, Br = reflib.expand_pema(P, G, pi, steps=4)
    assert np.array_equal(Z, Zr) and np.array_equal(bits(B), bits(Br))
    assert Z.shape[0] == 5


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("seed", [4, 5, 6])
def test_pema_selects_the_fp64_belief(mode, seed):
    P, G, pi = instance(seed)
    Z, B = product_pema(P, G, pi, steps=1, mode=mode)
    n = P["S"].shape[0]
    got = appended_dense(Z, B, n, Z.shape[0] - 1)
    want = pema_fp64(P, G, mode)
    np.testing.assert_allclose(got, want, atol=1e-5)


def test_pema_mode_changes_the_choice_somewhere():
    """The corrected alpha choice is a different algorithm: on some instance the appended
    belief differs from nova's choice (and each matches its own FP64 restatement above)."""
    differs = 0
    for seed in range(4, 12):
        P, G, pi = instance(seed)
        Z0, B0 = product_pema(P, G, pi, steps=1, mode=0)
        Z1, B1 = product_pema(P, G, pi, steps=1, mode=1)
        differs += int(not (np.array_equal(Z0, Z1) and np.array_equal(bits(B0), bits(B1))))
    assert differs > 0


def test_pema_mode_rejects_unknown():
    assert _abi.load().nova_amd_pema_mode(2) != 0


def sigma_instance(seed, n, r, rz, ties):
    P = synth.pomdp_arrays(n=n, m=2, z=2, ns=2, r=r, rz=rz, seed=seed)
    P["B"] = P["B"].copy()
    P["Z"] = P["Z"].copy()
    if ties:
        # uniform rows, rows with repeated values, and a ragged row
        P["B"][0] = 1.0 / rz
        P["B"][1, : rz // 2] = 0.5 / (rz // 2)
        P["B"][1, rz // 2:] = 0.5 / (rz - rz // 2)
        P["Z"][2, rz - 3:] = -1
        P["B"][2, rz - 3:] = 0.0
    return P


@pytest.mark.parametrize("n,r,rz,k,ties", [(10, 20, 6, 3, True), (12, 40, 12, 5, True),
                                           (30, 64, 24, 8, True), (30, 64, 24, 24, False),
                                           (40, 50, 20, 1, True)])
def test_sigma_bitwise_vs_reference(reflib, n, r, rz, k, ties):
    P = sigma_instance(n + r, n, r, rz, ties)
    p = POMDP.from_dict(P)
    sig = p.sigma_approximate(k)
    Z, B = p.beliefs()
    Zr, Br, sr = reflib.sigma(P, k)
    assert Z.shape == (r, k) and np.array_equal(Z, Zr)
    assert np.array_equal(bits(B), bits(Br)) and bits(np.float32(sig)) == bits(np.float32(sr))
    # each row keeps its k largest values, renormalised; sigma is the smallest kept mass
    mass = np.sort(P["B"], axis=1)[:, ::-1][:, :k].astype(np.float64).sum(axis=1)
    assert abs(sig - min(1.0, mass.min())) < 1e-6
    kept = np.where(Z >= 0, B, 0).astype(np.float64).sum(axis=1)
    np.testing.assert_allclose(kept, 1.0, atol=1e-5)


def test_sigma_more_entries_than_rz():
    """k > rz (nova reads past its arrays there): the extra entries are -1 / 0."""
    P = sigma_instance(3, 10, 12, 4, False)
    p = POMDP.from_dict(P)
    p.sigma_approximate(7)
    Z, B = p.beliefs()
    assert Z.shape == (12, 7) and np.all(Z[:, 4:] == -1) and np.all(B[:, 4:] == 0)
