"""Pins the oracle (oracle/oracle.c) before anything is compared against it:
  * the known answers of nova's own unit tests,
  * bit-for-bit agreement of the float twin with nova's CPU sources compiled where they lie
    (oracle/_ref/libnova_ref.so) on seeded and ragged inputs,
  * closed forms and FP64 brute force (dense numpy Bellman operators) for what those leave open.
CPU only."""
import numpy as np
import pytest

from conftest import bits
from nova_amd import synth


def three_state():
    S = np.array([[[0, 1], [0, -1]], [[1, -1], [2, -1]], [[2, -1], [2, -1]]], np.int32)
    T = np.array([[[0.5, 0.5], [1, 0]], [[1, 0], [1, 0]], [[1, 0], [1, 0]]], np.float32)
    R = np.array([[-1, -2], [-1, -2], [0, 0]], np.float32)
    return S, T, R


def test_vi_known_answers(oracle):
    # tests/unit/mdp/algorithms/mdp_vi_cpu_tests.cpp:148-253 and :383-488
    V, pi = oracle.vi_run(np.zeros((1, 1, 1), np.int32), np.ones((1, 1, 1), np.float32),
                          np.ones((1, 1), np.float32), 0.9, 3)
    assert abs(V[0] - 2.71) < 1e-5 and pi[0] == 0
    S, T, R = three_state()
    V, pi = oracle.vi_run(S, T, R, 0.9, 5)
    np.testing.assert_allclose(V, [-3.303775, -2, 0], atol=1e-5)
    assert pi[0] == 0 and pi[1] == 1
    for h, e in [(1, [-1, -1, 0]), (2, [-1.9, -1.9, 0]), (3, [-2.71, -2, 0])]:
        np.testing.assert_allclose(oracle.vi_run(S, T, R, 0.9, h)[0], e, atol=1e-5)


def test_vi_closed_form_and_contraction(oracle):
    # one state, R = 1: V_h = sum_{t<h} gamma^t (gamma is the float32 0.9 the model holds)
    g32 = float(np.float32(0.9))
    for h in (1, 2, 7, 30):
        V, _ = oracle.vi_run(np.zeros((1, 1, 1), np.int32), np.ones((1, 1, 1), np.float32),
                             np.ones((1, 1), np.float32), 0.9, h, dtype=np.float64)
        assert abs(V[0] - sum(g32 ** t for t in range(h))) < 1e-12
    w = synth.mdp_arrays(n=300, m=4, ns=8, seed=1, gamma=0.9)
    prev = np.zeros(300)
    d = []
    for h in range(1, 8):
        V, _ = oracle.vi_run(w["S"], w["T"], w["R"], w["gamma"], h, dtype=np.float64)
        d.append(np.max(np.abs(V - prev)))
        prev = V
    g = float(w["gamma"])
    assert all(b <= g * a + 1e-12 for a, b in zip(d[1:], d[2:]))


def test_vi_fp64_brute_force(oracle):
    """The FP64 sweep equals a dense numpy Bellman operator (plants no order assumptions)."""
    w = synth.mdp_arrays(n=40, m=3, ns=5, seed=2, gamma=0.95)
    S, T, R = w["S"], w["T"], w["R"]
    n, m, ns = S.shape
    P = np.zeros((m, n, n))
    for s in range(n):
        for a in range(m):
            for i in range(ns):
                P[a, s, S[s, a, i]] += T[s, a, i]
    V = np.zeros(n)
    for _ in range(25):
        Q = R.T.astype(np.float64) + float(w["gamma"]) * P @ V
        V = Q.max(axis=0)
    Vo, pio = oracle.vi_run(S, T, R, w["gamma"], 25, dtype=np.float64)
    np.testing.assert_allclose(Vo, V, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n,m,ns,h,cut", [(500, 8, 16, 20, 0.0), (300, 3, 7, 12, 0.3),
                                          (64, 2, 1, 9, 0.0)])
def test_vi_bitwise_vs_reference_build(oracle, reflib, n, m, ns, h, cut):
    w = synth.mdp_arrays(n=n, m=m, ns=ns, seed=n + m, gamma=0.97)
    S = w["S"].copy()
    rng = np.random.default_rng(n)
    for s, a in zip(*np.nonzero(rng.random((n, m)) < cut)):
        S[s, a, rng.integers(0, ns)] = -1
    v0 = rng.uniform(-3, 3, n).astype(np.float32)
    Vo, pio = oracle.vi_run(S, w["T"], w["R"], w["gamma"], h, V0=v0)
    Vr, pir = reflib.vi_execute(S, w["T"], w["R"], w["gamma"], h, V0=v0)
    assert np.array_equal(bits(Vo), bits(Vr)) and np.array_equal(pio, pir)


def test_pbvi_bitwise_vs_reference_build(oracle, reflib):
    P = synth.workload("pbvi_cpu_n64")
    Go, pio = oracle.pbvi_run(P, horizon=6)
    Gr, pir = reflib.pbvi_execute(P, horizon=6)
    assert np.array_equal(bits(Go), bits(Gr)) and np.array_equal(pio, pir)
    Q = synth.pomdp_arrays(n=30, m=3, z=5, ns=4, r=40, rz=6, seed=9, gamma=0.9, horizon=5)
    Q["S"] = Q["S"].copy()
    Q["S"][3, 1, 2] = -1
    Q["S"][7, 0, 1] = -1
    Q["Z"] = Q["Z"].copy()
    Q["Z"][5, 3] = -1
    Go, pio = oracle.pbvi_run(Q)
    Gr, pir = reflib.pbvi_execute(Q)
    assert np.array_equal(bits(Go), bits(Gr)) and np.array_equal(pio, pir)


def test_pbvi_fp64_brute_force(oracle):
    """One FP64 backup equals a dense enumeration: for each (a, o) the best alpha by
    b . (gamma * P_{a,o} Gamma_j), then alpha_a = R_a + sum_o gamma P_{a,o} Gamma_{j*}."""
    P = synth.pomdp_arrays(n=4, m=2, z=3, ns=3, r=5, rz=3, seed=4, gamma=0.9)
    S, T, O, R, Z, B = P["S"], P["T"], P["O"], P["R"], P["Z"], P["B"]
    n, m, ns = S.shape
    z = O.shape[2]
    rng = np.random.default_rng(0)
    G = rng.uniform(-5, 5, (5, n))
    Pao = np.zeros((m, z, n, n))
    for s in range(n):
        for a in range(m):
            for l in range(ns):
                sp = S[s, a, l]
                Pao[a, :, s, sp] += np.float64(T[s, a, l]) * O[a, sp, :].astype(np.float64)
    g = float(P["gamma"])
    out = oracle.pbvi_backup(P, G, dtype=np.float64)
    for i in range(Z.shape[0]):
        b = np.zeros(n)
        b[Z[i]] = B[i]
        best_v, best_a, best_alpha = -np.inf, None, None
        for a in range(m):
            alpha = R[:, a].astype(np.float64).copy()
            for o in range(z):
                proj = g * (Pao[a, o] @ G.T)           # n x r
                j = int(np.argmax(b @ proj))
                assert out["jstar"][i, a, o] == j
                alpha += proj[:, j]
            v = b @ alpha
            if v > best_v:
                best_v, best_a, best_alpha = v, a, alpha
        assert out["pi"][i] == best_a
        np.testing.assert_allclose(out["GammaPrime"][i], best_alpha, rtol=1e-12, atol=1e-12)


def test_pbvi_simple_closed_form(oracle):
    P = dict(S=np.zeros((1, 1, 1), np.int32), T=np.ones((1, 1, 1), np.float32),
             O=np.ones((1, 1, 1), np.float32), R=np.ones((1, 1), np.float32),
             Z=np.zeros((1, 1), np.int32), B=np.ones((1, 1), np.float32),
             gamma=np.float32(0.9), horizon=3, Gamma0=np.zeros((1, 1), np.float32))
    G, pi = oracle.pbvi_run(P)
    assert abs(G[0, 0] - 2.71) < 1e-6 and pi[0] == 0


def test_belief_update_vs_reference_and_invariants(oracle, reflib):
    P = synth.pomdp_arrays(n=50, m=3, z=4, ns=5, r=3, rz=7, seed=5)
    rng = np.random.default_rng(1)
    for _ in range(10):
        b = np.zeros(50, np.float32)
        supp = rng.choice(50, 7, replace=False)
        b[supp] = rng.dirichlet(np.ones(7)).astype(np.float32)
        a, o = int(rng.integers(3)), int(rng.integers(4))
        rc, bp = oracle.belief_update(P, b, a, o)
        rc2, bp2 = reflib.belief_update(P, b, a, o)
        assert rc == rc2 == 0 and np.array_equal(bits(bp), bits(bp2))
        assert abs(bp.astype(np.float64).sum() - 1) < 1e-5
        reach = set(P["S"][supp, a].ravel().tolist())
        assert set(np.nonzero(bp)[0].tolist()) <= reach


def test_expand_distinct_vs_reference(oracle, reflib):
    P = synth.pomdp_arrays(n=20, m=3, z=3, ns=4, r=6, rz=5, seed=8)
    Bnew, ba, bo = oracle.expand_distinct(P)
    Z2, B2 = reflib.expand_distinct(P)
    r = P["Z"].shape[0]
    # the reference appends the r chosen beliefs as sparse rows after the original r
    for i in range(r):
        dense = np.zeros(20, np.float32)
        k = Z2[r + i] >= 0
        dense[Z2[r + i][k]] = B2[r + i][k]
        assert np.array_equal(bits(dense), bits(Bnew[i]))


def test_alpha_value_known_answers(oracle):
    # tests/unit/pomdp/policies/pomdp_alpha_vectors_tests.cpp:35-105
    G = np.array([[1, 0], [0, 1]], np.float32)
    pi = np.array([1, 0], np.uint32)
    for b, v, a in [([0, 1], 1, 0), ([1, 0], 1, 1), ([0.25, 0.75], 0.75, 0), ([0.67, 0.33], 0.67, 1)]:
        Vb, act, _ = oracle.alpha_value(G, pi, np.array(b, np.float32))
        assert abs(Vb - v) < 1e-5 and act == a
    Vb, _, _ = oracle.alpha_value(G, pi, np.array([0.5, 0.5], np.float32))
    assert abs(Vb - 0.5) < 1e-5


def test_perseus_oracle_bitwise_vs_reference_build(oracle, reflib):
    """The oracle's Perseus step, driven by nova's loop over the same libc rand() stream, equals
    nova's own Perseus (oracle/_ref) bit for bit."""
    import ctypes as ct
    from nova_amd import _abi
    libc = ct.CDLL("libc.so.6")
    for params in [dict(n=25, m=3, z=3, ns=4, r=20, rz=5, seed=43, horizon=40),
                   dict(n=80, m=4, z=6, ns=5, r=60, rz=8, seed=44, horizon=30)]:
        P = synth.pomdp_arrays(gamma=0.9, **params)
        libc.srand(77)
        G, pi = oracle.perseus_run(P, libc.rand)
        pr, keep = reflib._pomdp(P)
        g0 = np.ascontiguousarray(P["Gamma0"], np.float32)
        per = _abi.NovaPOMDPPerseusCPU()
        per.GammaInitial = ct.cast(g0.ctypes.data, _abi.c_float_p)
        pol = _abi.NovaPOMDPAlphaVectors()
        libc.srand(77)
        assert reflib.lib.pomdp_perseus_execute_cpu(ct.byref(pr), ct.byref(per),
                                                    ct.byref(pol)) == 0
        n = P["S"].shape[0]
        Gr = np.ctypeslib.as_array(pol.Gamma, (pol.r * n,)).reshape(pol.r, n).copy()
        pir = np.ctypeslib.as_array(pol.pi, (pol.r,)).copy()
        reflib.lib.pomdp_alpha_vectors_uninitialize(ct.byref(pol))
        assert G.shape == Gr.shape and G.shape[0] > 3
        assert np.array_equal(bits(G), bits(Gr)) and np.array_equal(pi, pir)


def _dense_model(w):
    S, T = w["S"], w["T"].astype(np.float64)
    n, m, ns = S.shape
    P = np.zeros((m, n, n))
    for s in range(n):
        for a in range(m):
            for i in range(ns):
                P[a, s, S[s, a, i]] += T[s, a, i]
    return P


def _policy_iteration(P, R, g):
    """V* and pi* by exact policy iteration (dense linear solves)."""
    m, n, _ = P.shape
    pi = np.zeros(n, int)
    for _ in range(100):
        Ppi = P[pi, np.arange(n)]
        V = np.linalg.solve(np.eye(n) - g * Ppi, R[np.arange(n), pi])
        Q = R.T + g * P @ V
        new = Q.argmax(axis=0)
        if np.array_equal(new, pi):
            return V, pi
        pi = new
    raise AssertionError("policy iteration did not settle")


def test_fp64_solve_adjoint_and_gradients_vs_dense_algebra(oracle):
    """orc_vi_solve_f64 against exact policy iteration, orc_adjoint_f64 against a dense solve
    of (I - gamma P_pi)^T lambda = rho, and orc_vi_grads_f64 against central differences of
    the exact (policy-iteration) V* on every R coordinate and 40 T coordinates."""
    w = synth.mdp_arrays(n=60, m=3, ns=5, seed=21, gamma=0.9)
    g = float(w["gamma"])
    S, T, R = w["S"], w["T"], w["R"].astype(np.float64)
    n, m, ns = S.shape
    P = _dense_model(w)
    Vx, pix = _policy_iteration(P, R, g)
    V, pi, it, res = oracle.vi_solve_f64(S, T, w["R"], w["gamma"], tol=1e-13)
    assert res <= 1e-13
    np.testing.assert_allclose(V, Vx, rtol=0, atol=1e-11)
    assert np.array_equal(pi.astype(int), pix)
    rho = np.random.default_rng(3).uniform(0, 1, n)
    lam = oracle.adjoint_f64(S, T, w["gamma"], pi, rho)
    Ppi = P[pix, np.arange(n)]
    lamx = np.linalg.solve((np.eye(n) - g * Ppi).T, rho)
    np.testing.assert_allclose(lam, lamx, rtol=1e-11, atol=1e-11)
    dR, dT = oracle.vi_grads_f64(S, w["gamma"], pi, lam, V)

    def L(Rm, Tm):
        wm = dict(w, T=Tm)
        Vm, pim = _policy_iteration(_dense_model(wm), Rm, g)
        assert np.array_equal(pim, pix)   # the step stays inside the policy's region
        return rho @ Vm

    h = 1e-6
    fdR = np.empty((n, m))
    for s in range(n):
        for a in range(m):
            Rp, Rm_ = R.copy(), R.copy()
            Rp[s, a] += h
            Rm_[s, a] -= h
            fdR[s, a] = (L(Rp, T.astype(np.float64)) - L(Rm_, T.astype(np.float64))) / (2 * h)
    np.testing.assert_allclose(dR, fdR, rtol=1e-6, atol=1e-7)
    rng = np.random.default_rng(4)
    for _ in range(40):
        s, a, i = int(rng.integers(n)), int(rng.integers(m)), int(rng.integers(ns))
        Tp, Tm = T.astype(np.float64), T.astype(np.float64)
        Tp[s, a, i] += h
        Tm[s, a, i] -= h
        fd = (L(R, Tp) - L(R, Tm)) / (2 * h)
        assert abs(dT[s, a, i] - fd) <= 1e-6 * max(1.0, abs(fd))
