"""Differentiable value iteration (no reference counterpart): the residual-stopped solve, the
adjoint (I - gamma P_pi)^T lambda = rho and the gradients dL/dR, dL/dT, against the FP64
oracle and against FP64 central finite differences of device FP64 solves."""
import numpy as np
import pytest
import torch

from nova_amd import synth
from nova_amd.autograd import BellmanFixedPoint, adjoint, solve

pytestmark = pytest.mark.gpu


def tensors(w):
    dev = "cuda"
    return (torch.from_numpy(w["R"]).to(dev), torch.from_numpy(w["T"]).to(dev),
            torch.from_numpy(w["S"]).to(dev))


def test_solve_adjoint_vs_oracle(oracle):
    w = synth.mdp_arrays(n=600, m=4, ns=8, seed=51, gamma=0.95)
    R, T, S = tensors(w)
    V64, pi, sweeps, res = solve(R, T, S, float(w["gamma"]), tol=1e-10)
    assert res <= 1e-10
    Vo, pio, _, _ = oracle.vi_solve_f64(w["S"], w["T"], w["R"], w["gamma"], tol=1e-13)
    np.testing.assert_allclose(V64.cpu().numpy(), Vo, rtol=0, atol=1e-8)
    assert np.array_equal(pi.cpu().numpy().astype(np.uint32), pio)
    rho = np.random.default_rng(0).uniform(0, 1, 600)
    lam, dR, dT, it, r64 = adjoint(R, T, S, float(w["gamma"]), pi, V64,
                                   torch.from_numpy(rho.astype(np.float32)).cuda())
    lo = oracle.adjoint_f64(w["S"], w["T"], w["gamma"], pio, rho.astype(np.float32))
    dRo, dTo = oracle.vi_grads_f64(w["S"], w["gamma"], pio, lo, Vo)
    scale = np.max(np.abs(lo))
    assert np.max(np.abs(lam.cpu().numpy() - lo)) <= 1e-5 * scale
    assert np.max(np.abs(dR.cpu().numpy() - dRo)) <= 1e-5 * np.max(np.abs(dRo))
    assert np.max(np.abs(dT.cpu().numpy() - dTo)) <= 1e-5 * np.max(np.abs(dTo))
    print("adjoint: %d iterations, FP64 residual %.3g (scale %.3g)" % (it, r64, scale))
    assert r64 <= 1e-5 * scale


@pytest.mark.parametrize("ragged", [False, True], ids=["full_rows", "ragged_rows"])
def test_shifted_solve_vs_plain_and_oracle(oracle, ragged):
    """The solve's shifts along the constant vector (taken when the sweep difference is of one
    sign and nearly uniform) reach the fixed point of plain sweeps: n = 2,000, m = 4, ns = 8,
    gamma = 0.99, to a residual of 1e-10, both within 1e-7 of the FP64 oracle's plain solve
    (gamma / (1 - gamma) * 1e-10 = 1e-8 apart at most) with its greedy policy. Full rows
    (T sums to 1) take at most a fifth of the plain sweeps. Ragged rows (successor lists cut
    short with -1, so a row's valid T sums below 1 and the constant vector is no longer the
    slow mode) still converge to the same point. The adjoint, whose steps are extrapolated
    along the slow mode when the last two steps agree on it, matches the oracle's FP64
    adjoint within 1e-5 of its scale in both cases."""
    w = synth.mdp_arrays(n=2000, m=4, ns=8, seed=53, gamma=0.99)
    if ragged:
        rng = np.random.default_rng(53)
        S = w["S"].copy()
        cut = rng.random(S.shape[:2]) < 0.3
        pos = rng.integers(1, S.shape[2], size=S.shape[:2])
        for s, a in zip(*np.nonzero(cut)):
            S[s, a, pos[s, a]] = -1
        w = dict(w, S=S)
    R, T, S = tensors(w)
    g = float(w["gamma"])
    V64, pi, sweeps, res = solve(R, T, S, g, tol=1e-10)
    V64p, pip, sweepsp, resp = solve(R, T, S, g, tol=1e-10, plain=True)
    assert res <= 1e-10 and resp <= 1e-10
    Vo, pio, _, _ = oracle.vi_solve_f64(w["S"], w["T"], w["R"], w["gamma"], tol=1e-13)
    err = np.max(np.abs(V64.cpu().numpy() - Vo))
    errp = np.max(np.abs(V64p.cpu().numpy() - Vo))
    print("%s: %d sweeps with shifts (|V - V_oracle| %.3g), %d plain (%.3g)"
          % ("ragged" if ragged else "full", sweeps, err, sweepsp, errp), flush=True)
    assert err <= 1e-7 and errp <= 1e-7
    assert np.array_equal(pi.cpu().numpy().astype(np.uint32), pio)
    assert np.array_equal(pip.cpu().numpy().astype(np.uint32), pio)
    if not ragged:
        assert sweeps * 5 <= sweepsp
    rho = np.random.default_rng(1).uniform(0, 1, 2000).astype(np.float32)
    lam, _, _, it, r64 = adjoint(R, T, S, g, pi, V64, torch.from_numpy(rho).cuda(), tol=1e-8)
    lo = oracle.adjoint_f64(w["S"], w["T"], w["gamma"], pio, rho)
    scale = np.max(np.abs(lo))
    aerr = np.max(np.abs(lam.cpu().numpy() - lo)) / scale
    print("adjoint: %d iterations, error %.3g of its scale" % (it, aerr), flush=True)
    assert aerr <= 1e-5 and r64 <= 1e-5 * scale


def test_autograd_matches_adjoint():
    w = synth.mdp_arrays(n=300, m=3, ns=5, seed=52, gamma=0.9)
    R, T, S = tensors(w)
    R.requires_grad_(True)
    T.requires_grad_(True)
    V = BellmanFixedPoint.apply(R, T, S, float(w["gamma"]))
    rho = torch.full((300,), 1.0 / 300, device="cuda")
    (rho * V).sum().backward()
    V64, pi, _, _ = solve(R.detach(), T.detach(), S, float(w["gamma"]))
    _, dR, dT, _, _ = adjoint(R.detach(), T.detach(), S, float(w["gamma"]), pi, V64, rho)
    assert torch.equal(R.grad, dR) and torch.equal(T.grad, dT)


def test_finite_differences_full_size():
    """n = 200,000, m = 8, ns = 16, gamma = 0.99, L = rho.V* with rho uniform: dL/dR and dL/dT
    against FP64 central differences (h = 1e-4, 64 random coordinates, half of them on the
    greedy action, none whose step crosses a kink of V*) of warm-started FP64 device solves."""
    w = synth.workload("vi_diff_n2e5")
    g = float(w["gamma"])
    R, T, S = tensors(w)
    n, m, ns = w["S"].shape
    V64, pi, sweeps, res = solve(R, T, S, g, tol=1e-6)
    print("forward: %d sweeps, FP64 residual %.3g" % (sweeps, res))
    assert res <= 1e-6
    rho = torch.full((n,), 1.0 / n, device="cuda")
    _, dR, dT, it, r64 = adjoint(R, T, S, g, pi, V64, rho, tol=1e-8)
    print("adjoint: %d iterations, FP64 residual %.3g" % (it, r64))
    V64t, _, _, _ = solve(R, T, S, g, tol=1e-15, warm64=V64)
    rng = np.random.default_rng(5)
    pic = pi.cpu().numpy()
    h = 1e-4
    errs = []
    rows = []

    def L(Rp, Tp):
        Vx, _, _, r = solve(Rp, Tp, S, g, tol=1e-15, warm64=V64t)
        return float((Vx * rho.double()).sum())

    Vh = V64.cpu().numpy()
    Sh, Th, Rh = w["S"], w["T"].astype(np.float64), w["R"].astype(np.float64)

    def qrows(ps):
        """Q(p, .) at the FP64 fixed point for the states ps."""
        return Rh[ps] + g * np.sum(Th[ps] * Vh[Sh[ps]], axis=2)

    def clear_of_kinks(s, a, dq):
        """True when no greedy choice can switch within the step: at s itself (Q(s, a) moves by
        dq) and at every (p, a') that has s among its successors (Q(p, a') moves by
        gamma T(p, a', k) dV(s), with dV(s) <= 2 dq when a is the greedy action at s)."""
        Q = qrows(np.array([s]))[0]
        own = Q[pic[s]] - np.max(np.delete(Q, pic[s]))
        if own <= 4 * dq:
            return False
        if a != pic[s]:
            return True
        flat = np.nonzero(Sh.reshape(-1) == s)[0]
        ps, aps, ks = flat // (m * ns), (flat // ns) % m, flat % ns
        Qp = qrows(ps)
        best = Qp[np.arange(len(ps)), pic[ps]]
        Qo = Qp.copy()
        Qo[np.arange(len(ps)), pic[ps]] = -np.inf
        second = Qo.max(axis=1)
        gp = np.where(aps == pic[ps], best - second, best - Qp[np.arange(len(ps)), aps])
        return bool(np.all(gp > 4 * g * Th[ps, aps, ks] * 2 * dq))

    L0 = float((V64t * rho.double()).sum())
    skipped, kinked = 0, 0
    c = 0
    attempts = 0
    while c < 64:
        attempts += 1
        assert attempts <= 400, "too many coordinates rejected (%d at action ties, %d kinked)" % (
            skipped, kinked)
        # V* has a kink wherever two actions tie; a central difference whose step moves some
        # Q across a tie averages two one-sided slopes instead of measuring the derivative.
        # Coordinates whose step reaches a tie at s or at a predecessor of s are redrawn before
        # solving; ties further upstream show as one-sided slopes that disagree, and those
        # coordinates are redrawn too (for a smooth L they agree to about h * L''/L' ~ 1e-7 relative,
        # plus the solves' rounding floor).
        s = int(rng.integers(n))
        on = c % 2 == 0
        a = int(pic[s]) if on else int((pic[s] + 1 + rng.integers(m - 1)) % m)
        i = int(rng.integers(ns))
        dq = h if c < 32 else g * h * abs(Vh[Sh[s, a, i]])
        if not clear_of_kinks(s, a, dq):
            skipped += 1
            continue
        if c < 32:
            base = float(R[s, a])
            Rp, Rm = R.clone(), R.clone()
            Rp[s, a] = base + h
            Rm[s, a] = base - h
            dp = float(Rp[s, a]) - base
            dm = base - float(Rm[s, a])
            Lp, Lm = L(Rp, T), L(Rm, T)
            an = float(dR[s, a])
        else:
            base = float(T[s, a, i])
            Tp, Tm = T.clone(), T.clone()
            Tp[s, a, i] = base + h
            Tm[s, a, i] = base - h
            dp = float(Tp[s, a, i]) - base
            dm = base - float(Tm[s, a, i])
            Lp, Lm = L(R, Tp), L(R, Tm)
            an = float(dT[s, a, i])
        fwd, bwd = (Lp - L0) / dp, (L0 - Lm) / dm
        # the FP64 solves stop at their rounding floor (L to ~1e-11), so one-sided slopes carry
        # ~1e-7 of noise; a kink inside the step shows as a much larger gap
        if abs(fwd - bwd) > 5e-4 * max(abs(fwd), abs(bwd)) + 2e-7:
            kinked += 1
            print("  attempt %d: one-sided slopes apart (%.6g vs %.6g), redrawn" % (
                attempts, fwd, bwd), flush=True)
            continue
        fd = (Lp - Lm) / (dp + dm)
        scale = max(abs(fd), abs(an), 1e-12)
        errs.append(abs(fd - an) / scale if max(abs(fd), abs(an)) > 1e-9 else abs(fd - an))
        rows.append((c, "R" if c < 32 else "T", s, a, on, fd, an, errs[-1]))
        print("  coordinate %d done (error %.3g)" % (c, errs[-1]), flush=True)
        c += 1
    for row in rows:
        if row[-1] > 1e-3:
            print("  coord %d %s s=%d a=%d on=%s fd=%.6g analytic=%.6g err=%.3g" % row)
    print("finite differences: max relative error %.3g over 64 coordinates (redrawn: %d within "
          "the step of an action tie, %d with one-sided slopes apart)" % (max(errs), skipped,
                                                                           kinked))
    assert max(errs) <= 1e-3
