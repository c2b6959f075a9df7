"""PyTorch autograd for the Bellman fixed point V* = max_a (R + gamma T V*).

BellmanFixedPoint.apply(R, T, S, gamma) returns V* (float32, n) for R (n x m) and T
(n x m x ns) float32 CUDA tensors and the successor lists S (n x m x ns) int32 CUDA tensor.
The tensors are used in place by the library (zero copy): MDP.d_S/d_T/d_R point at them.
backward(g) solves the adjoint (I - gamma P_pi)^T lambda = g on the device and returns
dL/dR and dL/dT (the successor lists and gamma get no gradient).
"""
import ctypes as ct

import numpy as np
import torch

from . import _abi


def _mdp_view(R, T, S, gamma):
    n, m, ns = S.shape
    mdp = _abi.NovaMDP()
    mdp.n, mdp.m, mdp.ns = n, m, ns
    mdp.gamma = float(gamma)
    mdp.horizon = 1
    mdp.d_S = ct.cast(S.data_ptr(), _abi.c_int_p)
    mdp.d_T = ct.cast(T.data_ptr(), _abi.c_float_p)
    mdp.d_R = ct.cast(R.data_ptr(), _abi.c_float_p)
    return mdp


def solve(R, T, S, gamma, tol=1e-6, max_sweeps=100000, V0=None, warm64=None, plain=False):
    """(V64, pi, sweeps, residual): the FP64 fixed point on the device and its greedy policy.
    warm64: optional (n,) float64 CUDA tensor to start the FP64 sweeps from. plain: sweeps
    only, without the shifts along the constant vector that remove the slow part of the
    error (same fixed point, about 1/(1 - gamma) times more sweeps)."""
    lib = _abi.load()
    n = S.shape[0]
    dev = R.device
    mdp = _mdp_view(R, T, S, gamma)
    vi = _abi.NovaMDPVIGPU()
    V = torch.zeros(n, dtype=torch.float32, device=dev) if V0 is None else V0.clone()
    Vp = V.clone()
    pi = torch.zeros(n, dtype=torch.int32, device=dev)
    vi.d_V = ct.cast(V.data_ptr(), _abi.c_float_p)
    vi.d_VPrime = ct.cast(Vp.data_ptr(), _abi.c_float_p)
    vi.d_pi = ct.cast(pi.data_ptr(), _abi.c_uint_p)
    V64 = torch.empty(2 * n, dtype=torch.float64, device=dev)
    flags = 2 if plain else 0
    if warm64 is not None:
        V64[:n].copy_(warm64)
        flags |= 1
    torch.cuda.synchronize()
    sweeps = ct.c_uint(0)
    res = ct.c_double(0)
    rc = lib.mdp_vi_solve_gpu(ct.byref(mdp), ct.byref(vi), V64.data_ptr(), tol, max_sweeps, flags,
                              ct.byref(sweeps), ct.byref(res))
    if rc != 0:
        raise RuntimeError("mdp_vi_solve_gpu failed with code %d" % rc)
    return V64[:n].clone(), pi, sweeps.value, res.value


def adjoint(R, T, S, gamma, pi, V64, rho, tol=1e-7, max_iters=200000, grads=True):
    """(lambda, dR, dT, iterations, fp64 residual) for the adjoint at (pi, V64)."""
    lib = _abi.load()
    n, m, ns = S.shape
    dev = R.device
    mdp = _mdp_view(R, T, S, gamma)
    lam = torch.empty(n, dtype=torch.float32, device=dev)
    dR = torch.empty((n, m), dtype=torch.float32, device=dev) if grads else None
    dT = torch.empty((n, m, ns), dtype=torch.float32, device=dev) if grads else None
    rho = rho.to(torch.float32).contiguous()
    torch.cuda.synchronize()
    it = ct.c_uint(0)
    res = ct.c_double(0)
    rc = lib.mdp_vi_adjoint_gpu(ct.byref(mdp), pi.data_ptr(), V64.data_ptr(), rho.data_ptr(),
                                lam.data_ptr(), dR.data_ptr() if grads else None,
                                dT.data_ptr() if grads else None, tol, max_iters,
                                ct.byref(it), ct.byref(res))
    if rc != 0:
        raise RuntimeError("mdp_vi_adjoint_gpu failed with code %d" % rc)
    return lam, dR, dT, it.value, res.value


class BellmanFixedPoint(torch.autograd.Function):
    """V* of an MDP as a differentiable function of (R, T)."""

    @staticmethod
    def forward(ctx, R, T, S, gamma, tol=1e-6):
        R = R.contiguous()
        T = T.contiguous()
        S = S.contiguous()
        V64, pi, sweeps, res = solve(R, T, S, gamma, tol=tol)
        ctx.save_for_backward(R, T, S, pi, V64)
        ctx.gamma = gamma
        ctx.solve_info = (sweeps, res)
        return V64.to(torch.float32)

    @staticmethod
    def backward(ctx, g):
        R, T, S, pi, V64 = ctx.saved_tensors
        _, dR, dT, _, _ = adjoint(R, T, S, ctx.gamma, pi, V64, g)
        return dR, dT, None, None, None


# ---------------------------------------------------------------- PBVI argmax replay

def pbvi_objective(P, R, horizon=None, mode=None):
    """Runs PBVI on the device with recorded choices. Returns a dict with L = sum_b max_alpha
    alpha . b (float64 of the float32 values), the per-belief values and maximisers, the final
    Gamma, and the (device) records, plus the POMDP object that holds the device model."""
    from .pomdp import POMDP
    from .pomdp_pbvi import POMDPPBVIGPU, set_mode
    lib = _abi.load()
    Rn = R.detach().to("cpu", torch.float32).numpy() if torch.is_tensor(R) else np.asarray(R, np.float32)
    H = int(P["horizon"] if horizon is None else horizon)
    p = POMDP.from_arrays(P["S"], P["T"], P["O"], Rn, P["Z"], P["B"], P["gamma"], H)
    p.initialize_gpu()
    old = set_mode(mode) if mode is not None else None
    try:
        s = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"])
        r, z = p.r, p.z
        recJ = torch.empty(H * r * z, dtype=torch.int32, device="cuda")
        recA = torch.empty(H * r, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        rc = lib.pomdp_pbvi_run_recorded_gpu(ct.byref(p), ct.byref(s), H, recJ.data_ptr(),
                                             recA.data_ptr())
        if rc != 0:
            raise RuntimeError("pomdp_pbvi_run_recorded_gpu failed with code %d" % rc)
        G = torch.from_numpy(s.get_policy().Gamma_array()).cuda()
    finally:
        if old is not None:
            set_mode(old)
    vals = torch.empty(p.r, dtype=torch.float32, device="cuda")
    arg = torch.empty(p.r, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = lib.pomdp_pbvi_belief_values_gpu(ct.byref(p), G.data_ptr(), p.r, vals.data_ptr(),
                                          arg.data_ptr())
    if rc != 0:
        raise RuntimeError("pomdp_pbvi_belief_values_gpu failed with code %d" % rc)
    return dict(L=float(vals.double().sum()), values=vals, arg=arg, Gamma=G, recJ=recJ,
                recA=recA, pomdp=p, solver=s, horizon=H)


def pbvi_replay_grad(res):
    """dL/dR (n x m float32 CUDA tensor) for a result of pbvi_objective."""
    lib = _abi.load()
    p = res["pomdp"]
    dR = torch.empty((p.n, p.m), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = lib.pomdp_pbvi_replay_grad_gpu(ct.byref(p), res["horizon"], res["recJ"].data_ptr(),
                                        res["recA"].data_ptr(), res["arg"].data_ptr(),
                                        dR.data_ptr())
    if rc != 0:
        raise RuntimeError("pomdp_pbvi_replay_grad_gpu failed with code %d" % rc)
    return dR


class PBVIBeliefValue(torch.autograd.Function):
    """L(R) = sum over the belief set of max_alpha alpha . b after `horizon` PBVI iterations
    from GammaInitial, differentiable in R by argmax replay. P holds the other model arrays
    (S, T, O, Z, B, gamma, horizon, Gamma0) as numpy arrays."""

    @staticmethod
    def forward(ctx, R, P):
        res = pbvi_objective(P, R)
        ctx.res = res
        return torch.tensor(res["L"], dtype=torch.float64, device=R.device)

    @staticmethod
    def backward(ctx, g):
        dR = pbvi_replay_grad(ctx.res)
        return (g.to(torch.float32) * dR).to(ctx.res["Gamma"].device), None
