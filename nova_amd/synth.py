"""Seeded synthetic MDP/POMDP instances: the same bytes feed the CPU and GPU paths.

Distribution rules (one numpy PCG64 generator per instance, drawn in the order below):
  * successors: ns distinct states per (s, a), uniform without replacement from [0, n);
  * T rows and O[a, s', :] rows: Dirichlet(1), realised as Exp(1) samples normalised in FP64,
    then cast to float32;
  * R: uniform on [-1, 1] in float32;
  * beliefs: rz distinct support states (uniform, stored in ascending order) with Dirichlet(1)
    weights.
"""
import numpy as np

WORKLOADS = {
    # name: parameters
    "pbvi_cpu_n64": dict(kind="pomdp", n=64, m=4, z=8, ns=4, r=128, rz=8, gamma=0.95,
                         horizon=100, seed=0),
    "vi_n1e6": dict(kind="mdp", n=1_000_000, m=8, ns=16, gamma=0.99, horizon=200, seed=2),
    "pbvi_n20000": dict(kind="pomdp", n=20_000, m=16, z=32, ns=8, r=4096, rz=64, gamma=0.95,
                        horizon=50, seed=3),
    "vi_diff_n2e5": dict(kind="mdp", n=200_000, m=8, ns=16, gamma=0.99, horizon=100000, seed=5),
    "pbvi_diff_n5000": dict(kind="pomdp", n=5000, m=8, z=16, ns=8, r=1024, rz=32, gamma=0.95,
                            horizon=30, seed=6),
}


def successors(rng, rows, ns, n):
    """rows x ns int32, each row ns distinct states drawn uniformly from [0, n)."""
    if ns > n:
        raise ValueError("ns > n")
    if 4 * ns > n:
        keys = rng.random((rows, n))
        return np.argsort(keys, axis=1)[:, :ns].astype(np.int32)
    S = rng.integers(0, n, size=(rows, ns), dtype=np.int32)
    while True:
        Ss = np.sort(S, axis=1)
        dup = (Ss[:, 1:] == Ss[:, :-1]).any(axis=1)
        k = int(dup.sum())
        if k == 0:
            return S
        S[dup] = rng.integers(0, n, size=(k, ns), dtype=np.int32)


def dirichlet1(rng, shape):
    """Dirichlet(1) along the last axis: normalised Exp(1) in FP64, cast to float32."""
    e = rng.standard_exponential(size=shape, dtype=np.float64)
    e /= e.sum(axis=-1, keepdims=True)
    return e.astype(np.float32)


def uniform_r(rng, shape):
    return rng.uniform(-1.0, 1.0, size=shape).astype(np.float32)


def mdp_arrays(n, m, ns, seed, gamma=0.95, horizon=1, **_):
    """dict(S[n,m,ns] int32, T[n,m,ns] f32, R[n,m] f32, gamma, horizon)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    S = successors(rng, n * m, ns, n).reshape(n, m, ns)
    T = dirichlet1(rng, (n, m, ns))
    R = uniform_r(rng, (n, m))
    return dict(S=S, T=T, R=R, gamma=np.float32(gamma), horizon=int(horizon))


def beliefs(rng, r, rz, n):
    """Z[r,rz] int32 ascending supports, B[r,rz] f32 Dirichlet(1) weights."""
    Z = np.sort(successors(rng, r, rz, n), axis=1).astype(np.int32)
    B = dirichlet1(rng, (r, rz))
    return Z, B


def pomdp_arrays(n, m, z, ns, r, rz, seed, gamma=0.95, horizon=1, **_):
    """dict(S, T, O[m,n,z], R, Z, B, gamma, horizon, Gamma0[r,n]) with
    Gamma0 = min(R)/(1-gamma) everywhere (the default of python/nova/pomdp_pbvi.py:57-67)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    S = successors(rng, n * m, ns, n).reshape(n, m, ns)
    T = dirichlet1(rng, (n, m, ns))
    O = dirichlet1(rng, (m, n, z))
    R = uniform_r(rng, (n, m))
    Z, B = beliefs(rng, r, rz, n)
    g = np.float32(gamma)
    fill = np.float32(R.min() / (1.0 - float(g))) if g < 1 else np.float32(0)
    Gamma0 = np.full((r, n), fill, dtype=np.float32)
    return dict(S=S, T=T, O=O, R=R, Z=Z, B=B, gamma=g, horizon=int(horizon), Gamma0=Gamma0)


def dense_beliefs(P, seed):
    """Replaces P's belief set by r dense beliefs: Z[i] = 0..n-1, B[i] ~ Dirichlet(1) over all n
    states (the dense-belief variant of the PBVI workloads)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r, n = P["Gamma0"].shape
    P = dict(P)
    P["Z"] = np.tile(np.arange(n, dtype=np.int32), (r, 1))
    P["B"] = dirichlet1(rng, (r, n))
    return P


def workload(name, **overrides):
    p = dict(WORKLOADS[name])
    p.update(overrides)
    kind = p.pop("kind")
    return (mdp_arrays if kind == "mdp" else pomdp_arrays)(**p)


def batch_instances(count=256, seed=4, n_range=(500, 8000), r_range=(256, 2048), rz=16,
                    gamma=0.95, horizon=50, m_choices=range(4, 17), z_choices=range(4, 33),
                    ns_choices=(4, 8)):
    """The uneven batch: per instance n and r log-uniform, m, z, ns uniform over their choices,
    rz fixed; each instance drawn by its own generator seeded (seed, index)."""
    meta = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        n = int(round(np.exp(meta.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        r = int(round(np.exp(meta.uniform(np.log(r_range[0]), np.log(r_range[1])))))
        m = int(meta.choice(list(m_choices)))
        z = int(meta.choice(list(z_choices)))
        ns = int(meta.choice(list(ns_choices)))
        out.append(pomdp_arrays(n=n, m=m, z=z, ns=ns, r=r, rz=min(rz, n), gamma=gamma,
                                horizon=horizon, seed=(seed << 20) + k))
    return out


def mdp_batch_instances(count=256, seed=5, n_range=(500, 8000), m_choices=range(4, 17),
                        ns_choices=(4, 8), gamma=0.95, horizon=50):
    """The uneven MDP batch (the batched-solve workload's MDP counterpart): per instance n
    log-uniform, m and ns uniform over their choices, T Dirichlet(1), R uniform[-1, 1]; each
    instance drawn by its own generator seeded (seed, index)."""
    meta = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        n = int(round(np.exp(meta.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        m = int(meta.choice(list(m_choices)))
        ns = int(meta.choice(list(ns_choices)))
        out.append(mdp_arrays(n=n, m=m, ns=min(ns, n), gamma=gamma, horizon=horizon,
                              seed=(seed << 20) + k))
    return out
