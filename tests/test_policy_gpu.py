"""Executing alpha-vector policies over many beliefs on the GPU, through the C ABI, against the
oracle: nova_amd_alpha_vectors_value_and_action_gpu (pomdp_alpha_vectors.cpp:52-79 per belief)
and nova_amd_belief_update_batch_gpu (pomdp_model_cpu.cpp:89-137 per triple).

exact value/action and the belief update: bit for bit against the oracle's float twin.
fp32 (matrix cores): the chosen alpha is the FP64 maximiser except at near-ties (its FP64
      value within 1e-5 * max(1, |V|) of the maximum) and V within 1e-5 relative.
bf16 (matrix cores): reported (alpha disagreement rate, value error against FP64); the value
      of the chosen alpha stays within 1e-2 * max |V| of the FP64 maximum.
"""
import numpy as np
import pytest

from conftest import bits
from nova_amd import synth
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_alpha_vectors import POMDPAlphaVectors

pytestmark = pytest.mark.gpu


def policy(G, pi):
    import ctypes as ct
    from nova_amd import _abi
    p = POMDPAlphaVectors()
    r, n = G.shape
    lib = _abi.load()
    assert lib.pomdp_alpha_vectors_initialize(ct.byref(p), n, int(pi.max()) + 1, r) == 0
    np.ctypeslib.as_array(p.Gamma, (r * n,))[:] = G.reshape(-1)
    np.ctypeslib.as_array(p.pi, (r,))[:] = pi
    return p


def instance(k, n, r, seed, m=5):
    rng = np.random.default_rng(seed)
    G = rng.uniform(-3, 3, (r, n)).astype(np.float32)
    pi = rng.integers(0, m, r).astype(np.uint32)
    B = rng.exponential(size=(k, n))
    B[rng.random((k, n)) < 0.5] = 0.0          # sparse-ish dense rows
    B = (B / np.maximum(B.sum(axis=1, keepdims=True), 1e-30)).astype(np.float32)
    return G, pi, B


SHAPES = [(1, 1, 1), (70, 33, 65), (300, 257, 130), (1000, 600, 500)]


@pytest.mark.parametrize("k,n,r", SHAPES)
def test_exact_bitwise_vs_oracle(oracle, k, n, r):
    G, pi, B = instance(k, n, r, seed=k + n + r)
    # duplicated alphas: the first of equal maxima must win
    if r > 3:
        G[r // 2] = G[1]
        G[r - 1] = G[1]
    p = policy(G, pi)
    V, act, j = p.value_and_action_batch(B, "exact", return_alpha=True)
    V, act, j = V.cpu().numpy(), act.cpu().numpy(), j.cpu().numpy()
    for i in range(k):
        vo, ao, jo = oracle.alpha_value(G, pi, B[i])
        assert bits(np.float32(vo)) == bits(V[i]) and ao == act[i] and jo == j[i], i
    print("exact value/action: k=%d n=%d r=%d bit-identical" % (k, n, r), flush=True)


def test_exact_no_alpha_above_sentinel():
    """Every value at or below -1e10: V = -1e10, action 0 and alpha -1 (the reference leaves
    Vb and a at their initial values). Exact and fp32 only: in bf16 the sentinel itself rounds
    to -9.9992e9, so alphas holding it count as ordinary values."""
    G = np.full((5, 4), -1e10, np.float32)
    pi = np.array([3, 1, 2, 4, 1], np.uint32)
    B = np.array([[1, 0, 0, 0], [0.25, 0.25, 0.25, 0.25]], np.float32)
    p = policy(G, pi)
    for prec in ("exact", "fp32"):
        V, act, j = p.value_and_action_batch(B, prec, return_alpha=True)
        assert np.all(V.cpu().numpy() == np.float32(-1e10)), prec
        assert np.all(act.cpu().numpy() == 0) and np.all(j.cpu().numpy() == -1), prec


@pytest.mark.parametrize("k,n,r", SHAPES[1:] + [(4096, 2000, 1024)])
def test_fp32_matrix_cores_vs_fp64(k, n, r):
    G, pi, B = instance(k, n, r, seed=3 * k + r)
    p = policy(G, pi)
    V, act, j = p.value_and_action_batch(B, "fp32", return_alpha=True)
    V, act, j = V.cpu().numpy(), act.cpu().numpy(), j.cpu().numpy()
    C = B.astype(np.float64) @ G.astype(np.float64).T
    best = C.max(axis=1)
    scale = np.maximum(1.0, np.abs(best))
    chosen = C[np.arange(k), j]
    assert np.all(best - chosen <= 1e-5 * scale)
    assert np.all(np.abs(V - best) <= 1e-5 * scale)
    assert np.array_equal(act, pi[j])
    differ = int((j != C.argmax(axis=1)).sum())
    print("fp32: k=%d n=%d r=%d, %d of %d alpha choices differ from FP64 (near-ties)"
          % (k, n, r, differ, k), flush=True)


@pytest.mark.parametrize("k,n,r", [(1000, 600, 500), (4096, 2000, 1024)])
def test_bf16_matrix_cores_reported(k, n, r):
    G, pi, B = instance(k, n, r, seed=5 * k + r)
    p = policy(G, pi)
    V, act, j = p.value_and_action_batch(B, "bf16", return_alpha=True)
    V, act, j = V.cpu().numpy(), act.cpu().numpy(), j.cpu().numpy()
    C = B.astype(np.float64) @ G.astype(np.float64).T
    best = C.max(axis=1)
    chosen = C[np.arange(k), j]
    vmax = np.max(np.abs(best))
    rate = float(np.mean(j != C.argmax(axis=1)))
    err = float(np.max(best - chosen)) / vmax
    print("bf16: k=%d n=%d r=%d, alpha disagreement %.3g, value loss of the chosen alpha "
          "%.3g of max|V|, |V - V_fp64| %.3g" % (k, n, r, rate, err,
                                                 np.max(np.abs(V - best)) / vmax), flush=True)
    assert err <= 1e-2 and np.array_equal(act, pi[j])


def test_large_fp32_against_exact_sample():
    """At n = 20,000 and r = 4,096 (the dense PBVI size) with k = 8,192 beliefs: fp32 against
    exact on the first 256 beliefs (same choice or a near-tie)."""
    k, n, r = 8192, 20000, 4096
    G, pi, B = instance(k, n, r, seed=99)
    p = policy(G, pi)
    V, act, j = p.value_and_action_batch(B, "fp32", return_alpha=True)
    Ve, acte, je = p.value_and_action_batch(B[:256], "exact", return_alpha=True)
    V, j = V.cpu().numpy()[:256], j.cpu().numpy()[:256]
    Ve, je = Ve.cpu().numpy(), je.cpu().numpy()
    scale = np.maximum(1.0, np.abs(Ve))
    assert np.all(np.abs(V - Ve) <= 1e-5 * scale)
    print("large: %d of 256 sampled choices differ between fp32 and exact"
          % int((j != je).sum()), flush=True)


@pytest.mark.parametrize("n,m,z,ns,k", [(40, 3, 4, 5, 200), (700, 4, 6, 9, 300)])
def test_belief_update_batch_bitwise(oracle, n, m, z, ns, k):
    P = synth.pomdp_arrays(n=n, m=m, z=z, ns=ns, r=4, rz=3, seed=n + k)
    # a ragged successor list and an observation impossible under one action
    P["S"] = P["S"].copy()
    P["S"][0, 0, 2] = -1
    P["O"] = P["O"].copy()
    P["O"][1, :, 0] = 0.0
    rng = np.random.default_rng(k)
    B = rng.exponential(size=(k, n))
    B[rng.random((k, n)) < 0.7] = 0.0
    B[:, 0] += 1e-3
    B = (B / B.sum(axis=1, keepdims=True)).astype(np.float32)
    a = rng.integers(0, m, k).astype(np.uint32)
    o = rng.integers(0, z, k).astype(np.uint32)
    a[:5], o[:5] = 1, 0                       # impossible observations
    a[5] = m                                  # an action out of range
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    Bp, status = p.belief_update_batch(B, a, o)
    Bp, status = Bp.cpu().numpy(), status.cpu().numpy()
    assert status[5] == 1
    for i in range(k):
        if i == 5:
            continue
        rc, bp = oracle.belief_update(P, B[i], int(a[i]), int(o[i]))
        assert status[i] == rc, i
        assert np.array_equal(bits(Bp[i]), bits(bp)), i
    assert np.all(status[:5] == 10)
    print("belief update batch: n=%d k=%d bit-identical (%d impossible observations)"
          % (n, k, int((status == 10).sum())), flush=True)


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("k,n,r", [(1, 1, 1), (70, 33, 65), (300, 257, 130), (513, 100, 300),
                                   (1000, 600, 500)])
def test_bf16_blocked_kernel_equals_legacy(monkeypatch, k, n, r):
    """The k-blocked direct-to-LDS bf16 GEMM and the row-major register-staged one round the
    same operands and accumulate the k-steps in the same order: V, action and alpha are
    bit-identical (ragged k, n and r; 1 to 10 column tiles)."""
    G, pi, B = instance(k, n, r, seed=11 * k + n)
    p = policy(G, pi)
    monkeypatch.delenv("NOVA_AMD_DENSE_KERNEL", raising=False)
    V, act, j = (x.cpu().numpy() for x in p.value_and_action_batch(B, "bf16", return_alpha=True))
    monkeypatch.setenv("NOVA_AMD_DENSE_KERNEL", "legacy")
    Vl, actl, jl = (x.cpu().numpy() for x in p.value_and_action_batch(B, "bf16",
                                                                      return_alpha=True))
    monkeypatch.delenv("NOVA_AMD_DENSE_KERNEL")
    assert np.array_equal(bits(V), bits(Vl)) and np.array_equal(act, actl)
    assert np.array_equal(j, jl)
    # against FP64 on the bf16-rounded operands: only fp32 accumulation error remains
    C = bf16_round(B).astype(np.float64) @ bf16_round(G).astype(np.float64).T
    mag = np.abs(bf16_round(B)).astype(np.float64) @ np.abs(bf16_round(G)).astype(np.float64).T
    tol = 1e-6 * n * mag.max(axis=1) + 1e-30
    assert np.all(np.abs(V - C[np.arange(k), j]) <= tol)
    assert np.all(C.max(axis=1) - C[np.arange(k), j] <= 2 * tol)
    print("bf16 blocked == legacy: k=%d n=%d r=%d" % (k, n, r), flush=True)


def test_belief_update_batch_long_lists_and_global_belief(oracle):
    """n = 30,000 (past the LDS copy of b: b is read from global memory) and successors skewed
    onto states 0-63, whose block then holds more predecessor entries than a wave's product
    scratch (several sub-ranges per block): bit for bit against the oracle's float twin."""
    n, m, z, ns, k = 30000, 2, 3, 4, 12
    P = synth.pomdp_arrays(n=n, m=m, z=z, ns=ns, r=2, rz=2, seed=77)
    P["S"] = P["S"].copy()
    rng = np.random.default_rng(77)
    hot = rng.random((n, m)) < 0.2                  # about 24,000 entries onto states 0-63
    P["S"][:, :, 0] = np.where(hot, rng.integers(0, 64, (n, m)), P["S"][:, :, 0])
    B = rng.exponential(size=(k, n)).astype(np.float32)
    B = (B / B.sum(axis=1, keepdims=True)).astype(np.float32)
    a = rng.integers(0, m, k).astype(np.uint32)
    o = rng.integers(0, z, k).astype(np.uint32)
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    Bp, status = p.belief_update_batch(B, a, o)
    Bp, status = Bp.cpu().numpy(), status.cpu().numpy()
    for i in range(k):
        rc, bp = oracle.belief_update(P, B[i], int(a[i]), int(o[i]))
        assert status[i] == rc and np.array_equal(bits(Bp[i]), bits(bp)), i
    print("belief update batch: n=30000, skewed lists, bit-identical", flush=True)


@pytest.mark.parametrize("n,m,z,ns,k", [(65536, 2, 3, 4, 6), (65537, 2, 3, 4, 6),
                                        (8, 4097, 2, 2, 40)],
                         ids=["n65536_16bit_lists", "n65537_32bit_lists", "m4097_unordered"])
def test_belief_update_batch_index_width_and_order_bounds(oracle, n, m, z, ns, k):
    """The boundaries of the batched belief update's internal forms, bit for bit against the
    oracle's float twin: predecessor lists with 16-bit indices (n <= 65,536) and 32-bit ones
    (n = 65,537), and triples run in call order instead of grouped by action (m > 4,096)."""
    P = synth.pomdp_arrays(n=n, m=m, z=z, ns=ns, r=2, rz=2, seed=n + m)
    rng = np.random.default_rng(n + m)
    B = rng.exponential(size=(k, n)).astype(np.float32)
    B = (B / B.sum(axis=1, keepdims=True)).astype(np.float32)
    a = rng.integers(0, m, k).astype(np.uint32)
    o = rng.integers(0, z, k).astype(np.uint32)
    a[0] = m - 1                                  # the last action and state indices in use
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    Bp, status = p.belief_update_batch(B, a, o)
    Bp, status = Bp.cpu().numpy(), status.cpu().numpy()
    for i in range(k):
        rc, bp = oracle.belief_update(P, B[i], int(a[i]), int(o[i]))
        assert status[i] == rc and np.array_equal(bits(Bp[i]), bits(bp)), i
    print("belief update batch: n=%d m=%d bit-identical" % (n, m), flush=True)
