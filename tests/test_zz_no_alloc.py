"""The allocation-free exact value/action entry point
(nova_amd_alpha_vectors_value_and_action_noalloc_gpu) against the allocating one
(nova_amd_alpha_vectors_value_and_action_gpu, precision 0), through the C ABI: two calls in a
row, the second with a smaller input, each bit for bit equal (values, actions and alpha
indices). The second call reuses the static scratch array that still holds the first call's
partials. A call beyond the scratch array's limit is refused with an error and no kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instance(k, n, r, seed, m=6):
    rng = np.random.default_rng(seed)
    G = rng.uniform(-3, 3, (r, n)).astype(np.float32)
    if r > 3:
        G[r - 1] = G[1]                        # a tie: the first maximiser must win
    pi = rng.integers(0, m, r).astype(np.int32)
    B = rng.exponential(size=(k, n))
    B[rng.random((k, n)) < 0.4] = 0.0
    B = (B / np.maximum(B.sum(axis=1, keepdims=True), 1e-30)).astype(np.float32)
    return G, pi, B


def _call(fn, n, r, G, pi, k, B, precision=None):
    import torch
    V = torch.full((k,), float("nan"), dtype=torch.float32, device="cuda")
    act = torch.full((k,), -7, dtype=torch.int32, device="cuda")
    j = torch.full((k,), -7, dtype=torch.int32, device="cuda")
    args = [n, r, G.data_ptr(), pi.data_ptr(), k, B.data_ptr()]
    if precision is not None:
        args.append(precision)
    args += [V.data_ptr(), act.data_ptr(), j.data_ptr()]
    rc = fn(*args)
    torch.cuda.synchronize()
    return rc, V.cpu().numpy(), act.cpu().numpy(), j.cpu().numpy()


def test_noalloc_equals_allocating_entry_point_twice():
    import torch
    from nova_amd import _abi
    lib = _abi.load()
    dev = torch.device("cuda")
    for k, n, r, seed in ((700, 300, 450, 1), (65, 97, 70, 2)):   # the second call is smaller
        G, pi, B = _instance(k, n, r, seed)
        Gd, pid, Bd = (torch.from_numpy(x).to(dev) for x in (G, pi, B))
        rc0, V0, a0, j0 = _call(lib.nova_amd_alpha_vectors_value_and_action_gpu, n, r, Gd, pid,
                                k, Bd, precision=0)
        rc1, V1, a1, j1 = _call(lib.nova_amd_alpha_vectors_value_and_action_noalloc_gpu, n, r,
                                Gd, pid, k, Bd)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(V0.view(np.uint32), V1.view(np.uint32)), (k, n, r)
        assert np.array_equal(a0, a1) and np.array_equal(j0, j1), (k, n, r)
        # the values are real maxima of the alphas' products with the beliefs
        best = (G.astype(np.float64) @ B.T.astype(np.float64)).max(axis=0)
        assert np.max(np.abs(V1 - best)) <= 1e-5 * max(1.0, float(np.abs(best).max()))
        print("no-alloc value/action k=%d n=%d r=%d: equal bit for bit" % (k, n, r), flush=True)


def test_noalloc_refuses_calls_beyond_its_scratch():
    import torch
    from nova_amd import _abi
    lib = _abi.load()
    n, r, k = 4, 64 * 2049, 1024          # 2,049 alpha tiles x 1,024 beliefs > 2,097,152
    G = torch.zeros((r, n), dtype=torch.float32, device="cuda")
    pi = torch.zeros(r, dtype=torch.int32, device="cuda")
    B = torch.zeros((k, n), dtype=torch.float32, device="cuda")
    rc, V, _, _ = _call(lib.nova_amd_alpha_vectors_value_and_action_noalloc_gpu, n, r, G, pi,
                        k, B)
    assert rc != 0
    assert np.isnan(V).all()                   # nothing was written
