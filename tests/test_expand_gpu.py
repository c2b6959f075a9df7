"""Belief expansion with the trajectories stepped on the GPU (pomdp_expand_random_gpu and its
seeded variant), against the reference's loop (src/pomdp/utilities/pomdp_expand_gpu.cu:137-220)
restated here over the oracle's float tau and Pr(o|b,a), consuming the process's libc rand()
in the same order (horizon draw per trajectory, then action and target per step)."""
import ctypes as ct

import numpy as np
import pytest

from nova_amd import _abi, synth
from nova_amd.pomdp import POMDP

pytestmark = pytest.mark.gpu
libc = ct.CDLL("libc.so.6")
RAND_MAX = 2147483647


def _rand01():
    return np.float32(np.float32(libc.rand()) / np.float32(RAND_MAX))


def oracle_expand(oracle, P, num, seed):
    """The reference's trajectory loop, with the observation guard Pr(o) > 0 of its host twin."""
    libc.srand(seed)
    S = P["S"]
    n, m, _ = S.shape
    z = P["O"].shape[2]
    b0 = np.zeros(n, np.float32)
    k = P["Z"][0] >= 0
    b0[P["Z"][0][k]] = P["B"][0][k]
    rows = [b0.copy()]
    obs = []
    while len(rows) < num:
        h = int(np.float32(_rand01() * np.float32(P["horizon"] + 1)))
        b = b0.copy()
        for _ in range(h):
            if len(rows) >= num:
                break
            a = min(int(np.float32(_rand01() * np.float32(m))), m - 1)
            target = _rand01()
            cum = np.float32(0)
            o = None
            for op in range(z):
                pr = np.float32(oracle.prob_obs(P, b, a, op))
                cum = np.float32(cum + pr)
                if cum >= target and pr > 0:
                    o = op
                    break
            o = 0 if o is None else o
            _, b = oracle.belief_update(P, b, a, o)
            rows.append(b.copy())
            obs.append((a, o))
    return np.array(rows), obs


@pytest.mark.parametrize("n,m,z,ns,num,seed", [(30, 3, 4, 4, 40, 5), (200, 4, 6, 5, 120, 6)])
def test_seeded_expansion_vs_oracle(oracle, lib, n, m, z, ns, num, seed):
    P = synth.pomdp_arrays(n=n, m=m, z=z, ns=ns, r=3, rz=min(6, n), seed=seed, horizon=12)
    p = POMDP.from_dict(P)
    maxnz = ct.c_uint(0)
    Bp = _abi.c_float_p()
    rc = lib.pomdp_expand_random_seeded_gpu(ct.byref(p), 777, num, ct.byref(maxnz), ct.byref(Bp))
    assert rc == 0
    got = np.ctypeslib.as_array(Bp, (num * n,)).reshape(num, n).copy()
    lib.nova_amd_free_floats(Bp)
    want, _ = oracle_expand(oracle, P, num, 777)
    # the GPU forms tau by a different summation order: beliefs agree to float rounding, and an
    # observation draw can differ only when the cumulative sum is within rounding of the target
    close = np.all(np.abs(got - want) <= 1e-5, axis=1)
    assert close.mean() > 0.9, "too many beliefs differ from the oracle's trajectories"
    assert np.allclose(got.sum(axis=1), 1.0, atol=1e-5)
    assert maxnz.value == max(int((row > 0).sum()) for row in got[1:])


def test_unseeded_expansion_api(lib):
    """nova's signature (pomdp_expand_random_gpu): Bnew allocated by the library, row 0 is
    belief 0, every row a distribution."""
    P = synth.pomdp_arrays(n=50, m=2, z=3, ns=4, r=2, rz=5, seed=7, horizon=8)
    p = POMDP.from_dict(P)
    maxnz = ct.c_uint(0)
    Bp = _abi.c_float_p()
    assert lib.pomdp_expand_random_gpu(ct.byref(p), 256, 30, ct.byref(maxnz), ct.byref(Bp)) == 0
    got = np.ctypeslib.as_array(Bp, (30 * 50,)).reshape(30, 50).copy()
    lib.nova_amd_free_floats(Bp)
    b0 = np.zeros(50, np.float32)
    b0[P["Z"][0]] = P["B"][0]
    assert np.array_equal(got[0], b0)
    assert np.allclose(got.sum(axis=1), 1.0, atol=1e-5)
    assert lib.pomdp_expand_random_gpu(ct.byref(p), 256, 1, ct.byref(maxnz), ct.byref(Bp)) == 1


@pytest.mark.parametrize("n,m,z,ns,r,rz,seed", [(20, 3, 3, 4, 6, 5, 8), (60, 4, 5, 6, 40, 9, 9)])
def test_distinct_beliefs_gpu_bitwise(reflib, lib, n, m, z, ns, r, rz, seed):
    """pomdp_expand_distinct_beliefs_gpu appends exactly the beliefs nova's host function does."""
    P = synth.pomdp_arrays(n=n, m=m, z=z, ns=ns, r=r, rz=rz, seed=seed)
    P["S"] = P["S"].copy()
    P["S"][1, 0, 1] = -1          # a ragged successor list
    P["Z"] = P["Z"].copy()
    P["Z"][2, 3] = -1             # a ragged support
    p = POMDP.from_dict(P)
    assert lib.pomdp_expand_distinct_beliefs_gpu(ct.byref(p)) == 0
    Z, B = p.beliefs()
    Z2, B2 = reflib.expand_distinct(P)
    assert np.array_equal(Z, Z2)
    assert np.array_equal(B.view(np.uint32), B2.view(np.uint32))
