"""The sigma-approximation on the device copy of the belief set (pomdp_sigma_gpu) against
nova's own build (oracle/_ref: pomdp_sigma_cpu, src/pomdp/utilities/pomdp_sigma_cpu.cpp:50-118)
and the library's host restatement: rows of distinct values give the same Z, B and sigma bit
for bit; rows with equal values (whose order nova's std::sort leaves to the library's
algorithm) give the same kept masses and sigma; k above rz pads with -1 / 0 as the host does.
Both the device and the host copies of the set are replaced."""
import numpy as np
import pytest

from conftest import bits
from nova_amd import synth
from nova_amd.pomdp import POMDP

pytestmark = pytest.mark.gpu


def _gpu_sigma(P, k):
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    sig = p.sigma_approximate(k, gpu=True)
    Z, B = p.beliefs()
    assert Z.shape[1] == k
    return Z, B, sig


def _host_sigma(P, k):
    p = POMDP.from_dict(P)
    sig = p.sigma_approximate(k)
    Z, B = p.beliefs()
    return Z, B, sig


@pytest.mark.parametrize("n,r,rz,k,seed", [(40, 60, 12, 5, 1), (100, 130, 64, 64, 2),
                                           (300, 400, 700 // 2, 17, 3)])
def test_sigma_gpu_equals_nova(reflib, n, r, rz, k, seed):
    P = synth.pomdp_arrays(n=n, m=2, z=2, ns=3, r=r, rz=min(rz, n), seed=seed)
    Zg, Bg, sg = _gpu_sigma(P, k)
    Zr, Br, sr = reflib.sigma(P, k)
    assert np.array_equal(Zg, Zr) and np.array_equal(bits(Bg), bits(Br))
    assert bits(np.float32(sg)) == bits(np.float32(sr))
    print("sigma gpu n=%d r=%d rz=%d k=%d: equal to nova's build, sigma %.6f" % (n, r, rz, k, sg),
          flush=True)


def test_sigma_gpu_ties_and_long_k():
    P = synth.pomdp_arrays(n=50, m=2, z=2, ns=3, r=30, rz=10, seed=9)
    B = P["B"].copy()
    B[:, 3] = B[:, 4]                  # equal values in every row
    B[0, :] = 0.1                      # a row of ten equal values
    B[1, 5:] = 0.0                     # zeros: never kept
    P["B"] = B
    for k in (4, 10, 13):              # k > rz pads with -1 / 0
        Zg, Bg, sg = _gpu_sigma(P, k)
        Zh, Bh, sh = _host_sigma(P, k)
        assert bits(np.float32(sg)) == bits(np.float32(sh))
        assert np.array_equal(np.sort(bits(Bg), axis=1), np.sort(bits(Bh), axis=1))
        assert np.array_equal((Zg < 0), (Zh < 0))
        print("sigma gpu ties k=%d: masses and sigma equal the host's" % k, flush=True)
