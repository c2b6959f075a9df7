"""Shared fixtures. Tests marked `gpu` need an MI355X; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle  # oracle/oracle.py
    return Oracle()


@pytest.fixture(scope="session")
def reflib():
    from oracle import RefLib
    if not RefLib.available():
        pytest.skip("oracle/_ref/libnova_ref.so is not built (needs the reference tree)")
    return RefLib()


@pytest.fixture(scope="session")
def lib():
    from nova_amd import _abi
    return _abi.load()


def bits(a):
    """float32 array as its raw bit patterns, for bit-exact comparisons."""
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


# ---- the near-tie rule for alpha and action choices made in another rounding order
#
# A score is a float sum of K products. Summed in two different orders (the oracle's left to
# right, and FMA or matrix-core partial sums), the two results differ by rounding: for K terms
# of total magnitude A the typical difference is about sqrt(K) * u * A (u = 2^-24, random-walk
# rounding), the worst case K * u * A. An absolute tie rule |gap| < 1e-6 lies below one
# float32 ulp once |value| >= 8 (ulp(8) = 9.5e-7, ulp(20) = 1.9e-6), so any second rounding
# order breaks it at these value scales. The rule used instead: two choices count as tied when
# the oracle's best-minus-chosen gap is at most
#       tie_rel(K) * max(1, |best value|),  tie_rel(K) = max(1e-5, 4 * sqrt(K) * u)
# which is 1e-5 for K <= 1,745 terms (sparse supports: K = rz * ns <= 512 at these workloads)
# and 3.4e-5 for the dense K = n = 20,000. Values are compared at 1e-4 relative.
UNIT_ROUNDOFF = 2.0 ** -24


def tie_rel(K):
    return max(1e-5, 4.0 * np.sqrt(float(K)) * UNIT_ROUNDOFF)


def check_choices(chosen, best, gap, scale, K, what):
    """Asserts that every choice differing from the reference's is a near-tie under tie_rel(K)
    and prints how many differ and their gaps relative to the rule. chosen/best: index arrays;
    gap: reference best-minus-runner-up (or best-minus-chosen) gaps, same shape; scale: the
    reference's |best value| (broadcastable). Returns the number of differing choices."""
    chosen, best = np.asarray(chosen), np.asarray(best)
    gap = np.asarray(gap, dtype=np.float64)
    lim = tie_rel(K) * np.maximum(1.0, np.abs(np.asarray(scale, dtype=np.float64)))
    lim = np.broadcast_to(lim, gap.shape)
    bad = chosen != best
    out = bad & ~(gap <= lim)
    nb = int(bad.sum())
    if nb:
        ratio = gap[bad] / lim[bad]
        hist = np.histogram(np.clip(ratio, 0, 1), bins=[0, 1e-3, 1e-2, 1e-1, 0.5, 1.0])[0]
        print("%s: %d of %d choices differ; gap / tie limit: max %.3g, histogram over "
              "[0,1e-3,1e-2,0.1,0.5,1] = %s" % (what, nb, bad.size, float(ratio.max()),
                                                 hist.tolist()), flush=True)
    else:
        print("%s: all %d choices equal the reference's" % (what, bad.size), flush=True)
    assert not out.any(), "%s: %d choices differ outside near-ties (tie_rel(%d) = %.3g)" % (
        what, int(out.sum()), K, tie_rel(K))
    return nb
