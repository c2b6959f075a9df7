"""Batched PBVI (pomdp_pbvi_batch_execute_gpu): many POMDPs of uneven size in one captured graph.
Every instance must equal its plain GPU solve bit for bit (same kernels, same arithmetic), and
in exact mode the float oracle's run."""
import numpy as np
import pytest

from conftest import bits
from nova_amd import synth
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_pbvi import EXACT, FAST, POMDPPBVIGPU, batch_execute, set_mode

pytestmark = pytest.mark.gpu


def small_batch():
    return synth.batch_instances(count=10, seed=9, n_range=(20, 300), r_range=(8, 700),
                                 horizon=6, m_choices=range(2, 6), z_choices=range(2, 20))


def solvers_for(batch):
    out = []
    for P in batch:
        p = POMDP.from_dict(P)
        p.initialize_gpu()
        out.append(POMDPPBVIGPU(p, GammaInitial=P["Gamma0"]))
    return out


@pytest.mark.parametrize("mode", [FAST, EXACT])
def test_batch_equals_plain(mode):
    set_mode(mode)
    batch = small_batch()
    # uneven horizons too
    for k, P in enumerate(batch):
        P["horizon"] = 3 + (k % 4)
    plain = []
    for s in solvers_for(batch):
        pol = s.solve()
        plain.append((pol.Gamma_array(), pol.pi_array()))
    res, ms = batch_execute(solvers_for(batch))
    set_mode(FAST)
    for (G, pi), (Gb, pib) in zip(plain, res):
        assert np.array_equal(bits(G), bits(Gb)) and np.array_equal(pi, pib)


def test_batch_exact_vs_oracle(oracle):
    set_mode(EXACT)
    batch = small_batch()[:4]
    res, _ = batch_execute(solvers_for(batch))
    set_mode(FAST)
    for P, (G, pi) in zip(batch, res):
        Go, pio = oracle.pbvi_run(P)
        assert np.array_equal(bits(G), bits(Go)) and np.array_equal(pi, pio)
