"""Concurrent calls from several host threads (ctypes releases the GIL during each call): the
library's GPU entry points share per-process caches and one stream per device and serialise
on a process-wide lock, and graphs are captured on a stream of their own, which model uploads
from other threads (synchronous copies on the legacy null stream) cannot invalidate. Each
thread's VI and PBVI solves equal the same solves run one after another, bit for bit."""
import threading

import numpy as np
import pytest

from conftest import bits
from nova_amd import synth
from nova_amd.mdp import MDP
from nova_amd.mdp_vi import MDPValueIterationGPU
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_pbvi import FAST, POMDPPBVIGPU, set_mode

pytestmark = pytest.mark.gpu


def vi_solve(seed):
    w = synth.mdp_arrays(n=3000 + 100 * seed, m=4, ns=8, seed=seed, gamma=0.95)
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=30)
    mdp.initialize_gpu()
    vi = MDPValueIterationGPU(mdp, Vinitial=np.zeros(mdp.n, np.float32))
    pol = vi.solve()
    return pol.V_array().copy(), pol.pi_array().copy()


def pbvi_solve(seed):
    P = synth.pomdp_arrays(n=200 + 10 * seed, m=3, z=4, ns=4, r=64, rz=8, seed=seed, gamma=0.95,
                           horizon=6)
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    s = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"])
    pol = s.solve()
    return pol.Gamma_array().copy(), pol.pi_array().copy()


def test_concurrent_solves_equal_sequential():
    set_mode(FAST)
    seeds = [1, 2, 3, 4]
    want = {("vi", k): vi_solve(k) for k in seeds}
    want.update({("pbvi", k): pbvi_solve(k) for k in seeds})
    got, errors = {}, []

    def work(k):
        try:
            for _ in range(4):
                got[("vi", k)] = vi_solve(k)
                got[("pbvi", k)] = pbvi_solve(k)
        except Exception as e:          # reported by the main thread
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=work, args=(k,)) for k in seeds]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in threads), "a solve did not return"
    assert not errors, errors
    for key, (a, b) in want.items():
        ga, gb = got[key]
        assert np.array_equal(bits(ga), bits(a)) and np.array_equal(gb, b), key
    print("4 threads x (VI + PBVI) x 4: equal to the sequential solves", flush=True)
