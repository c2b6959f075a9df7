"""pomdp_expand_pema_gpu against the library's host PEMA (itself bit for bit nova's build in
nova's alpha choice, and pinned by an FP64 restatement in both choices: test_pema_sigma.py)
and, in nova's alpha choice, directly against nova's own build (oracle/_ref): the appended
beliefs must be equal bit for bit, in both alpha choices, over repeated calls."""
import ctypes as ct
import os

import numpy as np
import pytest

from conftest import bits
from nova_amd import _abi, synth
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_alpha_vectors import POMDPAlphaVectors

pytestmark = pytest.mark.gpu

DOMAINS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "domains")


def policy(G, pi, m):
    p = POMDPAlphaVectors()
    r, n = G.shape
    assert _abi.load().pomdp_alpha_vectors_initialize(ct.byref(p), n, m, r) == 0
    np.ctypeslib.as_array(p.Gamma, (r * n,))[:] = np.asarray(G, np.float32).reshape(-1)
    np.ctypeslib.as_array(p.pi, (r,))[:] = pi
    return p


def run(P, G, pi, steps, mode, gpu):
    lib = _abi.load()
    assert lib.nova_amd_pema_mode(mode) == 0
    try:
        p = POMDP.from_dict(P)
        if gpu:
            p.initialize_gpu()       # expand() keeps the device copy of the set current
        pol = policy(G, pi, P["S"].shape[1])
        for _ in range(steps):
            p.expand("pema", pemaPolicy=pol, gpu=gpu)
        return p.beliefs()
    finally:
        lib.nova_amd_pema_mode(0)


CASES = [dict(n=30, m=3, z=4, ns=4, r=10, rz=5, seed=1),
         dict(n=100, m=4, z=6, ns=6, r=200, rz=12, seed=2),
         dict(n=257, m=2, z=9, ns=5, r=70, rz=20, seed=3)]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_pema_gpu_bitwise_vs_host(mode, case):
    c = CASES[case]
    P = synth.pomdp_arrays(gamma=0.9, horizon=5, **c)
    rng = np.random.default_rng(c["seed"])
    G = rng.uniform(-5, 5, (9, c["n"])).astype(np.float32)
    if mode == 0 and case == 1:
        G[4, 7] = -3e10          # a product below -1e10: nova's loop then picks alpha 4
    pi = rng.integers(0, c["m"], 9).astype(np.uint32)
    print("PEMA case %d mode %d: r=%d n=%d" % (case, mode, c["r"], c["n"]), flush=True)
    Zh, Bh = run(P, G, pi, 2, mode, gpu=False)
    Zg, Bg = run(P, G, pi, 2, mode, gpu=True)
    assert np.array_equal(Zh, Zg) and np.array_equal(bits(Bh), bits(Bg))


def test_pema_gpu_tiger_and_tiger_grid(oracle):
    """tiger_95 with four calls (tests/tiger/tiger.py:62-66) and tiger_grid after a seeded
    expansion to 64 beliefs, against the oracle's PBVI policy."""
    for name, grow in (("tiger_95", 0), ("tiger_grid", 63)):
        p = POMDP().load(os.path.join(DOMAINS, name + ".pomdp"))
        if grow:
            p.initialize_gpu()
            p.expand("random", grow, gpu=True, seed=5)
        P = p.arrays()
        P["Gamma0"] = np.full((P["Z"].shape[0], P["S"].shape[0]),
                              float(P["R"].min()) / (1 - float(P["gamma"])), np.float32)
        G, pi = oracle.pbvi_run(P, horizon=20)
        for mode in (0, 1):
            Zh, Bh = run(P, G, pi, 4 if name == "tiger_95" else 2, mode, gpu=False)
            Zg, Bg = run(P, G, pi, 4 if name == "tiger_95" else 2, mode, gpu=True)
            assert np.array_equal(Zh, Zg) and np.array_equal(bits(Bh), bits(Bg)), (name, mode)
        print("%s: GPU PEMA equals the host's in both alpha choices" % name, flush=True)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_pema_gpu_bitwise_vs_nova_build(reflib, case):
    """nova's alpha choice (mode 0): the GPU PEMA appends the beliefs nova's own build
    (oracle/_ref, pomdp_expand_pema_cpu, src/pomdp/utilities/pomdp_expand_cpu.cpp:269-466)
    appends, bit for bit, over two calls."""
    c = CASES[case]
    P = synth.pomdp_arrays(gamma=0.9, horizon=5, **c)
    rng = np.random.default_rng(c["seed"] + 10)
    G = rng.uniform(-5, 5, (7, c["n"])).astype(np.float32)
    pi = rng.integers(0, c["m"], 7).astype(np.uint32)
    Zg, Bg = run(P, G, pi, 2, 0, gpu=True)
    Zr, Br = reflib.expand_pema(P, G, pi, steps=2)
    assert np.array_equal(Zg, Zr) and np.array_equal(bits(Bg), bits(Br))
    print("PEMA case %d: GPU equals nova's build over two calls" % case, flush=True)
