"""MDP value iteration on the GPU, through the C ABI, against the oracle.

The HIP sweep keeps nova's evaluation order (src/mdp/algorithms/mdp_vi_cpu.cpp:36-63), so V
and pi must equal the oracle's float twin bit for bit; the FP64 twin states the float error.
"""
import ctypes as ct

import numpy as np
import pytest

from conftest import bits
from nova_amd import synth
from nova_amd.mdp import MDP
from nova_amd.mdp_vi import MDPValueIterationGPU

pytestmark = pytest.mark.gpu


def three_state_mdp():
    """tests/unit/mdp/mdp_tests.cpp:63-153 (non-SSP rewards)."""
    S = np.array([[[0, 1], [0, -1]], [[1, -1], [2, -1]], [[2, -1], [2, -1]]], np.int32)
    T = np.array([[[0.5, 0.5], [1.0, 0.0]], [[1.0, 0.0], [1.0, 0.0]], [[1.0, 0.0], [1.0, 0.0]]],
                 np.float32)
    R = np.array([[-1.0, -2.0], [-1.0, -2.0], [0.0, 0.0]], np.float32)
    return MDP.from_arrays(S, T, R, gamma=0.9, horizon=5)


def solve_gpu(mdp, v0=None, threads=512):
    vi = MDPValueIterationGPU(mdp, numThreads=threads,
                              Vinitial=np.zeros(mdp.n, np.float32) if v0 is None else v0)
    pol = vi.solve()
    return pol.V_array(), pol.pi_array()


def test_known_answers():
    # tests/unit/mdp/algorithms/mdp_vi_gpu_tests.cpp:135-253: V = 2.71 (simple), and
    # V = [-3.303775, -2, 0], pi0 = 0, pi1 = 1 (three-state), numThreads = 512.
    simple = MDP.from_arrays(np.zeros((1, 1, 1), np.int32), np.ones((1, 1, 1), np.float32),
                             np.ones((1, 1), np.float32), gamma=0.9, horizon=3)
    simple.initialize_gpu()
    V, pi = solve_gpu(simple)
    assert abs(V[0] - 2.71) < 1e-5 and pi[0] == 0
    mdp = three_state_mdp()
    mdp.initialize_gpu()
    V, pi = solve_gpu(mdp)
    np.testing.assert_allclose(V, [-3.303775, -2.0, 0.0], atol=1e-5)
    assert pi[0] == 0 and pi[1] == 1


def test_per_step_values():
    # mdp_vi_gpu_tests.cpp:383-516 step-by-step V after each update (three-state MDP).
    mdp = three_state_mdp()
    mdp.initialize_gpu()
    vi = MDPValueIterationGPU(mdp, numThreads=512, Vinitial=np.zeros(3, np.float32))
    expect = [[-1, -1, 0], [-1.9, -1.9, 0], [-2.71, -2.0, 0]]
    for h, e in enumerate(expect):
        vi.update()
        assert vi.currentHorizon == h + 1
        np.testing.assert_allclose(vi.get_policy().V_array(), e, atol=1e-5)


@pytest.mark.parametrize("n,m,ns,horizon,seed", [
    (1000, 8, 16, 40, 11),      # the C2 shape, small n
    (777, 3, 12, 25, 12),       # m not a power of two, ns % 4 == 0, G*E > ns
    (513, 5, 7, 30, 13),        # ns odd: one entry per lane
    (300, 2, 1, 20, 14),        # ns = 1
    (2000, 1, 64, 10, 15),      # m = 1, wide rows
    (400, 4, 300, 5, 16),       # ns > 256: the one-thread-per-state kernel
])
def test_bitwise_vs_oracle(oracle, n, m, ns, horizon, seed):
    w = synth.mdp_arrays(n=n, m=m, ns=ns, seed=seed, gamma=0.95)
    S = w["S"].copy()
    # ragged lists: cut some rows short with a -1 (entries after it must be ignored, even >= 0)
    rng = np.random.default_rng(seed)
    cut = rng.random((n, m)) < 0.3
    pos = rng.integers(0, ns, size=(n, m))
    for s, a in zip(*np.nonzero(cut)):
        S[s, a, pos[s, a]] = -1
    v0 = rng.uniform(-5, 5, n).astype(np.float32)
    mdp = MDP.from_arrays(S, w["T"], w["R"], gamma=w["gamma"], horizon=horizon)
    mdp.initialize_gpu()
    V, pi = solve_gpu(mdp, v0)
    Vo, pio = oracle.vi_run(S, w["T"], w["R"], w["gamma"], horizon, V0=v0)
    assert np.array_equal(bits(V), bits(Vo)), "V differs from the float oracle"
    assert np.array_equal(pi, pio)
    V64, _ = oracle.vi_run(S, w["T"], w["R"], w["gamma"], horizon, V0=v0, dtype=np.float64)
    err = np.max(np.abs(V - V64)) / max(1.0, np.max(np.abs(V64)))
    print("n=%d m=%d ns=%d: ||V - V_fp64||inf/||V||inf = %.3g" % (n, m, ns, err))
    assert err < 1e-4


def test_run_gpu_matches_updates(oracle):
    """mdp_vi_run_gpu (graph of sweeps, restart from a device V0) equals repeated updates."""
    import torch
    w = synth.mdp_arrays(n=5000, m=8, ns=16, seed=21, gamma=0.99)
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=7)
    mdp.initialize_gpu()
    vi = MDPValueIterationGPU(mdp, Vinitial=np.zeros(mdp.n, np.float32))
    d_v0 = torch.zeros(mdp.n, dtype=torch.float32, device="cuda")
    for sweeps in (7, 8, 7):
        vi.run(sweeps, d_V0=d_v0.data_ptr())
        assert vi.currentHorizon == sweeps
        V, pi = vi.device_values()
        Vo, pio = oracle.vi_run(w["S"], w["T"], w["R"], w["gamma"], sweeps)
        assert np.array_equal(bits(V), bits(Vo)) and np.array_equal(pi, pio)


def test_full_size_properties():
    """n = 1e6, m = 8, ns = 16: properties that do not need the CPU oracle at this size.
    (1) Bellman contraction: ||V_{k+1} - V_k|| <= gamma ||V_k - V_{k-1}||; (2) V stays within
    [min R/(1-gamma), max R/(1-gamma)]; (3) re-running gives identical bits (no atomics)."""
    w = synth.workload("vi_n1e6")
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=1)
    mdp.initialize_gpu()
    vi = MDPValueIterationGPU(mdp, Vinitial=np.zeros(mdp.n, np.float32))
    prev = np.zeros(mdp.n, np.float32)
    diffs = []
    for k in range(6):
        vi.update()
        V, _ = vi.device_values()
        diffs.append(np.max(np.abs(V.astype(np.float64) - prev)))
        prev = V
    g = float(w["gamma"])
    for a, b in zip(diffs[1:], diffs[2:]):
        assert b <= g * a * (1 + 1e-5) + 1e-6
    vi2 = MDPValueIterationGPU(mdp, Vinitial=np.zeros(mdp.n, np.float32))
    for _ in range(6):
        vi2.update()
    V2, _ = vi2.device_values()
    assert np.array_equal(bits(V2), bits(prev))
    lo, hi = w["R"].min() / (1 - g), w["R"].max() / (1 - g)
    assert prev.min() >= lo - 1e-3 and prev.max() <= hi + 1e-3


def test_oracle_sample_at_full_size(oracle):
    """Bitwise check of one full-size sweep on a sample of states, against the oracle sweep
    restricted to those states (the sweep of one state depends only on V)."""
    w = synth.workload("vi_n1e6")
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=1)
    mdp.initialize_gpu()
    rng = np.random.default_rng(3)
    v0 = rng.uniform(-50, 0, mdp.n).astype(np.float32)
    vi = MDPValueIterationGPU(mdp, Vinitial=v0)
    vi.update()
    V, pi = vi.device_values()
    idx = np.sort(rng.choice(mdp.n, 20000, replace=False))
    Vo, pio = oracle.vi_sweep(w["S"][idx], w["T"][idx], w["R"][idx], w["gamma"], v0)
    assert np.array_equal(bits(V[idx]), bits(Vo)) and np.array_equal(pi[idx], pio)


@pytest.mark.parametrize("n,m,ns,horizon,seed", [
    (1000, 8, 16, 40, 31),      # 128 states per tile, eight tiles
    (2000, 3, 12, 25, 32),      # 455 states per tile: a partial fifth tile
    (2000, 1, 64, 10, 33),      # m = 1, wide rows
    (300, 16, 200, 6, 34),      # m * ns = 3,200: five states per tile
])
def test_sorted_tiles_bitwise_vs_oracle(oracle, monkeypatch, n, m, ns, horizon, seed):
    """The successor-sorted layout, forced at small sizes: same V and pi bit for bit, with
    ragged successor lists (entries after a -1 must stay out of the layout)."""
    monkeypatch.setenv("NOVA_AMD_VI_KERNEL", "sorted")
    w = synth.mdp_arrays(n=n, m=m, ns=ns, seed=seed, gamma=0.95)
    S = w["S"].copy()
    rng = np.random.default_rng(seed)
    cut = rng.random((n, m)) < 0.3
    pos = rng.integers(0, ns, size=(n, m))
    for s, a in zip(*np.nonzero(cut)):
        S[s, a, pos[s, a]] = -1
    S[0, 0, :] = -1                                  # an empty row
    v0 = rng.uniform(-5, 5, n).astype(np.float32)
    mdp = MDP.from_arrays(S, w["T"], w["R"], gamma=w["gamma"], horizon=horizon)
    mdp.initialize_gpu()
    vi = MDPValueIterationGPU(mdp, numThreads=512, Vinitial=v0)
    assert vi.sweep_kernel() == 2
    pol = vi.solve()
    Vo, pio = oracle.vi_run(S, w["T"], w["R"], w["gamma"], horizon, V0=v0)
    assert np.array_equal(bits(pol.V_array()), bits(Vo)) and np.array_equal(pol.pi_array(), pio)


@pytest.mark.parametrize("n,m,ns,horizon,seed,chunk,piece,parts,tile", [
    (1000, 8, 16, 40, 41, 4096, 8192, 1, 16384),     # one successor chunk
    (9000, 3, 12, 25, 42, 4096, 8192, 1, 16384),     # three chunks, a partial last tile and chunk
    (20000, 8, 16, 12, 43, 4096, 16384, 1, 16384),   # five chunks, several pieces per chunk
    (20000, 8, 16, 12, 43, 4096, 16384, 4, 16384),   # the same in four row parts
    (5000, 1, 5, 10, 44, 4096, 8192, 3, 16384),      # m = 1, ns = 5 (rows not a multiple of 4)
    (300, 16, 200, 6, 45, 4096, 8192, 1, 16384),     # m * ns = 3,200: five states per tile
    (70000, 2, 8, 5, 46, 32768, 131072, 1, 16384),   # the default chunk and piece sizes
    (20000, 8, 16, 12, 43, 4096, 16384, 1, 8192),    # tiles of 8,192 slots
    (9000, 3, 12, 25, 42, 4096, 8192, 3, 12288),     # tiles of 12,288 slots, three parts
    (5000, 4, 8, 10, 48, 4096, 8192, 1, 4096),       # tiles of 4,096 slots
    (300, 16, 200, 6, 45, 4096, 8192, 1, 4096),      # m * ns = 3,200: one state per tile
])
def test_chunked_layout_bitwise_vs_oracle(oracle, monkeypatch, n, m, ns, horizon, seed, chunk,
                                          piece, parts, tile):
    """The successor-chunked layout (V chunks in LDS, products through HBM), forced at small
    sizes with small chunks, whole or in row parts, with tiles of several slot counts: same V
    and pi bit for bit, with ragged successor lists and an empty row."""
    monkeypatch.setenv("NOVA_AMD_VI_KERNEL", "chunked")
    monkeypatch.setenv("NOVA_AMD_VI_CHUNK", str(chunk))
    monkeypatch.setenv("NOVA_AMD_VI_PIECE", str(piece))
    monkeypatch.setenv("NOVA_AMD_VI_PARTS", str(parts))
    monkeypatch.setenv("NOVA_AMD_VI_TILE", str(tile))
    print("chunked layout n=%d m=%d ns=%d chunk=%d piece=%d parts=%d tile=%d"
          % (n, m, ns, chunk, piece, parts, tile), flush=True)
    w = synth.mdp_arrays(n=n, m=m, ns=ns, seed=seed, gamma=0.95)
    S = w["S"].copy()
    rng = np.random.default_rng(seed)
    cut = rng.random((n, m)) < 0.3
    pos = rng.integers(0, ns, size=(n, m))
    for s, a in zip(*np.nonzero(cut)):
        S[s, a, pos[s, a]] = -1
    S[0, 0, :] = -1                                  # an empty row
    v0 = rng.uniform(-5, 5, n).astype(np.float32)
    mdp = MDP.from_arrays(S, w["T"], w["R"], gamma=w["gamma"], horizon=horizon)
    mdp.initialize_gpu()
    vi = MDPValueIterationGPU(mdp, numThreads=512, Vinitial=v0)
    assert vi.sweep_kernel() == 3
    pol = vi.solve()
    Vo, pio = oracle.vi_run(S, w["T"], w["R"], w["gamma"], horizon, V0=v0)
    assert np.array_equal(bits(pol.V_array()), bits(Vo)) and np.array_equal(pol.pi_array(), pio)


def test_full_size_layouts_equal_tiled(monkeypatch):
    """At n = 1e6, m = 8, ns = 16 the default is the successor-chunked layout; four sweeps of
    it, of the successor-sorted layout and of the wave-tiled kernel are equal bit for bit."""
    w = synth.workload("vi_n1e6")
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=1)
    mdp.initialize_gpu()
    rng = np.random.default_rng(5)
    v0 = rng.uniform(-50, 0, mdp.n).astype(np.float32)
    out = {}
    for kernel in ("default", "sorted", "tiled"):
        if kernel != "default":
            monkeypatch.setenv("NOVA_AMD_VI_KERNEL", kernel)
        vi = MDPValueIterationGPU(mdp, Vinitial=v0)
        out[kernel] = vi.sweep_kernel()
        for _ in range(4):
            vi.update()
        V, pi = vi.device_values()
        out[kernel + "_V"], out[kernel + "_pi"] = V, pi
        del vi
        print("n=1e6: %s layout (kernel %d) done" % (kernel, out[kernel]), flush=True)
    assert out["default"] == 3 and out["sorted"] == 2 and out["tiled"] == 1
    for k in ("sorted", "tiled"):
        assert np.array_equal(bits(out["default_V"]), bits(out[k + "_V"]))
        assert np.array_equal(out["default_pi"], out[k + "_pi"])


def test_layouts_rebuilt_with_other_settings(oracle, monkeypatch):
    """Solvers of one model built one after another with other layout settings (sorted, chunks
    of 32,768 and of 4,096 states, several row parts), each freed before the next is built, as
    a tuning loop does: each run equals the oracle bit for bit. A freed layout's memory can be
    reused by the next one at the same address, so the cached sweep graph must be keyed by the
    layout itself (a graph captured on another layout's chunking would read past V)."""
    w = synth.mdp_arrays(n=70000, m=2, ns=8, seed=47, gamma=0.95)
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=6)
    mdp.initialize_gpu()
    v0 = np.random.default_rng(47).uniform(-5, 5, mdp.n).astype(np.float32)
    Vo, pio = oracle.vi_run(w["S"], w["T"], w["R"], w["gamma"], 6, V0=v0)
    import torch
    d0 = torch.from_numpy(v0).cuda()
    for setting in (dict(NOVA_AMD_VI_KERNEL="sorted"),
                    dict(NOVA_AMD_VI_KERNEL="chunked", NOVA_AMD_VI_CHUNK="32768"),
                    dict(NOVA_AMD_VI_KERNEL="chunked", NOVA_AMD_VI_CHUNK="4096"),
                    dict(NOVA_AMD_VI_KERNEL="chunked", NOVA_AMD_VI_CHUNK="4096",
                         NOVA_AMD_VI_PARTS="3"),
                    dict(NOVA_AMD_VI_KERNEL="chunked", NOVA_AMD_VI_CHUNK="4096",
                         NOVA_AMD_VI_TILE="8192"),
                    dict(NOVA_AMD_VI_KERNEL="chunked", NOVA_AMD_VI_CHUNK="32768")):
        for k in ("NOVA_AMD_VI_KERNEL", "NOVA_AMD_VI_CHUNK", "NOVA_AMD_VI_PARTS",
                  "NOVA_AMD_VI_TILE"):
            monkeypatch.delenv(k, raising=False)
        for k, v in setting.items():
            monkeypatch.setenv(k, v)
        print("layout setting %s" % setting, flush=True)
        vi = MDPValueIterationGPU(mdp, Vinitial=v0)
        for _ in range(2):
            vi.run(6, d_V0=d0.data_ptr())
            V, pi = vi.device_values()
            assert np.array_equal(bits(V), bits(Vo)) and np.array_equal(pi, pio), setting
        del vi
