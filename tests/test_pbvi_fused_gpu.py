"""The one-launch PBVI iteration of small instances (pbvi_fused), through the C ABI.

Small sparse instances run each iteration as one workgroup per belief instead of the transpose,
score, action and materialize kernels, with every sum and product rounded in nova's order
(src/pomdp/algorithms/pomdp_pbvi_cpu.cpp:37-162). The tests lift the path's work limit so that
every instance within its LDS bounds takes it. Checked here:
  * exact mode: Gamma, pi and the per-iteration alpha choices equal the four-kernel path
    (NOVA_AMD_PBVI_FUSED=0) and the float oracle bit for bit, on ragged successor lists and
    supports, an empty belief, one alpha and one observation, and n above the workgroup size;
  * fast mode: the run equals the float oracle bit for bit (the kernel keeps nova's order);
  * recorded runs (the argmax records of the replay gradient) equal the four-kernel path's;
  * the profiled iteration reports the whole iteration as its score stage and equals an update;
  * a batch mixing one-launch and four-kernel instances equals the plain solves in both modes.
"""
import numpy as np
import pytest

from conftest import bits
from nova_amd import synth
from nova_amd.autograd import pbvi_objective
from nova_amd.pomdp import POMDP
from nova_amd.pomdp_pbvi import EXACT, FAST, POMDPPBVIGPU, batch_execute, set_mode

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def one_launch_for_every_instance_that_fits(monkeypatch):
    """Lifts the work limit of the one-launch path (its LDS bounds stay): each test then sets
    NOVA_AMD_PBVI_FUSED=0 for the four-kernel side of its comparison."""
    monkeypatch.setenv("NOVA_AMD_PBVI_FUSED", "1000000")


def fused_fits(P):
    """The LDS bounds of the one-launch path."""
    n, m, ns = P["S"].shape
    z = P["O"].shape[2]
    r, rz = P["Z"].shape
    return (n <= 2048 and z <= 256 and z * r <= 6144 and rz <= 512 and rz * ns <= 512
            and r * n + m * (z + 1) * n * ns + n * m + 2 * rz <= 24576)


def ragged(P, seed):
    rng = np.random.default_rng(seed)
    S, Z = P["S"].copy(), P["Z"].copy()
    n, m, ns = S.shape
    cut = rng.random((n, m)) < 0.3
    pos = rng.integers(0, ns, size=(n, m))
    for s, a in zip(*np.nonzero(cut)):
        S[s, a, pos[s, a]:] = -1
    r, rz = Z.shape
    for i in np.nonzero(rng.random(r) < 0.3)[0]:
        if rz > 1:
            Z[i, rng.integers(1, rz):] = -1
    Z[0, :] = -1                                   # a belief with no support
    P["S"], P["Z"] = S, Z
    return P


CASES = {
    "tiny": dict(n=5, m=2, z=3, ns=5, r=7, rz=5, seed=32, gamma=0.9),
    "tiger_like": dict(n=2, m=3, z=2, ns=2, r=16, rz=2, seed=5, gamma=0.95),
    "one_alpha": dict(n=12, m=1, z=1, ns=3, r=1, rz=4, seed=6, gamma=0.9),
    "mid": dict(n=60, m=3, z=12, ns=6, r=60, rz=12, seed=31, gamma=0.95),
    "n_above_block": dict(n=300, m=2, z=4, ns=3, r=20, rz=8, seed=7, gamma=0.9),
}


def instance(name):
    P = synth.pomdp_arrays(horizon=1, **CASES[name])
    if name != "one_alpha":
        P = ragged(P, CASES[name]["seed"])
    assert fused_fits(P)
    return P


def updates(P, mode, steps):
    """Gamma, pi and the argmax records after each of `steps` single updates."""
    set_mode(mode)
    try:
        p = POMDP.from_dict(P)
        p.initialize_gpu()
        s = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"])
        out = []
        for _ in range(steps):
            s.update()
            js, act = s.argmax_record()
            pol = s.get_policy()
            out.append((pol.Gamma_array().copy(), pol.pi_array().copy(), js.copy(), act.copy()))
        del s
    finally:
        set_mode(FAST)
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_exact_fused_equals_four_kernels_and_oracle(oracle, monkeypatch, name):
    P = instance(name)
    steps = 6
    fused = updates(P, EXACT, steps)
    monkeypatch.setenv("NOVA_AMD_PBVI_FUSED", "0")
    four = updates(P, EXACT, steps)
    for t, (a, b) in enumerate(zip(fused, four)):
        assert np.array_equal(bits(a[0]), bits(b[0])), "Gamma differs at update %d" % t
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3])
        assert np.array_equal(a[2], b[2]), "alpha choices differ at update %d" % t
    Go, pio = oracle.pbvi_run(P, horizon=steps)
    assert np.array_equal(bits(fused[-1][0]), bits(Go)) and np.array_equal(fused[-1][1], pio)
    print("%s: %d exact updates, one-launch path = four kernels = oracle, bit for bit"
          % (name, steps), flush=True)


@pytest.mark.parametrize("name", ["tiny", "mid", "n_above_block"])
def test_fast_fused_equals_float_oracle(oracle, name):
    P = instance(name)
    P["horizon"] = 10
    set_mode(FAST)
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    s = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"])
    pol = s.solve()
    Go, pio = oracle.pbvi_run(P, horizon=10)
    assert np.array_equal(bits(pol.Gamma_array()), bits(Go))
    assert np.array_equal(pol.pi_array(), pio)


def test_recorded_run_fused_equals_four_kernels(monkeypatch):
    P = instance("mid")
    P["horizon"] = 8
    res = pbvi_objective(P, P["R"], mode=EXACT)
    recJ, recA, G = res["recJ"].cpu().numpy(), res["recA"].cpu().numpy(), res["Gamma"].cpu().numpy()
    monkeypatch.setenv("NOVA_AMD_PBVI_FUSED", "0")
    ref = pbvi_objective(P, P["R"], mode=EXACT)
    assert np.array_equal(recJ, ref["recJ"].cpu().numpy())
    assert np.array_equal(recA, ref["recA"].cpu().numpy())
    assert np.array_equal(bits(G), bits(ref["Gamma"].cpu().numpy()))


def test_profile_of_fused_iteration():
    P = instance("mid")
    set_mode(EXACT)
    try:
        p = POMDP.from_dict(P)
        p.initialize_gpu()
        s = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"])
        ms = s.profile()
        G = s.get_policy().Gamma_array().copy()
        del s
    finally:
        set_mode(FAST)
    assert ms[1] > 0 and ms[0] + ms[2] + ms[3] <= 0.05   # nothing runs between those events
    ref = updates(P, EXACT, 1)[0]
    assert np.array_equal(bits(G), bits(ref[0]))


@pytest.mark.parametrize("mode", [FAST, EXACT])
def test_batch_mixing_fused_and_four_kernel_instances(mode):
    batch = [instance("tiny"), instance("mid"), instance("n_above_block"),
             synth.pomdp_arrays(n=200, m=4, z=12, ns=5, r=400, rz=10, seed=44, gamma=0.95,
                                horizon=1),
             synth.pomdp_arrays(n=3000, m=2, z=3, ns=4, r=50, rz=6, seed=45, gamma=0.9,
                                horizon=1)]
    kinds = [fused_fits(P) for P in batch]
    assert kinds == [True, True, True, False, False]
    for k, P in enumerate(batch):
        P["horizon"] = 3 + k
    set_mode(mode)
    try:
        plain = []
        for P in batch:
            p = POMDP.from_dict(P)
            p.initialize_gpu()
            pol = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"]).solve()
            plain.append((pol.Gamma_array(), pol.pi_array()))
        solvers = []
        for P in batch:
            p = POMDP.from_dict(P)
            p.initialize_gpu()
            solvers.append(POMDPPBVIGPU(p, GammaInitial=P["Gamma0"]))
        res, _ = batch_execute(solvers)
    finally:
        set_mode(FAST)
    for (G, pi), (Gb, pib) in zip(plain, res):
        assert np.array_equal(bits(G), bits(Gb)) and np.array_equal(pi, pib)
