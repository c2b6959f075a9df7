"""The .pomdp and raw CSV loaders (nova_amd/file_loader.py; the reference's loader is
python/nova/file_loader.py:31-587). Known answers derived by hand from the files, the format's
invariants on every shipped domain, the entry-override semantics on a small synthetic file, and
a raw-format round trip. CPU only."""
import math
import os

import numpy as np
import pytest

from nova_amd.file_loader import FileLoader

DOMAINS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "domains")
FILES = sorted(f for f in os.listdir(DOMAINS) if f.endswith(".pomdp"))


def load(name):
    f = FileLoader()
    f.load_cassandra(os.path.join(DOMAINS, name))
    return f


def test_tiger_known_answers():
    f = load("tiger_95.pomdp")
    assert (f.n, f.m, f.z, f.ns, f.r, f.rz) == (2, 3, 2, 2, 1, 2)
    assert f.gamma == pytest.approx(0.95)
    # listen: identity; open-left / open-right: uniform over both states
    np.testing.assert_array_equal(f.S[:, 0], [[0, -1], [1, -1]])
    np.testing.assert_array_equal(f.T[:, 0], [[1, 0], [1, 0]])
    for a in (1, 2):
        np.testing.assert_array_equal(f.S[:, a], [[0, 1], [0, 1]])
        np.testing.assert_allclose(f.T[:, a], 0.5)
    np.testing.assert_allclose(f.O[0], [[0.85, 0.15], [0.15, 0.85]], rtol=1e-7)
    np.testing.assert_allclose(f.O[1:], 0.5)
    np.testing.assert_array_equal(f.R, [[-1, -100, 10], [-1, 10, -100]])
    np.testing.assert_array_equal(f.Z, [[0, 1]])
    np.testing.assert_allclose(f.B, [[0.5, 0.5]])
    # horizon bringing the values within (Rmax - Rmin) / 1000 of optimal
    assert f.horizon == int(math.log(1e-3) / math.log(0.95)) + 1 == 135


def test_4x3_known_answers():
    f = load("4x3_95.pomdp")
    assert (f.n, f.m, f.z) == (11, 4, 6)
    expect = np.full(11, -0.04, np.float32)
    expect[3], expect[6] = 1.0, -1.0
    for a in range(4):
        np.testing.assert_allclose(f.R[:, a], expect, rtol=1e-7)
    assert f.r == 1 and f.rz == 11


@pytest.mark.parametrize("name", FILES)
def test_domain_invariants(name):
    f = load(name)
    n, m, ns = f.S.shape
    assert (n, m) == (f.n, f.m) and ns == f.ns
    valid = f.S >= 0
    # a row is a valid prefix followed by -1 padding (with zero probability)
    assert np.all(valid[..., :-1] | ~valid[..., 1:])
    assert np.all(f.S[valid] < n) and np.all(f.T[~valid] == 0)
    np.testing.assert_allclose(np.where(valid, f.T, 0).sum(axis=2), 1.0, atol=1e-5)
    assert f.O.shape == (f.m, f.n, f.z)
    np.testing.assert_allclose(f.O.sum(axis=2), 1.0, atol=1e-5)
    assert f.Z.shape == (1, f.rz) and np.all((f.Z >= 0) & (f.Z < n))
    assert float(f.B.sum()) == pytest.approx(1.0, abs=1e-5)
    assert f.Rmin == float(f.R.min()) and f.Rmax == float(f.R.max())


SYNTHETIC = """# wildcards, overrides and shorthands
discount: 0.5
values: reward
states: a b c
actions: 2
observations: x y
start exclude: b

T: * : a
uniform
T: 0 : b : c 1.0
T: 1 : b : b 1.0
T: 0 : c : a 0.25
T: 0 : c : c 0.75
T: 0 : c : a 0.125
T: 0 : c : b 0.125
T: 1 : c : c 1.0

O: *
uniform
O: 1 : c : x 0.9
O: 1 : c : y 0.1

R: * : * : * : * -1
R: 0 : a : * : * 2
R: 1 : c
3 5
"""


def test_synthetic_semantics(tmp_path):
    path = tmp_path / "synthetic.pomdp"
    path.write_text(SYNTHETIC)
    f = FileLoader()
    f.load_cassandra(str(path))
    assert (f.n, f.m, f.z) == (3, 2, 2) and f.gamma == 0.5
    # start exclude b: uniform over a, c
    np.testing.assert_array_equal(f.Z, [[0, 2]])
    np.testing.assert_allclose(f.B, [[0.5, 0.5]])
    # T: * : a uniform -> both actions from a to every state with 1/3
    for a in range(2):
        np.testing.assert_array_equal(f.S[0, a, :3], [0, 1, 2])
        np.testing.assert_allclose(f.T[0, a, :3], 1 / 3)
    # T: 0 : c : a repeats: the key keeps its first position and its last value (0.125)
    np.testing.assert_array_equal(f.S[2, 0, :3], [0, 2, 1])
    np.testing.assert_array_equal(f.T[2, 0, :3], [0.125, 0.75, 0.125])
    np.testing.assert_array_equal(f.S[1, :, 0], [2, 1])
    np.testing.assert_array_equal(f.S[2, 1, :2], [2, -1])
    # O: * uniform, then (1, c) overridden
    np.testing.assert_allclose(f.O[0], 0.5)
    np.testing.assert_allclose(f.O[1, :2], 0.5)
    np.testing.assert_allclose(f.O[1, 2], [0.9, 0.1], rtol=1e-6)
    # R: -1 everywhere, then R[a, 0] = 2, then R[c, 1] = mean(3, 5) = 4
    np.testing.assert_array_equal(f.R, [[2, -1], [-1, -1], [-1, 4]])
    assert f.horizon == int(math.log(1e-3) / math.log(0.5)) + 1 == 10


def test_identity_shorthand(tmp_path):
    path = tmp_path / "identity.pomdp"
    path.write_text("discount: 0.9\nstates: 3\nactions: 1\nobservations: 1\nstart: uniform\n"
                    "T: 0\nidentity\nO: 0\nuniform\nR: 0 : * : * : * 1\n")
    f = FileLoader()
    f.load_cassandra(str(path))
    np.testing.assert_array_equal(f.S[:, 0, 0], [0, 1, 2])
    np.testing.assert_array_equal(f.T[:, 0, 0], 1.0)
    np.testing.assert_array_equal(f.Z, [[0, 1, 2]])
    np.testing.assert_allclose(f.B, 1 / 3)
    assert f.R.min() == f.R.max() == 1.0 and f.horizon == 0


def _write_raw_pomdp(path, P, k=2):
    S, T, O, R, Z, B = P["S"], P["T"], P["O"], P["R"], P["Z"], P["B"]
    n, m, ns = S.shape
    z = O.shape[2]
    r, rz = Z.shape
    rows = [[n, ns, m, z, r, rz, k, 0, P["horizon"], float(P["gamma"])]]
    rows += [list(S[s, a]) for a in range(m) for s in range(n)]
    rows += [[repr(float(x)) for x in T[s, a]] for a in range(m) for s in range(n)]
    rows += [[repr(float(x)) for x in O[a, :, o]] for a in range(m) for o in range(z)]
    for i in range(k):
        rows += [[repr(float(x) * (i + 1)) for x in R[:, a]] for a in range(m)]
    rows += [list(Z[i]) for i in range(r)]
    rows += [[repr(float(x)) for x in B[i]] for i in range(r)]
    path.write_text("\n".join(",".join(str(x) for x in row) for row in rows) + "\n")


def test_raw_pomdp_round_trip(tmp_path):
    from nova_amd import synth
    P = synth.pomdp_arrays(n=12, m=3, z=4, ns=3, r=5, rz=4, seed=9, gamma=0.9, horizon=7)
    path = tmp_path / "p.csv"
    _write_raw_pomdp(path, P)
    f = FileLoader()
    f.load_raw_pomdp(str(path))
    for key in ("S", "T", "O", "R", "Z", "B"):
        np.testing.assert_array_equal(getattr(f, key), P[key])
    assert (f.horizon, f.k) == (7, 2) and f.gamma == pytest.approx(0.9)
    g = FileLoader()
    g.load_raw_pomdp(str(path), scalarize=lambda x: x[1])     # the second reward
    np.testing.assert_array_equal(g.R, P["R"] * 2)


def test_raw_mdp_round_trip(tmp_path):
    from nova_amd import synth
    w = synth.mdp_arrays(n=9, m=2, ns=3, seed=3, gamma=0.8, horizon=5)
    S, T, R = w["S"], w["T"], w["R"]
    n, m, ns = S.shape
    rows = [[n, ns, m, 1, 4, 2, 5, 0.8], [1, 7]]
    rows += [list(S[s, a]) for a in range(m) for s in range(n)]
    rows += [[repr(float(x)) for x in T[s, a]] for a in range(m) for s in range(n)]
    rows += [[repr(float(x)) for x in R[:, a]] for a in range(m)]
    path = tmp_path / "m.csv"
    path.write_text("\n".join(",".join(str(x) for x in row) for row in rows) + "\n")
    f = FileLoader()
    f.load_raw_mdp(str(path))
    np.testing.assert_array_equal(f.S, S)
    np.testing.assert_array_equal(f.T, T)
    np.testing.assert_array_equal(f.R, R)
    np.testing.assert_array_equal(f.goals, [1, 7])
    assert (f.s0, f.ng, f.horizon) == (4, 2, 5)


def test_bad_file_is_reported(tmp_path):
    path = tmp_path / "bad.pomdp"
    path.write_text("discount: 0.9\nstates: 2\nactions: 1\nobservations: 1\nstart: uniform\n"
                    "T: 0 : nowhere : 0 1.0\n")
    with pytest.raises(ValueError):
        FileLoader().load_cassandra(str(path))


# ------------------------------------------------------------- nova's raw CSV test models

RAW = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "raw")


def test_raw_tiger_is_the_cassandra_tiger():
    """tests/tiger/tiger_pomdp.raw and tiger_95.pomdp describe one model in two formats."""
    raw = FileLoader()
    raw.load_raw_pomdp(os.path.join(RAW, "tiger_pomdp.raw"))
    cas = load("tiger_95.pomdp")
    for key in ("S", "T", "O", "R"):
        np.testing.assert_array_equal(getattr(raw, key), getattr(cas, key), err_msg=key)
    assert (raw.n, raw.m, raw.z, raw.r, raw.rz, raw.horizon) == (2, 3, 2, 21, 2, 250)
    assert raw.gamma == pytest.approx(0.95)
    # 21 beliefs sweeping b(s1) from 0 to 1: the first and last as one-state supports
    np.testing.assert_array_equal(raw.Z[0], [0, -1])
    np.testing.assert_array_equal(raw.Z[20], [1, -1])
    np.testing.assert_array_equal(raw.Z[1:20], np.tile([0, 1], (19, 1)))
    i = np.arange(1, 20)
    np.testing.assert_allclose(raw.B[1:20, 1], 0.05 * i, atol=1e-7)
    np.testing.assert_allclose(raw.B.sum(axis=1), 1.0, atol=1e-6)


def test_raw_grid_world_header():
    f = FileLoader()
    f.load_raw_mdp(os.path.join(RAW, "grid_world_mdp.raw"))
    assert (f.n, f.ns, f.m, f.k, f.s0, f.ng, f.horizon) == (13, 3, 4, 1, 8, 1, 250)
    assert f.gamma == pytest.approx(0.9)
    np.testing.assert_array_equal(f.goals, [12])
    valid = f.S >= 0
    np.testing.assert_allclose(np.where(valid, f.T, 0).sum(axis=2), 1.0, atol=1e-6)
    g = FileLoader()
    g.load_raw_pomdp(os.path.join(RAW, "grid_world_pomdp.raw"))
    assert (g.n, g.ns, g.m, g.z, g.r, g.rz, g.s0) == (13, 3, 4, 3, 21, 4, 8)
    np.testing.assert_array_equal(g.S, f.S)
    np.testing.assert_allclose(g.O.sum(axis=2), 1.0, atol=1e-6)


@pytest.mark.parametrize("name", ["grid_world_mdp.raw", "another_grid_world_mdp.raw"])
def test_grid_world_vi_cpu(reflib, oracle, name):
    """The library's CPU value iteration on nova's grid worlds: bit-identical to nova's own
    CPU path, and after 250 sweeps at gamma = 0.9 (contraction 0.9^250 ~ 4e-12) at the FP64
    Bellman fixed point within float rounding."""
    from conftest import bits
    from nova_amd.mdp import MDP
    from nova_amd.mdp_vi import MDPValueIterationCPU, _default_v0
    mdp = MDP().load(os.path.join(RAW, name), filetype="raw")
    S, T, R = mdp.arrays()
    pol = MDPValueIterationCPU(mdp).solve()
    V, pi = pol.V_array(), pol.pi_array()
    Vr, pir = reflib.vi_execute(S, T, R, mdp.gamma, mdp.horizon, V0=_default_v0(mdp))
    assert np.array_equal(bits(V), bits(Vr)) and np.array_equal(pi, pir)
    Vs, _, _, res = oracle.vi_solve_f64(S, T, R, mdp.gamma)
    assert res < 1e-10
    assert np.max(np.abs(V - Vs)) <= 1e-5 * max(1.0, np.max(np.abs(Vs)))


def test_tiger_raw_pbvi_cpu(reflib):
    """PBVI on the raw tiger's 21 beliefs over its 250 iterations: bit-identical to nova's CPU
    path, and the values mirror under swapping the two tiger positions (b_i and b_{20-i})."""
    from conftest import bits
    from nova_amd.pomdp import POMDP
    from nova_amd.pomdp_pbvi import POMDPPBVICPU
    p = POMDP().load(os.path.join(RAW, "tiger_pomdp.raw"), filetype="raw")
    P = p.arrays()
    g0 = np.full((P["Z"].shape[0], 2), float(P["R"].min()) / (1 - float(P["gamma"])),
                 np.float32)
    P["Gamma0"] = g0
    pol = POMDPPBVICPU(p, GammaInitial=g0).solve()
    G, pi = pol.Gamma_array(), pol.pi_array()
    Gr, pir = reflib.pbvi_execute(P)
    assert np.array_equal(bits(G), bits(Gr)) and np.array_equal(pi, pir)
    b = np.zeros((21, 2))
    for i in range(21):
        for k in range(2):
            if P["Z"][i, k] >= 0:
                b[i, P["Z"][i, k]] = P["B"][i, k]
    V = (b @ G.T.astype(np.float64)).max(axis=1)
    np.testing.assert_allclose(V, V[::-1], rtol=1e-4)
