"""Benchmark of nova's Bellman-backup path on one MI355X.

Default workload (vi_n1e6): MDP value iteration, n = 1,000,000 states, m = 8 actions, ns = 16
successors per (s, a), T Dirichlet(1), R uniform[-1, 1] float32, gamma = 0.99, V0 = 0.
One step = one solve = 200 Bellman sweeps from V0 (the device V is reset from a resident V0
at the start of each step). The model arrays are resident in HBM before timing starts.

value       = whole-job state-updates/s = n * 200 * steps / wall time of the timed steps
roofline    = algorithmic bytes of one sweep (8*n*m*ns + 4*n*m + 12*n) / the sweep kernel's
              average duration, from HIP events on the library stream, against 8 TB/s
cpu_baseline= nova's own CPU sweep (oracle/_ref, compiled from the reference sources) on one
              pinned host core, timed over a bounded number of sweeps of the same instance

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "PBVI backups/s, VI updates/s (%HBM peak); batched vs plain solves/s @1e-4"


def host_facts():
    """CPU model, logical CPUs and the CPUs this process may run on, for every cpu_baseline."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}


def baseline(value, unit, kind, sample, cores=1):
    out = {"value": value, "unit": unit, "cores": cores, "kind": kind, "sample": sample}
    out.update(host_facts())
    return out


def vi_bytes_per_sweep(n, m, ns):
    # S and T (int32 + float32 per entry), R, then V read + V' and pi written per state
    return 8 * n * m * ns + 4 * n * m + 12 * n


def cpu_baseline_vi(w, sweeps):
    """nova's CPU sweep on one pinned core (kind 'reference' when oracle/_ref is built)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import RefLib, Oracle
    n = w["S"].shape[0]
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        if RefLib.available():
            dt = RefLib().vi_time_sweeps(w["S"], w["T"], w["R"], w["gamma"], sweeps)
            kind = "reference"
        else:
            o = Oracle()
            v = np.zeros(n, np.float32)
            t = time.perf_counter()
            for _ in range(sweeps):
                v, _ = o.vi_sweep(w["S"], w["T"], w["R"], w["gamma"], v)
            dt = time.perf_counter() - t
            kind = "port"
    finally:
        os.sched_setaffinity(0, old)
    return baseline(n * sweeps / dt, "state-updates/s", kind,
                    "%d full sweeps of the n=%d instance (of the 200 per solve), 1 pinned core"
                    % (sweeps, n))


def traffic_from_profiles(name):
    """HBM bytes per launch measured with rocprofv3 PMC counters (committed summary), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def bench_vi(args):
    import torch
    from nova_amd import synth
    from nova_amd.mdp import MDP
    from nova_amd.mdp_vi import MDPValueIterationGPU

    p = synth.WORKLOADS["vi_n1e6"]
    w = synth.workload("vi_n1e6")
    n, m, ns = p["n"], p["m"], p["ns"]
    sweeps = args.sweeps or p["horizon"]
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=sweeps)
    mdp.initialize_gpu()
    torch.cuda.synchronize()
    t_init = time.perf_counter()
    vi = MDPValueIterationGPU(mdp, Vinitial=np.zeros(n, np.float32))
    init_ms = (time.perf_counter() - t_init) * 1e3     # includes the sorted layout when built
    kernel = {0: "vi_sweep_plain", 1: "vi_sweep_tiled<4,4>", 2: "vi_sweep_sorted",
              3: "vi_chunk_products + vi_chunk_rows (one sweep)"}[vi.sweep_kernel()]
    d_v0 = torch.zeros(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        vi.run(sweeps, d_V0=d_v0.data_ptr())
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(vi.run(sweeps, d_V0=d_v0.data_ptr()))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    sweep_ms = float(np.mean(kernel_ms)) / sweeps
    alg = vi_bytes_per_sweep(n, m, ns)
    achieved = alg / (sweep_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles("vi_sweep_n1e6_traffic.txt")
    out = {
        "metric": METRIC,
        "value": n * sweeps * args.steps / wall,
        "unit": "state-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded; T Dirichlet(1), R U[-1,1])",
        "config": {"workload": "vi_n1e6", "n": n, "m": m, "ns": ns, "gamma": 0.99,
                   "sweeps_per_step": sweeps, "V0": 0},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_rate_GBs": (traffic / (sweep_ms * 1e-3) / 1e9) if traffic else None,
                     "traffic_note": "measured HBM bytes per sweep (rocprofv3 FETCH_SIZE + "
                                     "WRITE_SIZE, profiles/vi_sweep_n1e6_traffic.txt) over this "
                                     "run's sweep time; a plain 512 MiB device copy moves "
                                     "4,680 GB/s of reads plus writes",
                     "kernel": kernel, "kernel_avg_us": sweep_ms * 1e3,
                     "solver_initialize_ms": init_ms,
                     "algorithmic_bytes_per_launch": alg},
    }
    del vi
    mdp.uninitialize_gpu()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_vi(w, args.cpu_sweeps)
    else:
        out["cpu_baseline"] = None
    print("vi leg done", file=sys.stderr, flush=True)
    return out


def pbvi_flops_per_iteration(P):
    """2 * sum over (belief, action) of valid (support, successor) pairs * z * r: the
    multiply-adds of the scoring kernel (the dominant one)."""
    S, Z = P["S"], P["Z"]
    n, m, ns = S.shape
    z = P["O"].shape[2]
    r, rz = Z.shape
    valid_len = np.where((S < 0).any(axis=2), np.argmax(S < 0, axis=2), ns)   # n x m
    total = 0
    for i in range(r):
        k = int(np.argmax(Z[i] < 0)) if (Z[i] < 0).any() else rz
        total += int(valid_len[Z[i, :k]].sum())
    return 2.0 * total * z * r


def pbvi_algorithmic_flops(P):
    """Flops of one PBVI iteration in the projection formulation: every Gamma_{a,o} row
    (2*r*m*z*n*ns), the sparse scoring of every belief against it (2*r*m*z*r*rz) and the
    materialisation of the winners (2*r*z*n)."""
    n, m, ns = P["S"].shape
    z = P["O"].shape[2]
    r, rz = P["Z"].shape
    return 2.0 * r * m * z * (n * ns + r * rz) + 2.0 * r * z * n


def bench_c1(args):
    """nova's own CPU PBVI (oracle/_ref, compiled from the reference sources) on the n = 64
    POMDP (m = 4, z = 8, ns = 4, r = 128, rz = 8, gamma = 0.95, 100 iterations, seed 0), one
    pinned core, the whole pomdp_pbvi_execute_cpu call; beside it the GPU solve of the same
    instance in exact mode (bit for bit nova's result, checked) and fast mode."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, RefLib
    from nova_amd import synth
    from nova_amd.pomdp import POMDP
    from nova_amd.pomdp_pbvi import EXACT, FAST, POMDPPBVIGPU, set_mode
    P = synth.workload("pbvi_cpu_n64")
    r, h = P["Z"].shape[0], P["horizon"]
    out = {"workload": "pbvi_cpu_n64 (n=64, m=4, z=8, ns=4, r=128, rz=8, 100 iterations)",
           "backups_per_solve": r * h}
    ref = None
    if not args.no_cpu:
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old)})
        try:
            if RefLib.available():
                lib = RefLib()
                times = []
                for _ in range(3):
                    dt, Gc, pic = lib.pbvi_time_execute(P)
                    times.append(dt)
                kind = "reference"
            else:
                o = Oracle()
                times = []
                for _ in range(3):
                    t0 = time.perf_counter()
                    Gc, pic = o.pbvi_run(P)
                    times.append(time.perf_counter() - t0)
                kind = "port"
        finally:
            os.sched_setaffinity(0, old)
        dt = min(times)
        ref = (Gc, pic)
        out["cpu_baseline"] = baseline(r * h / dt, "backups/s", kind,
                                       "the whole 100-iteration solve (best of 3), 1 pinned "
                                       "core")
        out["cpu_solve_ms"] = dt * 1e3
    for mode, label in ((EXACT, "exact"), (FAST, "fast")):
        set_mode(mode)
        try:
            p = POMDP.from_dict(P)
            p.initialize_gpu()
            times = []
            for _ in range(4):
                s = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"])
                t0 = time.perf_counter()
                pol = s.solve()
                times.append(time.perf_counter() - t0)
            G, pi = pol.Gamma_array(), pol.pi_array()
            del s
        finally:
            set_mode(FAST)
        dt = min(times[1:])
        leg = {"solve_ms": dt * 1e3, "backups_per_s": r * h / dt}
        if ref is not None:
            leg["gpu_over_cpu"] = leg["backups_per_s"] / out["cpu_baseline"]["value"]
            if label == "exact":
                leg["equals_cpu_bitwise"] = bool(
                    np.array_equal(G.view(np.uint32), ref[0].view(np.uint32)) and
                    np.array_equal(pi, ref[1]))
        out["gpu_" + label] = leg
    print("c1 leg done", file=sys.stderr, flush=True)
    return out


def bench_pbvi(args):
    """PBVI at n=20,000, m=16, z=32, ns=8, r=4,096, rz=64, 50 iterations (fast mode), and one
    exact-mode iteration."""
    from nova_amd import synth
    from nova_amd.pomdp import POMDP
    from nova_amd.pomdp_pbvi import EXACT, FAST, POMDPPBVIGPU, set_mode
    set_mode(FAST)
    P = synth.workload("pbvi_n20000")
    r, iters = P["Z"].shape[0], P["horizon"]
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    s = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"])
    stages = s.profile()                      # warm-up iteration, per-stage times
    variant = s.scoring_variant()
    ms = s.run(iters)                         # one full solve's iterations
    t0 = time.perf_counter()
    s.run(iters)
    wall = time.perf_counter() - t0
    issued = pbvi_flops_per_iteration(P)
    alg = pbvi_algorithmic_flops(P)
    score_ms = stages[1]
    it_ms = ms / iters
    achieved = alg / (it_ms * 1e-3) / 1e12
    issued_tf = issued / (score_ms * 1e-3) / 1e12
    out = {"workload": "pbvi_n20000", "backups_per_s": r * iters / wall,
           "ms_per_iteration": it_ms, "stage_ms": dict(zip(
               ["transpose", "score", "action", "materialize"], stages)),
           "roofline": {"bound": "mfma" if variant >= 8 else "valu",
                        "precision": "fp32",
                        "kernel": {8: "pbvi_score_mfma<4 waves>", 16: "pbvi_score_mfma<8 waves>",
                                   32: "pbvi_score_mfma<8 waves, 8-step prefetch>",
                                   64: "pbvi_score_mfma16 (16 observation rows)",
                                   128: "pbvi_score_mfma_all (every alpha tile per workgroup)",
                                   256: "pbvi_score_mfma<4 waves, ring of 4 operand sets>",
                                   512: "pbvi_score_mfma<8 waves, ring of 4 operand sets>",
                                   1024: "pbvi_score_mfma16 (ring of 4 operand sets)",
                                   2048: "pbvi_score_mfma_all (ring of 4 operand sets)",
                                   4096: "pbvi_score_mfma16_all (ring of 4 operand sets)"}.get(
                                       variant, "pbvi_score<fast,%d>" % variant),
                        "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s",
                        "frac": achieved / 157.3,
                        "algorithmic_flops_per_iteration": alg,
                        "algorithmic_note": "2*r*m*z*(n*ns + r*rz) + 2*r*z*n per iteration "
                                            "(projection formulation) over the whole "
                                            "iteration's wall time",
                        "issued": {"flops_per_launch": issued, "achieved": issued_tf,
                                   "frac": issued_tf / 157.3,
                                   "note": "multiply-adds the direct scoring kernel issues "
                                           "(2 * valid pairs * z * r), over its own duration"}}}
    del s
    # exact mode (nova's rounding order, bit for bit): one iteration from the same Gamma after a
    # warm-up iteration (which also tunes the exact-mode variants)
    set_mode(EXACT)
    s = POMDPPBVIGPU(p, GammaInitial=P["Gamma0"])
    s.profile()
    ex = s.profile()
    out["exact"] = {"stage_ms": dict(zip(["transpose", "score", "action", "materialize"], ex)),
                    "ms_per_iteration": float(sum(ex)), "backups_per_s": r / (sum(ex) * 1e-3),
                    "scoring_variant": s.scoring_variant(),
                    "note": "one iteration's four stages (HIP events); every product and sum "
                            "rounded as nova's CPU backup, so Gamma' and pi equal its bit for bit"}
    del s
    set_mode(FAST)
    p.uninitialize_gpu()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        o = Oracle()
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old)})
        try:
            k = args.cpu_beliefs
            t = time.perf_counter()
            o.pbvi_backup(P, P["Gamma0"], 0, k, diag=False)
            dt = time.perf_counter() - t
        finally:
            os.sched_setaffinity(0, old)
        out["cpu_baseline"] = baseline(k / dt, "backups/s", "port",
                                       "%d beliefs backed up against all %d alphas (one "
                                       "iteration's per-belief work), 1 pinned core" % (k, r))
    return out


def bench_dense(args):
    """Dense-belief scoring on the matrix cores: n=20,000, m=16, z=32, ns=8, r=4,096 beliefs
    with full support (rz = n), one backup from the same Gamma with bf16 operands and with
    fp32 operands (fp32 accumulation in both); the bf16 choices are compared with fp32's."""
    from nova_amd import synth
    from nova_amd.pomdp import POMDP
    from nova_amd.pomdp_pbvi import FAST, POMDPPBVIGPU, set_mode
    base = synth.workload("pbvi_n20000", rz=4)
    P = synth.dense_beliefs(base, seed=8)
    del base
    r, n = P["Gamma0"].shape
    m, z = P["R"].shape[1], P["O"].shape[2]
    rng = np.random.default_rng(9)
    G = (P["Gamma0"] + rng.uniform(-1, 1, P["Gamma0"].shape)).astype(np.float32)
    flops = 2.0 * m * z * r * r * n
    peaks = {2: 2500.0, 3: 157.3}
    res = {}
    for mode, label in ((2, "bf16"), (3, "fp32")):
        set_mode(mode)
        try:
            p = POMDP.from_dict(P)
            p.initialize_gpu()
            s = POMDPPBVIGPU(p, GammaInitial=G)
            s.profile()                                   # warm-up
            s2 = POMDPPBVIGPU(p, GammaInitial=G)
            del s
            stages = s2.profile()
            js, act = s2.argmax_record()
            G1 = s2.get_policy().Gamma_array()
            del s2
            p.uninitialize_gpu()
        finally:
            set_mode(FAST)
        tf = flops / (stages[1] * 1e-3) / 1e12
        res[label] = dict(stage_ms=dict(zip(["convert", "score", "action", "materialize"],
                                            stages)),
                          achieved=tf, peak=peaks[mode], frac=tf / peaks[mode],
                          js=js, act=act, G1=G1)
        print("dense %s leg done" % label, file=sys.stderr, flush=True)
    b, f = res["bf16"], res["fp32"]
    Bd = P["B"].astype(np.float64)
    vb = (b["G1"].astype(np.float64) @ Bd.T).max(axis=0)
    vf = (f["G1"].astype(np.float64) @ Bd.T).max(axis=0)
    out = {"workload": "pbvi_n20000_dense (r=4096, rz=n)", "algorithmic_flops_per_launch": flops,
           "alpha_choice_disagreement_bf16_vs_fp32": float(np.mean(b["js"] != f["js"])),
           "action_disagreement_bf16_vs_fp32": float(np.mean(b["act"] != f["act"])),
           "value_rel_error_bf16_vs_fp32": float(np.max(np.abs(vb - vf)) /
                                                 max(1.0, np.max(np.abs(vf))))}
    kernels = {"bf16": ("dense_score_bf16", "score stage: per chunk of (a, o) pairs the scaled "
                        "Gamma (dense_scale_pairs) and the GEMM"),
               "fp32": ("dense_score<fp32>", "score stage: the GEMM")}
    for label in ("bf16", "fp32"):
        x = res[label]
        out[label] = {"stage_ms": x["stage_ms"],
                      "roofline": {"bound": "mfma", "kernel": kernels[label][0],
                                   "achieved": x["achieved"], "peak": x["peak"],
                                   "unit": "TFLOP/s", "frac": x["achieved"] / x["peak"],
                                   "note": kernels[label][1]}}
    return out


def bench_batch(args):
    """256 uneven POMDPs: batched graph vs 256 sequential plain GPU solves."""
    from nova_amd import synth
    from nova_amd.pomdp import POMDP
    from nova_amd.pomdp_pbvi import FAST, POMDPPBVIGPU, batch_execute, set_mode
    set_mode(FAST)
    batch = synth.batch_instances(count=args.batch_count)
    models = []
    for P in batch:
        p = POMDP.from_dict(P)
        p.initialize_gpu()
        models.append(p)
    solvers = [POMDPPBVIGPU(p, GammaInitial=P["Gamma0"]) for p, P in zip(models, batch)]
    # one untimed batch first: the tuner times each size class's scorers once per process, and
    # neither timed side should pay for it
    batch_execute(solvers)
    t0 = time.perf_counter()
    plain = []
    for s in solvers:
        pol = s.solve()
        plain.append((pol.Gamma_array(), pol.pi_array()))
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    res, dev_ms = batch_execute(solvers)
    t_batch = time.perf_counter() - t0
    same = all(np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and
               np.array_equal(a[1], b[1]) for a, b in zip(plain, res))
    count = len(batch)
    flops = sum(pbvi_algorithmic_flops(P) * P["horizon"] for P in batch)
    out = {"workload": "batch_%d_uneven" % count, "plain_solves_per_s": count / t_plain,
           "batched_solves_per_s": count / t_batch, "ratio": t_plain / t_batch,
           "batched_s": t_batch, "plain_s": t_plain,
           "batched_graph_ms": dev_ms, "batched_equals_plain_bitwise": bool(same),
           "roofline": {"bound": "mfma", "precision": "fp32",
                        "algorithmic_flops": flops, "achieved": flops / t_batch / 1e12,
                        "peak": 157.3, "unit": "TFLOP/s",
                        "frac": flops / t_batch / 1e12 / 157.3,
                        "note": "sum over instances of horizon * (2*r*m*z*(n*ns + r*rz) + "
                                "2*r*z*n), over the batched call's wall time"}}
    del solvers
    for p in models:
        p.uninitialize_gpu()
    return out


def bench_vi_batch(args):
    """256 MDPs of uneven size (n log-uniform in [500, 8000], m in 4..16, ns in {4, 8},
    gamma = 0.95, 50 sweeps): one batched graph (mdp_vi_batch_execute_gpu) against 256
    sequential plain GPU solves (mdp_vi_execute_gpu each), results compared bit for bit."""
    from nova_amd import synth
    from nova_amd.mdp import MDP
    from nova_amd.mdp_vi import MDPValueIterationGPU, batch_execute
    batch = synth.mdp_batch_instances(count=args.batch_count)
    models = []
    for w in batch:
        mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=w["horizon"])
        mdp.initialize_gpu()
        models.append(mdp)
    solvers = [MDPValueIterationGPU(mdp, Vinitial=np.zeros(mdp.n, np.float32))
               for mdp in models]
    solvers[0].solve()                                      # warm-up
    t0 = time.perf_counter()
    plain = []
    for s in solvers:
        pol = s.solve()
        plain.append((pol.V_array(), pol.pi_array()))
    t_plain = time.perf_counter() - t0
    batch_execute(solvers)                                  # warm-up
    t0 = time.perf_counter()
    res, dev_ms = batch_execute(solvers)
    t_batch = time.perf_counter() - t0
    same = all(np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and
               np.array_equal(a[1], b[1]) for a, b in zip(plain, res))
    count = len(batch)
    states = sum(w["S"].shape[0] for w in batch) * batch[0]["horizon"]
    out = {"workload": "mdp_batch_%d_uneven (n 500-8000, m 4-16, ns 4/8, 50 sweeps)" % count,
           "plain_solves_per_s": count / t_plain, "batched_solves_per_s": count / t_batch,
           "ratio": t_plain / t_batch, "batched_graph_ms": dev_ms,
           "batched_state_updates_per_s": states / t_batch,
           "batched_equals_plain_bitwise": bool(same)}
    del solvers
    for mdp in models:
        mdp.uninitialize_gpu()
    print("vi batch leg done", file=sys.stderr, flush=True)
    return out


def bench_policy(args):
    """Executing a policy over many beliefs: k = 16,384 dense beliefs against r = 4,096 alphas
    at n = 20,000 (value and action of every belief: a k x r x n GEMM with a fused row argmax),
    in each precision; and k = 16,384 belief updates on the n = 20,000, m = 16, z = 32, ns = 8
    model. Each timed call is the synchronous entry point from device-resident inputs."""
    import torch
    from nova_amd import _abi, synth
    from nova_amd.pomdp import POMDP
    lib = _abi.load()
    k, n, r = args.policy_beliefs, 20000, 4096
    rng = np.random.default_rng(21)
    dev = torch.device("cuda")
    G = torch.from_numpy(rng.uniform(-3, 3, (r, n)).astype(np.float32)).to(dev)
    pi = torch.from_numpy(rng.integers(0, 16, r).astype(np.int32)).to(dev)
    B = torch.from_numpy(rng.exponential(size=(k, n)).astype(np.float32)).to(dev)
    B /= B.sum(dim=1, keepdim=True)
    V = torch.empty(k, dtype=torch.float32, device=dev)
    act = torch.empty(k, dtype=torch.int32, device=dev)
    J = {}
    flops = 2.0 * k * r * n
    out = {"workload": "policy_k%d_r%d_n%d" % (k, r, n), "algorithmic_flops_per_call": flops}
    peaks = {"fp32": 157.3, "bf16": 2500.0}
    for prec, code in (("exact", 0), ("fp32", 1), ("bf16", 2)):
        kk = k if code else args.policy_exact_beliefs        # exact: a leading sample of rows
        j = torch.empty(kk, dtype=torch.int32, device=dev)
        call = lambda: lib.nova_amd_alpha_vectors_value_and_action_gpu(
            n, r, G.data_ptr(), pi.data_ptr(), kk, B.data_ptr(), code, V.data_ptr(),
            act.data_ptr(), j.data_ptr())
        assert call() == 0
        torch.cuda.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            assert call() == 0
        dt = (time.perf_counter() - t0) / reps
        J[prec] = j.cpu().numpy()
        tf = 2.0 * kk * r * n / dt / 1e12
        leg = {"beliefs": kk, "ms_per_call": dt * 1e3, "beliefs_per_s": kk / dt,
               "tflops": tf}
        if prec in peaks:
            leg["roofline"] = {"bound": "mfma",
                               "kernel": "dense_score_bf16" if prec == "bf16" else "dense_score<fp32>",
                               "achieved": tf, "peak": peaks[prec], "unit": "TFLOP/s",
                               "frac": tf / peaks[prec],
                               "note": "whole call: operand conversion, GEMM and merge"}
        out[prec] = leg
        print("policy %s leg done" % prec, file=sys.stderr, flush=True)
    ke = args.policy_exact_beliefs
    out["alpha_disagreement_fp32_vs_exact"] = float(np.mean(J["fp32"][:ke] != J["exact"]))
    out["alpha_disagreement_bf16_vs_exact"] = float(np.mean(J["bf16"][:ke] != J["exact"]))
    del G, B
    # batched belief updates on the PBVI model
    P = synth.workload("pbvi_n20000")
    p = POMDP.from_dict(P)
    p.initialize_gpu()
    kb = args.policy_beliefs
    Bb = torch.from_numpy(rng.exponential(size=(kb, P["S"].shape[0])).astype(np.float32)).to(dev)
    Bb /= Bb.sum(dim=1, keepdim=True)
    a = torch.from_numpy(rng.integers(0, P["S"].shape[1], kb).astype(np.int32)).to(dev)
    o = torch.from_numpy(rng.integers(0, P["O"].shape[2], kb).astype(np.int32)).to(dev)
    Bp = torch.empty_like(Bb)
    st = torch.empty(kb, dtype=torch.int32, device=dev)
    call = lambda: lib.nova_amd_belief_update_batch_gpu(p, kb, Bb.data_ptr(), a.data_ptr(),
                                                        o.data_ptr(), Bp.data_ptr(), st.data_ptr())
    assert call() == 0
    t0 = time.perf_counter()
    for _ in range(3):
        assert call() == 0
    dt = (time.perf_counter() - t0) / 3
    out["belief_update"] = {"updates": kb, "ms_per_call": dt * 1e3, "updates_per_s": kb / dt,
                            "impossible_observations": int((st != 0).sum().item())}
    del p
    print("policy belief-update leg done", file=sys.stderr, flush=True)
    return out


# nova's benchmark domains (tests/benchmarks/parallel/parallel.py:43-61): file, expansion steps
DOMAINS = [("tiger_95", 4), ("4x3_95", 5), ("shuttle_95", 4), ("paint_95", 4),
           ("aloha_10", 6), ("tiger_grid", 6), ("drive_san_francisco", 8), ("hallway2", 7),
           ("tag", 8)]


def bench_autograd(args):
    """The differentiable paths (no counterpart in the reference) on the workloads of their
    tests. Differentiable VI on vi_diff_n2e5 (n = 200,000, m = 8, ns = 16, gamma = 0.99): the
    residual-stopped solve to 1e-6 (FP64 sweeps, the iterate moved along the slow error mode
    between checks), the same solve with plain sweeps for comparison (float, then FP64), then
    the adjoint (I - gamma P_pi)^T lambda = rho with dL/dR and dL/dT, each
    timed after a warm-up; CPU: the oracle's FP64 sweep on one pinned core over 10 sweeps of
    the same instance, and that rate over the GPU solve's sweep count. PBVI
    argmax replay on pbvi_diff_n5000 (n = 5,000, r = 1,024, 30 iterations): the recorded
    forward run and the backward replay of dL/dR."""
    import torch
    from nova_amd import synth
    from nova_amd.autograd import adjoint, pbvi_objective, pbvi_replay_grad, solve
    out = {}
    w = synth.workload("vi_diff_n2e5")
    dev = torch.device("cuda")
    R, T, S = (torch.from_numpy(w[k]).to(dev) for k in ("R", "T", "S"))
    g = float(w["gamma"])
    n, m, ns = w["S"].shape

    def timed(f):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = f()
        torch.cuda.synchronize()
        return res, time.perf_counter() - t0

    (V64, pi, sweeps, res), ts = timed(lambda: solve(R, T, S, g))
    (V64p, _, sweepsp, resp), tp = timed(lambda: solve(R, T, S, g, plain=True))
    rho = torch.full((n,), 1.0 / n, device=dev)
    (lam, dR, dT, it, r64), ta = timed(lambda: adjoint(R, T, S, g, pi, V64, rho))
    entry = {"workload": "vi_diff_n2e5 (n=200000, m=8, ns=16, gamma=0.99)",
             "solve_ms": ts * 1e3, "solve_sweeps": int(sweeps), "solve_residual": float(res),
             "solve_state_updates_per_s": sweeps * n / ts,
             "solve_algorithmic_GBs": vi_bytes_per_sweep(n, m, ns) * sweeps / ts / 1e9,
             "plain_solve_ms": tp * 1e3, "plain_solve_sweeps": int(sweepsp),
             "plain_solve_residual": float(resp),
             "plain_vs_moved_max_abs_diff": float((V64p - V64).abs().max()),
             "adjoint_ms": ta * 1e3, "adjoint_iterations": int(it),
             "adjoint_fp64_residual": float(r64)}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        o = Oracle()
        V = np.zeros(n)
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old)})
        try:
            t0 = time.perf_counter()
            for _ in range(10):
                V, _ = o.vi_sweep(w["S"], w["T"], w["R"], g, V, dtype=np.float64)
            tc = time.perf_counter() - t0
        finally:
            os.sched_setaffinity(0, old)
        entry["cpu_baseline"] = baseline(10 * n / tc, "state-updates/s (FP64 sweeps)", "port",
                                         "10 FP64 sweeps of the same instance")
        entry["cpu_ms_for_the_gpu_solve_sweeps"] = tc / 10 * sweeps * 1e3
    out["diff_vi"] = entry
    print("autograd VI leg done", file=sys.stderr, flush=True)
    P = synth.workload("pbvi_diff_n5000")
    res_, tf = timed(lambda: pbvi_objective(P, P["R"]))
    _, tb = timed(lambda: pbvi_replay_grad(res_))
    out["pbvi_replay"] = {"workload": "pbvi_diff_n5000 (n=5000, m=8, z=16, ns=8, r=1024, "
                                      "rz=32, 30 iterations)",
                          "forward_recorded_ms": tf * 1e3, "backward_replay_ms": tb * 1e3}
    print("autograd PBVI leg done", file=sys.stderr, flush=True)
    return out


def bench_domains(args):
    """PBVI on nova's shipped domains as its parallel benchmark runs them
    (tests/benchmarks/parallel/parallel.py:64-127): the file's model, 2^k - 1 beliefs added by
    random expansion (here the seeded GPU expansion), 50 iterations from Gamma0 = Rmin/(1-gamma).
    GPU: the fast-mode solve (wall clock, after one warm-up solve), and the exact-mode solve
    (nova's rounding, bit for bit) where the supports fit it. CPU: the oracle port's
    backup of a leading sample of beliefs against the whole initial alpha set on one pinned
    core (per-backup work does not depend on the belief), as backups/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    from nova_amd.pomdp import POMDP
    from nova_amd.pomdp_pbvi import EXACT, FAST, POMDPPBVIGPU, set_mode
    o = Oracle()
    set_mode(FAST)
    horizon = 50
    out = {"workload": "nova's benchmark domains, PBVI, 50 iterations", "domains": {}}
    ddir = os.path.join(ROOT, "tests", "golden", "domains")
    for name, k in DOMAINS:
        p = POMDP().load(os.path.join(ddir, name + ".pomdp"))
        p.initialize_gpu()
        p.expand("random", 2 ** k - 1, gpu=True, seed=1)
        P = p.arrays()
        del p
        P["horizon"] = horizon
        r, n = P["Z"].shape[0], P["S"].shape[0]
        P["Gamma0"] = np.full((r, n), float(P["R"].min()) / (1.0 - float(P["gamma"])),
                              np.float32)
        q = POMDP.from_dict(P)
        q.initialize_gpu()
        times = []
        for _ in range(2):
            s = POMDPPBVIGPU(q, GammaInitial=P["Gamma0"])
            variant = s.scoring_variant()           # before solve, which releases the solver
            t0 = time.perf_counter()
            s.solve()
            times.append(time.perf_counter() - t0)
            del s
        # exact mode (nova's rounding, bit for bit); long supports are staged in chunks
        rz, ns = P["Z"].shape[1], P["S"].shape[2]
        exact_ms = None
        if ns <= 512 and rz <= 16384:
            set_mode(EXACT)
            try:
                for _ in range(2):
                    s = POMDPPBVIGPU(q, GammaInitial=P["Gamma0"])
                    t0 = time.perf_counter()
                    s.solve()
                    exact_ms = (time.perf_counter() - t0) * 1e3
                    del s
            finally:
                set_mode(FAST)
        del q
        gpu_bps = r * horizon / times[-1]
        entry = {"n": n, "m": P["S"].shape[1], "z": P["O"].shape[2], "ns": ns,
                 "r": r, "rz": rz, "gpu_solve_ms": times[-1] * 1e3,
                 "gpu_backups_per_s": gpu_bps, "scoring_variant": variant,
                 "gpu_exact_solve_ms": exact_ms,
                 "gpu_exact_backups_per_s": (r * horizon / (exact_ms * 1e-3)) if exact_ms
                 else None}
        if exact_ms is None:
            entry["gpu_exact_note"] = "successor lists too long for exact mode (ns > 512)"
        if not args.no_cpu:
            old = os.sched_getaffinity(0)
            os.sched_setaffinity(0, {min(old)})
            try:
                t0 = time.perf_counter()
                o.pbvi_backup(P, P["Gamma0"], 0, 1, diag=False)
                t1 = time.perf_counter() - t0
                kc = int(min(r, max(1, 0.3 / max(t1, 1e-6))))
                t0 = time.perf_counter()
                o.pbvi_backup(P, P["Gamma0"], 0, kc, diag=False)
                tc = time.perf_counter() - t0
            finally:
                os.sched_setaffinity(0, old)
            entry["cpu_baseline"] = baseline(kc / tc, "backups/s", "port",
                                             "%d of %d beliefs" % (kc, r))
            entry["gpu_over_cpu"] = gpu_bps / (kc / tc)
            if exact_ms:
                entry["gpu_exact_over_cpu"] = entry["gpu_exact_backups_per_s"] / (kc / tc)
        out["domains"][name] = entry
        print("domain %s done" % name, file=sys.stderr, flush=True)
        if name in PEMA_DOMAINS:
            out["pema"] = out.get("pema", {})
            out["pema"][name] = bench_pema_domain(P, horizon)
    return out


PEMA_DOMAINS = ("tiger_grid", "hallway2")


def bench_pema_domain(P, horizon):
    """One PEMA call (pomdp_expand_pema_cpu, nova's alpha choice) on the expanded set against
    the GPU PBVI policy of that set: the host entry point on one core vs the GPU one; the
    appended beliefs must be equal bit for bit."""
    import ctypes as ct
    from nova_amd import _abi
    from nova_amd.pomdp import POMDP
    from nova_amd.pomdp_pbvi import POMDPPBVIGPU
    q = POMDP.from_dict(P)
    q.initialize_gpu()
    pol = POMDPPBVIGPU(q, GammaInitial=P["Gamma0"]).solve()
    res = {}
    for gpu in (False, True):
        x = POMDP.from_dict(P)
        if gpu:
            x.initialize_gpu()
        old = os.sched_getaffinity(0)
        if not gpu:
            os.sched_setaffinity(0, {min(old)})
        try:
            t0 = time.perf_counter()
            x.expand("pema", pemaPolicy=pol, gpu=gpu)
            res[gpu] = (time.perf_counter() - t0, x.beliefs())
        finally:
            os.sched_setaffinity(0, old)
        del x
    (tc, (Zc, Bc)), (tg, (Zg, Bg)) = res[False], res[True]
    same = bool(np.array_equal(Zc, Zg) and np.array_equal(Bc.view(np.uint32), Bg.view(np.uint32)))
    r, n = P["Z"].shape[0], P["S"].shape[0]
    m, z = P["S"].shape[1], P["O"].shape[2]
    return {"r": r, "alphas": int(pol.r), "closest_search_ops": r * m * z * r * n,
            "gpu_ms": tg * 1e3, "cpu_ms": tc * 1e3, "gpu_over_cpu": tc / tg,
            "cpu_cores": 1, "equal_bitwise": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sweeps", type=int, default=100)
    ap.add_argument("--sweeps", type=int, default=0,
                    help="sweeps per step (default: the workload's 200; fewer only for profiling)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="run one plain device copy of this many MB first (counter scaling)")
    ap.add_argument("--no-vi", action="store_true",
                    help="skip the VI leg (the line then carries only the other legs; profiling)")
    ap.add_argument("--no-pbvi", action="store_true", help="skip the PBVI leg")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the n = 64 CPU-versus-GPU PBVI leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-solve leg")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense-belief MFMA leg")
    ap.add_argument("--cpu-beliefs", type=int, default=4)
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--no-policy", action="store_true", help="skip the policy-execution leg")
    ap.add_argument("--no-domains", action="store_true", help="skip nova's benchmark domains")
    ap.add_argument("--no-autograd", action="store_true",
                    help="skip the differentiable-VI and PBVI-replay leg")
    ap.add_argument("--policy-beliefs", type=int, default=16384)
    ap.add_argument("--policy-exact-beliefs", type=int, default=1024)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the Bellman-backup path here runs on one GPU (how it would shard is a "
              "design note, not code); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy_mb:
        # a plain device copy of known size, for scaling HBM byte counters
        import torch
        a = torch.ones(args.calibrate_copy_mb << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    out = {"metric": METRIC, "config": {"workload": "none (VI leg skipped)"}} if args.no_vi \
        else bench_vi(args)
    extras = {}
    if not args.no_c1:
        extras["c1"] = bench_c1(args)
    if not args.no_pbvi:
        extras["pbvi"] = bench_pbvi(args)
        print("pbvi leg done", file=sys.stderr, flush=True)
    if not args.no_dense:
        extras["dense"] = bench_dense(args)
    if not args.no_batch:
        extras["batch"] = bench_batch(args)
        print("batch leg done", file=sys.stderr, flush=True)
        extras["vi_batch"] = bench_vi_batch(args)
    if not args.no_policy:
        extras["policy"] = bench_policy(args)
    if not args.no_domains:
        extras["domains"] = bench_domains(args)
    if not args.no_autograd:
        extras["autograd"] = bench_autograd(args)
    if extras:
        out["extras"] = extras
    print(json.dumps(out))


if __name__ == "__main__":
    main()
