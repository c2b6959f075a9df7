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
    return {"value": n * sweeps / dt, "unit": "state-updates/s", "cores": 1, "kind": kind,
            "sample": "%d full sweeps of the n=%d instance (of the 200 per solve), 1 pinned core"
                      % (sweeps, n)}


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
    if args.calibrate_copy_mb:
        # a plain device copy of known size, for scaling HBM byte counters
        a = torch.ones(args.calibrate_copy_mb << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    mdp = MDP.from_arrays(w["S"], w["T"], w["R"], gamma=w["gamma"], horizon=sweeps)
    mdp.initialize_gpu()
    vi = MDPValueIterationGPU(mdp, Vinitial=np.zeros(n, np.float32))
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
                     "traffic": traffic_from_profiles("vi_sweep_n1e6_traffic.txt"),
                     "kernel": "vi_sweep_tiled<4,4>", "kernel_avg_us": sweep_ms * 1e3,
                     "algorithmic_bytes_per_launch": alg},
    }
    del vi
    mdp.uninitialize_gpu()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_vi(w, args.cpu_sweeps)
    else:
        out["cpu_baseline"] = None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sweeps", type=int, default=10)
    ap.add_argument("--sweeps", type=int, default=0,
                    help="sweeps per step (default: the workload's 200; fewer only for profiling)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="run one plain device copy of this many MB first (counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path is single-GPU; see the "
              "sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    out = bench_vi(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
