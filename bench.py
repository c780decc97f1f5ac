"""Benchmark: end-to-end stochastic block partitioning on one MI355X.

Workload (default): a directed DCSBM with N=200,000 vertices and 500 planted blocks of lognormal size
(sigma 0.5); out-degrees follow a power law with exponent 2.1 clipped to [1, 2000] with mean 20 (about
4M edges); 60% of edges are intra-block.  One step = one full agglomerative SBP run
(sbp::stochastic_block_partition with asynchronous-Gibbs finetune, block merges at rate 0.5 and the
golden-section search), with the graph already resident in device memory.

value = MCMC vertex-move proposals evaluated per second of whole-run wall time.  The JSON line also
carries accepted moves, merge proposals, MDL, NMI against the planted partition, the roofline of the
finetune proposal kernel (algorithmic bytes per launch / average launch time, measured with HIP events on
the kernel's stream) and a CPU baseline: the oracle's OpenMP asynchronous-Gibbs sweeps on the same graph
at the planted partition, timed for a bounded sample on the host cores.
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

WORKLOADS = {
    "c2": dict(desc="directed DCSBM N=200000 B=500 lognormal(0.5) sizes, power-law out-degree gamma=2.1 "
                    "on [1,2000] mean 20, p_in=0.6, seed 0; async_gibbs finetune + block_merge",
               N=200000, B=500, gamma=2.1, kmin=1, kmax=2000, mean_degree=20.0, p_in=0.6, sigma=0.5,
               algorithm="async_gibbs"),
    "c1": dict(desc="directed DCSBM N=1000 B=16 equal sizes, power-law out-degree gamma=2.5 on [1,100] "
                    "mean 10, p_in=0.75, seed 0; metropolis_hastings finetune + block_merge",
               N=1000, B=16, gamma=2.5, kmin=1, kmax=100, mean_degree=10.0, p_in=0.75, sigma=None,
               algorithm="metropolis_hastings"),
}


def nmi(a, b):
    """Normalised mutual information, MI / sqrt(H(a) H(b)) (src/evaluate.cpp:97-132)."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    n = len(a)
    _, ai = np.unique(a, return_inverse=True)
    _, bi = np.unique(b, return_inverse=True)
    cont = np.zeros((ai.max() + 1, bi.max() + 1), np.float64)
    np.add.at(cont, (ai, bi), 1.0)
    pa, pb, pab = cont.sum(1) / n, cont.sum(0) / n, cont / n
    nz = pab > 0
    mi = (pab[nz] * np.log(pab[nz] / np.outer(pa, pb)[nz])).sum()
    ha = -(pa[pa > 0] * np.log(pa[pa > 0])).sum()
    hb = -(pb[pb > 0] * np.log(pb[pb > 0])).sum()
    return float(mi / np.sqrt(ha * hb)) if ha > 0 and hb > 0 else 1.0


def cpu_baseline(g, algorithm, seconds, threads):
    import oracle as O
    og = O.OGraph.from_csr(g)
    ob = O.OBlockmodel.build(og, int(g.truth.max()) + 1, g.truth)
    prm = O.params(algorithm if algorithm != "metropolis_hastings" else "metropolis_hastings", seed=0,
                   threads=threads)
    props, t0, ep = 0, time.perf_counter(), 0
    while time.perf_counter() - t0 < seconds or ep == 0:
        if algorithm == "metropolis_hastings":
            _, st = ob.mh_sweep(prm, ep)
        else:
            _, st = ob.async_gibbs_sweep(prm, ep)
        props += st.proposals
        ep += 1
    dt = time.perf_counter() - t0
    return props / dt, ep, dt


def gpu_sweeps_at_truth(ctx, dg, g, algorithm, sweeps):
    from sbp_amd import native
    db = native.DeviceBlockmodel.build(dg, int(g.truth.max()) + 1, g.truth)
    prm = native.params(algorithm, seed=0)
    props, t0 = 0, time.perf_counter()
    for ep in range(sweeps):
        if algorithm == "metropolis_hastings":
            _, st = db.mh_sweep(prm, ep)
        else:
            st = db.async_gibbs_sweep(prm, ep)
        props += st.proposals
    return props / (time.perf_counter() - t0)


KERNEL_NAMES = {  # rocprofv3 names of the finetune proposal launches (greedy / non-greedy, parametric entropy)
    "ft": ("k_ft_propose<false, true, false>", "k_ft_propose<false, false, false>"),
    "ft_high": ("k_ft_propose<true, true, false>", "k_ft_propose<true, false, false>"),
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_ft_propose_c2_traffic.txt")


def committed_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE,
    scaled by a copy-kernel calibration), or None when there is none for this kernel."""
    try:
        with open(TRAFFIC_FILE) as f:
            lines = f.read().splitlines()
    except OSError:
        return None
    fields = dict(l.split(":", 1) for l in lines if ":" in l and not l.startswith("#"))
    if fields.get("kernel", "").strip() != kernel:
        return None
    try:
        return float(fields["traffic_bytes_per_launch"].split()[0])
    except (KeyError, ValueError, IndexError):
        return None


def roofline(tot, greedy=True):
    """Roofline of the dominant finetune proposal launch: algorithmic bytes per launch over its average
    device time (HIP events around that launch alone, on the library's stream)."""
    legs = {}
    for key, prefix in (("ft", "ft"), ("ft_high", "ft_high")):
        n = tot[f"{prefix}_launches"]
        if n:
            legs[key] = dict(kernel=KERNEL_NAMES[key][0 if greedy else 1], launches=n,
                             avg_launch_ms=tot[f"{prefix}_kernel_ms"] / n,
                             algorithmic_bytes_per_launch=tot[f"{prefix}_bytes"] / n,
                             total_ms=tot[f"{prefix}_kernel_ms"])
    if not legs:
        return None
    for leg in legs.values():
        leg["achieved_GBps"] = leg["algorithmic_bytes_per_launch"] / (leg["avg_launch_ms"] * 1e-3) / 1e9
    dom = max(legs.values(), key=lambda l: l["total_ms"])
    traffic = committed_traffic(dom["kernel"])
    return {"bound": "hbm", "achieved": dom["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom["achieved_GBps"] / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": dom["kernel"], "avg_launch_ms": dom["avg_launch_ms"],
            "algorithmic_bytes_per_launch": dom["algorithmic_bytes_per_launch"],
            "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if traffic is not None else None,
            "launches": {k: {kk: vv for kk, vv in v.items() if kk != "total_ms"} for k, v in legs.items()}}


def calibrate_copy(mib):
    """A plain device-to-device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20)
    src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB", default=0,
                    help="only copy MIB MiB device-to-device 3 times (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the path is single-device, "
              "see the sharding note in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from sbp_amd import native
    from sbp_amd.graph import dcsbm

    w = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    g = dcsbm(w["N"], w["B"], gamma=w["gamma"], kmin=w["kmin"], kmax=w["kmax"], mean_degree=w["mean_degree"],
              p_in=w["p_in"], sigma=w["sigma"], seed=0)
    gen_s = time.perf_counter() - t0
    print(f"[bench] graph N={g.N} E={g.E} generated in {gen_s:.1f}s", file=sys.stderr, flush=True)
    ctx = native.Context(0, seed=0)
    dg = ctx.upload(g)
    prm = native.params(w["algorithm"], seed=0)

    for i in range(args.warmup):
        a, nb, st = dg.partition(prm)
        print(f"[bench] warmup {i}: B={nb} {st.total_ms / 1e3:.2f}s", file=sys.stderr, flush=True)

    ctx.synchronize()
    t_start = time.perf_counter()
    tot = {k: 0.0 for k in native.Stats().as_dict()}
    results = []
    for i in range(args.steps):
        a, nb, st = dg.partition(prm)
        d = st.as_dict()
        for k in tot:
            tot[k] += d[k]
        results.append((a, nb, st.mdl))
        print(f"[bench] step {i}: B={nb} mdl={st.mdl:.6f} {st.total_ms / 1e3:.2f}s", file=sys.stderr, flush=True)
    ctx.synchronize()
    wall = time.perf_counter() - t_start

    a, nb, mdl = results[-1]
    value = tot["proposals"] / wall

    gpu_truth_rate = gpu_sweeps_at_truth(ctx, dg, g, w["algorithm"], 10)
    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        rate, sweeps, dt = cpu_baseline(g, w["algorithm"], args.cpu_seconds, threads)
        cpu = {"value": rate, "unit": "MCMC proposals/s", "cores": threads, "kind": "port",
               "sample": f"{sweeps} oracle {w['algorithm']} sweeps ({sweeps * g.N} proposals) over the same "
                         f"graph at its planted partition, {dt:.1f}s on {threads} OpenMP threads",
               "gpu_same_sample": gpu_truth_rate}

    line = {
        "metric": "MCMC moves/s, batched graphs/s & p99 latency vs plain GPU SBP at equal NMI",
        "value": value,
        "unit": "MCMC proposals/s (whole-run wall time)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 counts / fp64 delta-MDL",
        "data": "synthetic",
        "config": {"workload": w["desc"], "N": g.N, "E": g.E, "planted_blocks": w["B"],
                   "algorithm": w["algorithm"], "batches": 1, "merge_rate": 0.5},
        "accepted_moves_per_s": tot["moves"] / wall,
        "merge_proposals_per_s": tot["merge_proposals"] / wall,
        "graphs_per_s": args.steps / wall,
        "num_blocks": nb,
        "mdl": mdl,
        "nmi": nmi(g.truth, a),
        "phase_ms_per_step": {k: tot[k] / args.steps for k in
                              ("finetune_ms", "merge_ms", "carry_ms", "build_ms", "mdl_ms", "ft_kernel_ms",
                               "ft_high_kernel_ms", "mh_kernel_ms")},
        "sweeps_per_step": tot["sweeps"] / args.steps,
        "carry_per_step": {k: tot[k] / args.steps for k in
                           ("carry_load_ms", "carry_evaluations", "carry_speculative", "carry_reused",
                            "carry_requeues")},
        "roofline": roofline(tot, greedy=True),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
