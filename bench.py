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

Two more legs measure many graphs on the device (sbp_amd/workloads.py names the mixed-size workloads):
- "batched": graphs 0..G-1 of the batched workload (N log-uniform in [200, 50000]), resident in device
  memory, partitioned one after another on one context (the plain GPU path) and as one
  sbp_partition_batch call over worker slots; graphs/s of both, and whether every result is identical.
- "served": the first R requests of the served workload (N log-uniform in [1000, --served-max-n]) arriving
  as a Poisson process, replayed against a request server with one worker (the plain sequential path) and
  with several; p50/p99 latency from arrival to completion and graphs/s of both.

`--only sampled` runs the large single-graph workload instead (5,000,000 vertices, about 80M edges, 5,000
lognormal(1.0) blocks, p_in 0.4): per sampler (max_degree, degree_weighted) a 50% sample is partitioned,
extended to the full graph and finetuned there; wall time per phase, proposals/s, MDL, NMI and the roofline
of the full-graph finetune launches.
"""
import argparse
import concurrent.futures as cf
import json
import multiprocessing as mp
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
    if ha > 0 and hb > 0:
        return float(mi / np.sqrt(ha * hb))
    return None  # 0/0 when either side is one block: the reference's value is not finite


def mean_nmi(values):
    """Mean of the finite NMI values (None where a partition is one block)."""
    v = [x for x in values if x is not None]
    return float(np.mean(v)) if v else None


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
    "mh": ("k_mh<true, false>", "k_mh<false, false>"),  # the sequential walk (windowed)
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_ft_propose_c2_traffic.txt")


def committed_traffic(kernel, path=TRAFFIC_FILE):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE,
    scaled by a copy-kernel calibration), or None when there is none for this kernel."""
    try:
        with open(path) as f:
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


def roofline(tot, greedy=True, traffic_file=TRAFFIC_FILE):
    """Roofline of the dominant finetune proposal launch: algorithmic bytes per launch over its average
    device time (HIP events around that launch alone, on the library's stream).  traffic_file names the
    committed rocprofv3 summary of that launch on the same workload (None: no measured traffic)."""
    legs = {}
    for key, prefix in (("ft", "ft"), ("ft_high", "ft_high"), ("mh", "mh")):
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
    traffic = committed_traffic(dom["kernel"], traffic_file) if traffic_file else None
    return {"bound": "hbm", "achieved": dom["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom["achieved_GBps"] / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": dom["kernel"], "avg_launch_ms": dom["avg_launch_ms"],
            "algorithmic_bytes_per_launch": dom["algorithmic_bytes_per_launch"],
            "traffic_source": os.path.relpath(traffic_file, ROOT) if traffic is not None else None,
            "launches": {k: {kk: vv for kk, vv in v.items() if kk != "total_ms"} for k, v in legs.items()}}


def batched_leg(ctx, graphs, algorithm, slots, plain_max=None):
    """The batched workload's graphs one after another on ctx (plain path) and as one sbp_partition_batch.
    With plain_max below the graph count, the comparison with the plain path runs on graphs 0..plain_max-1,
    and one more sbp_partition_batch over all the graphs gives the throughput of the whole workload."""
    from sbp_amd import native
    if plain_max is not None and plain_max < len(graphs):
        full = graphs
        prms_full = [native.params(algorithm, seed=i) for i in range(len(full))]
        out = batched_leg(ctx, full[:plain_max], algorithm, slots)
        dgs = [ctx.upload(g) for g in full]
        ctx.synchronize()
        os.environ["SBP_AMD_BATCH_PROGRESS"] = "1"
        t0 = time.perf_counter()
        outs, nb, st, fin = ctx.partition_batch(dgs, prms_full, slots=slots)
        t_all = time.perf_counter() - t0
        del os.environ["SBP_AMD_BATCH_PROGRESS"]
        print(f"[bench] batched: all {len(full)} graphs on {slots} slots in {t_all:.2f}s", file=sys.stderr, flush=True)
        out["whole_workload"] = {
            "graphs": len(full), "vertices": int(sum(g.N for g in full)), "edges": int(sum(g.E for g in full)),
            "slots": slots, "seconds": t_all, "graphs_per_s": len(full) / t_all,
            "proposals_per_s": sum(x.proposals for x in st) / t_all,
            "p50_finish_s": float(np.percentile(fin, 50)) / 1e3, "last_finish_s": float(fin.max()) / 1e3,
            "same_as_prefix_run": bool(all(np.array_equal(outs[i], o) for i, o in enumerate(out.pop("_outs")))),
            "nmi_mean": mean_nmi(nmi(g.truth, o) for g, o in zip(full, outs)),
        }
        return out
    prms = [native.params(algorithm, seed=i) for i in range(len(graphs))]
    dgs = [ctx.upload(g) for g in graphs]
    ctx.synchronize()
    # creates the worker contexts and grows their buffers on the largest graphs
    big = sorted(range(len(graphs)), key=lambda i: -graphs[i].E)[:slots]
    ctx.partition_batch([dgs[i] for i in big], [prms[i] for i in big], slots=slots)
    ctx.debug_logf(0, 1)  # a one-thread kernel that marks the start of each phase in a kernel trace
    t0 = time.perf_counter()
    plain = [dg.partition(p) for dg, p in zip(dgs, prms)]
    t_plain = time.perf_counter() - t0
    ctx.debug_logf(0, 1)
    print(f"[bench] batched: plain path {len(graphs)} graphs in {t_plain:.2f}s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    outs, nb, st, fin = ctx.partition_batch(dgs, prms, slots=slots)
    t_batch = time.perf_counter() - t0
    ctx.debug_logf(0, 1)
    print(f"[bench] batched: {slots} slots in {t_batch:.2f}s", file=sys.stderr, flush=True)
    same = all(np.array_equal(outs[i], plain[i][0]) and int(nb[i]) == plain[i][1] and st[i].mdl == plain[i][2].mdl
               for i in range(len(graphs)))
    return {
        "workload": f"graphs 0..{len(graphs) - 1} of the batched mixed-size workload: directed DCSBM, N log-uniform "
                    "in [200, 50000], B = round(N^0.4), mean degree U[8, 32] (power law 2.5), p_in U[0.5, 0.8], "
                    f"graph i seed i; {algorithm}",
        "graphs": len(graphs), "vertices": int(sum(g.N for g in graphs)), "edges": int(sum(g.E for g in graphs)),
        "slots": slots,
        "graphs_per_s": len(graphs) / t_batch, "plain_graphs_per_s": len(graphs) / t_plain,
        "speedup_vs_plain": t_plain / t_batch,
        "proposals_per_s": sum(x.proposals for x in st) / t_batch,
        "identical_to_plain": bool(same),
        "nmi_mean": mean_nmi(nmi(g.truth, o) for g, o in zip(graphs, outs)),
        "plain_nmi_mean": mean_nmi(nmi(g.truth, p[0]) for g, p in zip(graphs, plain)),
        "_outs": outs,
    }


def serve_trace(graphs, arrivals, prms, slots):
    """Replays the arrival trace against a request server with `slots` workers; per-request results and
    latencies (ms) and the wall time from the first arrival to the last completion."""
    from sbp_amd import native
    srv = native.Server(0, slots=slots, hbm_budget_bytes=200 * 10 ** 9)
    t0 = time.perf_counter()
    tickets = []
    for g, at, p in zip(graphs, arrivals, prms):
        dt = at - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        tickets.append(srv.submit(g, p))
    res = []
    for j, t in enumerate(tickets):
        res.append(srv.wait(t))
        print(f"[bench] served ({slots} slots): request {j} (N={graphs[j].N}) in {res[-1][3]:.0f} ms",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    info = srv.info()
    srv.close()
    return res, wall, info


def served_leg(graphs, arrivals, algorithm, slots, rate, max_n):
    from sbp_amd import native
    prms = [native.params(algorithm, seed=j) for j in range(len(graphs))]
    out = {}
    runs = {}
    for name, k in (("plain", 1), ("served", slots)):
        res, wall, info = serve_trace(graphs, arrivals, prms, k)
        lat = np.array([r[3] for r in res])
        runs[name] = res
        out[name] = {"slots": k, "graphs_per_s": len(graphs) / wall, "p50_ms": float(np.percentile(lat, 50)),
                     "p99_ms": float(np.percentile(lat, 99)), "mean_queue_ms": float(np.mean([r[4] for r in res])),
                     "max_running": info["max_running"], "peak_reserved_bytes": info["peak_in_use"]}
        print(f"[bench] served: {k} slot(s) p50 {out[name]['p50_ms']:.0f} ms p99 {out[name]['p99_ms']:.0f} ms "
              f"{out[name]['graphs_per_s']:.2f} graphs/s", file=sys.stderr, flush=True)
    same = all(np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2].mdl == b[2].mdl
               for a, b in zip(runs["plain"], runs["served"]))
    return {
        "workload": f"requests 0..{len(graphs) - 1} of the served workload: directed DCSBM, N log-uniform in "
                    f"[1000, {max_n}], B = round(N^0.4), mean degree 16 (power law 2.1), p_in 0.55, request j seed "
                    f"j; Poisson arrivals at {rate} req/s (arrival seed 0); {algorithm}; 200 GB admission budget",
        "requests": len(graphs), "vertices": int(sum(g.N for g in graphs)), "edges": int(sum(g.E for g in graphs)),
        "rate_per_s": rate,
        "graphs_per_s": out["served"]["graphs_per_s"], "p50_ms": out["served"]["p50_ms"],
        "p99_ms": out["served"]["p99_ms"],
        "speedup_vs_plain": out["served"]["graphs_per_s"] / out["plain"]["graphs_per_s"],
        "identical_to_plain": bool(same),
        "nmi_mean": mean_nmi(nmi(g.truth, r[0]) for g, r in zip(graphs, runs["served"])),
        "plain": out["plain"], "served": out["served"],
    }


SAMPLED = dict(N=5_000_000, B=5000, gamma=2.1, kmin=1, kmax=20000, mean_degree=16.0, p_in=0.4, sigma=1.0)
SAMPLED_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_ft_propose_c3_traffic.txt")


def sampled_leg(algs, algorithm, n):
    """The large single-graph workload: a directed DCSBM with 5,000,000 vertices and 5,000 lognormal(1.0)
    blocks, out-degree power law 2.1 on [1, 20000] with mean 16 (about 80M edges), p_in 0.4, seed 0.  Per
    sampler: keep 50% of the vertices (sample::sample + from_vertices), partition the sample, extend its
    partition to the full graph (sample::extend) and finetune it there (finetune_assignment).  The graph is
    resident on the device before the clock starts."""
    from sbp_amd import native
    from sbp_amd.graph import dcsbm
    os.environ.setdefault("SBP_AMD_TRACE", "1")  # one line per merge phase, so a long partition shows progress
    w = dict(SAMPLED, N=n, B=max(1, round(SAMPLED["B"] * n / SAMPLED["N"])))
    t0 = time.perf_counter()
    g = dcsbm(w["N"], w["B"], gamma=w["gamma"], kmin=w["kmin"], kmax=w["kmax"], mean_degree=w["mean_degree"],
              p_in=w["p_in"], sigma=w["sigma"], seed=0)
    print(f"[bench] sampled: graph N={g.N} E={g.E} generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    ctx = native.Context(0, seed=0)
    warm = dcsbm(2000, 12, gamma=2.5, kmin=1, kmax=100, mean_degree=10, p_in=0.7, seed=1)
    ctx.upload(warm).partition(native.params(algorithm, seed=0))  # loads the code objects
    dg = ctx.upload(g)
    ctx.synchronize()
    prm = native.params(algorithm, seed=0)
    planted_mdl = native.DeviceBlockmodel.build(dg, w["B"], g.truth).mdl()[0]
    runs = {}
    for alg in algs:
        ctx.set_seed(0)
        t = [time.perf_counter()]
        sg, mapping = dg.sample(alg, 0.5, seed=0)
        ctx.synchronize()
        t.append(time.perf_counter())
        print(f"[bench] sampled/{alg}: sample N={sg.N} E={sg.E} in {t[1] - t[0]:.2f}s", file=sys.stderr, flush=True)
        a_s, nb, st_p = sg.partition(prm)
        t.append(time.perf_counter())
        print(f"[bench] sampled/{alg}: sample partitioned, B={nb} in {t[2] - t[1]:.1f}s", file=sys.stderr,
              flush=True)
        sbm = native.DeviceBlockmodel.build(sg, nb, a_s)
        full, _ = sbm.extend(dg, seed=0)
        ctx.synchronize()
        t.append(time.perf_counter())
        ctx.set_seed(0)
        _, mdl, st_f = full.finetune_assignment(prm, 0)
        a = full.export()["assignment"]
        t.append(time.perf_counter())
        print(f"[bench] sampled/{alg}: extended and finetuned in {t[3] - t[2]:.1f}s + {t[4] - t[3]:.1f}s, "
              f"mdl {mdl:.2f}", file=sys.stderr, flush=True)
        del sbm, full, sg
        wall = t[4] - t[0]
        ft = st_f.as_dict()
        sampled_truth = g.truth[mapping >= 0][np.argsort(mapping[mapping >= 0])]
        runs[alg] = {
            "sample_vertices": int((mapping >= 0).sum()), "num_blocks": nb, "mdl": mdl,
            "nmi": nmi(g.truth, a), "nmi_sample": nmi(sampled_truth, a_s),
            "wall_s": wall, "phase_s": {"sample": t[1] - t[0], "partition_sample": t[2] - t[1],
                                        "extend": t[3] - t[2], "finetune_full": t[4] - t[3]},
            "proposals_per_s": (st_p.proposals + st_f.proposals) / wall,
            "accepted_moves_per_s": (st_p.moves + st_f.moves) / wall,
            "partition_sample": dict({k: st_p.as_dict()[k] for k in ("proposals", "moves", "sweeps", "merge_phases",
                                                                     "finetune_ms", "merge_ms", "carry_ms", "build_ms",
                                                                     "mdl_ms")},
                                     roofline=roofline(st_p.as_dict(), greedy=True, traffic_file=None)),
            "finetune_full": {"proposals": ft["proposals"], "moves": ft["moves"], "sweeps": ft["sweeps"],
                              "finetune_ms": ft["finetune_ms"],
                              "proposals_per_s": ft["proposals"] / max(t[4] - t[3], 1e-9),
                              "roofline": roofline(ft, greedy=True, traffic_file=SAMPLED_TRAFFIC_FILE)},
        }
    return {"workload": f"directed DCSBM N={g.N} B={w['B']} lognormal(1.0) sizes, power-law out-degree gamma=2.1 "
                        f"on [1,20000] mean 16, p_in=0.4, seed 0; 50% sample, {algorithm} partition of the sample, "
                        "extend, finetune_assignment on the full graph",
            "N": g.N, "E": g.E, "algorithm": algorithm, "planted_mdl": planted_mdl, "runs": runs}


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
    ap.add_argument("--batched-graphs", type=int, default=64, help="graphs in the batched leg (0: skip it)")
    ap.add_argument("--batched-slots", type=int, default=None, help="default: the autotuned count, else 16")
    ap.add_argument("--batched-plain-max", type=int, default=None,
                    help="compare with the plain path on the first K graphs only, and time one batch over all "
                         "of them (e.g. --batched-graphs 4096 --batched-plain-max 512)")
    ap.add_argument("--served-requests", type=int, default=32, help="requests in the served leg (0: skip it)")
    ap.add_argument("--served-rate", type=float, default=32.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--served-max-n", type=int, default=200_000, help="upper bound of the requests' N")
    ap.add_argument("--served-slots", type=int, default=None, help="default: 16")
    ap.add_argument("--many-algorithm", default="async_gibbs", help="algorithm of the batched and served legs")
    ap.add_argument("--only", choices=["batched", "served", "sampled"],
                    help="run only that leg and print its JSON object (sampled: the 5,000,000-vertex sampled "
                         "workload, which the default run leaves out for its length)")
    ap.add_argument("--sampled-algs", default="max_degree,degree_weighted", help="samplers of the sampled leg")
    ap.add_argument("--sampled-algorithm", default="async_gibbs", help="finetune algorithm of the sampled leg")
    ap.add_argument("--sampled-n", type=int, default=SAMPLED["N"],
                    help="vertices of the sampled leg's graph (blocks scale with it)")
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

    if args.only == "sampled":
        print(json.dumps({"sampled": sampled_leg(args.sampled_algs.split(","), args.sampled_algorithm,
                                                 args.sampled_n)}))
        return

    from sbp_amd import workloads
    from sbp_amd.autotune import tuned_slots
    if args.batched_slots is None:
        args.batched_slots = tuned_slots(16)
    if args.served_slots is None:
        # the autotuner measures the batched workload; for Poisson-arriving requests 16 workers measured best
        # (7.53 graphs/s against 6.76 with 12 at 32 requests/s)
        args.served_slots = 16
    # the mixed-size graphs are generated in worker processes (forked before the device is touched) while
    # the single-graph workload is generated and warmed up
    pool = cf.ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1), mp_context=mp.get_context("fork"))
    fut_b = [pool.submit(workloads.make, workloads.batched_spec(i)) for i in range(args.batched_graphs)]
    fut_s = [pool.submit(workloads.make, workloads.served_spec(j, n_max=args.served_max_n))
             for j in range(args.served_requests)]

    from sbp_amd import native
    from sbp_amd.graph import dcsbm

    if args.only:
        many_b = []
        for i, f in enumerate(fut_b):
            many_b.append(f.result())
            if (i + 1) % 256 == 0:
                print(f"[bench] generated {i + 1} batched graphs", file=sys.stderr, flush=True)
        many_s = []
        for j, f in enumerate(fut_s):
            many_s.append(f.result())
            print(f"[bench] generated request {j}: N={many_s[-1].N} E={many_s[-1].E}", file=sys.stderr, flush=True)
        pool.shutdown()
        if args.only == "batched":
            ctx = native.Context(0, seed=0)
            out = batched_leg(ctx, many_b, args.many_algorithm, args.batched_slots, args.batched_plain_max)
            out.pop("_outs", None)
        else:
            arrivals = workloads.poisson_arrivals(args.served_rate, args.served_requests, seed=0)
            out = served_leg(many_s, arrivals, args.many_algorithm, args.served_slots, args.served_rate,
                             args.served_max_n)
        print(json.dumps({args.only: out}))
        return

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

    # the timed steps and the CPU baseline run with the generator processes finished
    many_b = [f.result() for f in fut_b]
    many_s = [f.result() for f in fut_s]
    pool.shutdown()
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

    batched = served = None
    if args.batched_graphs:
        batched = batched_leg(ctx, many_b, args.many_algorithm, args.batched_slots, args.batched_plain_max)
        batched.pop("_outs", None)
    if args.served_requests:
        arrivals = workloads.poisson_arrivals(args.served_rate, args.served_requests, seed=0)
        served = served_leg(many_s, arrivals, args.many_algorithm, args.served_slots, args.served_rate,
                            args.served_max_n)

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
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
