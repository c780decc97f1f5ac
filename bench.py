"""Benchmark: SBP's hot path on one MI355X.

Headline (`value`): MCMC moves/s on the 5,000,000-vertex workload -- a directed DCSBM with 5,000 lognormal(1.0)
blocks, power-law out-degree (exponent 2.1, clipped to [1, 20000], mean 16; about 80M edges), 40% intra-block
edges, seed 0, resident in device memory.  One step is one asynchronous-Gibbs sweep over all 5,000,000 vertices
(finetune::asynchronous_gibbs' iteration body: a proposal per vertex against the frozen blockmodel, the commit in
vertex-id order, the rebuild, and the MDL), starting from the planted 5,000-block partition and continuing from
sweep to sweep.  value = accepted vertex moves per second of wall time (the reference's timers::MCMC_moves
counts accepted moves); the line also gives proposals/s, the roofline of the dominant proposal kernel
(algorithmic bytes per launch over its HIP-event time, and the rocprofv3-measured HBM bytes over the same time)
and a CPU baseline: the oracle's OpenMP asynchronous-Gibbs sweeps on the same graph at the same partition.

Further legs, each once per run:
- "sampled": the same graph's whole sampled flow (degree-weighted 50% sample, partition of the sample, extension
  to the full graph, finetune_assignment) with wall time per phase.
- "split": divisive::split_communities (random splits) on the same graph, from one block (the top-down start,
  target 2) and from the planted partition (target twice its blocks), with the wall time of each call.
- "c2": the 200,000-vertex workload (500 lognormal(0.5) blocks, power law 2.1 on [1, 2000], mean 20, p_in 0.6):
  full agglomerative partitions with asynchronous-Gibbs finetune, against the oracle's committed partition.
- "c1": the 1,000-vertex workload (16 equal blocks, power law 2.5 on [1, 100], mean 10, p_in 0.75): full
  Metropolis-Hastings partitions, and the oracle's full run of the same on one host core.
- "batched": graphs of the batched mixed-size workload (N log-uniform in [200, 50000]) resident on the device:
  all of them in one sbp_partition_batch call, and the first few one after another on one context (the plain
  GPU path) and as a batch, with graphs/s of both and whether every result is identical.
- "served": requests of the served workload (N log-uniform in [1000, 2000000]) arriving as a Poisson process,
  replayed against a request server with one worker (the plain sequential path) and with several; p50/p99
  latency from arrival to completion and graphs/s of both.
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

GRAPHS = {
    "c3": dict(desc="directed DCSBM N=5000000 B=5000 lognormal(1.0) sizes, power-law out-degree gamma=2.1 on "
                    "[1,20000] mean 16, p_in=0.4, seed 0",
               N=5_000_000, B=5000, gamma=2.1, kmin=1, kmax=20000, mean_degree=16.0, p_in=0.4, sigma=1.0),
    "c2": dict(desc="directed DCSBM N=200000 B=500 lognormal(0.5) sizes, power-law out-degree gamma=2.1 on "
                    "[1,2000] mean 20, p_in=0.6, seed 0",
               N=200000, B=500, gamma=2.1, kmin=1, kmax=2000, mean_degree=20.0, p_in=0.6, sigma=0.5),
    "c1": dict(desc="directed DCSBM N=1000 B=16 equal sizes, power-law out-degree gamma=2.5 on [1,100] mean 10, "
                    "p_in=0.75, seed 0",
               N=1000, B=16, gamma=2.5, kmin=1, kmax=100, mean_degree=10.0, p_in=0.75, sigma=None),
}


def make_graph(key, n=None):
    from sbp_amd.graph import dcsbm
    w = dict(GRAPHS[key])
    if n is not None and n != w["N"]:  # a scaled-down copy (blocks scale with N), for quick checks
        w["B"] = max(1, round(w["B"] * n / w["N"]))
        w["N"] = n
    return dcsbm(w["N"], w["B"], gamma=w["gamma"], kmin=w["kmin"], kmax=w["kmax"], mean_degree=w["mean_degree"],
                 p_in=w["p_in"], sigma=w["sigma"], seed=0), w


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


def cpu_baseline(g, B, seconds, threads):
    """The oracle's OpenMP asynchronous-Gibbs sweeps over g from its planted partition, as many as fit in about
    `seconds` (at least one): (proposals/s, accepted moves/s, sweeps, seconds)."""
    import oracle as O
    og = O.OGraph.from_csr(g)
    ob = O.OBlockmodel.build(og, B, g.truth)
    prm = O.params("async_gibbs", seed=0, threads=threads)
    props = moves = 0
    t0, ep = time.perf_counter(), 0
    while time.perf_counter() - t0 < seconds or ep == 0:
        _, st = ob.async_gibbs_sweep(prm, ep)
        ob.mdl()  # each iteration of asynchronous_gibbs ends with the MDL (src/finetune.cpp:88)
        props += st.proposals
        moves += st.moves
        ep += 1
    dt = time.perf_counter() - t0
    return props / dt, moves / dt, ep, dt


CLASS_RANGES = ("k <= 8", "8 < k <= 32", "32 < k <= 128", "128 < k <= 256", "256 < k <= 1024", "k > 1024")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_ft_propose_c3_traffic.txt")


def class_kernels(shape, greedy=True, B=None):
    """rocprofv3 names of the finetune proposal launch of each degree class (k = |out(v)| + |in(v)|) under the
    launch shape (lanes, waves per SIMD) of Context.finetune_shape(): lane-group launches for classes 0-3,
    one workgroup per vertex for classes 4-5 (the last indexes its table by block id when B <= 6144)."""
    lanes, occ = shape
    g = "true" if greedy else "false"
    names = [f"k_ft_lanes<{lanes[b]}, {g}, false, {occ[b] if greedy else 4}>" for b in range(4)]
    names.append(f"k_ft_hub<{g}, false, false, {lanes[4]}, 11, 0>")
    mode = "*" if B is None else ("1" if B <= 6144 else "2")  # a whole partition run passes through both
    names.append(f"k_ft_hub<{g}, false, false, 512, 12, {mode}>")
    return names


def add_stats(tot, d):
    """Accumulates one Stats.as_dict() into tot (per-class fields are lists)."""
    for k, v in d.items():
        if isinstance(v, list):
            tot[k] = [a + b for a, b in zip(tot.get(k, [0] * len(v)), v)]
        else:
            tot[k] = tot.get(k, 0.0) + v
    return tot


def committed_traffic(path):
    """HBM bytes per launch of each degree class's proposal launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE scaled by a copy-kernel calibration): {class: (kernel, bytes)}, or {} when absent.
    Lines read: "class <b> traffic_bytes_per_launch=<x> ... kernel=<name>"."""
    out = {}
    try:
        with open(path) as f:
            for line in f:
                if not line.startswith("class ") or " kernel=" not in line:
                    continue
                head, kernel = line.rstrip("\n").split(" kernel=", 1)
                fields = dict(x.split("=", 1) for x in head.split()[2:] if "=" in x)
                out[int(head.split()[1])] = (kernel.strip(), float(fields["traffic_bytes_per_launch"]))
    except (OSError, KeyError, ValueError):
        return {}
    return out


def roofline(tot, shape, greedy=True, traffic_file=TRAFFIC_FILE, B=None):
    """Roofline of the finetune proposal launches, one entry per degree class: algorithmic bytes per launch over the
    launch's average device time (HIP events around each class's launch on the library's stream).  The headline
    figures are those of the dominant class (largest device time); traffic_file names the committed rocprofv3
    summary of the same launches on the same workload (None: no measured traffic)."""
    names = class_kernels(shape, greedy, B)
    traffic = committed_traffic(traffic_file) if traffic_file else {}
    legs = {}
    for b in range(6):
        n = tot["ft_class_launches"][b]
        if not n:
            continue
        ms = tot["ft_class_ms"][b] / n
        byts = tot["ft_class_bytes"][b] / n
        leg = dict(degree_class=CLASS_RANGES[b], kernel=names[b], launches=n, avg_launch_ms=ms,
                   algorithmic_bytes_per_launch=byts, achieved_GBps=byts / (ms * 1e-3) / 1e9,
                   total_ms=tot["ft_class_ms"][b])
        t = traffic.get(b)
        if t is not None and t[0] == names[b]:
            leg["traffic_bytes_per_launch"] = t[1]
            leg["hbm_achieved_GBps"] = t[1] / (ms * 1e-3) / 1e9
        legs[b] = leg
    if not legs:
        return None
    dom = max(legs.values(), key=lambda l: l["total_ms"])
    all_ms = sum(l["total_ms"] for l in legs.values())
    all_bytes = sum(tot["ft_class_bytes"])
    out = {"bound": "hbm", "achieved": dom["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": dom["achieved_GBps"] / HBM_PEAK_GBS, "traffic": dom.get("traffic_bytes_per_launch"),
           "hbm_achieved": dom.get("hbm_achieved_GBps"),
           "hbm_frac": dom["hbm_achieved_GBps"] / HBM_PEAK_GBS if "hbm_achieved_GBps" in dom else None,
           "kernel": dom["kernel"], "degree_class": dom["degree_class"], "avg_launch_ms": dom["avg_launch_ms"],
           "algorithmic_bytes_per_launch": dom["algorithmic_bytes_per_launch"],
           "traffic_source": os.path.relpath(traffic_file, ROOT) if dom.get("traffic_bytes_per_launch") else None,
           "all_proposal_launches": {"ms": all_ms, "algorithmic_GBps": all_bytes / (all_ms * 1e-3) / 1e9}}
    if all(("traffic_bytes_per_launch" in l) for l in legs.values()):
        hb = sum(l["traffic_bytes_per_launch"] * l["launches"] for l in legs.values())
        out["all_proposal_launches"]["hbm_GBps"] = hb / (all_ms * 1e-3) / 1e9
        out["all_proposal_launches"]["hbm_frac"] = hb / (all_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    out["classes"] = {str(b): {k: v for k, v in l.items() if k != "total_ms"} for b, l in legs.items()}
    return out


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


def headline_leg(ctx, dg, g, w, steps, warmup):
    """Asynchronous-Gibbs sweeps over the whole graph from its planted partition: `warmup` untimed, then `steps`
    timed, each with the MDL that closes an iteration of finetune::asynchronous_gibbs (src/finetune.cpp:50-88)."""
    from sbp_amd import native
    db = native.DeviceBlockmodel.build(dg, w["B"], g.truth)
    prm = native.params("async_gibbs", seed=0)
    mdl0 = db.mdl()[0]
    ep = 0
    for _ in range(warmup):
        db.async_gibbs_sweep(prm, ep)
        db.mdl()
        ep += 1
    ctx.synchronize()
    tot = {}
    t0 = time.perf_counter()
    for i in range(steps):
        st = db.async_gibbs_sweep(prm, ep)
        mdl = db.mdl()[0]
        ep += 1
        add_stats(tot, st.as_dict())
    ctx.synchronize()
    wall = time.perf_counter() - t0
    print(f"[bench] headline: {steps} sweeps in {wall:.2f}s, {tot['moves'] / wall:.0f} moves/s, "
          f"{tot['proposals'] / wall:.0f} proposals/s, mdl {mdl0:.1f} -> {mdl:.1f}", file=sys.stderr, flush=True)
    return tot, wall, mdl0, mdl, db.B


def sampled_leg(ctx, dg, g, w, sampler, algorithm="async_gibbs"):
    """The large graph's sampled flow: keep 50% of the vertices (sample::sample + from_vertices), partition the
    sample, extend its partition to the full graph (sample::extend) and finetune it there
    (finetune_assignment).  The graph is resident on the device before the clock starts."""
    from sbp_amd import native
    os.environ.setdefault("SBP_AMD_TRACE", "1")  # one line per merge phase, so a long partition shows progress
    prm = native.params(algorithm, seed=0)
    ctx.set_seed(0)
    t = [time.perf_counter()]
    sg, mapping = dg.sample(sampler, 0.5, seed=0)
    ctx.synchronize()
    t.append(time.perf_counter())
    print(f"[bench] sampled/{sampler}: sample N={sg.N} E={sg.E} in {t[1] - t[0]:.2f}s", file=sys.stderr, flush=True)
    a_s, nb, st_p = sg.partition(prm)
    t.append(time.perf_counter())
    print(f"[bench] sampled/{sampler}: sample partitioned, B={nb} in {t[2] - t[1]:.1f}s", file=sys.stderr, flush=True)
    sbm = native.DeviceBlockmodel.build(sg, nb, a_s)
    full, _ = sbm.extend(dg, seed=0)
    ctx.synchronize()
    t.append(time.perf_counter())
    ctx.set_seed(0)
    _, mdl, st_f = full.finetune_assignment(prm, 0)
    a = full.export()["assignment"]
    t.append(time.perf_counter())
    print(f"[bench] sampled/{sampler}: extended and finetuned in {t[3] - t[2]:.1f}s + {t[4] - t[3]:.1f}s, "
          f"mdl {mdl:.2f}", file=sys.stderr, flush=True)
    del sbm, full, sg
    wall = t[4] - t[0]
    ft = st_f.as_dict()
    p = st_p.as_dict()
    sampled_truth = g.truth[mapping >= 0][np.argsort(mapping[mapping >= 0])]
    return {
        "workload": f"{w['desc']}; {sampler} 50% sample, {algorithm} partition of the sample, extend, "
                    "finetune_assignment on the full graph",
        "sampler": sampler, "sample_vertices": int((mapping >= 0).sum()), "num_blocks": nb, "mdl": mdl,
        "nmi": nmi(g.truth, a), "nmi_sample": nmi(sampled_truth, a_s),
        "wall_s": wall, "phase_s": {"sample": t[1] - t[0], "partition_sample": t[2] - t[1],
                                    "extend": t[3] - t[2], "finetune_full": t[4] - t[3]},
        "proposals_per_s": (st_p.proposals + st_f.proposals) / wall,
        "accepted_moves_per_s": (st_p.moves + st_f.moves) / wall,
        "partition_sample": dict({k: p[k] for k in ("proposals", "moves", "sweeps", "merge_phases", "finetune_ms",
                                                    "merge_ms", "carry_ms", "carry_load_ms", "build_ms", "mdl_ms")},
                                 roofline=roofline(p, ctx.finetune_shape(), greedy=True, traffic_file=None)),
        "finetune_full": {"proposals": ft["proposals"], "moves": ft["moves"], "sweeps": ft["sweeps"],
                          "finetune_ms": ft["finetune_ms"]},
    }


def split_leg(ctx, dg, g, w, calls=3):
    """divisive::split_communities (src/top_down_sbp.cpp:488-549) on the large graph, from one block (the top-down
    start, target 2) and from the planted partition (target 2B); each start is rebuilt before every call, which
    uses its own epoch.  A call includes the vertex and edge counting passes, the host description lengths of
    10 B two-block models and the blockmodel rebuild."""
    from sbp_amd import native
    out = {}
    for name, B, assign, target in (("one_block", 1, np.zeros(g.N, np.int32), 2),
                                    ("planted", w["B"], np.asarray(g.truth, np.int32), 2 * w["B"])):
        ms, nb = [], None
        for call in range(calls):
            db = native.DeviceBlockmodel.build(dg, B, assign)
            ctx.synchronize()
            t0 = time.perf_counter()
            nb, delta = db.split_communities(target, call)
            ms.append((time.perf_counter() - t0) * 1e3)
            del db
        out[name] = {"blocks": B, "target": target, "blocks_after": nb, "ms_per_call": ms,
                     "best_ms": min(ms)}
        print(f"[bench] split {name}: B {B} -> {nb} in {min(ms):.1f} ms (best of {calls})", file=sys.stderr,
              flush=True)
    return out


def c2_leg(runs, cpu_full=False):
    """Full agglomerative partitions of the 200,000-vertex graph (asynchronous Gibbs), compared with the oracle's
    partition committed in tests/golden/partition_200k.npz.  With cpu_full, the oracle's full partition of the
    same graph is timed once on min(nproc, 16) OpenMP threads beside it (a few minutes)."""
    from sbp_amd import native
    g, w = make_graph("c2")
    ctx = native.Context(0, seed=0)
    dg = ctx.upload(g)
    prm = native.params("async_gibbs", seed=0)
    a, nb, st = dg.partition(prm)  # warm-up: loads the code objects, grows the buffer pool
    ctx.synchronize()
    tot = {}
    t0 = time.perf_counter()
    for _ in range(runs):
        a, nb, st = dg.partition(prm)
        add_stats(tot, st.as_dict())
    wall = time.perf_counter() - t0
    ctx_shape = ctx.finetune_shape()
    ctx.close()
    want = np.load(os.path.join(ROOT, "tests", "golden", "partition_200k.npz"))
    same = bool(nb == int(want["num_blocks"]) and np.array_equal(a, want["assignment"].astype(np.int32)) and
                st.mdl == float(want["mdl"]))
    print(f"[bench] c2: {runs} runs in {wall:.2f}s, B={nb}, identical to the oracle: {same}", file=sys.stderr,
          flush=True)
    cpu = None
    if cpu_full:
        import oracle as O
        threads = min(os.cpu_count() or 1, 16)
        t1 = time.perf_counter()
        ca, cnb, cst = O.sbp(O.OGraph.from_csr(g), O.params("async_gibbs", seed=0, threads=threads))
        cdt = time.perf_counter() - t1
        cpu = {"seconds": cdt, "cores": threads, "kind": "port", "num_blocks": int(cnb), "mdl": cst.mdl,
               "same_partition": bool(cnb == nb and np.array_equal(ca, a))}
        print(f"[bench] c2 cpu: oracle full partition {cdt:.1f}s on {threads} threads", file=sys.stderr, flush=True)
    return {
        "workload": w["desc"] + "; full SBP: async_gibbs finetune + block_merge, golden-section search",
        "runs": runs, "seconds_per_run": wall / runs, "graphs_per_s": runs / wall,
        "accepted_moves_per_s": tot["moves"] / wall, "proposals_per_s": tot["proposals"] / wall,
        "merge_proposals_per_s": tot["merge_proposals"] / wall,
        "num_blocks": nb, "mdl": st.mdl, "nmi": nmi(g.truth, a), "identical_to_oracle": same,
        "phase_ms_per_run": {k: tot[k] / runs for k in ("finetune_ms", "merge_ms", "carry_ms", "carry_load_ms",
                                                        "build_ms", "mdl_ms", "ft_kernel_ms", "ft_high_kernel_ms")},
        "roofline": roofline(tot, ctx_shape, greedy=True, traffic_file=None),
        "cpu_full_run": cpu,
    }


def c1_leg(runs):
    """Full Metropolis-Hastings partitions of the 1,000-vertex graph on the device, and the oracle's full run of the
    same (sequential by definition) on one host core."""
    import oracle as O
    from sbp_amd import native
    g, w = make_graph("c1")
    ctx = native.Context(0, seed=0)
    dg = ctx.upload(g)
    prm = native.params("metropolis_hastings", seed=0)
    dg.partition(prm)
    ctx.synchronize()
    t0 = time.perf_counter()
    props = 0
    for _ in range(runs):
        a, nb, st = dg.partition(prm)
        props += st.proposals
    wall = time.perf_counter() - t0
    ctx.close()
    og = O.OGraph.from_csr(g)
    t0 = time.perf_counter()
    a_o, nb_o, st_o = O.sbp(og, O.params("metropolis_hastings", seed=0, threads=1))
    cpu = time.perf_counter() - t0
    same = bool(nb == nb_o and np.array_equal(a, a_o) and st.mdl == st_o.mdl)
    print(f"[bench] c1: {wall / runs * 1e3:.0f} ms per run on the device, {cpu * 1e3:.0f} ms for the oracle on one "
          f"core, identical: {same}", file=sys.stderr, flush=True)
    return {"workload": w["desc"] + "; full SBP with metropolis_hastings finetune",
            "runs": runs, "ms_per_run": wall / runs * 1e3, "proposals_per_s": props / wall,
            "cpu_ms_per_run": cpu * 1e3, "cpu_proposals_per_s": st_o.proposals / cpu, "cpu_cores": 1,
            "num_blocks": nb, "mdl": st.mdl, "nmi": nmi(g.truth, a), "identical_to_oracle": same}


def calibrate_copy(mib):
    """A plain device-to-device copy of known size (the HIP runtime's copy kernel), for scaling the FETCH_SIZE /
    WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib * (1 << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_bytes": n, "copies": 3}))


LEGS = ("sampled", "split", "c2", "c1", "batched", "served")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed asynchronous-Gibbs sweeps of the headline")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=None, help="vertices of the large graph (default 5,000,000)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default=",".join(LEGS), help="legs to run besides the headline (comma list, or none)")
    ap.add_argument("--only", choices=["headline", *LEGS], help="run only that part and print its JSON object")
    ap.add_argument("--sampler", default="degree_weighted", help="sampler of the sampled leg")
    ap.add_argument("--c2-runs", type=int, default=2)
    ap.add_argument("--c2-cpu", action="store_true", help="also time the oracle's full partition of the c2 graph")
    ap.add_argument("--c1-runs", type=int, default=3)
    ap.add_argument("--batched-graphs", type=int, default=4096, help="graphs in the batched leg")
    ap.add_argument("--batched-plain-max", type=int, default=128,
                    help="graphs compared with the plain path (the batch over all graphs is timed on its own)")
    ap.add_argument("--batched-slots", type=int, default=None, help="default: the autotuned count, else 16")
    ap.add_argument("--served-requests", type=int, default=16, help="requests in the served leg")
    ap.add_argument("--served-rate", type=float, default=2.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--served-max-n", type=int, default=2_000_000, help="upper bound of the requests' N")
    ap.add_argument("--served-slots", type=int, default=None, help="default: 16")
    ap.add_argument("--many-algorithm", default="async_gibbs", help="algorithm of the batched and served legs")
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB", default=0,
                    help="only copy MIB MiB device-to-device 3 times (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the path is single-device; the "
              "design notes say how it would shard)", file=sys.stderr)
        sys.exit(2)
    legs = set() if args.legs == "none" else set(x for x in args.legs.split(",") if x)
    if args.only:
        legs = {args.only} - {"headline"}

    from sbp_amd import workloads
    from sbp_amd.autotune import tuned_slots
    if args.batched_slots is None:
        args.batched_slots = tuned_slots(16)
    if args.served_slots is None:
        # the autotuner measures the batched workload; for Poisson-arriving requests 16 workers measured best
        args.served_slots = 16
    # the mixed-size graphs are generated in worker processes (forked before the device is touched) while the
    # large graph is generated; every timed leg starts after they are all done, on an otherwise idle host
    pool = cf.ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1), mp_context=mp.get_context("fork"))
    nb_graphs = args.batched_graphs if "batched" in legs else 0
    ns_graphs = args.served_requests if "served" in legs else 0
    fut_b = [pool.submit(workloads.make, workloads.batched_spec(i)) for i in range(nb_graphs)]
    fut_s = [pool.submit(workloads.make, workloads.served_spec(j, n_max=args.served_max_n)) for j in range(ns_graphs)]

    waited = []

    def workloads_ready():
        if (fut_b or fut_s) and not waited:
            waited.append(True)
            t = time.perf_counter()
            cf.wait(fut_b + fut_s)
            print(f"[bench] batched and served graphs ready ({time.perf_counter() - t:.1f}s wait)", file=sys.stderr,
                  flush=True)

    from sbp_amd import native

    out = {}
    line = None
    if args.only in (None, "headline", "sampled", "split"):
        t0 = time.perf_counter()
        g, w = make_graph("c3", args.n)
        print(f"[bench] graph N={g.N} E={g.E} generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
              flush=True)
        ctx = native.Context(0, seed=0)
        dg = ctx.upload(g)
        ctx.synchronize()
        workloads_ready()
        if args.only in (None, "headline"):
            tot, wall, mdl0, mdl, B = headline_leg(ctx, dg, g, w, args.steps, args.warmup)
            cpu = None
            if not args.no_cpu:
                threads = min(os.cpu_count() or 1, 16)
                rate, mrate, sweeps, dt = cpu_baseline(g, w["B"], args.cpu_seconds, threads)
                cpu = {"value": mrate, "unit": "MCMC moves/s (accepted)", "cores": threads, "kind": "port",
                       "proposals_per_s": rate,
                       "sample": f"{sweeps} oracle asynchronous-Gibbs sweeps ({sweeps * g.N} proposals) with the MDL "
                                 f"after each, over the same graph from its planted partition, {dt:.1f}s on "
                                 f"{threads} OpenMP threads"}
                print(f"[bench] cpu baseline: {mrate:.0f} moves/s, {rate:.0f} proposals/s on {threads} threads",
                      file=sys.stderr, flush=True)
            line = {
                "metric": "MCMC moves/s, batched graphs/s & p99 latency vs plain GPU SBP at equal NMI",
                "value": tot["moves"] / wall,
                "unit": "MCMC moves/s (accepted vertex moves per second of wall time)",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "int32 counts / fp64 delta-MDL", "data": "synthetic",
                "config": {"workload": w["desc"] + "; one step = one asynchronous-Gibbs sweep over every vertex "
                                                   "(proposals, commit, rebuild, MDL) from the planted partition",
                           "N": g.N, "E": g.E, "planted_blocks": w["B"], "algorithm": "async_gibbs", "batches": 1},
                "proposals_per_s": tot["proposals"] / wall,
                "moves_per_sweep": tot["moves"] / args.steps,
                "mdl_start": mdl0, "mdl_end": mdl, "num_blocks": B,
                "phase_ms_per_step": {k: tot[k] / args.steps for k in
                                      ("finetune_ms", "build_ms", "mdl_ms", "ft_kernel_ms", "ft_high_kernel_ms")},
                "roofline": roofline(tot, ctx.finetune_shape(), greedy=True, traffic_file=TRAFFIC_FILE, B=B),
                "finetune_shape": {"lanes": ctx.finetune_shape()[0], "waves_per_simd": ctx.finetune_shape()[1]},
                "cpu_baseline": cpu,
            }
        if "split" in legs:
            out["split"] = split_leg(ctx, dg, g, w)
        if "sampled" in legs:
            out["sampled"] = sampled_leg(ctx, dg, g, w, args.sampler)
        del dg
        ctx.close()
        del g
    workloads_ready()
    if "c2" in legs:
        out["c2"] = c2_leg(args.c2_runs, args.c2_cpu)
    if "c1" in legs:
        out["c1"] = c1_leg(args.c1_runs)
    if "batched" in legs:
        many_b = []
        for i, f in enumerate(fut_b):
            many_b.append(f.result())
            if (i + 1) % 512 == 0:
                print(f"[bench] generated {i + 1} batched graphs", file=sys.stderr, flush=True)
        ctx = native.Context(0, seed=0)
        plain_max = min(args.batched_plain_max, len(many_b))
        out["batched"] = batched_leg(ctx, many_b, args.many_algorithm, args.batched_slots, plain_max)
        out["batched"].pop("_outs", None)
        ctx.close()
        del many_b
    if "served" in legs:
        many_s = [f.result() for f in fut_s]
        arrivals = workloads.poisson_arrivals(args.served_rate, len(many_s), seed=0)
        out["served"] = served_leg(many_s, arrivals, args.many_algorithm, args.served_slots, args.served_rate,
                                   args.served_max_n)
    pool.shutdown()
    if args.only and args.only != "headline":
        print(json.dumps({args.only: out[args.only]}))
        return
    line.update(out)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
