"""Autotuner of the batched mode's worker-slot count and of the finetune proposal launch shape.

Worker slots.  sbp_partition_batch runs one host thread and one stream per slot.  Too few slots leave the
device idle while graphs wait on the host; too many oversubscribe the host's cores and the device's four
hardware queues.  The best count depends on the host's cores and the graph mix, so it is measured: graphs of
the batched workload with seeds 1000.. (apart from the benchmark's 0..) are partitioned as one batch at each
candidate count, the best of two calls is kept, and the fastest count is written to autotune.json under a key
for the host (its usable CPUs).  Context.partition_batch(slots=None) and bench.py read it.

Launch shape (--finetune-shape).  The proposal launches run per degree class of k = |out(v)| + |in(v)| (k <= 8,
32, 128, 256 on lane groups of a wavefront; k <= 1,024 on one workgroup per vertex): how many lanes a vertex
gets, and whether the lane-group launch's registers are sized for 3 or 4 wavefronts per SIMD, trade lane
utilisation against loads in flight.  Each class is its own launch, timed with HIP events, so every class is
tuned at once: asynchronous-Gibbs sweeps of a 1,000,000-vertex graph shaped like the 5,000,000-vertex workload
(power law 2.1 on [1, 20000], mean 16, 40% intra-block, 1,000 lognormal(1.0) blocks, planted partition) run
under each candidate of every class, from the same state and epochs, and each class keeps its fastest.  The
result is written under "finetune_shape"; native.Context applies it.

    python -m sbp_amd.autotune [--graphs 32] [--slots 4,8,12,16,20,24] [--out FILE]
    python -m sbp_amd.autotune --finetune-shape [--n 1000000] [--out FILE]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "autotune.json")


def host_key():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return f"{n}cpu"


def tuned_slots(default=0):
    """The tuned slot count for this host, or `default` when none was measured."""
    try:
        with open(TUNED) as f:
            return int(json.load(f)[host_key()]["slots"])
    except (OSError, KeyError, ValueError, TypeError):
        return default


def tuned_shape():
    """(lanes[5], waves_per_simd[4]) measured by --finetune-shape, or None."""
    try:
        with open(TUNED) as f:
            t = json.load(f)["finetune_shape"]
        return [int(x) for x in t["lanes"]], [int(x) for x in t["waves_per_simd"]]
    except (OSError, KeyError, ValueError, TypeError):
        return None


# candidates per degree class: (lanes, waves per SIMD); class 4 varies its threads per vertex only
SHAPE_CANDIDATES = [
    [(2, 4), (2, 3), (4, 4), (4, 3), (8, 4), (8, 3), (16, 4), (32, 4)],
    [(4, 4), (4, 3), (8, 4), (8, 3), (16, 4), (16, 3), (32, 4), (64, 4)],
    [(16, 4), (16, 3), (32, 4), (32, 3), (64, 4), (64, 3)],
    [(32, 4), (32, 3), (64, 4), (64, 3)],
    [(256, 4), (128, 4)],
]


def tune_shape(n, sweeps=3):
    from sbp_amd import native
    from sbp_amd.graph import dcsbm

    g = dcsbm(n, max(2, n // 1000), gamma=2.1, kmin=1, kmax=20000, mean_degree=16, p_in=0.4, sigma=1.0, seed=7)
    ctx = native.Context(0, seed=7)
    dg = ctx.upload(g)
    prm = native.params("async_gibbs", seed=7)
    rounds = max(len(c) for c in SHAPE_CANDIDATES)
    times = [dict() for _ in SHAPE_CANDIDATES]
    for i in range(rounds):
        pick = [c[i % len(c)] for c in SHAPE_CANDIDATES]
        ctx.set_finetune_shape([p[0] for p in pick], [p[1] for p in pick[:4]])
        db = native.DeviceBlockmodel.build(dg, max(2, n // 1000), g.truth)
        db.async_gibbs_sweep(prm, 0)  # warm-up
        ms = [0.0] * 6
        for ep in range(1, sweeps + 1):
            st = db.async_gibbs_sweep(prm, ep).as_dict()
            ms = [a + b for a, b in zip(ms, st["ft_class_ms"])]
        for b, p in enumerate(pick):
            if i < len(SHAPE_CANDIDATES[b]):
                times[b][p] = ms[b] / sweeps
        print(f"[autotune] shape {pick}: class ms {[round(x / sweeps, 3) for x in ms]}", flush=True)
        del db
    ctx.close()
    best = [min(t, key=t.get) for t in times]
    return {"lanes": [p[0] for p in best], "waves_per_simd": [p[1] for p in best[:4]],
            "class_ms": [{f"{k[0]}x{k[1]}": round(v, 4) for k, v in t.items()} for t in times],
            "graph": f"DCSBM N={g.N} E={g.E}, power law 2.1 on [1,20000] mean 16, p_in 0.4, seed 7", "sweeps": sweeps}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--finetune-shape", action="store_true", help="tune the proposal launch shape instead")
    ap.add_argument("--n", type=int, default=1_000_000, help="vertices of the launch-shape calibration graph")
    ap.add_argument("--graphs", type=int, default=32)
    ap.add_argument("--slots", default="4,8,12,16,20,24")
    ap.add_argument("--algorithm", default="async_gibbs")
    ap.add_argument("--out", default=TUNED, help="the table to update (default: autotune.json in the package)")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(HERE))
    from sbp_amd import native, workloads

    if args.finetune_shape:
        shape = tune_shape(args.n)
        table = {}
        for src in (args.out, TUNED):
            if os.path.exists(src):
                with open(src) as f:
                    table = json.load(f)
                break
        table["finetune_shape"] = shape
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        print(json.dumps({"finetune_shape": {"lanes": shape["lanes"], "waves_per_simd": shape["waves_per_simd"]}}))
        return

    specs = [workloads.batched_spec(1000 + i) for i in range(args.graphs)]
    graphs = workloads.generate(specs)
    ctx = native.Context(0)
    dgs = [ctx.upload(g) for g in graphs]
    prms = [native.params(args.algorithm, seed=1000 + i) for i in range(args.graphs)]
    ctx.synchronize()
    measured = {}
    for k in [int(x) for x in args.slots.split(",")]:
        best = float("inf")
        for _ in range(2):
            t0 = time.perf_counter()
            ctx.partition_batch(dgs, prms, slots=k)
            best = min(best, time.perf_counter() - t0)
        measured[k] = args.graphs / best
        print(f"[autotune] slots={k}: {measured[k]:.1f} graphs/s", flush=True)
    ctx.close()
    slots = max(measured, key=measured.get)
    table = {}
    for src in (args.out, TUNED):
        if os.path.exists(src):
            with open(src) as f:
                table = json.load(f)
            break
    table[host_key()] = {"slots": slots, "graphs_per_s": {str(k): round(v, 2) for k, v in measured.items()},
                         "graphs": args.graphs, "algorithm": args.algorithm}
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({"host": host_key(), "slots": slots}))


if __name__ == "__main__":
    main()
