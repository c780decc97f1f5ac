"""Autotuner of the batched mode's worker-slot count.

sbp_partition_batch runs one host thread and one stream per slot.  Too few slots leave the device idle
while graphs wait on the host; too many oversubscribe the host's cores and the device's four hardware
queues.  The best count depends on the host's cores and the graph mix, so it is measured: graphs of the
batched workload with seeds 1000.. (apart from the benchmark's 0..) are partitioned as one batch at each
candidate count, the best of two calls is kept, and the fastest count is written to autotune.json under a
key for the host (its usable CPUs).  Context.partition_batch(slots=None) and bench.py read it.

    python -m sbp_amd.autotune [--graphs 32] [--slots 4,8,12,16,20,24] [--out FILE]
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=32)
    ap.add_argument("--slots", default="4,8,12,16,20,24")
    ap.add_argument("--algorithm", default="async_gibbs")
    ap.add_argument("--out", default=TUNED, help="the table to update (default: autotune.json in the package)")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(HERE))
    from sbp_amd import native, workloads

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
