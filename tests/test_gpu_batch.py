"""Batched (sbp_partition_batch) and served (sbp_server_*) partitioning on the GPU: every graph's result
(assignment, block count, description length) equals the one sbp_partition gives it alone, whatever the
number of workers, the order of completion or the other requests in flight; the server overlaps requests,
admits them by footprint and keeps the measured device footprint of a run under its estimate."""
import threading
import time

import numpy as np
import pytest

from sbp_amd import native, workloads
from sbp_amd.graph import dcsbm

pytestmark = pytest.mark.gpu

ALGS = ["async_gibbs", "metropolis_hastings", "hybrid_mcmc"]


def mixed_graphs(count, n_max=1500):
    specs = [workloads.batched_spec(i, n_max=n_max) for i in range(count)]
    return specs, [workloads.make(s) for s in specs]


def mixed_params(count):
    out = []
    for i in range(count):
        # degree correction on an async-Gibbs graph: with Metropolis-Hastings its histogram drift (the
        # reference's) can stop the early exit from firing, which makes a run of 100 sweeps per call
        out.append(native.params(ALGS[i % 3], seed=1000 + i, nonparametric=(i % 4 == 3),
                                 degreecorrected=(i == 3)))
    return out


@pytest.fixture(scope="module")
def workload():
    specs, graphs = mixed_graphs(10)
    prms = mixed_params(len(graphs))
    solo = []
    c = native.Context(0)
    for i, (g, p) in enumerate(zip(graphs, prms)):
        print(f"[solo] graph {i} N={g.N} E={g.E} alg={p.algorithm}", flush=True)
        t0 = time.perf_counter()
        a, nb, st = c.upload(g).partition(p)
        solo.append((a, nb, st.mdl))
        print(f"  B={nb} in {time.perf_counter() - t0:.2f}s", flush=True)
    c.close()
    return graphs, prms, solo


def same(got, want):
    a, nb, mdl = got
    wa, wnb, wmdl = want
    assert nb == wnb
    assert np.array_equal(a, wa)
    assert mdl == wmdl


@pytest.mark.parametrize("slots", [1, 4, 10])
def test_batch_equals_solo(workload, slots):
    graphs, prms, solo = workload
    print(f"[batch] slots={slots}", flush=True)
    c = native.Context(0)
    dgs = [c.upload(g) for g in graphs]
    t0 = time.perf_counter()
    outs, nb, st, fin = c.partition_batch(dgs, prms, slots=slots)
    print(f"  batch of {len(graphs)} in {time.perf_counter() - t0:.2f}s", flush=True)
    for i in range(len(graphs)):
        same((outs[i], int(nb[i]), st[i].mdl), solo[i])
    assert np.all(fin > 0)
    # a second call reuses the workers
    outs2, nb2, st2, _ = c.partition_batch(dgs[:3], prms[:3], slots=slots)
    for i in range(3):
        same((outs2[i], int(nb2[i]), st2[i].mdl), solo[i])
    c.close()


def test_batch_reports_failing_graph():
    c = native.Context(0)
    g = dcsbm(300, 3, seed=1)
    dg = c.upload(g)
    bad = native.params("async_gibbs", seed=1)
    bad.batches = 0
    with pytest.raises(native.SbpError):
        c.partition_batch([dg, dg], [native.params(seed=1), bad])
    c.close()


def test_server_overlaps_and_matches_solo(workload):
    graphs, prms, solo = workload
    srv = native.Server(0, slots=4)
    tickets = []
    for i, (g, p) in enumerate(zip(graphs, prms)):
        tickets.append(srv.submit(g, p))
        time.sleep(0.002 * (i % 3))  # staggered arrivals while earlier requests run
    res = [srv.wait(t) for t in tickets]
    info = srv.info()
    srv.close()
    for i, r in enumerate(res):
        same((r[0], r[1], r[2].mdl), solo[i])
        assert r[3] >= r[4] >= 0.0  # latency covers the time in the queue
    assert info["max_running"] > 1
    assert info["running"] == 0 and info["queued"] == 0 and info["in_use"] == 0


def test_server_waits_from_other_threads(workload):
    graphs, prms, solo = workload
    srv = native.Server(0, slots=3)
    tickets = [srv.submit(g, p) for g, p in zip(graphs[:6], prms[:6])]
    got = {}

    def waiter(k):
        got[k] = srv.wait(tickets[k])

    ths = [threading.Thread(target=waiter, args=(k,)) for k in range(6)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    srv.close()
    for k in range(6):
        same((got[k][0], got[k][1], got[k][2].mdl), solo[k])


def test_server_admission_by_footprint():
    """A budget that holds one request at a time runs them one by one with the same results; a request larger
    than the budget is refused at submission."""
    graphs = [dcsbm(2000, 20, mean_degree=10, kmax=200, seed=50 + k) for k in range(4)]
    prms = [native.params("hybrid_mcmc", seed=k) for k in range(4)]
    c = native.Context(0)
    solo = []
    for g, p in zip(graphs, prms):
        a, nb, st = c.upload(g).partition(p)
        solo.append((a, nb, st.mdl))
    c.close()
    fps = [native.footprint_bytes(g, "hybrid_mcmc") for g in graphs]
    assert min(fps) > 0.75 * max(fps)
    # the server sets aside min(4 GiB, budget / (4 slots)) per worker, a quarter of this budget; what is
    # left (1.5 x the largest footprint) admits any one of these requests but no two
    budget = 2 * max(fps)
    srv = native.Server(0, slots=4, hbm_budget_bytes=budget)
    left = srv.info()["budget"]
    assert max(fps) <= left < 2 * min(fps)
    tickets = [srv.submit(g, p) for g, p in zip(graphs, prms)]
    res = [srv.wait(t) for t in tickets]
    info = srv.info()
    huge = dcsbm(20000, 50, mean_degree=30, kmax=2000, seed=3)
    assert native.footprint_bytes(huge, "hybrid_mcmc") > left
    with pytest.raises(native.SbpError):
        srv.submit(huge, prms[0])
    srv.close()
    assert info["max_running"] == 1
    assert info["peak_in_use"] <= left
    for i, r in enumerate(res):
        same((r[0], r[1], r[2].mdl), solo[i])


@pytest.mark.parametrize("n,b,mean,kmax,alg", [(3000, 20, 12, 300, "hybrid_mcmc"),
                                               (40000, 70, 20, 1000, "async_gibbs"),
                                               (20000, 50, 16, 2000, "async_gibbs"),
                                               (6000, 30, 20, 1000, "hybrid_mcmc"),
                                               (2000, 20, 8, 1000, "metropolis_hastings")])
def test_footprint_bounds_measured_peak(n, b, mean, kmax, alg):
    g = dcsbm(n, b, gamma=2.1, kmin=1, kmax=kmax, mean_degree=mean, p_in=0.6, seed=n)
    print(f"[footprint] N={g.N} E={g.E} {alg}", flush=True)
    c = native.Context(0)
    c.upload(g).partition(native.params(alg, seed=2))
    _, peak = c.memory()
    c.close()
    est = native.footprint_bytes(g, alg)
    print(f"  peak {peak} + graph {native.graph_bytes(g.N, g.E)} <= estimate {est}", flush=True)
    assert peak + native.graph_bytes(g.N, g.E) <= est
