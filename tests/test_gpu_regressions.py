"""Regression tests on the GPU for paths that once read out of bounds or raced:
- a one-block blockmodel of a graph with an isolated vertex (common::propose_random_block draws from the empty
  range U[0, B-2] there; src/common.cpp:200-209): finetune_assignment and the MH sweep leave it unchanged;
- one graph handle shared by several workers of a batch with hybrid MCMC (the degree order is built once and
  published only when complete);
- a second wait on a server ticket that another thread is already waiting for is refused, and the first
  waiter still gets its result."""
import threading

import numpy as np
import pytest

import oracle as O
from sbp_amd import native
from sbp_amd.graph import CSRGraph, dcsbm

pytestmark = pytest.mark.gpu


def with_isolated_vertex(g):
    """g plus one vertex with no edges (the last id)."""
    out_off = np.append(g.out_off, g.out_off[-1]).astype(np.int64)
    in_off = np.append(g.in_off, g.in_off[-1]).astype(np.int64)
    return CSRGraph(g.N + 1, out_off, g.out_adj.copy(), in_off, g.in_adj.copy())


@pytest.mark.parametrize("entry", ["finetune_assignment", "mh_sweep", "mh_sweep_ordered"])
def test_one_block_with_isolated_vertex(entry):
    g = with_isolated_vertex(dcsbm(400, 4, seed=3))
    print(f"[one-block] entry={entry} N={g.N} E={g.E}", flush=True)
    assign = np.zeros(g.N, np.int32)
    ctx = native.Context(0, seed=5)
    dg = ctx.upload(g)
    db = native.DeviceBlockmodel.build(dg, 1, assign)
    prm = native.params("metropolis_hastings", seed=5)
    if entry == "finetune_assignment":
        _, mdl, _ = db.finetune_assignment(prm, 0)
    else:
        db.mh_sweep(prm, 0, ordered=(entry == "mh_sweep_ordered"))
        mdl = db.mdl()[0]
    x = db.export()
    assert np.array_equal(x["assignment"], assign)
    assert db.validate()
    ob = O.OBlockmodel.build(O.OGraph.from_csr(g), 1, assign)
    assert mdl == ob.mdl()
    ctx.close()


def test_batch_shares_one_graph_hybrid():
    g = dcsbm(3000, 10, seed=11)
    prms = [native.params("hybrid_mcmc", seed=s) for s in (21, 22, 23, 24)]
    ctx = native.Context(0)
    solo = []
    for p in prms:
        a, nb, st = ctx.upload(g).partition(p)
        solo.append((a, nb, st.mdl))
    dg = ctx.upload(g)
    print("[batch] one handle, four hybrid runs on four workers", flush=True)
    outs, nb, st, _ = ctx.partition_batch([dg] * len(prms), prms, slots=4)
    for i in range(len(prms)):
        assert int(nb[i]) == solo[i][1]
        assert np.array_equal(outs[i], solo[i][0])
        assert st[i].mdl == solo[i][2]
    ctx.close()


def test_server_second_wait_refused():
    g = dcsbm(2000, 8, seed=4)
    srv = native.Server(0, slots=1)
    t = srv.submit(g, native.params("async_gibbs", seed=9))
    got = {}

    def first():
        try:
            got["first"] = srv.wait(t)
        except native.SbpError as e:
            got["first_error"] = str(e)

    th = threading.Thread(target=first)
    th.start()
    # whichever call reaches the server second is refused; the other gets the result
    try:
        got["second"] = srv.wait(t)
    except native.SbpError as e:
        got["second_error"] = str(e)
    th.join(timeout=120)
    assert not th.is_alive()
    if "second" in got:  # this thread claimed the ticket first; the thread's call was refused
        assert "first" not in got and "first_error" in got
        res = got["second"]
    else:
        assert "already being waited for" in got["second_error"] or "unknown ticket" in got["second_error"]
        res = got["first"]
    a, nb, st, lat, q = res
    ctx = native.Context(0)
    a0, nb0, st0 = ctx.upload(g).partition(native.params("async_gibbs", seed=9))
    assert nb == nb0 and np.array_equal(a, a0) and st.mdl == st0.mdl
    ctx.close()
    srv.close()
