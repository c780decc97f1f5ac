"""GPU parity of the sampling path (src/sample.cpp) through the C-ABI against the oracle: the sampled
vertices and mapping, the induced subgraph's CSR (list order included), the extension to the full graph,
and the whole sampled run (sample -> partition the sample -> extend -> finetune_assignment on the full
graph), all bit-exact."""
import numpy as np
import pytest

import oracle as O
from sbp_amd import native
from sbp_amd.graph import dcsbm

pytestmark = pytest.mark.gpu

ALGS = ["max_degree", "random", "degree_product", "degree_weighted", "expansion_snowball", "degree_product_snowball"]


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0, seed=1234)
    yield c
    c.close()


def graph(N=3000, seed=21, kmax=300):
    return dcsbm(N, 12, gamma=2.1, kmin=1, kmax=kmax, mean_degree=10, p_in=0.5, sigma=1.0, seed=seed)


def same_graph(dgraph, ograph):
    oo, oa, io, ia = dgraph.download()
    assert dgraph.N == ograph.N and dgraph.E == ograph.E
    assert np.array_equal(oo, ograph.out_off) and np.array_equal(io, ograph.in_off)
    assert np.array_equal(oa, ograph.out_adj) and np.array_equal(ia, ograph.in_adj)


@pytest.mark.parametrize("alg", ALGS)
@pytest.mark.parametrize("fraction", [0.05, 0.5])
def test_sample_and_subgraph_bit_exact(ctx, alg, fraction):
    print(f"[sample] alg={alg} fraction={fraction}", flush=True)
    g = graph()
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    sg, m = dg.sample(alg, fraction, seed=3)
    om, ovs = O.sample(og, alg, fraction, 3)
    assert np.array_equal(m, om)
    same_graph(sg, O.subgraph(og, om, ovs))


@pytest.mark.parametrize("alg", ["max_degree", "random", "degree_weighted", "expansion_snowball",
                                 "degree_product_snowball"])
def test_extend_bit_exact(ctx, alg):
    print(f"[extend] alg={alg}", flush=True)
    g = graph(seed=22)
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    sg, m = dg.sample(alg, 0.4, seed=4)
    n = sg.N
    B = 15
    rng = np.random.default_rng(1)
    sa = rng.integers(0, B, n).astype(np.int32)
    sa[:B] = np.arange(B)
    sbm = native.DeviceBlockmodel.build(sg, B, sa)
    full, a = sbm.extend(dg, seed=5)
    om, _ = O.sample(og, alg, 0.4, 4)
    assert np.array_equal(a, O.extend(og, om, sa, B, 5))
    assert full.validate()
    assert np.array_equal(full.export()["assignment"], a)


def test_extend_high_degree_unsampled(ctx):
    """Random sampling leaves hubs (total degree > 256) unsampled: they take the global-table launch."""
    g = graph(N=4000, seed=23, kmax=1500)
    assert (g.degrees() > 256).sum() > 5
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    sg, m = dg.sample("random", 0.3, seed=8)
    B = 9
    sa = (np.arange(sg.N) % B).astype(np.int32)
    sbm = native.DeviceBlockmodel.build(sg, B, sa)
    _, a = sbm.extend(dg, seed=2)
    om, _ = O.sample(og, "random", 0.3, 8)
    assert np.array_equal(a, O.extend(og, om, sa, B, 2))


def test_sampled_run_bit_exact(ctx):
    """The large single-graph configuration's flow (max_degree sampling at 50%, partition, extend,
    finetune_assignment on the full graph) on a 20,000-vertex graph of the same shape."""
    g = dcsbm(20000, 20, gamma=2.1, kmin=1, kmax=800, mean_degree=16, p_in=0.4, sigma=1.0, seed=0)
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    sg, m = dg.sample("max_degree", 0.5)
    om, ovs = O.sample(og, "max_degree", 0.5)
    assert np.array_equal(m, om)
    osg = O.subgraph(og, om, ovs)
    print("[sampled run] partition of the sample", flush=True)
    a_d, nb_d, _ = sg.partition(native.params("hybrid_mcmc", seed=0))
    a_o, nb_o, st_o = O.sbp(osg, O.params("hybrid_mcmc", seed=0, threads=8))
    assert nb_d == nb_o and np.array_equal(a_d, a_o)
    print("[sampled run] extend + finetune_assignment", flush=True)
    sbm = native.DeviceBlockmodel.build(sg, nb_d, a_d)
    full, a = sbm.extend(dg, seed=0)
    assert np.array_equal(a, O.extend(og, om, a_o, nb_o, 0))
    ob = O.OBlockmodel.build(og, nb_o, a)
    ctx.set_seed(0)
    ep_d, mdl_d, _ = full.finetune_assignment(native.params("hybrid_mcmc", seed=0), 0)
    ep_o, _ = ob.finetune_assignment(O.params("hybrid_mcmc", seed=0), 0)
    assert ep_d == ep_o
    assert np.array_equal(full.export()["assignment"], ob.export()["assignment"])
    assert mdl_d == ob.mdl()
    from sklearn.metrics import normalized_mutual_info_score as nmi
    assert nmi(g.truth, full.export()["assignment"]) > 0.5


def test_sample_extend_500k_bit_exact(ctx):
    """Size-independent checks at 500,000 vertices (8M edges): sample, subgraph and extend bit-exact."""
    g = dcsbm(500000, 300, gamma=2.1, kmin=1, kmax=5000, mean_degree=16, p_in=0.4, sigma=1.0, seed=1)
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    for alg in ("max_degree", "degree_weighted"):
        print(f"[500k] alg={alg}", flush=True)
        sg, m = dg.sample(alg, 0.5, seed=9)
        om, ovs = O.sample(og, alg, 0.5, 9)
        assert np.array_equal(m, om)
        same_graph(sg, O.subgraph(og, om, ovs))
        B = 300
        sa = (np.arange(sg.N) % B).astype(np.int32)
        sbm = native.DeviceBlockmodel.build(sg, B, sa)
        full, a = sbm.extend(dg, seed=3)
        assert np.array_equal(a, O.extend(og, om, sa, B, 3))
        assert (a >= 0).all() and (a < B).all()


@pytest.mark.parametrize("alg", ["expansion_snowball", "degree_product_snowball"])
def test_snowball_sampled_run_bit_exact(ctx, alg):
    """A sampled run with a snowball sampler: sample (host frontier), subgraph (device), partition, extend."""
    g = dcsbm(6000, 12, gamma=2.2, kmin=1, kmax=300, mean_degree=10, p_in=0.6, sigma=1.0, seed=4)
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    print(f"[snowball run] {alg}", flush=True)
    sg, m = dg.sample(alg, 0.4, seed=6)
    om, ovs = O.sample(og, alg, 0.4, 6)
    assert np.array_equal(m, om)
    osg = O.subgraph(og, om, ovs)
    same_graph(sg, osg)
    a_d, nb_d, _ = sg.partition(native.params("async_gibbs", seed=0))
    a_o, nb_o, _ = O.sbp(osg, O.params("async_gibbs", seed=0, threads=8))
    assert nb_d == nb_o and np.array_equal(a_d, a_o)
    sbm = native.DeviceBlockmodel.build(sg, nb_d, a_d)
    _, a = sbm.extend(dg, seed=0)
    assert np.array_equal(a, O.extend(og, om, a_o, nb_o, 0))


def test_detached_run_bit_exact(ctx):
    """--detach (src/main.cpp:86-92,153-156): the vertices of degree > 1 are partitioned, the others
    reattached; the detached graph, its partition and the reattached assignment are bit-exact."""
    from test_detach_oracle import leafy_graph
    g = leafy_graph(5, N=3000, mean_degree=10)
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    dd, m = dg.detach()
    om, ovs = O.detach(og)
    assert np.array_equal(m, om)
    osg = O.subgraph(og, om, ovs)
    same_graph(dd, osg)
    print("[detached run] partition of the detached graph", flush=True)
    a_d, nb_d, _ = dd.partition(native.params("hybrid_mcmc", seed=3))
    a_o, nb_o, _ = O.sbp(osg, O.params("hybrid_mcmc", seed=3, threads=8))
    assert nb_d == nb_o and np.array_equal(a_d, a_o)
    full = native.reattach(g, m, a_d, nb_d, seed=3)
    assert np.array_equal(full, O.reattach(og, om, a_o, nb_o, seed=3))
    bm = native.DeviceBlockmodel.build(dg, nb_d, full)
    assert bm.validate()
