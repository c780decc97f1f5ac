"""GPU parity of the nonparametric description length (entropy::nonparametric, src/entropy.cpp:617-1225)
through the C-ABI against the oracle: the description length and its parts, proposals, sweeps, merges and
a whole run, bit-exact; the reference's toy known answers on the device."""
import json
import os

import numpy as np
import pytest

import oracle as O
from sbp_amd import native
from sbp_amd.graph import CSRGraph, dcsbm

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOY = json.load(open(os.path.join(HERE, "golden", "toy_example.json")))


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0, seed=99)
    yield c
    c.close()


def graph(seed=5, N=800, kmax=400):
    return dcsbm(N, 8, gamma=2.3, kmin=1, kmax=kmax, mean_degree=8, p_in=0.7, seed=seed)


def both(ctx, g, B, assign):
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    return (og, O.OBlockmodel.build(og, B, assign, nonparametric=True), dg,
            native.DeviceBlockmodel.build(dg, B, assign, nonparametric=True))


def same_state(ob, db):
    x, y = ob.export(), db.export()
    assert x["B"] == y["B"]
    for k in ("row_ptr", "cols", "cnts", "sizes", "assignment"):
        assert np.array_equal(x[k], y[k]), k


def random_assign(N, B, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, B, N).astype(np.int32)
    a[:B] = np.arange(B)
    return a


def test_toy_description_length_on_device(ctx):
    e = np.array(TOY["edges"])
    g = CSRGraph.from_edges(e[:, 0], e[:, 1], TOY["num_vertices"])
    t = TOY["nonparametric"]["toy"]
    og, ob, dg, db = both(ctx, g, 3, t["assignment"])
    mdl, sparse, part, edges = db.np_parts()
    osp, opart, oedges = ob.np_parts()
    assert (sparse, part, edges) == (osp, opart, oedges)
    assert mdl == ob.mdl()
    # the reference's check is EXPECT_FLOAT_EQ (4 single-precision ULPs)
    assert abs(int(np.float32(mdl).view(np.int32)) - int(np.float32(t["mdl"]).view(np.int32))) <= 4
    assert db.mdl()[0] == mdl


@pytest.mark.parametrize("B", [6, 60, 600])
def test_description_length_bit_exact(ctx, B):
    g = graph(seed=B)
    og, ob, dg, db = both(ctx, g, B, random_assign(g.N, B, B))
    assert db.np_parts()[0] == ob.mdl()


@pytest.mark.parametrize("greedy", [True, False])
@pytest.mark.parametrize("B", [4, 40, 400])
def test_propose_all_bit_exact(ctx, greedy, B):
    print(f"[np propose] B={B} greedy={greedy}", flush=True)
    g = graph(seed=B + 1)
    og, ob, dg, db = both(ctx, g, B, random_assign(g.N, B, B + 1))
    ctx.set_seed(7)
    s_d, dS_d, acc_d = db.propose_all(epoch=2, greedy=greedy)
    for v in range(g.N):
        s, dS, acc = ob.ft_propose(7, 2, v, greedy)
        assert (s_d[v], dS_d[v], acc_d[v]) == (s, dS, acc), v


@pytest.mark.parametrize("alg", ["async_gibbs", "metropolis_hastings", "hybrid_mcmc"])
def test_mcmc_phase_matches(ctx, alg):
    print(f"[np mcmc] {alg}", flush=True)
    g = graph(seed=31, N=900, kmax=600)
    B = 30
    og, ob, dg, db = both(ctx, g, B, random_assign(g.N, B, 4))
    ctx.set_seed(13)
    ep_o, st_o = ob.mcmc(O.params(alg, seed=13, greedy=False, nonparametric=True, mh_percent=0.1), True, 0)
    ep_d, mdl_d, st_d = db.mcmc(native.params(alg, seed=13, greedy=False, nonparametric=True, mh_percent=0.1),
                                True, 0)
    assert ep_o == ep_d
    same_state(ob, db)
    assert mdl_d == ob.mdl()


def test_mh_occupying_empty_blocks(ctx):
    """Blocks left empty by the initial assignment: a move into one changes the non-empty block count."""
    g = graph(seed=32, N=500)
    B = 20
    a = random_assign(g.N, B, 5) % 15  # labels 15..19 unused
    og, ob, dg, db = both(ctx, g, B, a)
    ctx.set_seed(21)
    tot_o, _ = ob.mh_sweep(O.params("metropolis_hastings", seed=21, greedy=False, nonparametric=True), 0)
    tot_d, _ = db.mh_sweep(native.params("metropolis_hastings", seed=21, greedy=False, nonparametric=True), 0)
    assert tot_o == tot_d
    same_state(ob, db)


@pytest.mark.parametrize("B", [40, 800])
def test_merge_propose_bit_exact(ctx, B):
    g = graph(seed=B + 7, N=800)
    og, ob, dg, db = both(ctx, g, B, random_assign(g.N, B, B))
    ctx.set_seed(3)
    s_d, dS_d = db.merge_propose(epoch=1)
    s_o, dS_o = ob.merge_propose(3, 1)
    assert np.array_equal(s_d, s_o) and np.array_equal(dS_d, dS_o)


@pytest.mark.parametrize("approximate", [False, True])
def test_merge_blocks_matches(ctx, approximate):
    g = graph(seed=41, N=1000)
    B = g.N
    og, ob, dg, db = both(ctx, g, B, np.arange(B, dtype=np.int32))
    ctx.set_seed(8)
    out_d, _ = db.merge_blocks(epoch=0, num_to_merge=B // 2, approximate=approximate)
    s_o, dS_o = ob.merge_propose(8, 0)
    out_o = ob.merge_apply(s_o, dS_o, B // 2, approximate)
    same_state(out_o, out_d)


@pytest.mark.parametrize("alg", ["async_gibbs", "hybrid_mcmc"])
def test_full_partition_bit_exact(ctx, alg):
    print(f"[np partition] {alg}", flush=True)
    g = dcsbm(1000, 16, gamma=2.5, kmin=1, kmax=100, mean_degree=10, p_in=0.75, seed=0)
    og = O.OGraph.from_csr(g)
    a_o, nb_o, st_o = O.sbp(og, O.params(alg, seed=0, greedy=False, nonparametric=True, threads=8))
    dg = ctx.upload(g)
    a_d, nb_d, st_d = dg.partition(native.params(alg, seed=0, greedy=False, nonparametric=True))
    assert nb_d == nb_o and np.array_equal(a_d, a_o)
    assert st_d.mdl == st_o.mdl
