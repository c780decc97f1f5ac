"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the same seeded inputs.

Integer state (block matrix, degrees, sizes, proposals, assignments) must be bit-exact; delta-MDL and
MDL are FP64 sums of glibc-logf terms accumulated in 128-bit fixed point on both sides, so they must be
bit-identical too (the stated tolerance, 1e-12 relative, is checked as well for the record).
"""
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
    c = native.Context(0, seed=1234)
    yield c
    c.close()


def toy_graph():
    e = np.array(TOY["edges"])
    return CSRGraph.from_edges(e[:, 0], e[:, 1], TOY["num_vertices"])


def small_graph(seed=3, N=500, B=8, kmax=60, mean=8):
    return dcsbm(N, B, gamma=2.5, kmin=1, kmax=kmax, mean_degree=mean, p_in=0.7, seed=seed)


def both(ctx, g, B, assign):
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    return og, O.OBlockmodel.build(og, B, assign), dg, native.DeviceBlockmodel.build(dg, B, assign)


def same_state(ob, db):
    x, y = ob.export(), db.export()
    assert x["B"] == y["B"]
    for k in ("row_ptr", "cols", "cnts", "sizes", "assignment"):
        assert np.array_equal(x[k], y[k]), k
    for k in ("d_out", "d_in", "d_tot"):
        assert np.array_equal(x[k], y[k].astype(np.int64)), k


# ------------------------------------------------------------------ primitives

def test_device_logf_matches_glibc(ctx):
    """Restated glibc logf on the device: every integer count up to 2^25 and, for every exponent of the
    positive floats (subnormals included), three windows of 65536 consecutive mantissas."""
    ints = np.arange(1, 1 << 25, dtype=np.float32).view(np.uint32)
    lo, hi = int(ints[0]), int(ints[-1]) + 1
    for start in range(lo, hi, 1 << 24):
        n = min(1 << 24, hi - start)
        assert np.array_equal(ctx.debug_logf(start, n).view(np.uint32), O.logf_range(start, n).view(np.uint32))
    for exp in range(0, 255):
        for off in (0, 0x2AAAAA, 0x7F0000):
            start = (exp << 23) | off
            if start == 0:
                start = 1
            n = 65536
            dev = ctx.debug_logf(start, n).view(np.uint32)
            ref = O.logf_range(start, n).view(np.uint32)
            assert np.array_equal(dev, ref), hex(start)


# ------------------------------------------------------------------ blockmodel + MDL

def test_toy_build_and_mdl(ctx):
    g = toy_graph()
    og, ob, dg, db = both(ctx, g, 3, TOY["assignment"])
    same_state(ob, db)
    m, lp = db.mdl()
    assert m == ob.mdl()
    assert abs(m - TOY["entropy"]) < 4 * 7.63e-6  # 4 float ULPs at 92.6, as the reference's test


@pytest.mark.parametrize("seed,B", [(1, 5), (2, 40), (3, 500)])
def test_build_and_mdl_random(ctx, seed, B):
    g = small_graph(seed=seed)
    rng = np.random.default_rng(seed)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    og, ob, dg, db = both(ctx, g, B, assign)
    same_state(ob, db)
    assert db.mdl()[0] == ob.mdl()
    assert db.validate()


# ------------------------------------------------------------------ finetune proposals

@pytest.mark.parametrize("greedy", [True, False])
@pytest.mark.parametrize("B", [3, 30, 300])
def test_propose_all_bit_exact(ctx, greedy, B):
    g = small_graph(seed=B, N=600, kmax=600, mean=10)  # includes vertices above the LDS-table threshold
    rng = np.random.default_rng(B)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    og, ob, dg, db = both(ctx, g, B, assign)
    ctx.set_seed(99)
    s, dS, acc = db.propose_all(epoch=7, greedy=greedy)
    mismatches = 0
    for v in range(g.N):
        os_, odS, oacc = ob.ft_propose(99, 7, v, greedy)
        assert s[v] == os_, v
        if s[v] != assign[v]:
            assert dS[v] == odS, (v, dS[v], odS)
            assert abs(dS[v] - odS) <= 1e-12 * max(1.0, abs(odS))
        mismatches += int(acc[v] != oacc)
    assert mismatches == 0


SHAPES = [((4, 8, 32, 64, 256), (4, 4, 4, 4)), ((4, 16, 32, 64, 128), (3, 3, 3, 3)),
          ((2, 4, 32, 64, 256), (4, 4, 4, 4)), ((2, 4, 16, 32, 128), (3, 3, 3, 3)),
          ((4, 8, 16, 32, 256), (4, 4, 4, 4)),
          ((8, 8, 32, 64, 128), (3, 3, 3, 3)), ((16, 16, 32, 64, 256), (4, 4, 4, 4)),
          ((32, 32, 64, 64, 128), (3, 4, 3, 4)), ((64, 64, 64, 64, 256), (4, 3, 4, 3))]


@pytest.mark.parametrize("greedy", [True, False])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "-".join(map(str, s[0] + s[1])))
def test_propose_all_every_launch_shape(greedy, shape):
    """Every allowed lanes-per-vertex / wavefronts-per-SIMD choice of the degree-class launches gives the oracle's
    proposals: degrees span all four lane-group classes and the workgroup class."""
    c = native.Context(0, seed=4321)
    c.set_finetune_shape(*shape)
    assert c.finetune_shape() == (list(shape[0]), list(shape[1]))
    B = 40
    g = with_hubs(small_graph(seed=11, N=2500, kmax=900, mean=14), seed=11, hubs=(0, 1), degree=700)
    deg = np.diff(g.out_off) + np.diff(g.in_off)
    assert (deg <= 8).any() and ((deg > 8) & (deg <= 32)).any() and ((deg > 32) & (deg <= 128)).any()
    assert ((deg > 128) & (deg <= 256)).any() and ((deg > 256) & (deg <= 1024)).any() and (deg > 1024).any()
    rng = np.random.default_rng(5)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    og, ob, dg, db = both(c, g, B, assign)
    s, dS, acc = db.propose_all(epoch=3, greedy=greedy)
    for v in range(g.N):
        os_, odS, oacc = ob.ft_propose(4321, 3, v, greedy)
        assert s[v] == os_, v
        if s[v] != assign[v]:
            assert dS[v] == odS, (v, dS[v], odS)
        assert acc[v] == oacc, v
    # a sweep in two batches groups its items on the device (not the graph's cached grouping)
    prm = native.params("async_gibbs", batches=2, seed=4321, greedy=greedy)
    db.async_gibbs_sweep(prm, 4)
    ob.async_gibbs_sweep(O.params("async_gibbs", batches=2, seed=4321, greedy=greedy), 4)
    same_state(ob, db)
    c.close()


@pytest.mark.parametrize("greedy", [True, False])
@pytest.mark.parametrize("k,shape,B", [(32, (4, 4, 32, 64, 256), 48), (128, (4, 8, 16, 32, 256), 160),
                                       (256, (4, 8, 16, 32, 256), 300)])
def test_full_lane_group_tables(greedy, k, shape, B):
    """Lane groups with fewer lanes than fit tables of twice the class's degree get tables of the class's degree
    (class 1 on 4 lanes: 32 slots; class 2 on 16: 128; class 3 on 32: 256): vertices with exactly k edges into k
    distinct blocks other than their own fill every slot, so looking up their own block (absent) must end without
    an empty slot."""
    c = native.Context(0, seed=99)
    c.set_finetune_shape(shape, (4, 4, 4, 4))
    base = small_graph(seed=12, N=1200, B=8, kmax=60, mean=6)
    rng = np.random.default_rng(12)
    assign = rng.integers(0, B, base.N).astype(np.int32)
    assign[:B] = np.arange(B)
    src = list(np.repeat(np.arange(base.N), np.diff(base.out_off)))
    dst = list(base.out_adj)
    full = []
    for j in range(12):  # new vertices, each with k out-edges only, one into each of k blocks other than its own
        v = base.N + j
        own = int(rng.integers(0, B))
        others = rng.choice([b for b in range(B) if b != own], k, replace=False)
        for b in others:
            src.append(v)
            dst.append(int(b))  # vertex b is in block b (assign[:B] = arange)
        full.append((v, own))
    g = CSRGraph.from_edges(np.array(src), np.array(dst), base.N + len(full))
    assign = np.concatenate([assign, np.array([o for _, o in full], np.int32)])
    og, ob, dg, db = both(c, g, B, assign)
    for epoch in range(3):
        s, dS, acc = db.propose_all(epoch=epoch, greedy=greedy)
        for v in [v for v, _ in full] + list(range(0, base.N, 37)):
            os_, odS, oacc = ob.ft_propose(99, epoch, v, greedy)
            assert s[v] == os_, (epoch, v)
            if s[v] != assign[v]:
                assert dS[v] == odS, (epoch, v)
            assert acc[v] == oacc, (epoch, v)
    c.close()


def test_launch_shape_rejects_unfit_choices(ctx):
    before = ctx.finetune_shape()
    for lanes, occ in [((1, 16, 32, 64, 256), (3, 3, 3, 3)), ((8, 2, 32, 64, 256), (3, 3, 3, 3)),
                       ((8, 16, 8, 64, 256), (3, 3, 3, 3)), ((8, 16, 32, 16, 256), (3, 3, 3, 3)),
                       ((8, 16, 32, 64, 256), (2, 3, 3, 3)), ((8, 12, 32, 64, 256), (3, 3, 3, 3)),
                       ((8, 16, 32, 64, 512), (3, 3, 3, 3)), ((8, 16, 32, 64, 64), (3, 3, 3, 3))]:
        with pytest.raises(native.SbpError):
            ctx.set_finetune_shape(lanes, occ)
        assert ctx.finetune_shape() == before


def hub_graph(seed=5, N=8000, hubs=(0, 1, 2), degree=3000):
    """A DCSBM plus a few hub vertices with `degree` out-edges and `degree` in-edges to distinct vertices."""
    return with_hubs(dcsbm(N, 8, gamma=2.5, kmin=1, kmax=60, mean_degree=6, p_in=0.7, seed=seed), seed, hubs, degree)


def with_hubs(g, seed, hubs, degree):
    """g plus `degree` out-edges and `degree` in-edges to distinct vertices at each vertex of `hubs`."""
    N = g.N
    src = np.repeat(np.arange(g.N), np.diff(g.out_off)).astype(np.int64)
    dst = g.out_adj.astype(np.int64)
    rng = np.random.default_rng(seed)
    s_extra, d_extra = [], []
    for h in hubs:
        outs = rng.choice(np.setdiff1d(np.arange(N), [h]), degree, replace=False)
        ins = rng.choice(np.setdiff1d(np.arange(N), [h]), degree, replace=False)
        s_extra += [np.full(degree, h), ins]
        d_extra += [outs, np.full(degree, h)]
    src = np.concatenate([src] + s_extra)
    dst = np.concatenate([dst] + d_extra)
    key = np.unique(src * N + dst)  # unweighted: duplicates dropped
    return CSRGraph.from_edges(key // N, key % N, N)


@pytest.mark.parametrize("greedy", [True, False])
@pytest.mark.parametrize("B,degree", [(1000, 3000), (3000, 3000), (7000, 3000), (7000, 900)])
def test_propose_all_hubs_bit_exact(ctx, greedy, B, degree):
    """Hub vertices of 2 * degree edges, above the lane-group classes' 256 and the middle class's 1,024.  Up to
    6,144 blocks the last class's table is indexed by block id in LDS (B = 1,000 and 3,000); above, a hub of 6,000
    edges can have more distinct neighbour blocks than the LDS hash table holds and uses a global-memory table
    (B = 7,000, degree 3,000), and one of 1,800 edges the LDS hash table (degree 900)."""
    g = hub_graph(degree=degree)
    rng = np.random.default_rng(B)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    og, ob, dg, db = both(ctx, g, B, assign)
    for epoch in range(4):
        ctx.set_seed(77)
        s, dS, acc = db.propose_all(epoch=epoch, greedy=greedy)
        for v in list(range(3)) + list(rng.choice(g.N, 300, replace=False)):
            os_, odS, oacc = ob.ft_propose(77, epoch, int(v), greedy)
            assert s[v] == os_, (epoch, v)
            if s[v] != assign[v]:
                assert dS[v] == odS, (epoch, v, dS[v], odS)
            assert bool(acc[v]) == bool(oacc), (epoch, v)


def test_async_gibbs_sweeps_match_with_refusals(ctx):
    g = small_graph(seed=21, N=800, mean=6)
    B = 400  # blocks of about two vertices: many commits would empty a block
    rng = np.random.default_rng(5)
    assign = np.concatenate([np.arange(B), rng.integers(0, B, g.N - B)]).astype(np.int32)
    og, ob, dg, db = both(ctx, g, B, assign)
    ctx.set_seed(5)
    prm_o = O.params("async_gibbs", seed=5, threads=4)
    prm_d = native.params("async_gibbs", seed=5)
    refused = 0
    for ep in range(4):
        _, st_o = ob.async_gibbs_sweep(prm_o, ep)
        st_d = db.async_gibbs_sweep(prm_d, ep)
        same_state(ob, db)
        assert st_d.moves == st_o.moves and st_d.refused == st_o.refused
        refused += st_o.refused
    assert refused > 0


@pytest.mark.parametrize("batches", [1, 3])
def test_async_gibbs_batches_and_shuffle(ctx, batches):
    g = small_graph(seed=8, N=700)
    B = 12
    assign = (np.arange(g.N) % B).astype(np.int32)
    og, ob, dg, db = both(ctx, g, B, assign)
    ctx.set_seed(3)
    prm_o = O.params("async_gibbs", seed=3, batches=batches)
    prm_d = native.params("async_gibbs", seed=3, batches=batches)
    for ep in range(3):
        ob.async_gibbs_sweep(prm_o, ep)
        db.async_gibbs_sweep(prm_d, ep)
        same_state(ob, db)


@pytest.mark.parametrize("greedy", [True, False])
def test_mh_sweeps_match(ctx, greedy):
    g = small_graph(seed=31, N=400, kmax=600)
    B = 25
    rng = np.random.default_rng(1)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    og, ob, dg, db = both(ctx, g, B, assign)
    ctx.set_seed(11)
    prm_o = O.params("metropolis_hastings", seed=11, greedy=greedy)
    prm_d = native.params("metropolis_hastings", seed=11, greedy=greedy)
    for ep in range(3):
        so, _ = ob.mh_sweep(prm_o, ep, ordered=(ep == 2))
        sd, _ = db.mh_sweep(prm_d, ep, ordered=(ep == 2))
        same_state(ob, db)
        assert so == sd
    assert db.validate()


@pytest.mark.parametrize("alg", ["async_gibbs", "metropolis_hastings", "hybrid_mcmc"])
def test_mcmc_phase_matches(ctx, alg):
    g = small_graph(seed=41, N=900, kmax=700)
    B = 40
    rng = np.random.default_rng(2)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    og, ob, dg, db = both(ctx, g, B, assign)
    ctx.set_seed(17)
    ep_o, st_o = ob.mcmc(O.params(alg, seed=17, mh_percent=0.1), True, 0)
    ep_d, mdl_d, st_d = db.mcmc(native.params(alg, seed=17, mh_percent=0.1), True, 0)
    assert ep_o == ep_d
    same_state(ob, db)
    assert mdl_d == ob.mdl()


# ------------------------------------------------------------------ block merge

@pytest.mark.parametrize("B", [50, 900])
def test_merge_propose_bit_exact(ctx, B):
    g = small_graph(seed=B, N=900)
    rng = np.random.default_rng(B)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    og, ob, dg, db = both(ctx, g, B, assign)
    ctx.set_seed(42)
    s_d, dS_d = db.merge_propose(epoch=3)
    s_o, dS_o = ob.merge_propose(42, 3)
    assert np.array_equal(s_d, s_o)
    assert np.array_equal(dS_d, dS_o)


@pytest.mark.parametrize("approximate", [False, True])
def test_merge_blocks_matches(ctx, approximate):
    g = small_graph(seed=77, N=1000)
    B = g.N
    assign = np.arange(B, dtype=np.int32)
    og, ob, dg, db = both(ctx, g, B, assign)
    ctx.set_seed(6)
    out_d, st = db.merge_blocks(epoch=0, num_to_merge=B // 2, approximate=approximate)
    s_o, dS_o = ob.merge_propose(6, 0)
    out_o = ob.merge_apply(s_o, dS_o, B // 2, approximate)
    same_state(out_o, out_d)


# ------------------------------------------------------------------ end to end

@pytest.mark.parametrize("alg", ["metropolis_hastings", "async_gibbs", "hybrid_mcmc"])
def test_full_partition_bit_exact(ctx, alg):
    g = dcsbm(1000, 16, gamma=2.5, kmin=1, kmax=100, mean_degree=10, p_in=0.75, seed=0)
    og = O.OGraph.from_csr(g)
    a_o, nb_o, st_o = O.sbp(og, O.params(alg, seed=0, threads=8))
    dg = ctx.upload(g)
    a_d, nb_d, st_d = dg.partition(native.params(alg, seed=0))
    assert nb_d == nb_o
    assert np.array_equal(a_d, a_o)
    assert st_d.mdl == st_o.mdl
    from sklearn.metrics import normalized_mutual_info_score as nmi
    assert nmi(g.truth, a_d) > 0.9
