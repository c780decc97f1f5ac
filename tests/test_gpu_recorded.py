"""Asynchronous-Gibbs sweeps replayed from a recorded HIP graph against the same sweeps run call by call and
against the oracle: every replay, with the program's own work and synchronizations between replays, must leave the
exact state (assignment, block matrix, degrees) and MDL of the eager sweep."""
import numpy as np
import pytest

import oracle as O
from sbp_amd import native
from sbp_amd.graph import dcsbm

pytestmark = pytest.mark.gpu


def same(x, y):
    for k in ("B", "row_ptr", "cols", "cnts", "sizes", "assignment", "d_out", "d_in", "d_tot"):
        assert np.array_equal(np.asarray(x[k]), np.asarray(y[k])), k


@pytest.mark.parametrize("greedy", [True, False])
@pytest.mark.parametrize("batches,B", [(1, 60), (3, 60), (1, 1500)])
def test_replays_equal_eager_sweeps_and_the_oracle(greedy, batches, B):
    g = dcsbm(3000, 12, gamma=2.5, kmin=1, kmax=300, mean_degree=10, p_in=0.7, seed=17)
    rng = np.random.default_rng(B)
    assign = np.concatenate([np.arange(B), rng.integers(0, B, g.N - B)]).astype(np.int32)
    ctx = native.Context(0, seed=9)
    dg = ctx.upload(g)
    eager = native.DeviceBlockmodel.build(dg, B, assign)
    rec = native.DeviceBlockmodel.build(dg, B, assign)
    og = O.OGraph.from_csr(g)
    ob = O.OBlockmodel.build(og, B, assign)
    prm = native.params("async_gibbs", batches=batches, greedy=greedy, seed=9)
    oprm = O.params("async_gibbs", batches=batches, greedy=greedy, seed=9)
    moved = 0
    for ep in range(6):
        st_e = eager.async_gibbs_sweep(prm, ep)
        m_e = eager.mdl()[0]
        st_r, m_r = rec.async_gibbs_sweep_recorded(prm, ep)
        assert st_r.recorded_sweeps == 1
        assert st_r.moves == st_e.moves and st_r.proposals == st_e.proposals and st_r.refused == st_e.refused
        assert list(st_r.ft_class_bytes) == list(st_e.ft_class_bytes)
        assert m_r == m_e
        x = rec.export()  # eager work between replays
        same(x, eager.export())
        if ep < 2:
            ob.async_gibbs_sweep(oprm, ep)
            same(x, ob.export())
            assert m_r == ob.mdl()
        moved += st_r.moves
        assert rec.validate()
    assert moved > 0
    ctx.close()


def test_recording_follows_a_changed_blockmodel():
    """A replay for another blockmodel (or after a merge phase gave new buffers) records again."""
    g = dcsbm(2000, 10, gamma=2.5, kmin=1, kmax=200, mean_degree=8, p_in=0.7, seed=3)
    ctx = native.Context(0, seed=1)
    dg = ctx.upload(g)
    rng = np.random.default_rng(1)
    prm = native.params("async_gibbs", seed=1)
    for B in (40, 25, 40):
        assign = np.concatenate([np.arange(B), rng.integers(0, B, g.N - B)]).astype(np.int32)
        a = native.DeviceBlockmodel.build(dg, B, assign)
        b = native.DeviceBlockmodel.build(dg, B, assign)
        for ep in range(3):
            a.async_gibbs_sweep(prm, ep)
            _, m = b.async_gibbs_sweep_recorded(prm, ep)
            same(a.export(), b.export())
            assert m == a.mdl()[0]
    ctx.close()


def test_recording_grows_the_cells_of_a_sparse_block_matrix():
    """A planted partition gives far fewer cells than min(E, B^2): recording first grows the cell arrays to that
    bound, keeping the current cells (a regression test: the growth once dropped them)."""
    g = dcsbm(3000, 150, gamma=2.5, kmin=1, kmax=200, mean_degree=10, p_in=0.97, seed=8)
    ctx = native.Context(0, seed=4)
    dg = ctx.upload(g)
    B = int(g.truth.max()) + 1
    eager = native.DeviceBlockmodel.build(dg, B, g.truth)
    rec = native.DeviceBlockmodel.build(dg, B, g.truth)
    assert len(rec.export()["cols"]) * 2 < min(g.E, B * B)
    prm = native.params("async_gibbs", seed=4)
    for ep in range(4):
        eager.async_gibbs_sweep(prm, ep)
        _, m = rec.async_gibbs_sweep_recorded(prm, ep)
        same(eager.export(), rec.export())
        assert m == eager.mdl()[0]
    ctx.close()


def test_partitions_with_replayed_sweeps_match_the_oracle_and_each_other():
    """Full partitions with every finetune phase replayed from recorded graphs (forced on for small graphs): alone
    they equal the oracle's partition; as a batch on several worker slots, where one worker records while others
    allocate, free and copy, each equals the graph alone."""
    from sbp_amd import workloads
    graphs = workloads.generate([workloads.batched_spec(i) for i in range(10)])
    ctx = native.Context(0)
    ctx.set_recorded_sweeps(1)
    dgs = [ctx.upload(g) for g in graphs]
    prms = [native.params("async_gibbs", seed=i) for i in range(len(graphs))]
    solo = [dg.partition(p) for dg, p in zip(dgs, prms)]
    assert all(s[2].recorded_sweeps > 0 for s in solo)
    for i in (2, 8):  # 4,585 and 3,367 vertices
        a_o, nb_o, st_o = O.sbp(O.OGraph.from_csr(graphs[i]), O.params("async_gibbs", seed=i, threads=8))
        assert nb_o == solo[i][1] and np.array_equal(a_o, solo[i][0]) and st_o.mdl == solo[i][2].mdl
    outs, nb, st, _ = ctx.partition_batch(dgs, prms, slots=6)
    for i in range(len(graphs)):
        assert int(nb[i]) == solo[i][1] and np.array_equal(outs[i], solo[i][0]) and st[i].mdl == solo[i][2].mdl, i
        assert st[i].recorded_sweeps > 0
    ctx.close()
