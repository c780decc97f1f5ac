"""The block matrix updated in place after an asynchronous-Gibbs commit (the moved vertices' edge deltas merged
into the sorted row-major and column-major cell lists, the cell hash updated in place) against the full rebuild
and the oracle: after every sweep the exported state (CSR, CSC, degrees, sizes, assignment) must be bit-exact, and
the next sweep's proposals, which read the updated hash, must equal the oracle's.  Forced on (SBP_AMD_INCREMENTAL=1)
and off (0), on a graph with hubs and self-loops, with blocks that empty and blocks that fill, under both
description lengths, and over enough sweeps for the hash to pass its load limit and be rebuilt.  Each case also runs
with the dense block matrix forced on (SBP_AMD_DENSE=1: lookups read a B x B copy that the update keeps) and off."""
import numpy as np
import pytest

import oracle as O
from sbp_amd import native
from sbp_amd.graph import dcsbm
from test_gpu_parity import hub_graph, same_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0, seed=1)
    yield c
    c.close()


def run(ctx, g, B, assign, sweeps, mode, nonparametric=False, seed=5, batches=1, dense="0"):
    import os
    os.environ["SBP_AMD_INCREMENTAL"] = mode
    os.environ["SBP_AMD_DENSE"] = dense
    try:
        og = O.OGraph.from_csr(g)
        ob = O.OBlockmodel.build(og, B, assign, nonparametric=nonparametric)
        dg = ctx.upload(g)
        db = native.DeviceBlockmodel.build(dg, B, assign, nonparametric=nonparametric)
        ctx.set_seed(seed)
        prm_o = O.params("async_gibbs", seed=seed, threads=4, batches=batches, nonparametric=nonparametric)
        prm_d = native.params("async_gibbs", seed=seed, batches=batches, nonparametric=nonparametric)
        moves, updates = [], 0
        for ep in range(sweeps):
            print(f"update mode={mode} N={g.N} B={B} np={nonparametric} sweep {ep}", flush=True)
            _, st_o = ob.async_gibbs_sweep(prm_o, ep)
            st_d = db.async_gibbs_sweep(prm_d, ep)
            assert st_d.moves == st_o.moves and st_d.refused == st_o.refused, ep
            same_state(ob, db)
            moves.append(st_d.moves)
            updates += st_d.cell_updates
        # the next proposals read the hash the updates left
        s, dS, acc = db.propose_all(epoch=sweeps, greedy=True)
        x = ob.export()["assignment"]
        for v in np.random.default_rng(0).choice(g.N, 200, replace=False):
            os_, odS, oacc = ob.ft_propose(seed, sweeps, int(v), True)
            assert s[v] == os_, v
            if s[v] != x[v]:
                assert dS[v] == odS, v
        assert db.validate()
        if mode == "1":  # every commit that moved a vertex updated the cells in place
            swept = sum(1 for m in moves if m > 0)
            assert updates == swept if batches == 1 else updates >= swept
        else:
            assert updates == 0
        return db.export(), moves
    finally:
        os.environ.pop("SBP_AMD_INCREMENTAL", None)
        os.environ.pop("SBP_AMD_DENSE", None)


@pytest.mark.parametrize("dense", ["0", "1"])
@pytest.mark.parametrize("nonparametric", [False, True])
def test_update_matches_rebuild_and_oracle(ctx, nonparametric, dense):
    g = hub_graph(seed=3, N=6000, degree=1500)
    B = 40
    assign = (np.arange(g.N) % B).astype(np.int32)
    a, moves_a = run(ctx, g, B, assign, 8, "1", nonparametric, dense=dense)
    b, moves_b = run(ctx, g, B, assign, 8, "0", nonparametric, dense=dense)
    assert moves_a == moves_b and sum(moves_a) > 0
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.parametrize("dense", ["0", "1"])
def test_update_with_emptying_blocks(ctx, dense):
    """Blocks of about two vertices: moves empty blocks (their rows and columns leave the lists) and refill them."""
    g = dcsbm(900, 8, gamma=2.5, kmin=1, kmax=40, mean_degree=6, p_in=0.7, seed=12)
    B = 420
    rng = np.random.default_rng(2)
    assign = np.concatenate([np.arange(B), rng.integers(0, B, g.N - B)]).astype(np.int32)
    run(ctx, g, B, assign, 6, "1", batches=3, dense=dense)


@pytest.mark.parametrize("dense", ["0", "1"])
def test_update_many_sweeps_hash_reload(ctx, dense):
    """Twenty sweeps from a random partition: the hash takes new cells each sweep until its load limit, then the
    next proposal launch rebuilds it (the dense copy, when kept, stays valid throughout)."""
    g = dcsbm(20000, 30, gamma=2.3, kmin=1, kmax=300, mean_degree=10, p_in=0.6, seed=4)
    B = 2000
    assign = np.random.default_rng(9).integers(0, B, g.N).astype(np.int32)
    run(ctx, g, B, assign, 20, "1", dense=dense)


@pytest.mark.parametrize("alg", ["async_gibbs", "hybrid_mcmc"])
def test_partition_dense_bit_exact(ctx, alg):
    """Full partitions with the dense block matrix forced on: every merge phase with B <= 8192 builds one, and
    the finetune sweeps read it."""
    import os
    g = dcsbm(3000, 16, gamma=2.3, kmin=1, kmax=150, mean_degree=10, p_in=0.7, seed=6)
    og = O.OGraph.from_csr(g)
    a_o, nb_o, st_o = O.sbp(og, O.params(alg, seed=2, threads=8))
    os.environ["SBP_AMD_DENSE"] = "1"
    try:
        dg = ctx.upload(g)
        a_d, nb_d, st_d = dg.partition(native.params(alg, seed=2))
    finally:
        os.environ.pop("SBP_AMD_DENSE", None)
    assert nb_d == nb_o
    assert np.array_equal(a_d, a_o)
    assert st_d.mdl == st_o.mdl


def test_mh_partition_without_dense(ctx):
    """Metropolis-Hastings walks keep a dense copy of the matrix by default (their commits write it); with it
    turned off (SBP_AMD_DENSE=0) lookups search the slack rows, and the partition is the same."""
    import os
    g = dcsbm(1000, 16, gamma=2.5, kmin=1, kmax=100, mean_degree=10, p_in=0.75, seed=0)
    og = O.OGraph.from_csr(g)
    a_o, nb_o, st_o = O.sbp(og, O.params("metropolis_hastings", seed=0, threads=8))
    os.environ["SBP_AMD_DENSE"] = "0"
    try:
        dg = ctx.upload(g)
        a_d, nb_d, st_d = dg.partition(native.params("metropolis_hastings", seed=0))
    finally:
        os.environ.pop("SBP_AMD_DENSE", None)
    assert nb_d == nb_o
    assert np.array_equal(a_d, a_o)
    assert st_d.mdl == st_o.mdl
