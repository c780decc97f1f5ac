"""GPU parity of divisive::split_communities (src/top_down_sbp.cpp:488-549, random splits) through the C-ABI
(sbp_split_communities) against the oracle's literal restatement: the new assignment, block count and every
block's kept split value must be bit-identical, and the rebuilt blockmodel must equal the oracle's blockmodel
of the new assignment.  Cases: the small graphs of the oracle's own test (singleton and empty blocks, exact
ties, both null models), then 100,000 vertices at one block (the top-down start), at 64 planted blocks (counters
in LDS) and at 1,000 blocks (global counters), and a degree-corrected blockmodel, which is refused."""
import numpy as np
import pytest

import oracle as O
from sbp_amd import native
from sbp_amd.graph import dcsbm
from test_split_oracle import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0, seed=1)
    yield c
    c.close()


def same_state(ob, db):
    x, y = ob.export(), db.export()
    assert x["B"] == y["B"]
    for k in ("row_ptr", "cols", "cnts", "sizes", "assignment"):
        assert np.array_equal(x[k], y[k]), k
    for k in ("d_out", "d_in", "d_tot"):
        assert np.array_equal(x[k], y[k].astype(np.int64)), k


def run_case(ctx, g, a, B, target, seed, epoch, nonparametric, dg=None):
    print(f"split case N={g.N} E={g.E} B={B} target={target} seed={seed} epoch={epoch} np={nonparametric}",
          flush=True)
    og = O.OGraph.from_csr(g)
    oa, onb, od = O.split_communities(og, a, B, target, seed, epoch, nonparametric)
    ctx.set_seed(seed)
    dg = dg or ctx.upload(g)
    db = native.DeviceBlockmodel.build(dg, B, a, nonparametric=nonparametric)
    nb, delta = db.split_communities(target, epoch)
    assert nb == onb
    np.testing.assert_array_equal(delta, od)
    np.testing.assert_array_equal(db.export()["assignment"], oa)
    assert db.validate()
    same_state(O.OBlockmodel.build(og, onb, oa), db)
    return onb


@pytest.mark.parametrize("i", range(9))
def test_split_small_cases(ctx, i):
    run_case(ctx, *cases()[i])


@pytest.fixture(scope="module")
def big():
    return dcsbm(100_000, 64, gamma=2.1, kmin=1, kmax=2000, mean_degree=16, p_in=0.6, seed=21)


@pytest.mark.parametrize("B,target,nonparametric", [(1, 2, True), (1, 2, False), (64, 96, True), (64, 200, False),
                                                    (1000, 1500, True)])
def test_split_100k(ctx, big, B, target, nonparametric):
    g = big
    if B == 64:
        a = np.asarray(g.truth, np.int32)
    else:
        a = np.random.default_rng(B).integers(0, B, g.N).astype(np.int32)
    run_case(ctx, g, a, B, target, 5, 7, nonparametric)


def test_split_rejects_degree_corrected(ctx):
    g = dcsbm(500, 4, gamma=2.5, kmin=1, kmax=30, mean_degree=6, p_in=0.7, seed=1)
    dg = ctx.upload(g)
    db = native.DeviceBlockmodel.build(dg, 4, np.asarray(g.truth, np.int32), nonparametric=True, degreecorrected=True)
    with pytest.raises(native.SbpError, match="degree-corrected"):
        db.split_communities(8, 0)
