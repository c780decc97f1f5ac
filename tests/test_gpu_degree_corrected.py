"""GPU parity of the degree-corrected nonparametric description length (args.degreecorrected with
args.nonparametric; src/entropy.cpp:617-1225, src/spence.cpp) through the C-ABI against the oracle.

The device evaluates log_q (get_v, spence) with the device exp / log / log1p / sqrt, which can differ from
glibc in the last place, so the terms that contain it are held to the stated FP64 tolerance (1e-12 relative)
and accept decisions may differ only within 1e-9 of the threshold; everything else (proposed blocks, the
sparse entropy, histogram bookkeeping, states after sweeps and merges) is compared exactly."""
import math

import numpy as np
import pytest

import oracle as O
from sbp_amd import native
from sbp_amd.graph import dcsbm

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0, seed=77)
    yield c
    c.close()


def graph(seed=5, N=800, kmax=400):
    return dcsbm(N, 8, gamma=2.3, kmin=1, kmax=kmax, mean_degree=8, p_in=0.7, seed=seed)


def both(ctx, g, B, assign):
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    return (og, O.OBlockmodel.build(og, B, assign, nonparametric=True, degreecorrected=True), dg,
            native.DeviceBlockmodel.build(dg, B, assign, nonparametric=True, degreecorrected=True))


def random_assign(N, B, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, B, N).astype(np.int32)
    a[:B] = np.arange(B)
    return a


def close(a, b, scale):
    if math.isinf(a) or math.isinf(b):
        return a == b
    return abs(a - b) <= TOL * max(1.0, abs(scale))


def same_state(ob, db):
    x, y = ob.export(), db.export()
    assert x["B"] == y["B"]
    for k in ("row_ptr", "cols", "cnts", "sizes", "assignment"):
        assert np.array_equal(x[k], y[k]), k


@pytest.mark.parametrize("B", [6, 60, 600])
def test_description_length(ctx, B):
    g = graph(seed=B)
    og, ob, dg, db = both(ctx, g, B, random_assign(g.N, B, B))
    mdl, sparse, part, edges, deg = db.np_parts_dc()
    osp, opart, oedges = ob.np_parts()
    assert (sparse, part, edges) == (osp, opart, oedges)  # no log_q in these: bit-exact
    assert close(deg, ob.np_deg_dl(), ob.np_deg_dl())
    assert close(mdl, ob.mdl(), ob.mdl())


@pytest.mark.parametrize("greedy", [True, False])
@pytest.mark.parametrize("B", [4, 40, 400])
def test_propose_all(ctx, greedy, B):
    print(f"[dc propose] B={B} greedy={greedy}", flush=True)
    g = graph(seed=B + 1)
    og, ob, dg, db = both(ctx, g, B, random_assign(g.N, B, B + 1))
    ctx.set_seed(7)
    s_d, dS_d, acc_d = db.propose_all(epoch=2, greedy=greedy)
    for v in range(g.N):
        s, dS, acc = ob.ft_propose(7, 2, v, greedy)
        assert s_d[v] == s, v
        assert close(dS_d[v], dS, max(abs(dS), 1e3)), (v, dS_d[v], dS)
        if acc_d[v] != acc:  # only at the threshold
            assert greedy and abs(dS) < 1e-9, (v, dS)


@pytest.mark.parametrize("B", [40, 800])
def test_merge_propose(ctx, B):
    g = graph(seed=B + 7)
    og, ob, dg, db = both(ctx, g, B, random_assign(g.N, B, B))
    ctx.set_seed(3)
    s_d, dS_d = db.merge_propose(epoch=1)
    s_o, dS_o = ob.merge_propose(3, 1)
    for r in range(B):
        assert close(dS_d[r], dS_o[r], max(abs(dS_o[r]), 1e3)), r
        if s_d[r] != s_o[r]:  # a near tie between two candidates
            assert abs(ob.merge_delta_mdl(r, int(s_d[r])) - dS_o[r]) <= 1e-9 * max(1.0, abs(dS_o[r])), r


@pytest.mark.parametrize("alg", ["async_gibbs", "metropolis_hastings", "hybrid_mcmc"])
def test_mcmc_phase(ctx, alg):
    """Metropolis-Hastings moves key the in-degree histogram without self-loops (the reference's drift), async
    Gibbs commits with the full in-degree; the graph has self-loops."""
    print(f"[dc mcmc] {alg}", flush=True)
    g = graph(seed=31, N=900, kmax=600)
    B = 30
    og, ob, dg, db = both(ctx, g, B, random_assign(g.N, B, 4))
    ctx.set_seed(13)
    prm = dict(seed=13, greedy=False, nonparametric=True, degreecorrected=True, mh_percent=0.1)
    ep_o, _ = ob.mcmc(O.params(alg, **prm), True, 0)
    ep_d, mdl_d, _ = db.mcmc(native.params(alg, **prm), True, 0)
    assert ep_o == ep_d
    same_state(ob, db)
    assert close(db.np_parts_dc()[4], ob.np_deg_dl(), ob.np_deg_dl())
    assert close(mdl_d, ob.mdl(), ob.mdl())


@pytest.mark.parametrize("approximate", [False, True])
def test_merge_blocks(ctx, approximate):
    g = graph(seed=41, N=1000)
    B = g.N
    og, ob, dg, db = both(ctx, g, B, np.arange(B, dtype=np.int32))
    ctx.set_seed(8)
    out_d, _ = db.merge_blocks(epoch=0, num_to_merge=B // 2, approximate=approximate)
    s_o, dS_o = ob.merge_propose(8, 0)
    out_o = ob.merge_apply(s_o, dS_o, B // 2, approximate)
    same_state(out_o, out_d)
    assert close(out_d.np_parts_dc()[4], out_o.np_deg_dl(), out_o.np_deg_dl())


@pytest.mark.parametrize("alg", ["async_gibbs", "hybrid_mcmc"])
def test_full_partition(ctx, alg):
    print(f"[dc partition] {alg}", flush=True)
    g = dcsbm(1000, 16, gamma=2.5, kmin=1, kmax=100, mean_degree=10, p_in=0.75, seed=0)
    og = O.OGraph.from_csr(g)
    prm = dict(seed=0, greedy=False, nonparametric=True, degreecorrected=True)
    a_o, nb_o, st_o = O.sbp(og, O.params(alg, threads=8, **prm))
    dg = ctx.upload(g)
    a_d, nb_d, st_d = dg.partition(native.params(alg, **prm))
    assert nb_d == nb_o and np.array_equal(a_d, a_o)
    assert close(st_d.mdl, st_o.mdl, st_o.mdl)
