"""The drop-in surface's per-move entry points on the GPU against the oracle.

- tests/dropin/api_check (the reference's finetune / block-merge call sequence over the C++ API: edge_weights,
  blockmodel_delta, entropy::delta_mdl, hastings_correction, block_merge_delta_mdl, move_vertex, merge_block) on a
  TSV graph: every delta-MDL, Hastings correction and description length equal to the oracle's, every Delta
  summing to zero (the reference's own invariant, test/finetune_test.cpp:106-127), the commits' refusals and the
  final assignment as the reference's rules give them.
- the C-ABI entries behind them (sbp_ft_evaluate, sbp_merge_evaluate, sbp_bm_move_vertex, sbp_bm_merge_block,
  sbp_merge_apply) on graphs with hub vertices and in the nonparametric model, bit-exact against the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from sbp_amd import native
from sbp_amd.graph import dcsbm, save_tsv

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
API_CHECK = os.path.join(ROOT, "tests", "dropin", "api_check")


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0, seed=3)
    yield c
    c.close()


def test_api_check_program_matches_oracle(tmp_path):
    g = dcsbm(1200, 10, gamma=2.3, kmin=1, kmax=400, mean_degree=9, p_in=0.7, seed=8)
    base = str(tmp_path / "g")
    save_tsv(g, base)
    B = 40
    print(f"[api_check] N={g.N} E={g.E} B={B}", flush=True)
    r = subprocess.run([API_CHECK, base, str(B), "0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.splitlines()]
    og = O.OGraph.from_csr(g)
    assign = ((np.arange(g.N) * 7919 + 13) % B).astype(np.int32)
    ob = O.OBlockmodel.build(og, B, assign)
    moves = merges = 0
    for l in lines:
        if l[0] == "mdl":
            assert float(l[1]) == ob.mdl()
        elif l[0] == "move":
            v, s, dS, H, total = int(l[1]), int(l[2]), float(l[3]), float(l[4]), int(l[5])
            assert dS == ob.vertex_delta_mdl(v, s), (v, s)
            assert H == ob.vertex_hastings(v, s), (v, s)
            assert total == 0
            moves += 1
        elif l[0] == "merge":
            r_, s, dS = int(l[1]), int(l[2]), float(l[3])
            assert dS == ob.merge_delta_mdl(r_, s), (r_, s)
            merges += 1
    assert moves > 20 and merges > 10
    # commits: refused when the vertex's block has one vertex (src/blockmodel/blockmodel.cpp:439-443)
    a = assign.copy()
    for l in lines:
        if l[0] == "commit":
            v, s, ok = int(l[1]), int(l[2]), int(l[3])
            want = int((a == a[v]).sum() > 1)
            assert ok == want, (v, s)
            if ok:
                a[v] = s
    after = dict((l[0], l[1:]) for l in lines if l[0] in ("after_moves", "after_merge", "assignment"))
    assert float(after["after_moves"][0]) == O.OBlockmodel.build(og, B, a).mdl()
    a[a == 0] = 1
    assert float(after["after_merge"][0]) == O.OBlockmodel.build(og, B, a).mdl()
    assert after["after_merge"][1] == "1"
    assert np.array_equal(np.array(after["assignment"], dtype=np.int32), a)
    # divisive::split_communities on the first blockmodel, parametric null model, epoch 0
    split = [l for l in lines if l[0] == "split"][0]
    sa, snb, _ = O.split_communities(og, assign, B, B + B // 2, seed=0, epoch=0, nonparametric=False)
    assert int(split[1]) == snb == B + B // 2
    assert np.array_equal(np.array(split[2:], dtype=np.int32), sa)


def hubby(seed=5):
    from test_gpu_parity import hub_graph
    return hub_graph(seed=seed, N=6000, degree=2500)


@pytest.mark.parametrize("B,nonparametric", [(60, False), (3000, False), (300, True)])
def test_evaluate_moves_and_merges(ctx, B, nonparametric):
    g = hubby()
    rng = np.random.default_rng(B)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    og = O.OGraph.from_csr(g)
    ob = O.OBlockmodel.build(og, B, assign, nonparametric=nonparametric)
    db = native.DeviceBlockmodel.build(ctx.upload(g), B, assign, nonparametric=nonparametric)
    vs = np.concatenate([[0, 1, 2], rng.choice(g.N, 300, replace=False)]).astype(np.int32)
    ts = rng.integers(0, B, len(vs)).astype(np.int32)
    dS, H = db.evaluate_moves(vs, ts)
    for i, (v, t) in enumerate(zip(vs, ts)):
        if t == assign[v]:
            assert dS[i] == 0.0 and H[i] == 1.0
            continue
        assert dS[i] == ob.vertex_delta_mdl(int(v), int(t)), (v, t)
        assert H[i] == (1.0 if nonparametric else ob.vertex_hastings(int(v), int(t))), (v, t)
    rs = rng.integers(0, B, 200).astype(np.int32)
    ss = ((rs + 1 + rng.integers(0, B - 1, 200)) % B).astype(np.int32)
    mdS = db.evaluate_merges(rs, ss)
    for i in range(len(rs)):
        assert mdS[i] == ob.merge_delta_mdl(int(rs[i]), int(ss[i])), (rs[i], ss[i])


def test_move_merge_and_carry_out(ctx):
    g = dcsbm(3000, 12, gamma=2.3, kmin=1, kmax=500, mean_degree=9, p_in=0.7, seed=12)
    B = 1500  # blocks of about two vertices: moves get refused
    rng = np.random.default_rng(4)
    assign = np.concatenate([np.arange(B), rng.integers(0, B, g.N - B)]).astype(np.int32)
    og = O.OGraph.from_csr(g)
    dg = ctx.upload(g)
    db = native.DeviceBlockmodel.build(dg, B, assign)
    a = assign.copy()
    refusals = 0
    for v in rng.choice(g.N, 400, replace=False):
        s = int(rng.integers(0, B))
        ok = db.move_vertex(int(v), s)
        want = (a == a[v]).sum() > 1
        assert ok == want
        refusals += int(not ok)
        if ok:
            a[v] = s
    assert refusals > 5
    db.merge_block(5, 7)
    a[a == 5] = 7
    ob = O.OBlockmodel.build(og, B, a)
    x, y = db.export(), ob.export()
    assert np.array_equal(x["assignment"], y["assignment"])
    assert np.array_equal(x["cols"], y["cols"]) and np.array_equal(x["cnts"], y["cnts"])
    assert db.mdl()[0] == ob.mdl()
    # the carry-out of given proposals against the oracle's, both carry-outs
    s_best, dS_best = ob.merge_propose(seed=3, epoch=2)
    for approximate in (False, True):
        merged = db.copy().merge_apply(s_best, dS_best, B // 3, approximate)
        om = O.OBlockmodel.build(og, B, a).merge_apply(s_best, dS_best, B // 3, approximate)
        assert merged.B == om.B
        assert np.array_equal(merged.export()["assignment"], om.export()["assignment"])
