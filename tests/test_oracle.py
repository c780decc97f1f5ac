"""Pins the CPU oracle before anything is compared against it.

Checks, in order: published Philox4x32-10 known-answer vectors; the 128-bit
fixed-point conversion against exact integer arithmetic; every known answer
of the reference's own unit tests on the 11-vertex toy graph (restated in
tests/golden/toy_example.json); dense brute-force recomputation of delta-MDL
with glibc logf; the zero-net Delta quirk; incremental-commit == rebuild; and
an end-to-end run that must recover a planted partition.
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from sbp_amd.graph import CSRGraph, dcsbm

HERE = os.path.dirname(os.path.abspath(__file__))
TOY = json.load(open(os.path.join(HERE, "golden", "toy_example.json")))

_libm = ctypes.CDLL("libm.so.6")
_libm.logf.restype = ctypes.c_float
_libm.logf.argtypes = [ctypes.c_float]


def glibc_logf(x):
    return float(_libm.logf(float(x)))


def float_eq(a, b, ulps=4):
    """gtest EXPECT_FLOAT_EQ: within 4 single-precision ULPs."""
    fa, fb = np.float32(a), np.float32(b)
    ia = np.frombuffer(fa.tobytes(), np.int32)[0]
    ib = np.frombuffer(fb.tobytes(), np.int32)[0]
    ia = ia if ia >= 0 else -(ia & 0x7FFFFFFF)
    ib = ib if ib >= 0 else -(ib & 0x7FFFFFFF)
    return abs(int(ia) - int(ib)) <= ulps


@pytest.fixture(scope="module")
def toy():
    e = np.array(TOY["edges"])
    g = CSRGraph.from_edges(e[:, 0], e[:, 1], TOY["num_vertices"])
    return g, O.OGraph.from_csr(g)


# ---------------------------------------------------------------- primitives

def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert list(O.philox(0, 0, 0, 0, 0)) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox(0xFFFFFFFFFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF)) == \
        [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox(0x299F31D0A4093822, 0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344)) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_below_and_unit_double():
    L = O.lib()
    assert L.orc_uniform_below(0, 7) == 0
    assert L.orc_uniform_below(0xFFFFFFFF, 7) == 6
    assert L.orc_uniform_below(0x80000000, 10) == 5
    assert L.orc_unit_double(0, 0) == 0.0
    assert L.orc_unit_double(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53


def test_fixed_point_matches_exact_truncation():
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.normal(0, 1e3, 2000), rng.normal(0, 1e-10, 200), [0.0, -0.0, 1.5, -2.75, 1e-30]])
    out = np.zeros(2, np.uint64)
    for d in vals:
        O.lib().orc_fx_from_double(float(d), out.ctypes.data)
        got = int(out[0]) | (int(out[1]) << 64)
        if got >= 1 << 127:
            got -= 1 << 128
        exact = int(float(d) * 2.0 ** 64)  # exact product, Python int() truncates toward zero
        assert got == exact


# ---------------------------------------------------------------- toy known answers

def test_toy_blockmodel_degrees(toy):
    g, og = toy
    bm = O.OBlockmodel.build(og, 3, TOY["assignment"])
    x = bm.export()
    assert list(x["d_out"]) == TOY["degrees_out"]
    assert list(x["d_in"]) == TOY["degrees_in"]
    assert list(x["d_tot"]) == TOY["degrees"]
    M = bm.dense()
    # closed form: count edges by (block(src), block(dst))
    a = np.array(TOY["assignment"])
    ref = np.zeros((3, 3), np.int64)
    for s, d in TOY["edges"]:
        ref[a[s], a[d]] += 1
    assert (M == ref).all()
    assert M.sum() == len(TOY["edges"])


def test_toy_mdl_matches_reference_golden(toy):
    _, og = toy
    bm = O.OBlockmodel.build(og, 3, TOY["assignment"])
    assert float_eq(bm.mdl(), TOY["entropy"])


def test_null_model_golden():
    E, V = TOY["null_model"]["E"], TOY["null_model"]["V"]
    mdl = O.null_mdl_v1(E)
    x = 1.0 / E
    h = (1 + x) * math.log(1 + x) - x * math.log(x)
    closed = E * h - E * math.log(E / (E * E))
    assert float_eq(mdl, closed)
    assert float_eq(mdl, TOY["null_model"]["mdl"])


def test_toy_vertex_delta_mdl_equals_mdl_difference(toy):
    _, og = toy
    mv = TOY["move"]
    bm = O.OBlockmodel.build(og, 3, TOY["assignment"])
    a2 = list(TOY["assignment"])
    a2[mv["vertex"]] = mv["to"]
    bm2 = O.OBlockmodel.build(og, 3, a2)
    dS = bm.vertex_delta_mdl(mv["vertex"], mv["to"])
    assert float_eq(dS, bm2.mdl() - bm.mdl())
    x2 = bm2.export()
    assert list(x2["d_out"]) == mv["new_block_degrees_out"]
    assert list(x2["d_in"]) == mv["new_block_degrees_in"]
    assert list(x2["d_tot"]) == mv["new_block_degrees"]
    # the fixture's Delta, applied to M, gives the rebuilt matrix
    M = bm.dense()
    for r, c, d in mv["delta"]:
        M[r, c] += d
    assert (M == bm2.dense()).all()


def test_toy_special_case_delta_mdl(toy):
    _, og = toy
    c = TOY["complex"]
    sc = c["special_case"]
    b3 = O.OBlockmodel.build(og, 6, c["assignment"])
    a = list(c["assignment"])
    a[sc["vertex"]] = sc["to"]
    b5 = O.OBlockmodel.build(og, 6, a)
    assert float_eq(b3.vertex_delta_mdl(sc["vertex"], sc["to"]), b5.mdl() - b3.mdl())
    x = b5.export()
    assert list(x["d_out"]) == c["degrees_after_move_6_to_0"]["out"]
    assert list(x["d_in"]) == c["degrees_after_move_6_to_0"]["in"]
    assert list(x["d_tot"]) == c["degrees_after_move_6_to_0"]["total"]


def test_toy_merge_delta_mdl(toy):
    _, og = toy
    m = TOY["merge"]
    bm = O.OBlockmodel.build(og, 3, TOY["assignment"])
    bm2 = O.OBlockmodel.build(og, 3, m["assignment_after"])
    assert float_eq(bm.merge_delta_mdl(m["from"], m["to"]), bm2.mdl() - bm.mdl())
    x2 = bm2.export()
    assert list(x2["d_out"]) == m["new_block_degrees_out"]
    assert list(x2["d_in"]) == m["new_block_degrees_in"]
    assert list(x2["d_tot"]) == m["new_block_degrees"]
    M = bm.dense()
    for r, c, d in m["delta"]:
        M[r, c] += d
    assert (M == bm2.dense()).all()


def test_toy_hastings_matches_dense_update_formula(toy):
    """entropy::hastings_correction with a Delta equals the dense-updates form (entropy_test.cpp:148-159)."""
    _, og = toy
    mv = TOY["move"]
    v, r, s = mv["vertex"], mv["from"], mv["to"]
    bm = O.OBlockmodel.build(og, 3, TOY["assignment"])
    H = bm.vertex_hastings(v, s)
    # dense form (src/entropy.cpp:406-448) from the fixture's updates and new degrees
    a = TOY["assignment"]
    e = np.array(TOY["edges"])
    counts = {}
    for d in e[e[:, 0] == v][:, 1]:
        counts[a[d]] = counts.get(a[d], 0) + 1
    for u in e[e[:, 1] == v][:, 0]:
        if u != v:
            counts[a[u]] = counts.get(a[u], 0) + 1
    M = bm.dense()
    x = bm.export()
    fwd = bwd = 0.0
    for blk in sorted(counts):
        c = counts[blk]
        pw = M[s, blk] + M[blk, s] + 1.0
        bd = x["d_tot"][blk] + 3
        bw = mv["updates_block_row"][blk] + mv["updates_block_col"][blk] + 1.0
        pd = mv["new_block_degrees"][blk] + 3
        fwd += c * pw / bd
        bwd += c * bw / pd
    assert abs(H - bwd / fwd) <= 1e-12 * abs(H)


def test_toy_sample_max_degree(toy):
    g, og = toy
    sm = TOY["sample_max_degree"]
    mapping, verts = O.sample(og, 0, sm["samplesize"])
    assert len(verts) == sm["num_vertices"]
    src, dst = g.edges()
    inside = (mapping[src] >= 0) & (mapping[dst] >= 0)
    assert inside.sum() == sm["num_edges"]
    deg = g.degrees()
    # stable descending-degree order (utils::argsort) picks the top vertices, ties by id
    order = sorted(range(g.N), key=lambda i: (-deg[i], i))
    assert list(verts) == order[: sm["num_vertices"]]


# ---------------------------------------------------------------- brute force with glibc logf

def _cell(v, din, dout):
    if v == 0:
        return 0.0
    fl = lambda x: 0.0 if x == 0 else glibc_logf(x)
    return v * (fl(v) - (fl(din) + fl(dout)))


def _dense_state(g, B, assign):
    src, dst = g.edges()
    a = np.asarray(assign)
    M = np.zeros((B, B), np.int64)
    np.add.at(M, (a[src], a[dst]), 1)
    return M, M.sum(1), M.sum(0)


def _brute_dS(g, B, assign_before, assign_after, rows_cols):
    M0, do0, di0 = _dense_state(g, B, assign_before)
    M1, do1, di1 = _dense_state(g, B, assign_after)
    terms = []
    cells = set()
    for x in rows_cols:
        for j in range(B):
            cells.add((x, j))
            cells.add((j, x))
    for (i, j) in cells:
        terms.append(_cell(M0[i, j], di0[j], do0[i]))
        terms.append(-_cell(M1[i, j], di1[j], do1[i]))
    return math.fsum(terms), sum(abs(t) for t in terms)


def _small_graph(seed=3, N=60, B=5):
    return dcsbm(N, B, gamma=2.5, kmin=1, kmax=12, mean_degree=5, p_in=0.6, seed=seed)


def _zero_net_cells(g, assign, v, s):
    """Delta cells whose adds and subs cancel: (s, r) and (r, s) (finetune.cpp:201-235)."""
    r = assign[v]
    src, dst = g.edges()
    outs = dst[(src == v) & (dst != v)]
    ins = src[(dst == v) & (src != v)]
    out_r = int((assign[outs] == r).sum())
    out_s = int((assign[outs] == s).sum())
    in_r = int((assign[ins] == r).sum())
    in_s = int((assign[ins] == s).sum())
    cells = []
    if out_r > 0 and out_r == in_s:
        cells.append((s, r))
    if out_s > 0 and out_s == in_r:
        cells.append((r, s))
    return cells


def test_vertex_delta_mdl_brute_force():
    g = _small_graph()
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(0)
    B = 5
    assign = rng.integers(0, B, g.N).astype(np.int32)
    bm = O.OBlockmodel.build(og, B, assign)
    checked = 0
    for v in range(g.N):
        r = assign[v]
        for s in range(B):
            if s == r or _zero_net_cells(g, assign, v, s):
                continue
            a2 = assign.copy()
            a2[v] = s
            want, scale = _brute_dS(g, B, assign, a2, [r, s])
            got = bm.vertex_delta_mdl(v, s)
            assert abs(got - want) <= 1e-12 * max(scale, 1.0), (v, s, got, want)
            checked += 1
    assert checked > 100


def test_zero_net_delta_cell_is_counted_twice():
    """delta_mdl counts a zero-net Delta entry once as an entry and once as an unchanged cell (entropy.cpp:350-392)."""
    g = _small_graph(seed=5)
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(2)
    B = 5
    assign = rng.integers(0, B, g.N).astype(np.int32)
    bm = O.OBlockmodel.build(og, B, assign)
    M, do, di = _dense_state(g, B, assign)
    found = 0
    for v in range(g.N):
        r = assign[v]
        for s in range(B):
            zero = _zero_net_cells(g, assign, v, s)
            if s == r or not zero:
                continue
            a2 = assign.copy()
            a2[v] = s
            want, scale = _brute_dS(g, B, assign, a2, [r, s])
            _, do1, di1 = _dense_state(g, B, a2)
            extra = sum(_cell(M[i, j], di[j], do[i]) - _cell(M[i, j], di1[j], do1[i]) for i, j in zero)
            got = bm.vertex_delta_mdl(v, s)
            assert abs(got - (want + extra)) <= 1e-12 * max(scale, 1.0)
            found += 1
    assert found > 0


def test_merge_delta_mdl_brute_force():
    g = _small_graph(seed=7, N=80, B=8)
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(4)
    B = 8
    assign = rng.integers(0, B, g.N).astype(np.int32)
    bm = O.OBlockmodel.build(og, B, assign)
    for r in range(B):
        for s in range(B):
            if r == s:
                continue
            a2 = np.where(assign == r, s, assign)
            want, scale = _brute_dS(g, B, assign, a2, [r, s])
            got = bm.merge_delta_mdl(r, s)
            assert abs(got - want) <= 1e-12 * max(scale, 1.0)


# ---------------------------------------------------------------- incremental == rebuild

@pytest.mark.parametrize("alg", ["async_gibbs", "metropolis_hastings", "hybrid_mcmc"])
def test_sweeps_keep_blockmodel_consistent(alg):
    g = dcsbm(300, 6, gamma=2.5, kmin=1, kmax=40, mean_degree=8, p_in=0.7, seed=11)
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(5)
    assign = rng.integers(0, 20, g.N).astype(np.int32)
    assign[:20] = np.arange(20)
    bm = O.OBlockmodel.build(og, 20, assign)
    prm = O.params(alg, seed=9, greedy=(alg != "hybrid_mcmc"))
    ep, st = bm.mcmc(prm, True, 0)
    assert st.moves > 0
    assert bm.validate()
    assert abs(bm.mdl() - O.OBlockmodel.build(og, 20, bm.export()["assignment"]).mdl()) == 0.0


def test_one_block_with_isolated_vertex_does_not_move():
    """At B = 1 no vertex has another block: an isolated vertex's random-block draw (U[0, B-2], empty) proposes
    the current block, and every other proposal is block 0; MH sweeps and finetune_assignment change nothing."""
    g = dcsbm(300, 4, seed=3)
    g = CSRGraph(g.N + 1, np.append(g.out_off, g.out_off[-1]).astype(np.int64), g.out_adj.copy(),
                 np.append(g.in_off, g.in_off[-1]).astype(np.int64), g.in_adj.copy())
    og = O.OGraph.from_csr(g)
    assign = np.zeros(g.N, np.int32)
    bm = O.OBlockmodel.build(og, 1, assign)
    m0 = bm.mdl()
    for ordered in (False, True):
        tot, st = bm.mh_sweep(O.params("metropolis_hastings", seed=2), 0, ordered=ordered)
        assert tot == 0.0 and st.moves == 0
    ep, st = bm.finetune_assignment(O.params("metropolis_hastings", seed=2), 0)
    assert st.moves == 0 and bm.mdl() == m0
    assert np.array_equal(bm.export()["assignment"], assign) and bm.validate()


def test_merge_apply_counts_and_labels():
    g = dcsbm(400, 8, gamma=2.5, kmin=1, kmax=40, mean_degree=8, p_in=0.7, seed=12)
    og = O.OGraph.from_csr(g)
    bm = O.OBlockmodel.build(og, g.N, np.arange(g.N, dtype=np.int32))
    for approximate in (False, True):
        best_s, best_dS = bm.merge_propose(seed=1, epoch=0)
        assert (best_s >= 0).all() and (best_s != np.arange(g.N)).all()
        out = bm.copy().merge_apply(best_s, best_dS, g.N // 2, approximate)
        x = out.export()
        assert out.B == g.N - g.N // 2
        assert sorted(set(x["assignment"].tolist())) == list(range(out.B))
        assert out.validate()


def test_full_sbp_recovers_planted_partition():
    from sklearn.metrics import normalized_mutual_info_score as nmi
    g = dcsbm(1000, 16, gamma=2.5, kmin=1, kmax=100, mean_degree=10, p_in=0.75, seed=0)
    og = O.OGraph.from_csr(g)
    a, nb, st = O.sbp(og, O.params("async_gibbs", seed=0, threads=4))
    assert nb == 16
    assert nmi(g.truth, a) > 0.9
    a2, nb2, _ = O.sbp(og, O.params("async_gibbs", seed=0, threads=1))
    assert (a == a2).all() and nb == nb2  # deterministic regardless of thread count


# ---------------------------------------------------------------- factored vs per-cell delta-MDL

@pytest.mark.parametrize("B", [4, 30, 200])
def test_factored_vertex_delta_mdl_matches_per_cell(B):
    """The replay contract's bulk form equals the reference's cell-by-cell delta_mdl within 1e-12 of the
    sum of |cell terms|, and never flips the sign of a delta that is not a rounding-level tie."""
    g = dcsbm(600, 6, gamma=2.3, kmin=1, kmax=120, mean_degree=9, p_in=0.6, seed=B)
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(B)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    bm = O.OBlockmodel.build(og, B, assign)
    n = 0
    for v in rng.choice(g.N, 150, replace=False):
        for s in rng.choice(B, 4, replace=False):
            if s == assign[v]:
                continue
            f = bm.vertex_delta_mdl(int(v), int(s))
            c, scale = bm.vertex_delta_mdl_cells(int(v), int(s))
            assert abs(f - c) <= 1e-12 * max(scale, 1.0), (v, s, f, c, scale)
            if abs(c) > 1e-9 * max(scale, 1.0):
                assert (f < 0) == (c < 0)
            n += 1
    assert n > 300


@pytest.mark.parametrize("B", [10, 120])
def test_factored_merge_delta_mdl_matches_per_cell(B):
    g = dcsbm(800, 8, gamma=2.3, kmin=1, kmax=120, mean_degree=9, p_in=0.6, seed=100 + B)
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(B)
    assign = rng.integers(0, B, g.N).astype(np.int32)
    assign[:B] = np.arange(B)
    bm = O.OBlockmodel.build(og, B, assign)
    for r in range(min(B, 40)):
        for s in rng.choice(B, 5, replace=False):
            if s == r:
                continue
            f = bm.merge_delta_mdl(r, int(s))
            c, scale = bm.merge_delta_mdl_cells(r, int(s))
            assert abs(f - c) <= 1e-12 * max(scale, 1.0), (r, s, f, c, scale)
