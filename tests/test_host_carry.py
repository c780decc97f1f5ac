"""The library's host-side merge carry-out (sbp_amd/csrc/carry.hpp, built without HIP as
libsbp_carry.so) against the oracle's carry_out_best_merges_advanced / carry_out_best_merges
restatement, up to the first merge phase of the 200,000-vertex benchmark graph."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from sbp_amd.graph import dcsbm

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sbp_amd", "libsbp_carry.so")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def hist_slots(g, assignment):
    """Degree histograms as (key, count) slots: key = block << 33 | direction << 32 | degree (1 = in)."""
    out_deg, in_deg = np.diff(g.out_off), np.diff(g.in_off)
    keys = np.concatenate([(assignment.astype(np.uint64) << np.uint64(33)) | out_deg.astype(np.uint64),
                           (assignment.astype(np.uint64) << np.uint64(33)) | np.uint64(1 << 32) | in_deg.astype(np.uint64)])
    u, c = np.unique(keys, return_counts=True)
    return u.astype(np.uint64), c.astype(np.int32)


def host_carry(bm, best_s, best_dS, num_to_merge, approximate, threads=0, seed=True, verify=False, np_model=False,
               hist=None):
    L = C.CDLL(LIB)
    L.sbp_host_carry_out.restype = C.c_int32
    L.sbp_host_carry_out_dc.restype = C.c_int32
    x = bm.export()
    B = x["B"]
    rp, cols, cnts = x["row_ptr"], x["cols"], x["cnts"]
    rows = np.repeat(np.arange(B), np.diff(rp))
    order = np.lexsort((rows, cols))
    cp = np.zeros(B + 1, np.int32)
    np.cumsum(np.bincount(cols, minlength=B), out=cp[1:])
    crow, ccnt = rows[order].astype(np.int32), cnts[order].astype(np.int32)
    dout, din = x["d_out"].astype(np.int32), x["d_in"].astype(np.int32)
    s = np.ascontiguousarray(best_s, np.int32)
    d = np.ascontiguousarray(best_dS, np.float64)
    bmap = np.zeros(B, np.int32)
    st = np.zeros(10, np.int64)
    if hist is None:
        nb = L.sbp_host_carry_out(B, _p(rp), _p(cols), _p(cnts), _p(cp), _p(crow), _p(ccnt), _p(dout), _p(din),
                                  _p(x["sizes"]), _p(s), _p(d), C.c_int64(num_to_merge), int(approximate),
                                  int(threads), int(seed), int(verify), int(np_model),
                                  C.c_int64(len(x["assignment"])), _p(bmap), _p(st))
    else:
        hk, hc = hist
        nb = L.sbp_host_carry_out_dc(B, _p(rp), _p(cols), _p(cnts), _p(cp), _p(crow), _p(ccnt), _p(dout), _p(din),
                                     _p(x["sizes"]), _p(s), _p(d), C.c_int64(num_to_merge), int(approximate),
                                     int(threads), int(seed), int(verify), int(np_model),
                                     C.c_int64(len(x["assignment"])), _p(hk), _p(hc), C.c_int64(len(hk)), _p(bmap),
                                     _p(st))
    return nb, bmap, st


@pytest.mark.parametrize("approximate", [False, True])
@pytest.mark.parametrize("N,B", [(3000, 3000), (3000, 400), (200000, 200000)])
def test_host_carry_matches_oracle(N, B, approximate):
    if not os.path.exists(LIB):
        pytest.fail("libsbp_carry.so not built (run __graft_entry__.build())")
    if N == 200000:
        g = dcsbm(N, 500, gamma=2.1, kmin=1, kmax=2000, mean_degree=20, p_in=0.6, sigma=0.5, seed=0)
    else:
        g = dcsbm(N, 20, gamma=2.3, kmin=1, kmax=150, mean_degree=12, p_in=0.65, seed=N + B)
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(B)
    assign = np.arange(N, dtype=np.int32) if B == N else np.concatenate(
        [np.arange(B), rng.integers(0, B, N - B)]).astype(np.int32)
    bm = O.OBlockmodel.build(og, B, assign)
    best_s, best_dS = bm.merge_propose(seed=3, epoch=1)
    nb, bmap, st = host_carry(bm, best_s, best_dS, B // 2, approximate)
    out = bm.copy().merge_apply(best_s, best_dS, B // 2, approximate)
    assert nb == out.B == B - B // 2
    assert np.array_equal(bmap[assign], out.export()["assignment"])


@pytest.mark.parametrize("N,B", [(3000, 400), (60000, 60000)])
def test_host_carry_speculation_is_exact(N, B):
    """The speculative, parallel carry-out takes the sequential loop's decisions: seeded from the proposal
    values or not, on 1 thread or several, the block map is the same, and every proposal value equals the
    host re-evaluation on the pre-merge state bit for bit (what makes the seeding exact)."""
    g = dcsbm(N, 40, gamma=2.1, kmin=1, kmax=600, mean_degree=16, p_in=0.6, sigma=0.5, seed=7)
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(5)
    assign = np.arange(N, dtype=np.int32) if B == N else np.concatenate(
        [np.arange(B), rng.integers(0, B, N - B)]).astype(np.int32)
    bm = O.OBlockmodel.build(og, B, assign)
    best_s, best_dS = bm.merge_propose(seed=11, epoch=2)
    ref = None
    for threads, seed in [(1, False), (1, True), (4, True), (8, False)]:
        nb, bmap, st = host_carry(bm, best_s, best_dS, B // 2, False, threads=threads, seed=seed, verify=True)
        assert st[9] == 0, "proposal value differs from the host re-evaluation"
        assert st[5] == B // 2
        if seed:
            assert st[3] > 0
        if ref is None:
            ref = (nb, bmap)
        assert nb == ref[0] and np.array_equal(bmap, ref[1]), (threads, seed)
    out = bm.copy().merge_apply(best_s, best_dS, B // 2, False)
    assert np.array_equal(ref[1][assign], out.export()["assignment"])


@pytest.mark.parametrize("N,B", [(2000, 2000), (3000, 300)])
def test_host_carry_nonparametric_matches_oracle(N, B):
    """Nonparametric re-evaluations (block_merge_delta_mdl with the changing non-empty block count)."""
    g = dcsbm(N, 20, gamma=2.3, kmin=1, kmax=150, mean_degree=10, p_in=0.65, seed=N + 1)
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(N)
    assign = np.arange(N, dtype=np.int32) if B == N else np.concatenate(
        [np.arange(B), rng.integers(0, B, N - B)]).astype(np.int32)
    bm = O.OBlockmodel.build(og, B, assign, nonparametric=True)
    best_s, best_dS = bm.merge_propose(seed=2, epoch=3)
    for threads in (1, 8):
        nb, bmap, st = host_carry(bm, best_s, best_dS, B // 2, False, threads=threads, verify=True, np_model=True)
        assert st[9] == 0
        out = bm.copy().merge_apply(best_s, best_dS, B // 2, False)
        assert nb == out.B
        assert np.array_equal(bmap[assign], out.export()["assignment"])


@pytest.mark.parametrize("split", ["1,1,1,1", "2,3,1,2", "64,16,32,8"])
def test_host_carry_split_rows_are_exact(split, monkeypatch):
    """One evaluation or one merge split over the pool (fixed-point partial sums; renames in distinct maps)
    takes the same decisions as one thread, down to one-entry pieces."""
    g = dcsbm(4000, 30, gamma=2.1, kmin=1, kmax=800, mean_degree=24, p_in=0.6, sigma=0.5, seed=9)
    og = O.OGraph.from_csr(g)
    B = 600
    rng = np.random.default_rng(3)
    assign = np.concatenate([np.arange(B), rng.integers(0, B, g.N - B)]).astype(np.int32)
    bm = O.OBlockmodel.build(og, B, assign)
    best_s, best_dS = bm.merge_propose(seed=4, epoch=1)
    ref = host_carry(bm, best_s, best_dS, B // 2, False, threads=1)
    monkeypatch.setenv("SBP_AMD_CARRY_SPLIT", split)
    for seed in (False, True):
        nb, bmap, st = host_carry(bm, best_s, best_dS, B // 2, False, threads=8, seed=seed)
        assert nb == ref[0] and np.array_equal(bmap, ref[1]), seed
    out = bm.copy().merge_apply(best_s, best_dS, B // 2, False)
    assert np.array_equal(ref[1][assign], out.export()["assignment"])


@pytest.mark.parametrize("N,B", [(2000, 2000), (3000, 300)])
def test_host_carry_degree_corrected_matches_oracle(N, B):
    """Degree-corrected re-evaluations: the histograms of the merged blocks add up as merges happen."""
    g = dcsbm(N, 20, gamma=2.3, kmin=1, kmax=150, mean_degree=10, p_in=0.65, seed=N + 5)
    og = O.OGraph.from_csr(g)
    rng = np.random.default_rng(N + 1)
    assign = np.arange(N, dtype=np.int32) if B == N else np.concatenate(
        [np.arange(B), rng.integers(0, B, N - B)]).astype(np.int32)
    bm = O.OBlockmodel.build(og, B, assign, nonparametric=True, degreecorrected=True)
    best_s, best_dS = bm.merge_propose(seed=2, epoch=3)
    for threads in (1, 8):
        nb, bmap, st = host_carry(bm, best_s, best_dS, B // 2, False, threads=threads, verify=True, np_model=True,
                                  hist=hist_slots(g, assign))
        assert st[9] == 0
        out = bm.copy().merge_apply(best_s, best_dS, B // 2, False)
        assert nb == out.B
        assert np.array_equal(bmap[assign], out.export()["assignment"])


def test_host_carry_pipelined_merges_are_exact(monkeypatch):
    """Merges applied on their own thread while the loop goes on deciding (parametric, several threads,
    >= 32768 blocks) give the block map of the loop that applies each merge itself."""
    N = 40000
    g = dcsbm(N, 40, gamma=2.1, kmin=1, kmax=600, mean_degree=16, p_in=0.6, sigma=0.5, seed=9)
    og = O.OGraph.from_csr(g)
    bm = O.OBlockmodel.build(og, N, np.arange(N, dtype=np.int32))
    best_s, best_dS = bm.merge_propose(seed=4, epoch=1)
    maps = {}
    for pipe in ("0", "1"):
        monkeypatch.setenv("SBP_AMD_CARRY_PIPE", pipe)
        nb, bmap, st = host_carry(bm, best_s, best_dS, N // 2, False, threads=4)
        maps[pipe] = (nb, bmap, tuple(st[:6]))
    assert maps["0"][0] == maps["1"][0] and np.array_equal(maps["0"][1], maps["1"][1])
    assert maps["0"][2] == maps["1"][2]  # the same pops, evaluations and requeues
    out = bm.copy().merge_apply(best_s, best_dS, N // 2, False)
    assert np.array_equal(maps["1"][1], out.export()["assignment"])
