"""--detach (sample::detach / sample::reattach, src/sample.cpp:12-26,367-405): the oracle's restatement and the
library's host reattach against a literal Python restatement of the reference's loops under the replay
draws, on graphs with isolated vertices, self-loops, chains of degree-1 vertices and edges between them.
No device is used."""
import numpy as np
import pytest

import oracle as O
from sbp_amd import native
from sbp_amd.graph import CSRGraph, dcsbm

TAG_REATTACH = 0x52410001


def degrees(g):
    src, dst = g.edges()
    self_flag = np.zeros(g.N, np.int64)
    self_flag[src[src == dst]] = 1
    return np.diff(g.out_off) + np.diff(g.in_off) - self_flag


def py_reattach(g, mapping, detached_assignment, B, seed):
    """The reference's loop, statement by statement (sample.cpp:367-405)."""
    out_n = [g.out_adj[g.out_off[v]:g.out_off[v + 1]] for v in range(g.N)]
    in_n = [g.in_adj[g.in_off[v]:g.in_off[v + 1]] for v in range(g.N)]
    assignment = [-1] * g.N
    for v in range(g.N):
        if mapping[v] == -1:
            continue
        assignment[v] = int(detached_assignment[mapping[v]])
    draws = 0
    for v in range(g.N):
        if assignment[v] != -1:
            continue
        w = O.philox(seed, draws, 0, TAG_REATTACH, 0)
        draws += 1
        random_community = (int(w[0]) * B) >> 32
        for u in out_n[v]:
            community = assignment[u]
            if community == -1:
                assignment[v] = random_community
                assignment[u] = random_community
                break
            assignment[v] = community
        for u in in_n[v]:
            community = assignment[u]
            if community == -1:
                assignment[v] = random_community
                assignment[u] = random_community
                break
            assignment[v] = community
        if assignment[v] < 0:
            assignment[v] = random_community
    return np.array(assignment, np.int32)


def leafy_graph(seed, N=600, mean_degree=4):
    """A DCSBM core plus pendant chains, isolated vertices, self-loops and leaf-to-leaf edges."""
    core = dcsbm(N // 2, 4, gamma=2.5, kmin=1, kmax=20, mean_degree=mean_degree, p_in=0.8, seed=seed)
    rng = np.random.default_rng(seed)
    src, dst = [list(x) for x in core.edges()]
    nxt = core.N
    for _ in range(N // 6):  # chains hanging off the core, either direction
        a = int(rng.integers(core.N))
        for _ in range(int(rng.integers(1, 4))):
            if rng.random() < 0.5:
                src.append(a), dst.append(nxt)
            else:
                src.append(nxt), dst.append(a)
            a = nxt
            nxt += 1
    for _ in range(10):  # isolated pairs and self-loop-only vertices
        src.append(nxt), dst.append(nxt + 1)
        nxt += 2
        src.append(nxt), dst.append(nxt)
        nxt += 1
    nxt += 5  # isolated vertices
    return CSRGraph.from_edges(np.array(src), np.array(dst), nxt)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_detach_keeps_degree_above_one(seed):
    g = leafy_graph(seed)
    og = O.OGraph.from_csr(g)
    mapping, verts = O.detach(og)
    keep = np.flatnonzero(degrees(g) > 1)
    assert np.array_equal(verts, keep)
    assert np.array_equal(mapping[keep], np.arange(len(keep)))
    assert np.all(mapping[degrees(g) <= 1] == -1)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_reattach_matches_the_reference_loop(seed):
    g = leafy_graph(seed)
    og = O.OGraph.from_csr(g)
    mapping, verts = O.detach(og)
    B = 7
    rng = np.random.default_rng(seed)
    detached = rng.integers(0, B, len(verts)).astype(np.int32)
    want = py_reattach(g, mapping, detached, B, seed=seed)
    assert np.array_equal(O.reattach(og, mapping, detached, B, seed=seed), want)
    assert np.array_equal(native.reattach(g, mapping, detached, B, seed=seed), want)
    assert want.min() >= 0 and want.max() < B
    assert np.array_equal(want[verts], detached)  # kept vertices keep their blocks
