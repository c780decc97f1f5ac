"""divisive::split_communities (src/top_down_sbp.cpp:488-549) with random splits (propose_random_split :228-240,
extract_subgraph :98-132, apply_best_splits :30-66): the oracle's restatement against a literal Python
restatement of the reference's loops under the replay draws, on graphs with singleton and empty blocks, ties,
both null models and several targets.  Each deliberate mistake in the restatement must change at least one
result.  The description lengths come from the oracle's nonparametric::mdl on blockmodels of the Python-built
subgraphs (that function is pinned on its own by the reference's fixtures).  No device is used."""
import numpy as np
import pytest

import oracle as O
from sbp_amd.graph import dcsbm

TAG_SPLIT_SIZE = 0x53500001
TAG_SPLIT_SIDE = 0x53500002
NUM_AGG_PROPOSALS_PER_BLOCK = 10  # include/top_down_sbp.hpp:28
DBL_MAX = np.finfo(np.float64).max

MUTATIONS = ["small_le", "hi_exclusive", "side_swap", "ascending_walk", "tie_order", "no_break", "parametric_null",
             "keep_last_min", "size_unit_proposal"]


def uniform_below(w, n):
    return (int(w) * n) >> 32


def extract_subgraph(g, assignment, cluster):
    """extract_subgraph (:98-132): vertices ascending, translator, induced out-edges in list order."""
    translator = {}
    cluster_vertices = []
    for vertex in range(g.N):
        if assignment[vertex] != cluster:
            continue
        translator[vertex] = len(cluster_vertices)
        cluster_vertices.append(vertex)
    n = len(cluster_vertices)
    out_lists = [[] for _ in range(n)]
    in_lists = [[] for _ in range(n)]
    if n >= 2:
        for vertex in cluster_vertices:
            for neighbor in g.out_adj[g.out_off[vertex]:g.out_off[vertex + 1]]:
                if int(neighbor) not in translator:
                    continue
                out_lists[translator[vertex]].append(translator[int(neighbor)])  # Graph::add_edge
                in_lists[translator[int(neighbor)]].append(translator[vertex])
    out_off = np.concatenate([[0], np.cumsum([len(x) for x in out_lists])]).astype(np.int64)
    in_off = np.concatenate([[0], np.cumsum([len(x) for x in in_lists])]).astype(np.int64)
    out_adj = np.array([u for x in out_lists for u in x], np.int32)
    in_adj = np.array([u for x in in_lists for u in x], np.int32)
    return O.OGraph(out_off, out_adj, in_off, in_adj), cluster_vertices


def propose_random_split(sub_n, vertices, block, p, seed, epoch, mutate):
    """propose_random_split (:228-240) under the replay draws."""
    sub = p if "size_unit_proposal" in mutate else p // 4
    w = O.philox(seed, block, epoch, TAG_SPLIT_SIZE, sub)
    lo, hi = int(sub_n * 0.2), int(sub_n * 0.8)
    span = hi - lo + (0 if "hi_exclusive" in mutate else 1)
    split_size = lo + uniform_below(w[p % 4], max(span, 1))
    assignment = []
    for i in range(sub_n):
        w = O.philox(seed, vertices[i], epoch, TAG_SPLIT_SIDE, p // 4)
        side0 = uniform_below(w[p % 4], sub_n) < split_size
        if "side_swap" in mutate:
            side0 = not side0
        assignment.append(0 if side0 else 1)
    return assignment


def py_split_communities(g, assignment, B, target, seed, epoch, nonparametric, mutate=()):
    """split_communities and apply_best_splits, statement by statement."""
    assignment = [int(x) for x in assignment]
    N = g.N
    sizes = [0] * B
    for b in assignment:
        sizes[b] += 1
    best = [None] * B  # Split: (num_vertices, translator vertices, split assignment)
    delta = [DBL_MAX] * B
    for block in range(B):
        sub, vertices = extract_subgraph(g, assignment, block)
        for p in range(NUM_AGG_PROPOSALS_PER_BLOCK):
            small = sizes[block] <= 0.001 * N if "small_le" in mutate else sizes[block] < 0.001 * N
            if small:
                best[block] = None
                delta[block] = DBL_MAX
                continue
            n = len(vertices)
            if n < 2:
                split_assignment = [0] * n
                bm = O.OBlockmodel.build(sub, 1, np.zeros(max(n, 1), np.int32)[:n], nonparametric=True)
            else:
                split_assignment = propose_random_split(n, vertices, block, p, seed, epoch, mutate)
                bm = O.OBlockmodel.build(sub, 2, np.array(split_assignment, np.int32), nonparametric=True)
            new_entropy = bm.mdl()
            if nonparametric and "parametric_null" not in mutate:
                null_model = O.OBlockmodel.build(sub, 1, np.zeros(n, np.int32), nonparametric=True)
                old_entropy = null_model.mdl()
            else:
                old_entropy = O.null_mdl_v1(sub.E)
            delta_entropy = new_entropy - old_entropy
            better = delta_entropy <= delta[block] if "keep_last_min" in mutate else delta_entropy < delta[block]
            if better:
                delta[block] = delta_entropy
                best[block] = (n, vertices, split_assignment)
    # apply_best_splits: argsort is stable by descending value, walked from its end
    if "tie_order" in mutate:
        sorted_splits = sorted(range(B), key=lambda b: (-delta[b], -b))
    else:
        sorted_splits = sorted(range(B), key=lambda b: -delta[b])  # Python's sort is stable
    walk = range(B) if "ascending_walk" in mutate else range(B - 1, -1, -1)
    num_blocks = B
    for index_of_split in walk:
        block = sorted_splits[index_of_split]
        split = best[block]
        num_vertices = split[0] if split is not None else 2 ** 63 - 1
        if num_vertices > N:
            continue
        _, vertices, split_assignment = split
        for index in range(num_vertices):
            if split_assignment[index] == 1:
                assignment[vertices[index]] = num_blocks
        num_blocks += 1
        if num_blocks >= target and "no_break" not in mutate:
            break
    return np.array(assignment, np.int32), num_blocks, np.array(delta)


def cases():
    """(graph, assignment, B, target, seed, epoch, nonparametric)."""
    out = []
    g = dcsbm(600, 6, gamma=2.5, kmin=1, kmax=30, mean_degree=6, p_in=0.7, seed=3)
    rng = np.random.default_rng(0)
    a = rng.integers(0, 6, g.N).astype(np.int32)
    out.append((g, a, 6, 9, 11, 0, True))
    out.append((g, a, 6, 9, 11, 1, False))
    out.append((g, a, 6, 100, 5, 2, True))
    out.append((g, np.zeros(g.N, np.int32), 1, 2, 7, 0, True))  # the top-down start: one block
    out.append((g, np.zeros(g.N, np.int32), 1, 2, 7, 0, False))
    # singletons (active when N <= 1000: 1 >= 0.001 N) and an empty block; delta ties at 0
    g2 = dcsbm(1000, 4, gamma=2.5, kmin=1, kmax=20, mean_degree=4, p_in=0.8, seed=9)
    a2 = np.array(g2.truth if hasattr(g2, "truth") else rng.integers(0, 4, g2.N), np.int32) % 4
    a2 = a2.copy()
    for i, v in enumerate([5, 17, 900]):
        a2[v] = 5 + i  # blocks 5, 6, 7 singletons; block 4 empty
    out.append((g2, a2, 8, 16, 2, 3, True))
    out.append((g2, a2, 8, 10, 2, 4, True))
    out.append((g2, a2, 8, 16, 2, 5, False))
    # exact ties between real splits: six blocks of two non-adjacent vertices without self-loops, each of whose
    # best splits has value 0, and a target that stops inside the tie
    a3 = (np.array(g2.truth) % 4).astype(np.int32)
    src, dst = g2.edges()
    adj = set(zip(src.tolist(), dst.tolist()))
    free = [v for v in range(g2.N) if (v, v) not in adj]
    pairs, i = [], 0
    while len(pairs) < 6:
        u, v = free[i], free[i + 1]
        i += 2
        if (u, v) not in adj and (v, u) not in adj:
            pairs.append((u, v))
    for j, (u, v) in enumerate(pairs):
        a3[u] = a3[v] = 4 + j
    out.append((g2, a3, 10, 12, 1, 0, True))
    return out


@pytest.fixture(scope="module")
def all_cases():
    return cases()


@pytest.mark.parametrize("i", range(9))
def test_oracle_matches_literal_restatement(all_cases, i):
    g, a, B, target, seed, epoch, np_ = all_cases[i]
    og = O.OGraph.from_csr(g)
    oa, onb, od = O.split_communities(og, a, B, target, seed, epoch, np_)
    pa, pnb, pd = py_split_communities(g, a, B, target, seed, epoch, np_)
    assert onb == pnb
    np.testing.assert_array_equal(oa, pa)
    np.testing.assert_array_equal(od, pd)  # bit-exact, DBL_MAX included
    # invariants of the step: old blocks keep their side-0 vertices, new blocks are numbered from B up
    assert onb >= B and oa.max() < onb
    moved = oa != a
    assert np.all(oa[moved] >= B)
    if onb > B:
        assert np.all(np.bincount(oa, minlength=onb)[B:] >= 0)


@pytest.mark.parametrize("mutation", MUTATIONS)
def test_each_mistake_is_caught(all_cases, mutation):
    for g, a, B, target, seed, epoch, np_ in all_cases:
        og = O.OGraph.from_csr(g)
        oa, onb, od = O.split_communities(og, a, B, target, seed, epoch, np_)
        pa, pnb, pd = py_split_communities(g, a, B, target, seed, epoch, np_, mutate={mutation})
        if onb != pnb or not np.array_equal(oa, pa) or not np.array_equal(od, pd):
            return
    pytest.fail(f"mutation {mutation} not detected by any case")


def test_split_values_are_two_block_description_lengths(all_cases):
    """The kept value of a block equals nonparametric::mdl of the two-block model of its subgraph, rebuilt here
    from the applied split, minus the null model."""
    g, a, B, target, seed, epoch, np_ = all_cases[2]
    og = O.OGraph.from_csr(g)
    oa, onb, od = O.split_communities(og, a, B, target, seed, epoch, np_)
    assert onb == 2 * B  # every block is split once and the target is not reached
    for b in range(B):
        sub, vertices = extract_subgraph(g, a, b)
        side = (oa[vertices] != b).astype(np.int32)
        two = O.OBlockmodel.build(sub, 2, side, nonparametric=True).mdl()
        null = O.OBlockmodel.build(sub, 1, np.zeros(len(vertices), np.int32), nonparametric=True).mdl()
        assert two - null == od[b]
        assert np.isfinite(od[b])
