"""Pins the oracle's sampling, induced-subgraph and extension functions (src/sample.cpp) before the device
path is compared against them.  Each check restates the reference loop independently in Python, with
library routines where the reference uses them (Python's sorted() is a stable sort, as std::stable_sort
and utils::argsort are), on small graphs."""
import ctypes

import numpy as np
import pytest

import oracle as O
from sbp_amd.graph import dcsbm

_libm = ctypes.CDLL("libm.so.6")
_libm.logf.restype = ctypes.c_float
_libm.logf.argtypes = [ctypes.c_float]

TAG_SAMPLE = 0x53410001
TAG_EXTEND = 0x45580001


def _graph(seed=4, N=300, B=6):
    return dcsbm(N, B, gamma=2.3, kmin=1, kmax=40, mean_degree=6, p_in=0.7, seed=seed)


def _lists(g):
    """Per-vertex out/in lists in file order, as Graph holds them."""
    outs = [list(g.out_adj[g.out_off[v]:g.out_off[v + 1]]) for v in range(g.N)]
    ins = [list(g.in_adj[g.in_off[v]:g.in_off[v + 1]]) for v in range(g.N)]
    return outs, ins


def _degrees(g):
    """Graph::degrees (src/graph.cpp:24-31): |out| + |in| - self-loop flag."""
    outs, ins = _lists(g)
    return [len(outs[v]) + len(ins[v]) - (1 if v in outs[v] else 0) for v in range(g.N)]


def _take(order, n, N):
    mapping = np.full(N, -1, np.int32)
    for i, v in enumerate(order[:n]):
        mapping[v] = i
    return mapping, np.array(order[:n], np.int32)


@pytest.mark.parametrize("fraction", [0.1, 0.5, 0.93])
def test_max_degree_matches_stable_argsort(fraction):
    g = _graph()
    deg = _degrees(g)
    order = sorted(range(g.N), key=lambda v: -deg[v])  # stable: ties keep ascending ids
    n = int(fraction * g.N)
    m, vs = O.sample(O.OGraph.from_csr(g), "max_degree", fraction)
    em, evs = _take(order, n, g.N)
    assert np.array_equal(vs, evs) and np.array_equal(m, em)


@pytest.mark.parametrize("seed", [0, 7])
def test_random_is_ascending_philox_key(seed):
    g = _graph()
    keys = {v: int(O.philox(seed, v, 0, TAG_SAMPLE, 0)[0]) for v in range(g.N)}
    order = sorted(range(g.N), key=lambda v: (keys[v], v))
    m, vs = O.sample(O.OGraph.from_csr(g), "random", 0.4, seed)
    em, evs = _take(order, int(0.4 * g.N), g.N)
    assert np.array_equal(vs, evs) and np.array_equal(m, em)


@pytest.mark.parametrize("fraction", [0.0, 0.05, 0.5, 1.0])
def test_degree_product_matches_reference_loop(fraction):
    """sample::degree_product (src/sample.cpp:28-49) over Graph::sorted_edge_list (src/graph.cpp:302-316)."""
    g = _graph(seed=9)
    deg = _degrees(g)
    outs, _ = _lists(g)
    edges = [(s, d, deg[s] * deg[d]) for s in range(g.N) for d in outs[s]]
    edges = sorted(edges, key=lambda e: -e[2])  # std::stable_sort by descending information
    target = int(fraction * g.N)
    mapping = [-1] * g.N
    sampled = []
    for s, d, _ in edges:
        for x in (s, d):
            if mapping[x] == -1:
                mapping[x] = len(sampled)
                sampled.append(x)
        if len(sampled) >= target:
            break
    m, vs = O.sample(O.OGraph.from_csr(g), "degree_product", fraction)
    assert list(vs) == sampled and list(m) == mapping


def test_degree_weighted_keys_and_order():
    """Efraimidis-Spirakis keys -logf(u)/degree in float, u = (2 (w0 >> 9) + 1) 2^-24, ascending."""
    g = _graph(seed=5)
    deg = _degrees(g)
    keys = {}
    for v in range(g.N):
        w0 = int(O.philox(3, v, 0, TAG_SAMPLE, 0)[0])
        u = np.float32((((w0 >> 9) << 1) | 1)) * np.float32(2.0 ** -24)
        k = np.float32(np.inf) if deg[v] == 0 else np.float32(-np.float32(_libm.logf(float(u)))) / np.float32(deg[v])
        assert np.float32(O.degree_weight_key(w0, deg[v])) == k
        keys[v] = float(k)
    order = sorted(range(g.N), key=lambda v: (keys[v], v))
    m, vs = O.sample(O.OGraph.from_csr(g), "degree_weighted", 0.3, 3)
    em, evs = _take(order, int(0.3 * g.N), g.N)
    assert np.array_equal(vs, evs) and np.array_equal(m, em)


def test_degree_weighted_prefers_high_degree():
    """Inclusion frequency over seeds grows with degree (weighted sampling without replacement)."""
    g = _graph(seed=6, N=400)
    deg = np.array(_degrees(g))
    og = O.OGraph.from_csr(g)
    hits = np.zeros(g.N)
    for seed in range(40):
        m, _ = O.sample(og, "degree_weighted", 0.2, seed)
        hits += m >= 0
    lo, hi = deg <= np.percentile(deg, 25), deg >= np.percentile(deg, 90)
    assert hits[hi].mean() > 2 * hits[lo].mean()


@pytest.mark.parametrize("alg", ["max_degree", "random", "degree_product"])
def test_subgraph_matches_from_vertices(alg):
    """sample::from_vertices (src/sample.cpp:202-217): add_edge in ascending full source id."""
    g = _graph(seed=8)
    og = O.OGraph.from_csr(g)
    m, vs = O.sample(og, alg, 0.5, 1)
    n = len(vs)
    outs, _ = _lists(g)
    s_out = [[] for _ in range(n)]
    s_in = [[] for _ in range(n)]
    for v in range(g.N):
        if m[v] == -1:
            continue
        for u in outs[v]:
            if m[u] == -1:
                continue
            s_out[m[v]].append(int(m[u]))
            s_in[m[u]].append(int(m[v]))
    sg = O.subgraph(og, m, vs)
    assert sg.N == n
    for i in range(n):
        assert list(sg.out_adj[sg.out_off[i]:sg.out_off[i + 1]]) == s_out[i]
        assert list(sg.in_adj[sg.in_off[i]:sg.in_off[i + 1]]) == s_in[i]


@pytest.mark.parametrize("alg", ["max_degree", "random"])
def test_extend_matches_reference_loop(alg):
    """sample::extend (src/sample.cpp:157-200): sequential plurality vote over out- and in-neighbours
    (in-neighbour self-loops skipped), ties to the smallest block, Philox random block when none."""
    g = _graph(seed=10, N=400)
    og = O.OGraph.from_csr(g)
    m, vs = O.sample(og, alg, 0.5, 2)
    B = 7
    rng = np.random.default_rng(0)
    sample_assign = rng.integers(0, B, len(vs)).astype(np.int32)
    outs, ins = _lists(g)
    a = [int(sample_assign[m[v]]) if m[v] >= 0 else -1 for v in range(g.N)]
    for v in range(g.N):
        if a[v] != -1:
            continue
        counts = {}
        for u in outs[v]:
            if a[u] != -1:
                counts[a[u]] = counts.get(a[u], 0) + 1
        for u in ins[v]:
            if a[u] != -1 and u != v:
                counts[a[u]] = counts.get(a[u], 0) + 1
        if not counts:
            w0 = int(O.philox(5, v, 0, TAG_EXTEND, 0)[0])
            a[v] = (w0 * B) >> 32
        else:
            best = max(counts.values())
            a[v] = min(c for c, k in counts.items() if k == best)
    out = O.extend(og, m, sample_assign, B, 5)
    assert list(out) == a
