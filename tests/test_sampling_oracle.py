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


TAG_SNOWBALL = 0x53410002
_HOST = None


def _host_lib():
    """The library's host-only build (std::priority_queue, segment and Fenwick trees)."""
    global _HOST
    if _HOST is None:
        import os
        path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sbp_amd", "libsbp_carry.so")
        _HOST = ctypes.CDLL(path)
        _HOST.sbp_host_snowball.restype = ctypes.c_int64
        _HOST.sbp_host_snowball.argtypes = [ctypes.c_int32, ctypes.c_int32] + [ctypes.c_void_p] * 5 + [
            ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
    return _HOST


def _host_snowball(g, alg, fraction, seed):
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    deg = np.array(_degrees(g), np.int32)
    oo, oa = np.ascontiguousarray(g.out_off, np.int64), np.ascontiguousarray(g.out_adj, np.int32)
    io, ia = np.ascontiguousarray(g.in_off, np.int64), np.ascontiguousarray(g.in_adj, np.int32)
    mapping = np.zeros(g.N, np.int32)
    verts = np.zeros(g.N + 1, np.int32)
    n = _host_lib().sbp_host_snowball(O.SAMPLING[alg], g.N, p(oo), p(oa), p(io), p(ia), p(deg), fraction, seed,
                                      p(mapping), p(verts))
    return mapping, verts[:n]


def _components_graph(seed):
    """Three small DCSBM graphs side by side plus isolated vertices: the frontier empties and restarts."""
    from sbp_amd.graph import CSRGraph
    parts, src, dst, off = [], [], [], 0
    for i, n in enumerate((60, 90, 40)):
        h = dcsbm(n, 3, gamma=2.3, kmin=1, kmax=12, mean_degree=4, p_in=0.8, seed=seed + i)
        s = np.repeat(np.arange(h.N), np.diff(h.out_off))
        src.append(s + off)
        dst.append(np.asarray(h.out_adj) + off)
        off += h.N
    N = off + 7
    return CSRGraph.from_edges(np.concatenate(src), np.concatenate(dst), N)


def _ring_graph(N=120, k=3):
    """Every vertex has the same degree: all degree products tie and only the heap order decides."""
    from sbp_amd.graph import CSRGraph
    src = np.repeat(np.arange(N), k)
    dst = (src + np.tile(np.arange(1, k + 1), N)) % N
    return CSRGraph.from_edges(src, dst, N)


def _es_reference(g, fraction, seed):
    """sample::expansion_snowball (src/sample.cpp:94-155) restated: std::set as a Python set read in ascending
    order, utils::argmax as the first maximum over all vertices."""
    outs, ins = _lists(g)
    N, target = g.N, int(fraction * g.N)
    nxt = [0]

    def rint(lo, hi):
        w = int(O.philox(seed, nxt[0], 0, TAG_SNOWBALL, 0)[0])
        nxt[0] += 1
        return lo + ((w * (hi - lo + 1)) >> 32)

    contribution, flag, neighbors = [0] * N, [False] * N, set()
    sampled, mapping, total = [], [-1] * N, [0]

    def update(u):
        for nb in outs[u]:
            if flag[nb] or mapping[nb] >= 0:
                continue
            contribution[u] += 1
            total[0] += 1
        for nb in ins[u]:
            if contribution[nb] > 0:
                contribution[nb] -= 1
                total[0] -= 1

    def add(v):
        sampled.append(v)
        mapping[v] = len(sampled) - 1
        for nb in outs[v]:
            if flag[nb] or mapping[nb] >= 0:
                continue
            neighbors.add(nb)
            flag[nb] = True
            if contribution[nb] > 0:
                continue
            update(nb)
        neighbors.discard(v)
        flag[v] = False

    add(rint(0, N - 1))
    while len(sampled) < target:
        if not neighbors:
            v = rint(0, N - 1)
            while mapping[v] >= 0:
                v = rint(0, N - 1)
            add(v)
        elif total[0] == 0:
            add(sorted(neighbors)[rint(0, len(neighbors) - 1)])
        else:
            add(max(range(N), key=lambda u: (contribution[u], -u)))
    return np.array(mapping, np.int32), np.array(sampled, np.int32)


@pytest.mark.parametrize("graph", ["dcsbm", "components"])
@pytest.mark.parametrize("fraction", [0.05, 0.3, 0.9])
def test_expansion_snowball_matches_reference_loop(graph, fraction):
    g = _graph(seed=11) if graph == "dcsbm" else _components_graph(3)
    for seed in (0, 5):
        em, ev = _es_reference(g, fraction, seed)
        m, vs = O.sample(O.OGraph.from_csr(g), "expansion_snowball", fraction, seed)
        assert np.array_equal(vs, ev) and np.array_equal(m, em), seed


@pytest.mark.parametrize("alg", ["expansion_snowball", "degree_product_snowball"])
@pytest.mark.parametrize("graph", ["dcsbm", "components", "ring"])
@pytest.mark.parametrize("fraction", [0.0, 0.1, 0.5, 1.0])
def test_snowball_library_matches_oracle(alg, graph, fraction):
    """The oracle's heap restatement against libstdc++'s std::priority_queue (ring: every product ties), and
    the literal argmax / set scans against the library's segment and Fenwick trees."""
    g = {"dcsbm": lambda: _graph(seed=12, N=600), "components": lambda: _components_graph(8),
         "ring": _ring_graph}[graph]()
    for seed in (1, 9):
        m, vs = O.sample(O.OGraph.from_csr(g), alg, fraction, seed)
        hm, hv = _host_snowball(g, alg, fraction, seed)
        assert np.array_equal(vs, hv) and np.array_equal(m, hm), seed


@pytest.mark.parametrize("graph", ["dcsbm", "components"])
def test_degree_product_snowball_invariants(graph):
    """Sizes, distinct vertices, and every vertex after a restart joins through an edge of the sample."""
    g = _graph(seed=13, N=500) if graph == "dcsbm" else _components_graph(4)
    outs, ins = _lists(g)
    deg = _degrees(g)
    for fraction in (0.1, 0.6):
        m, vs = O.sample(O.OGraph.from_csr(g), "degree_product_snowball", fraction, 2)
        target = int(fraction * g.N)
        assert target <= len(vs) <= max(target, 1) + 1
        assert len(set(vs.tolist())) == len(vs)
        assert all(m[v] == i for i, v in enumerate(vs))
        seen = {int(vs[0])}
        for v in vs[1:]:
            v = int(v)
            linked = any(u in seen for u in outs[v]) or any(u in seen for u in ins[v])
            frontier_empty = all(all(u in seen for u in outs[w]) and all(u in seen for u in ins[w]) for w in seen)
            assert linked or frontier_empty, v
            seen.add(v)
        assert max(deg) >= 1
