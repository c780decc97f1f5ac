"""Host-side parts of the batched and served modes (no GPU): the mixed-size workload draws and the footprint
bound the server admits requests by."""
import math

import numpy as np

from sbp_amd import native, workloads


def test_batched_specs_follow_the_stated_distribution():
    specs = [workloads.batched_spec(i) for i in range(2000)]
    assert specs[7] == workloads.batched_spec(7)  # graph i is a function of i
    N = np.array([s["N"] for s in specs])
    assert N.min() >= 200 and N.max() <= 50_000
    # log-uniform: log N is uniform on [log 200, log 50000]
    u = (np.log(N) - math.log(200)) / (math.log(50_000) - math.log(200))
    assert abs(u.mean() - 0.5) < 0.03 and abs(np.quantile(u, 0.25) - 0.25) < 0.04
    for s in specs[:50]:
        assert s["B"] == max(1, round(s["N"] ** 0.4))
        assert 8 <= s["mean_degree"] <= 32 and 0.5 <= s["p_in"] <= 0.8 and s["gamma"] == 2.5
    assert [s["seed"] for s in specs[:5]] == [0, 1, 2, 3, 4]


def test_served_specs_and_arrivals():
    specs = [workloads.served_spec(j) for j in range(500)]
    N = np.array([s["N"] for s in specs])
    assert N.min() >= 1000 and N.max() <= 2_000_000
    capped = [workloads.served_spec(j, n_max=200_000)["N"] for j in range(500)]
    assert max(capped) <= 200_000
    assert all(s["mean_degree"] == 16.0 and s["p_in"] == 0.55 and s["gamma"] == 2.1 for s in specs)
    t = workloads.poisson_arrivals(10.0, 4000, seed=0)
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    assert abs(np.diff(t).mean() - 0.1) < 0.01
    assert np.array_equal(t, workloads.poisson_arrivals(10.0, 4000, seed=0))


def test_generated_graph_matches_its_spec():
    s = workloads.batched_spec(3, n_max=3000)
    g = workloads.make(s)
    assert g.N == s["N"] and len(np.unique(g.truth)) == s["B"]
    deg = np.diff(g.out_off)
    assert deg.min() >= 1 and deg.max() <= s["kmax"]


def test_footprint_bound_is_monotone_and_covers_the_graph():
    L = native.lib()  # host-only functions: no device needed
    prev = 0
    for N, E in [(200, 2000), (1000, 16000), (50_000, 800_000), (2_000_000, 32_000_000)]:
        g, f = native.graph_bytes(N, E), int(L.sbp_footprint_bytes(N, E, 100, 0, 0))
        assert 0 < g < f and f > prev
        assert g >= 16 * (N + 1) + 8 * E
        prev = f
    base = int(L.sbp_footprint_bytes(1000, 20000, 100, 0, 0))
    assert int(L.sbp_footprint_bytes(1000, 40000, 100, 0, 0)) > base
    # high-degree vertices add one neighbour-block table each; Metropolis-Hastings adds its slack layout
    assert int(L.sbp_footprint_bytes(1000, 20000, 900, 10, 0)) > base
    assert int(L.sbp_footprint_bytes(1000, 20000, 900, 20, 0)) > int(L.sbp_footprint_bytes(1000, 20000, 900, 10, 0))
    assert int(L.sbp_footprint_bytes(1000, 20000, 100, 0, 1)) > base
    assert int(L.sbp_footprint_bytes(1000, 20000, 100, 0, 2)) == int(L.sbp_footprint_bytes(1000, 20000, 100, 0, 1))
    # the tables are capped at 1 GiB: the largest served request stays a small part of a 200 GB budget (its
    # blockmodels with their sorted cell keys, their dense copies and the cell update's work areas included)
    big = int(L.sbp_footprint_bytes(2_000_000, 32_000_000, 2000, 100_000, 0))
    assert big < 37e9
    # graphs of 2^22 edges or more add a dense copy of up to 8,192^2 counts for each of six blockmodels
    e = 1 << 22
    assert (int(L.sbp_footprint_bytes(100_000, e, 100, 0, 0)) - int(L.sbp_footprint_bytes(100_000, e - 1, 100, 0, 0))
            >= 6 * 4 * 8192 ** 2)
    # so do Metropolis-Hastings walks on any graph
    assert (int(L.sbp_footprint_bytes(100_000, e - 1, 100, 0, 1)) - int(L.sbp_footprint_bytes(100_000, e - 1, 100, 0, 0))
            >= 6 * 4 * 8192 ** 2)
    assert big - int(L.sbp_footprint_bytes(2_000_000, 32_000_000, 2000, 1, 0)) <= (1 << 30) * 5 // 4 + (1 << 20)


def test_footprint_of_a_graph():
    from sbp_amd.graph import dcsbm
    g = dcsbm(3000, 20, gamma=2.1, kmin=1, kmax=1000, mean_degree=12, seed=1)
    k = np.diff(g.out_off) + np.diff(g.in_off)
    assert native.footprint_bytes(g, "async_gibbs") == int(native.lib().sbp_footprint_bytes(
        g.N, g.E, int(k.max()), int((k > 256).sum()), 0))


def test_autotuned_slots_lookup(tmp_path, monkeypatch):
    import json
    from sbp_amd import autotune
    path = tmp_path / "autotune.json"
    monkeypatch.setattr(autotune, "TUNED", str(path))
    assert autotune.tuned_slots(16) == 16  # nothing measured
    path.write_text(json.dumps({autotune.host_key(): {"slots": 12}, "9999cpu": {"slots": 3}}))
    assert autotune.tuned_slots(16) == 12
    path.write_text("not json")
    assert autotune.tuned_slots(7) == 7


# sha256 over out_off, out_adj, in_off, in_adj and truth of graphs the parity suite and the benchmark use; the
# generator's output for a seed is part of what the recorded measurements refer to
GENERATOR_DIGESTS = [
    ((2000, 12, dict(gamma=2.5, kmin=1, kmax=100, mean_degree=10, p_in=0.7, seed=1)), 19832,
     "4fd1698d969183b72f309056002bc48beda8b11b40cdbd104dde7cb1f848cc4a"),
    ((50000, 80, dict(gamma=2.1, kmin=1, kmax=2000, mean_degree=20.0, p_in=0.6, sigma=0.5, seed=3)), 1018389,
     "bdd9bea3cae7b7ed3e9e7a56f7dcbb6a68722b2620ad6777aef3df684032b84a"),
    ((20000, 20, dict(gamma=2.1, kmin=1, kmax=800, mean_degree=16, p_in=0.4, sigma=1.0, seed=0)), 327247,
     "a3f849bd6f4b8b41bf11d4992c28d5f460a556701583e00ea3079341e50e64e1"),
]


def test_generator_output_is_pinned():
    import hashlib

    from sbp_amd.graph import dcsbm
    for (N, B, kw), E, digest in GENERATOR_DIGESTS:
        g = dcsbm(N, B, **kw)
        assert g.E == E
        h = hashlib.sha256()
        for a in (g.out_off, g.out_adj, g.in_off, g.in_adj, g.truth):
            h.update(np.ascontiguousarray(a).tobytes())
        assert h.hexdigest() == digest, (N, B, kw)
        src, dst = g.edges()
        assert len(np.unique(src * N + dst)) == g.E  # no duplicate edges
