"""The command-line driver (the reference's src/main.cpp flow over the C++ drop-in API) on the device,
against the oracle on the same TSV graph: a plain run and a sampled run (--samplesize), bit-exact."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from sbp_amd.graph import dcsbm, load_tsv, save_tsv

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sbp_amd", "sbp_cli")


def run_cli(base, out, *flags):
    """Runs the driver and returns its JSON record (utils::write_json's keys, src/utils.cpp:258-292), written to
    <json dir>/<output_file>, plus num_blocks from its Results."""
    d, name = os.path.split(out)
    cmd = [CLI, "-f", base, "-j", d, "--output_file", name, "--seed", "0", *flags]
    print("[cli]", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    text = open(out).read()
    try:
        rec = json.loads(text)
    except json.JSONDecodeError:
        raise AssertionError(f"not JSON: {text[:300]}")
    for key in ("Algorithm", "Description Length", "Filepath", "Greedy", "Num. MCMC Iterations", "Num. Vertex Moves",
                "Results", "Runtime (s)", "Sample Size", "Sampling Algorithm", "Tag", "Undirected"):
        assert key in rec, key
    rec["num_blocks"] = len(set(rec["Results"]))
    return rec


@pytest.fixture(scope="module")
def tsv(tmp_path_factory):
    base = str(tmp_path_factory.mktemp("graph") / "dcsbm")
    save_tsv(dcsbm(1500, 12, gamma=2.5, kmin=1, kmax=100, mean_degree=10, p_in=0.75, seed=5), base)
    return base


def test_cli_partition_matches_oracle(tsv, tmp_path):
    csv = str(tmp_path / "eval" / "run")
    res = run_cli(tsv, str(tmp_path / "r.json"), "-a", "async_gibbs", "--evaluate", "-c", csv, "-n", "1500")
    g = load_tsv(tsv)
    a, nb, st = O.sbp(O.OGraph.from_csr(g), O.params("async_gibbs", seed=0, threads=8))
    assert res["num_blocks"] == nb
    assert res["Results"] == a.tolist()
    assert res["Description Length"] == st.mdl
    assert res["Algorithm"] == "async_gibbs" and res["Num. MCMC Iterations"] == st.sweeps
    # evaluate::write_results (src/evaluate.cpp:205-326): header, then the run's row
    rows = open(csv + "1500/static.csv").read().splitlines()
    head, row = rows[0].split(","), rows[1].split(",")
    assert head[:4] == ["tag", "numvertices", "numedges", "overlap"] and len(head) == len(row) == 40
    rec = dict(zip(head, row))
    assert int(rec["numvertices"]) == g.N and int(rec["numedges"]) == g.E and int(rec["num_blocks"]) == nb
    assert float(rec["nmi"]) > 0.9 and float(rec["mdl"]) == st.mdl


def test_cli_sampled_run_matches_oracle(tsv, tmp_path):
    res = run_cli(tsv, str(tmp_path / "s.json"), "-a", "hybrid_mcmc", "--samplesize", "0.5", "--samplingalg",
                  "max_degree")
    g = load_tsv(tsv)
    og = O.OGraph.from_csr(g)
    m, vs = O.sample(og, "max_degree", 0.5)
    sg = O.subgraph(og, m, vs)
    sa, snb, _ = O.sbp(sg, O.params("hybrid_mcmc", seed=0, threads=8))
    a = O.extend(og, m, sa, snb, 0)
    bm = O.OBlockmodel.build(og, snb, a)
    # the full-graph finetune_assignment starts its own epoch count at 0 (a different graph's vertices)
    bm.finetune_assignment(O.params("hybrid_mcmc", seed=0), 0)
    assert res["num_blocks"] == snb
    assert res["Results"] == bm.export()["assignment"].tolist()
    assert res["Description Length"] == bm.mdl()


def test_cli_detached_run_matches_oracle(tmp_path):
    """--detach: partition the vertices of degree > 1, reattach the rest (src/main.cpp:86-92,153-156)."""
    from test_detach_oracle import leafy_graph
    base = str(tmp_path / "leafy")
    save_tsv(leafy_graph(7, N=2000, mean_degree=10), base)
    res = run_cli(base, str(tmp_path / "d.json"), "-a", "hybrid_mcmc", "--detach")
    g = load_tsv(base)
    og = O.OGraph.from_csr(g)
    m, vs = O.detach(og)
    sa, snb, _ = O.sbp(O.subgraph(og, m, vs), O.params("hybrid_mcmc", seed=0, threads=8))
    a = O.reattach(og, m, sa, snb, seed=0)
    assert res["num_blocks"] == snb
    assert res["Results"] == a.tolist()
    # reattach returns a blockmodel built from the assignment, whose description length is the constructor's
    # initial value (src/sample.cpp:404; include/blockmodel/blockmodel.hpp:53-58)
    assert res["Description Length"] == np.finfo(np.float64).max
