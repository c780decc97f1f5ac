"""The drop-in C++ surface compiles from the reference's own header names: the command-line driver
(sbp_amd/host/main.cpp, the reference's src/main.cpp single-rank flow) and tests/dropin/api_check.cpp (the Delta,
entropy and Blockmodel calls as the reference's finetune and block-merge code makes them) build with g++ against
include/sbp/ and link to libsbp_amd.so.  No device is touched."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sbp_amd", "libsbp_amd.so")


def gxx(src, out):
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include", "sbp"), src, "-o", out,
           "-L", os.path.join(ROOT, "sbp_amd"), "-lsbp_amd", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
           "-Wl,-rpath," + os.path.join(ROOT, "sbp_amd")]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libsbp_amd.so is not built")
@pytest.mark.parametrize("src", ["sbp_amd/host/main.cpp", "tests/dropin/api_check.cpp"])
def test_compiles_against_reference_header_names(src, tmp_path):
    r = gxx(os.path.join(ROOT, src), str(tmp_path / "prog"))
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.exists(tmp_path / "prog")


# --- Graph::load: the reference's readers (src/graph.cpp:33-72 text, :75-129 matrix market; utils::insert and
# insert_nodup, src/utils.cpp:100-120), restated: edges in file order, each appended to its source's out-list and
# its target's in-list; --undirected (and a "symmetric" matrix market header) also appends the reverse of every
# non-self edge right after it; --noduplicates skips an edge already in its source's out-list.


def expected_lists(N, edges, undirected, noduplicates):
    out = [[] for _ in range(N)]
    inn = [[] for _ in range(N)]

    def insert(lst, a, b):
        if noduplicates and b in lst[a]:
            return
        lst[a].append(b)

    for a, b in edges:
        insert(out, a, b)
        insert(inn, b, a)
        if undirected and a != b:
            insert(out, b, a)
            insert(inn, a, b)
    return out, inn


def run_load_check(prog, base, *flags):
    r = subprocess.run([prog, base, *flags], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    N, E = map(int, lines[0].split())
    out = [list(map(int, lines[1 + 2 * v].split(":")[1].split())) for v in range(N)]
    inn = [list(map(int, lines[2 + 2 * v].split(":")[1].split())) for v in range(N)]
    truth = list(map(int, lines[1 + 2 * N].split(":")[1].split()))
    return N, E, out, inn, truth


@pytest.fixture(scope="module")
def load_check(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.skip("libsbp_amd.so is not built")
    prog = str(tmp_path_factory.mktemp("load") / "load_check")
    r = gxx(os.path.join(ROOT, "tests", "dropin", "load_check.cpp"), prog)
    assert r.returncode == 0, r.stderr[-3000:]
    return prog


# 1-indexed; a repeated edge (2->3 twice), a pair of opposite edges (1->2, 2->1), a self-loop (4->4) and a
# vertex (6) that only appears as a target
EDGES_1 = [(1, 2), (2, 3), (2, 1), (3, 1), (2, 3), (4, 4), (4, 6), (5, 4), (1, 5)]


@pytest.mark.parametrize("flags", [(), ("--undirected",), ("--noduplicates",), ("--undirected", "--noduplicates")])
def test_load_tsv(load_check, tmp_path, flags):
    base = str(tmp_path / "g")
    with open(base + ".tsv", "w") as f:
        for a, b in EDGES_1:
            f.write(f"{a}\t{b}\n")
    with open(base + "_truePartition.tsv", "w") as f:
        for v, c in [(1, 1), (2, 1), (3, 2), (4, 2), (5, 3), (6, 3)]:
            f.write(f"{v}\t{c}\n")
    N, E, out, inn, truth = run_load_check(load_check, base, *flags)
    want_out, want_in = expected_lists(6, [(a - 1, b - 1) for a, b in EDGES_1], "--undirected" in flags,
                                       "--noduplicates" in flags)
    assert N == 6
    assert out == want_out and inn == want_in
    assert E == sum(map(len, want_out))
    assert truth == [0, 0, 1, 1, 2, 2]


@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("noduplicates", [False, True])
def test_load_matrix_market(load_check, tmp_path, symmetric, noduplicates):
    base = str(tmp_path / "m")
    kind = "symmetric" if symmetric else "general"
    with open(base + ".mtx", "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate pattern {kind}\n% a comment line\n7 7 {len(EDGES_1)}\n")
        for a, b in EDGES_1:
            f.write(f"{a} {b}\n")
    flags = ("--noduplicates",) if noduplicates else ()
    N, E, out, inn, truth = run_load_check(load_check, base, *flags)
    want_out, want_in = expected_lists(7, [(a - 1, b - 1) for a, b in EDGES_1], symmetric, noduplicates)
    assert N == 7  # the header's vertex count, above the largest id
    assert out == want_out and inn == want_in
    assert truth == [0] * 7  # no true partition: every vertex in community 0
