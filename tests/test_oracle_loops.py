"""Pins the oracle's sequential loops against a literal Python restatement of the reference's (reference_loops.py).

For each loop -- an asynchronous-Gibbs sweep, a Metropolis-Hastings sweep, hybrid MCMC, finetune_assignment, the
merge proposals, both merge carry-outs with label compaction, and the whole driver with its golden-section
triplet -- the oracle and the restatement start from the same blockmodel and replay stream, and must end in the
same assignment (and, where the loop has them, the same sum of accepted delta-MDL, move and refusal counts,
epochs used, description length and search trace).

Each loop also has deliberate mistakes that the comparison must reject, run on inputs where the mistake
matters: committing in visit order instead of vertex-id order (src/finetune.cpp:80-85), a missing or wrong
refusal (src/blockmodel/blockmodel.cpp:439-443), processing the batch tail the reference skips
(src/finetune.cpp:57-58), dropping MH's singleton skip (:845-847), swapping hybrid's vertex lists or its split
bound (src/graph.cpp:251-259), swapping the early-stop thresholds (:265-283) or giving the 3-argument form a
zero exit (:285-294), never re-queueing a merge that got worse (src/block_merge.cpp:79-83), compacting labels in
descending order (src/blockmodel/blockmodel.cpp:34-46), an unstable sort or a one-step translate in the
approximate carry-out (:55-122), and a wrong golden-section branch or ratio
(src/blockmodel/blockmodel_triplet.cpp:27-35).
"""
import numpy as np
import pytest

import oracle as O
import reference_loops as R
from sbp_amd.graph import CSRGraph, dcsbm


def small_graph(seed, N=120):
    return dcsbm(N, 4, gamma=2.5, kmin=1, kmax=30, mean_degree=6, p_in=0.7, seed=seed)


def with_isolated_vertex(g):
    return CSRGraph(g.N + 1, np.append(g.out_off, g.out_off[-1]).astype(np.int64), g.out_adj.copy(),
                    np.append(g.in_off, g.in_off[-1]).astype(np.int64), g.in_adj.copy())


def start(g, B, seed):
    """Random assignment over B blocks, every block non-empty (the first B vertices seed one block each)."""
    a = np.random.default_rng(seed).integers(0, B, g.N).astype(np.int32)
    a[:B] = np.arange(B)
    return a


def same_state(ob, rs):
    return np.array_equal(ob.export()["assignment"], np.asarray(rs.assign, np.int32))


# ---------------------------------------------------------------------------------------------- sweeps

def async_case(g, B, seed, batches, ordered=False, mutate=()):
    og = O.OGraph.from_csr(g)
    assign = start(g, B, seed)
    prm = O.params("async_gibbs", seed=seed, batches=batches, ordered=ordered)
    ob = O.OBlockmodel.build(og, B, assign)
    _, st = ob.async_gibbs_sweep(prm, 2)
    rs = R.State(g, B, assign)
    moved, refused = R.async_gibbs_sweep(rs, og, prm, 2, None, mutate)
    return same_state(ob, rs) and st.moves == moved and st.refused == refused, st


@pytest.mark.parametrize("N,B,batches,ordered", [(120, 40, 1, False), (101, 90, 3, False), (101, 90, 3, True),
                                                 (121, 30, 4, False)])
def test_async_gibbs_sweep_matches_reference_loop(N, B, batches, ordered):
    g = with_isolated_vertex(small_graph(0, N))
    ok, st = async_case(g, B, 0, batches, ordered)
    assert ok
    assert st.moves > 0


def test_async_gibbs_case_has_refusals():
    """The mutation inputs below exercise a refusal (a block emptied at commit time)."""
    _, st = async_case(small_graph(0, 101), 90, 0, 3)
    assert st.refused > 0


@pytest.mark.parametrize("mistake", ["commit_in_visit_order", "no_refusal", "refuse_size_2", "process_tail"])
def test_async_gibbs_mistakes_are_caught(mistake):
    ok, _ = async_case(small_graph(0, 101), 90, 0, 3, mutate={mistake})
    assert not ok


def mh_case(g, B, seed, greedy, ordered, nonparametric=False, mutate=()):
    og = O.OGraph.from_csr(g)
    assign = start(g, B, seed)
    prm = O.params("metropolis_hastings", seed=seed, greedy=greedy, nonparametric=nonparametric)
    ob = O.OBlockmodel.build(og, B, assign, nonparametric=nonparametric)
    tot, st = ob.mh_sweep(prm, 3, ordered)
    rs = R.State(g, B, assign, nonparametric)
    tot2, moves = R.mh_sweep(rs, og, prm, 3, ordered, mutate)
    return same_state(ob, rs) and tot == tot2 and st.moves == moves, st


@pytest.mark.parametrize("greedy", [True, False])
@pytest.mark.parametrize("ordered", [False, True])
def test_mh_sweep_matches_reference_loop(greedy, ordered):
    ok, st = mh_case(with_isolated_vertex(small_graph(1)), 40, 1, greedy, ordered)
    assert ok and st.moves > 0


def test_mh_sweep_nonparametric_matches_reference_loop():
    ok, st = mh_case(small_graph(2), 60, 2, True, False, nonparametric=True)
    assert ok and st.moves > 0


def test_mh_singleton_skip_is_caught():
    """The nonparametric description length charges for non-empty blocks, so a walk that also proposes moves out
    of one-vertex blocks empties some and diverges."""
    ok, _ = mh_case(small_graph(2), 60, 2, True, False, nonparametric=True, mutate={"no_singleton_skip"})
    assert not ok


def mcmc_case(alg, seed, grnr, greedy=True, mutate=()):
    g = small_graph(seed)
    og = O.OGraph.from_csr(g)
    assign = start(g, 30, seed)
    prm = O.params(alg, seed=seed, mh_percent=0.2, greedy=greedy)
    ob = O.OBlockmodel.build(og, 30, assign)
    ep, st = ob.mcmc(prm, grnr, 4)
    rs = R.State(g, 30, assign)
    ep2 = R.mcmc(rs, og, prm, grnr, 4, mutate)
    return same_state(ob, rs) and ep == ep2 and ob.mdl() == rs.overall_entropy, ep - 4


@pytest.mark.parametrize("alg", ["async_gibbs", "metropolis_hastings", "hybrid_mcmc"])
@pytest.mark.parametrize("grnr", [True, False])
def test_mcmc_matches_reference_loop(alg, grnr):
    ok, sweeps = mcmc_case(alg, 0, grnr)
    assert ok and sweeps > 3


def test_hybrid_nongreedy_matches_reference_loop():
    ok, _ = mcmc_case("hybrid_mcmc", 3, True, greedy=False)
    assert ok


@pytest.mark.parametrize("mistake,alg", [("hybrid_swap_lists", "hybrid_mcmc"), ("split_le", "hybrid_mcmc"),
                                         ("swap_thresholds", "async_gibbs"), ("swap_thresholds", "metropolis_hastings")])
def test_mcmc_mistakes_are_caught(mistake, alg):
    ok, _ = mcmc_case(alg, 0, True, mutate={mistake})
    assert not ok


def finetune_assignment_case(seed, mutate=()):
    g = small_graph(seed)
    og = O.OGraph.from_csr(g)
    assign = start(g, 30, seed)
    prm = O.params("metropolis_hastings", seed=seed)
    ob = O.OBlockmodel.build(og, 30, assign)
    ep, _ = ob.finetune_assignment(prm, 7)
    rs = R.State(g, 30, assign)
    ep2 = R.finetune_assignment(rs, og, prm, 7, mutate)
    return same_state(ob, rs) and ep == ep2 and ob.mdl() == rs.overall_entropy


def test_finetune_assignment_matches_reference_loop():
    assert finetune_assignment_case(0)
    assert finetune_assignment_case(3)


def test_finetune_assignment_zero_exit_is_caught():
    assert not finetune_assignment_case(0, {"zero_exit"})


# ---------------------------------------------------------------------------------------------- block merge

@pytest.mark.parametrize("nonparametric", [False, True])
def test_merge_proposals_match_reference_loop(nonparametric):
    g = with_isolated_vertex(small_graph(4, 150))
    og = O.OGraph.from_csr(g)
    assign = start(g, 60, 4)
    ob = O.OBlockmodel.build(og, 60, assign, nonparametric=nonparametric)
    s, dS = ob.merge_propose(seed=4, epoch=9)
    s2, dS2 = R.merge_propose(R.State(g, 60, assign, nonparametric), og, 4, 9)
    assert np.array_equal(s, s2) and np.array_equal(dS, dS2)


def carry_case(approximate, seed, mutate=(), B=60):
    g = small_graph(seed, 150)
    og = O.OGraph.from_csr(g)
    assign = start(g, B, seed)
    ob = O.OBlockmodel.build(og, B, assign)
    s, dS = ob.merge_propose(seed=seed, epoch=9)
    out = O.OBlockmodel.build(og, B, assign).merge_apply(s, dS, B // 2, approximate)
    rs = R.State(g, B, assign)
    if approximate:
        R.carry_out_best_merges(rs, dS, s, mutate)
    else:
        R.carry_out_best_merges_advanced(rs, og, dS, s, mutate)
    return out.B == rs.B and same_state(out, rs)


@pytest.mark.parametrize("approximate", [False, True])
@pytest.mark.parametrize("seed", [0, 1])
def test_carry_out_matches_reference_loop(approximate, seed):
    assert carry_case(approximate, seed)


@pytest.mark.parametrize("mistake,approximate", [("no_requeue", False), ("descending_labels", False),
                                                 ("unstable_order", True), ("translate_once", True),
                                                 ("descending_labels", True)])
def test_carry_out_mistakes_are_caught(mistake, approximate):
    assert not carry_case(approximate, 0, {mistake})


# ---------------------------------------------------------------------------------------------- driver

def sbp_case(alg, seed, mutate=(), approximate=False, batches=1, nonparametric=False, N=200, B=8):
    g = dcsbm(N, B, gamma=2.5, kmin=1, kmax=40, mean_degree=10, p_in=0.9, seed=seed)
    og = O.OGraph.from_csr(g)
    prm = O.params(alg, seed=seed, approximate=approximate, batches=batches, nonparametric=nonparametric)
    a, nb, st = O.sbp(og, prm)
    rs, tr = R.stochastic_block_partition(g, og, prm, mutate, max_outer=60)
    same = (not tr["stopped"] and nb == rs.B and np.array_equal(a, np.asarray(rs.assign, np.int32)) and
            st.mdl == rs.overall_entropy and st.merge_phases == tr["merge_phases"] and
            st.merge_proposals == tr["merge_proposals"] and st.sweeps == tr["sweeps"])
    return same, nb, tr


@pytest.mark.parametrize("alg,approximate,batches", [("async_gibbs", False, 1), ("metropolis_hastings", False, 1),
                                                     ("hybrid_mcmc", False, 1), ("async_gibbs", True, 2)])
def test_driver_matches_reference_loop(alg, approximate, batches):
    same, nb, tr = sbp_case(alg, 0, approximate=approximate, batches=batches)
    assert same
    assert nb > 1 and len(tr["block_counts"]) > 5  # the search brackets and narrows


def test_driver_nonparametric_matches_reference_loop():
    same, _, _ = sbp_case("async_gibbs", 1, nonparametric=True, N=100, B=6)
    assert same


@pytest.mark.parametrize("mistake", ["golden_index_gt", "golden_ratio_0382"])
def test_golden_section_mistakes_are_caught(mistake):
    same, _, _ = sbp_case("async_gibbs", 0, {mistake})
    assert not same
