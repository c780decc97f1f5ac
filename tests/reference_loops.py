"""A literal Python restatement of the reference's sequential loops, used to pin the oracle's versions of them.

The reference's control flow is restated here line by line, on a dense block matrix kept in Python, under the
replay contract (Philox draws keyed by (seed; unit, epoch, tag, sub), multinomials over row cells by ascending
column then column cells by ascending row, keyed shuffles, heap ties to the lower block id):

- vertex proposals, common::propose_new_block in vertex mode (src/common.cpp:138-198, choose_neighbor :14-33,
  propose_random_block :200-209) with the neighbour lists of edge_weights (src/finetune.cpp:419-430);
- merge proposals, block_merge::merge_blocks' loop and propose_merge_sparse (src/block_merge.cpp:189-217,251-285)
  with propose_new_block in merge mode;
- finetune::asynchronous_gibbs (src/finetune.cpp:34-107) with Blockmodel::move_vertex(const VertexMove_v3&)
  (src/blockmodel/blockmodel.cpp:437-494);
- finetune::metropolis_hastings, propose_move and move_vertex (src/finetune.cpp:773-869) with
  Blockmodel::move_vertex(Vertex, const Delta&, ...) (src/blockmodel/blockmodel.cpp:408-435);
- finetune::hybrid_mcmc (src/finetune.cpp:563-646) with Graph::sort_vertices (src/graph.cpp:235-268);
- finetune::finetune_assignment (src/finetune.cpp:903-932);
- finetune::early_stop, both forms (src/finetune.cpp:265-294), and finetune::mcmc's dispatch (:754-771);
- carry_out_best_merges_advanced with its O(B) relabel of block_map (src/block_merge.cpp:40-117),
  Blockmodel::carry_out_best_merges (src/blockmodel/blockmodel.cpp:55-122) and Blockmodel::build_mapping
  (:34-46);
- BlockmodelTriplet::{update, get_next_blockmodel, is_done} (src/blockmodel/blockmodel_triplet.cpp:4-105) and
  sbp::stochastic_block_partition's loop (src/sbp.cpp:16-86).

The numbers the loops compare -- delta-MDL of a vertex move or a merge, the Hastings correction and the MDL --
come from the oracle evaluated on a blockmodel rebuilt from this restatement's current assignment.  Those
functions are pinned on their own (the reference's toy fixtures, brute-force MDL differences with glibc logf,
the per-cell form); what this module pins is everything around them: draws, proposals, visit and commit
orders, refusals, skips, batch bounds, early stops, the heap and its re-queue rule, relabelling, label
compaction and the golden-section search.

Every loop takes `mutate`, a set of names of deliberate mistakes, so the tests can show that the comparison
would catch each of them.
"""
import heapq
import math

import numpy as np

import oracle as O

TAG_FT = 0x46540001
TAG_MERGE = 0x4D470001
TAG_SHUFFLE = 0x53480001
DBL_MAX = np.finfo(np.float64).max
MAX_NUM_ITERATIONS = 100  # include/finetune.hpp:34
BLOCK_REDUCTION_RATE = 0.5  # include/blockmodel/blockmodel.hpp:24
NUM_AGG_PROPOSALS_PER_BLOCK = 10  # include/block_merge.hpp:23


def draws(seed, unit, epoch, tag, sub=0):
    return [int(x) for x in O.philox(seed, unit, epoch, tag, sub)]


def uniform_below(w, n):
    return (w * n) >> 32


def unit_double(a, b):
    return (((a << 32) | b) >> 11) * 2.0 ** -53


def random_block(current, B, w):
    """common::propose_random_block (src/common.cpp:200-209): U[0, B-2], shifted past the current block."""
    if B < 2:
        return current
    p = uniform_below(w, B - 1)
    return p + 1 if p >= current else p


def keyed_shuffle(items, seed, epoch):
    """The replay stream's std::shuffle: ascending Philox key, ties by vertex id."""
    return sorted(items, key=lambda v: (draws(seed, v, epoch, TAG_SHUFFLE)[0], v))


class State:
    """A blockmodel as the reference keeps it (include/blockmodel/blockmodel.hpp:46-219), with a dense matrix."""

    def __init__(self, g, B, assignment, nonparametric=False):
        self.g = g
        self.np = bool(nonparametric)  # args.nonparametric: entropy::nonparametric's description length
        self.N = g.N
        self.B = int(B)
        self.assign = [int(x) for x in assignment]
        self.out = [list(map(int, g.out_adj[g.out_off[v]:g.out_off[v + 1]])) for v in range(g.N)]
        self.inn = [list(map(int, g.in_adj[g.in_off[v]:g.in_off[v + 1]])) for v in range(g.N)]
        self.overall_entropy = DBL_MAX
        self.num_to_merge = int(self.B * BLOCK_REDUCTION_RATE)  # Blockmodel ctor (blockmodel.hpp:70-76)
        self.initialize_edge_counts()

    def initialize_edge_counts(self):
        """Blockmodel::initialize_edge_counts (src/blockmodel/blockmodel.cpp:193-276)."""
        B = self.B
        M = np.zeros((B, B), np.int64)
        for v in range(self.N):
            for u in self.out[v]:
                M[self.assign[v], self.assign[u]] += 1
        self.M = M
        self.d_out = M.sum(axis=1)
        self.d_in = M.sum(axis=0)
        self.d_tot = self.d_out + self.d_in - np.diag(M)
        self.size = np.bincount(np.asarray(self.assign, np.int64), minlength=B)[:B].astype(np.int64)
        self._ob = None

    def copy(self):
        """Blockmodel::copy (src/blockmodel/blockmodel.cpp:139-153)."""
        c = State.__new__(State)
        c.__dict__.update(self.__dict__)
        c.assign = list(self.assign)
        c.M, c.d_out, c.d_in, c.d_tot, c.size = (x.copy() for x in (self.M, self.d_out, self.d_in, self.d_tot,
                                                                     self.size))
        c._ob = None
        return c

    def changed(self):
        self._ob = None

    def oracle_bm(self, og):
        if self._ob is None:
            self._ob = O.OBlockmodel.build(og, self.B, np.asarray(self.assign, np.int32), nonparametric=self.np)
        return self._ob

    # ---- vertex neighbour lists (edge_weights, src/finetune.cpp:419-430): in-edges without the self-loop
    def out_edges(self, v):
        return self.out[v]

    def in_edges(self, v):
        return [u for u in self.inn[v] if u != v]

    def row_cells(self, b):
        return [(int(c), int(self.M[b, c])) for c in np.nonzero(self.M[b])[0]]

    def col_cells(self, b):
        return [(int(x), int(self.M[x, b])) for x in np.nonzero(self.M[:, b])[0]]


def multinomial_pick(cells, target):
    prefix = 0
    for key, cnt in cells:
        prefix += cnt
        if target < prefix:
            return key
    raise AssertionError("multinomial target past its cells")


def propose_vertex(st, v, w):
    """propose_new_block, vertex mode (src/common.cpp:138-198): returns (s, k_out, k_in)."""
    r = st.assign[v]
    outs, ins = st.out_edges(v), st.in_edges(v)
    k_out, k_in = len(outs), len(ins)
    k = k_out + k_in
    if k == 0:
        return random_block(r, st.B, w[1]), k_out, k_in
    u = (outs + ins)[uniform_below(w[0], k)]
    t = st.assign[u]
    # neighbors_weights(t) (dict_transpose_matrix.cpp:176-188), the current block not excluded
    cells = st.row_cells(t) + [(x, c) for x, c in st.col_cells(t) if x != t]
    total = sum(c for _, c in cells)
    return multinomial_pick(cells, uniform_below(w[1], total)), k_out, k_in


def propose_merge(st, r, w):
    """propose_new_block, merge mode (src/common.cpp:138-198): the neighbour block t is drawn from row r and column
    r; the multinomial over neighbors_weights(t) has edges[r] = 0 (:176-177); an empty one proposes a random block."""
    cells_r = st.row_cells(r) + st.col_cells(r)
    k = sum(c for _, c in cells_r)
    if k == 0:
        return random_block(r, st.B, w[1])
    t = multinomial_pick(cells_r, uniform_below(w[0], k))
    cells = [(c, n) for c, n in st.row_cells(t) if c != r] + [(x, n) for x, n in st.col_cells(t) if x not in (t, r)]
    total = sum(n for _, n in cells)
    if total == 0:
        return random_block(r, st.B, w[1])
    return multinomial_pick(cells, uniform_below(w[1], total))


def accept(dS, H, greedy, w):
    """finetune::accept (src/finetune.cpp:22-32)."""
    if greedy:
        return dS < 0.0
    p = math.exp(-3.0 * dS) * H
    p = 1.0 if p >= 1.0 else p
    return unit_double(w[2], w[3]) <= p


def evaluate(st, og, v, seed, epoch, greedy):
    """propose_gibbs_move_v3 / propose_move's proposal and evaluation: (s, dS, accepted, k_out, k_in)."""
    w = draws(seed, v, epoch, TAG_FT)
    r = st.assign[v]
    s, k_out, k_in = propose_vertex(st, v, w)
    if s == r:
        return s, 0.0, False, k_out, k_in
    ob = st.oracle_bm(og)
    dS = ob.vertex_delta_mdl(v, s)
    # hastings_correction is 1 for greedy or nonparametric runs (src/entropy.cpp:497-499)
    H = 1.0 if (greedy or st.np) else ob.vertex_hastings(v, s)
    return s, dS, accept(dS, H, greedy, w), k_out, k_in


def move_vertex_v3(st, v, s, mutate=()):
    """Blockmodel::move_vertex(const VertexMove_v3&) (src/blockmodel/blockmodel.cpp:437-494): refused when the
    current block has one vertex; the cells follow the neighbours' current blocks."""
    r = st.assign[v]
    limit = 2 if "refuse_size_2" in mutate else 1
    if "no_refusal" not in mutate and st.size[r] <= limit:
        return False
    for u in st.out_edges(v):
        ob = st.assign[u]
        st.M[r, ob] -= 1
        st.d_out[r] -= 1
        st.d_out[s] += 1
        if u == v:
            st.M[s, s] += 1
            st.d_in[r] -= 1
            st.d_in[s] += 1
        else:
            st.M[s, ob] += 1
    for u in st.in_edges(v):
        ib = st.assign[u]
        st.M[ib, r] -= 1
        st.d_in[r] -= 1
        st.M[ib, s] += 1
        st.d_in[s] += 1
    st.d_tot[r] = st.d_in[r] + st.d_out[r] - st.M[r, r]
    st.d_tot[s] = st.d_in[s] + st.d_out[s] - st.M[s, s]
    st.assign[v] = s
    st.size[r] -= 1
    st.size[s] += 1
    st.changed()
    return True


def move_vertex_delta(st, v, s):
    """Blockmodel::move_vertex(Vertex, const Delta&, ...) (src/blockmodel/blockmodel.cpp:408-435): the matrix after
    update_edge_counts(delta) is the matrix of the new assignment, and so are the degrees the function sets."""
    st.assign[v] = s
    st.initialize_edge_counts()


# ---------------------------------------------------------------------------------------------- finetune sweeps

def async_gibbs_sweep(st, og, prm, epoch, items=None, mutate=()):
    """One iteration's body of finetune::asynchronous_gibbs (src/finetune.cpp:50-86), over `items` (default every
    vertex; hybrid_mcmc passes its low-degree list).  Returns (moves committed, moves refused)."""
    if items is None:
        items = list(range(st.N))
        if not prm.ordered:
            items = keyed_shuffle(items, prm.seed, epoch)
    n = len(items)
    batch_size = int(math.ceil(n / prm.batches))
    nbatches = n // batch_size if batch_size else 0
    if "process_tail" in mutate and batch_size and n % batch_size:
        nbatches += 1
    moved = refused = 0
    for batch in range(nbatches):
        start, end = batch * batch_size, min(n, (batch + 1) * batch_size)
        moves = {}
        for index in range(start, end):  # against the frozen blockmodel
            v = items[index]
            s, dS, ok, _, _ = evaluate(st, og, v, prm.seed, epoch, prm.greedy)
            if ok:
                moves[v] = s
        order = sorted(moves) if "commit_in_visit_order" not in mutate else list(moves)
        for v in order:  # for (const VertexMove_v3 &move : moves): vertex-id order
            if move_vertex_v3(st, v, moves[v], mutate):
                moved += 1
            else:
                refused += 1
    return moved, refused


def mh_walk(st, og, prm, epoch, order, mutate=()):
    """The sequential walk of metropolis_hastings / hybrid_mcmc / finetune_assignment: propose_move skips a vertex
    whose block has one vertex (src/finetune.cpp:845-847) and commits an accepted move at once.  Returns (sum of
    accepted delta-MDL in visit order, moves)."""
    total, moves = 0.0, 0
    for v in order:
        r = st.assign[v]
        if "no_singleton_skip" not in mutate and st.size[r] == 1:
            continue
        s, dS, ok, _, _ = evaluate(st, og, v, prm.seed, epoch, prm.greedy)
        if ok:
            move_vertex_delta(st, v, s)
            total += dS
            moves += 1
    return total, moves


def mh_sweep(st, og, prm, epoch, ordered, mutate=()):
    order = list(range(st.N))
    if not ordered:
        order = keyed_shuffle(order, prm.seed, epoch)
    return mh_walk(st, og, prm, epoch, order, mutate)


def early_stop4(iteration, golden_ratio_not_reached, initial, de, mutate=()):
    """finetune::early_stop, 4-argument form (src/finetune.cpp:265-283)."""
    if de[-1] == 0.0:
        return True
    if iteration < 3:
        return False
    average = (de[-1] + de[-2] + de[-3]) / -3.0
    hi, lo = (1e-4, 5e-4) if "swap_thresholds" in mutate else (5e-4, 1e-4)
    return average < (hi * initial if golden_ratio_not_reached else lo * initial)


def early_stop3(iteration, initial, de, mutate=()):
    """finetune::early_stop, 3-argument form (src/finetune.cpp:285-294): no zero-delta exit."""
    if "zero_exit" in mutate and de[-1] == 0.0:
        return True
    if iteration < 3:
        return False
    return (de[-1] + de[-2] + de[-3]) / -3.0 < 1e-4 * initial


def mdl(st, og):
    return st.oracle_bm(og).mdl()


def sort_vertices(st, mh_percent, mutate=()):
    """Graph::sort_vertices (src/graph.cpp:235-268): stable descending degree (out + in - self-loop flag); the first
    indices with index < mh_percent * N are the high-degree vertices."""
    deg = [len(st.out[v]) + len(st.inn[v]) - (1 if v in st.out[v] else 0) for v in range(st.N)]
    idx = sorted(range(st.N), key=lambda v: -deg[v])  # Python's sort is stable
    cut = [i for i in range(st.N) if (i <= mh_percent * st.N if "split_le" in mutate else i < mh_percent * st.N)]
    nh = len(cut)
    return idx[:nh], idx[nh:]


def mcmc(st, og, prm, golden_ratio_not_reached, epoch, mutate=(), trace=None):
    """finetune::mcmc (src/finetune.cpp:754-771) with asynchronous_gibbs (:34-107), hybrid_mcmc (:563-646) or
    metropolis_hastings (:773-820).  prm.algorithm: 0 async_gibbs, 1 metropolis_hastings, 2 hybrid_mcmc.
    Returns the next epoch (one per sweep)."""
    if st.B == 1:
        return epoch
    de = []
    if prm.algorithm in (0, 2):
        st.overall_entropy = mdl(st, og)
        initial = last = st.overall_entropy
        if prm.algorithm == 2:
            hdv, ldv = sort_vertices(st, prm.mh_percent, mutate)
        for it in range(prm.max_iterations or MAX_NUM_ITERATIONS):
            ep = epoch
            epoch += 1
            if prm.algorithm == 0:
                async_gibbs_sweep(st, og, prm, ep, None, mutate)
            else:
                h, l = (hdv, ldv) if prm.ordered else (keyed_shuffle(hdv, prm.seed, ep), keyed_shuffle(ldv, prm.seed, ep))
                if "hybrid_swap_lists" in mutate:
                    h, l = l, h
                mh_walk(st, og, prm, ep, h, mutate)
                async_gibbs_sweep(st, og, prm, ep, l, mutate)
            if trace is not None:
                trace["sweeps"] += 1
            e = mdl(st, og)
            de.append(e - last)
            last = e
            if early_stop4(it, golden_ratio_not_reached, initial, de, mutate):
                break
    else:
        st.overall_entropy = mdl(st, og)
        for it in range(prm.max_iterations or MAX_NUM_ITERATIONS):
            ep = epoch
            epoch += 1
            total, _ = mh_sweep(st, og, prm, ep, prm.ordered, mutate)
            if trace is not None:
                trace["sweeps"] += 1
            de.append(total)
            if early_stop4(it, golden_ratio_not_reached, st.overall_entropy, de, mutate):
                break
    st.overall_entropy = mdl(st, og)
    return epoch


def finetune_assignment(st, og, prm, epoch, mutate=()):
    """finetune::finetune_assignment (src/finetune.cpp:903-932): ordered MH walks, 3-argument early stop."""
    st.overall_entropy = mdl(st, og)
    de = []
    for it in range(prm.max_iterations or MAX_NUM_ITERATIONS):
        ep = epoch
        epoch += 1
        total, _ = mh_sweep(st, og, prm, ep, True, mutate)
        de.append(total)
        if early_stop3(it, st.overall_entropy, de, mutate):
            break
    st.overall_entropy = mdl(st, og)
    return epoch


# ---------------------------------------------------------------------------------------------- block merge

def merge_propose(st, og, seed, epoch):
    """block_merge::merge_blocks' proposal loop (src/block_merge.cpp:189-217) with propose_merge_sparse
    (:251-285): ten proposals per block, repeats skipped, the lowest delta-MDL kept (first on ties)."""
    ob = st.oracle_bm(og)
    best_s = np.full(st.B, -1, np.int32)
    best_dS = np.full(st.B, DBL_MAX)
    for r in range(st.B):
        past = set()
        for i in range(NUM_AGG_PROPOSALS_PER_BLOCK):
            s = propose_merge(st, r, draws(seed, r, epoch, TAG_MERGE, i))
            dS = DBL_MAX
            if s not in past:
                dS = ob.merge_delta_mdl(r, s)
                past.add(s)
            if dS < best_dS[r]:
                best_dS[r] = dS
                best_s[r] = s
    return best_s, best_dS


def build_mapping(values, mutate=()):
    """Blockmodel::build_mapping (src/blockmodel/blockmodel.cpp:34-46): labels in ascending order (std::map)."""
    keys = sorted(set(values), reverse=("descending_labels" in mutate))
    mapping = [-1] * len(values)
    for i, k in enumerate(keys):
        mapping[k] = i
    return mapping


def merge_vertices(st, frm, to):
    """Blockmodel::merge_block (src/blockmodel/blockmodel.cpp:337-363): every vertex of `frm` moves to `to`; the
    matrix and degrees after update_edge_counts(delta) are those of the new assignment."""
    st.assign = [to if b == frm else b for b in st.assign]
    st.initialize_edge_counts()


def carry_out_best_merges_advanced(st, og, best_dS, best_s, mutate=()):
    """src/block_merge.cpp:40-117: a min-heap of (delta-MDL, block) (ties to the lower block id); each popped merge
    is re-evaluated on the current blockmodel and pushed back when it got worse than the top; block_map is
    relabelled in O(B) per merge; labels are compacted after the loop."""
    heap = [(float(best_dS[i]), i, int(best_s[i])) for i in range(st.B)]
    heapq.heapify(heap)
    num_merged = 0
    block_map = list(range(st.B))
    while num_merged < st.num_to_merge and heap:
        dS, frm, newblock = heapq.heappop(heap)
        to = block_map[newblock]
        if frm == to:
            continue
        actual = st.oracle_bm(og).merge_delta_mdl(frm, to)
        if "no_requeue" not in mutate and heap and actual > heap[0][0]:
            heapq.heappush(heap, (actual, frm, newblock))
            continue
        for i in range(len(block_map)):
            if block_map[i] == frm:
                block_map[i] = to
        merge_vertices(st, frm, to)
        num_merged += 1
    mapping = build_mapping(st.assign, mutate)
    st.assign = [mapping[b] for b in st.assign]
    st.B = st.B - st.num_to_merge


def carry_out_best_merges(st, best_dS, best_s, mutate=()):
    """Blockmodel::carry_out_best_merges (src/blockmodel/blockmodel.cpp:55-122): stable ascending order of the
    proposals' delta-MDL, targets translated through the merge chains, no re-evaluation."""
    if "unstable_order" in mutate:
        best_merges = sorted(range(st.B), key=lambda i: (best_dS[i], -i))
    else:
        best_merges = sorted(range(st.B), key=lambda i: best_dS[i])  # utils::partial_sort_indices: stable_sort
    block_map = list(range(st.B))

    def translate(block):
        b = block
        while True:
            b = block_map[b]
            if block_map[b] == b or "translate_once" in mutate:
                return b

    num_merged = counter = 0
    while num_merged < st.num_to_merge:
        frm = best_merges[counter]
        to = translate(int(best_s[frm]))
        counter += 1
        if to != frm:
            block_map[frm] = to
            num_merged += 1
    st.assign = [translate(b) for b in st.assign]
    mapping = build_mapping(st.assign, mutate)
    st.assign = [mapping[b] for b in st.assign]
    st.B -= st.num_to_merge


def merge_blocks(st, og, seed, epoch, approximate, mutate=()):
    """block_merge::merge_blocks (src/block_merge.cpp:178-228); modifies st in place, as the reference does."""
    best_s, best_dS = merge_propose(st, og, seed, epoch)
    if approximate:
        carry_out_best_merges(st, best_dS, best_s, mutate)
    else:
        carry_out_best_merges_advanced(st, og, best_dS, best_s, mutate)
    st.initialize_edge_counts()
    return st


# ---------------------------------------------------------------------------------------------- driver

class Triplet:
    """BlockmodelTriplet (src/blockmodel/blockmodel_triplet.cpp:4-105)."""

    def __init__(self, mutate=()):
        self.bm = [None, None, None]
        self.optimal_num_blocks_found = False
        self.mutate = mutate

    def golden_ratio_not_reached(self):
        return self.bm[2] is None

    def update(self, b):
        if self.bm[1] is None:
            index = 1
        elif b.overall_entropy <= self.bm[1].overall_entropy:
            old_index = 0 if self.bm[1].B > b.B else 2
            self.bm[old_index] = self.bm[1].copy()
            index = 1
        else:
            index = 2 if self.bm[1].B > b.B else 0
        self.bm[index] = b.copy()

    def is_done(self):
        if ((self.bm[0] is not None and self.bm[0].B - self.bm[2].B == 2) or
                (self.bm[0] is None and self.bm[1].B - self.bm[2].B == 1)):
            self.optimal_num_blocks_found = True
        return self.optimal_num_blocks_found

    def get_next_blockmodel(self, old):
        old.num_to_merge = 0
        self.update(old)
        if self.golden_ratio_not_reached():
            b = self.bm[1].copy()
            b.num_to_merge = int(b.B * BLOCK_REDUCTION_RATE)
            if b.num_to_merge == 0:
                self.optimal_num_blocks_found = True
            return b
        if self.is_done():
            return self.bm[1].copy()
        if self.bm[0] is None and self.bm[1].B > self.bm[2].B:
            index = 1
        else:
            upper = (self.bm[0].B if self.bm[0] is not None else 0) - self.bm[1].B
            lower = self.bm[1].B - self.bm[2].B
            if "golden_index_gt" in self.mutate:
                index = 0 if upper > lower else 1
            else:
                index = 0 if upper >= lower else 1
        ratio = 0.382 if "golden_ratio_0382" in self.mutate else 0.618
        nxt = self.bm[index + 1].B + int((self.bm[index].B - self.bm[index + 1].B) * ratio)
        b = self.bm[index].copy()
        b.num_to_merge = b.B - nxt
        return b


def stochastic_block_partition(g, og, prm, mutate=(), max_outer=None):
    """sbp::stochastic_block_partition (src/sbp.cpp:16-72): merges, then finetune, then the triplet's next
    blockmodel, until the optimal block count is found.  Epochs: one per merge phase, one per sweep; async_gibbs
    runs Metropolis-Hastings from iteration `asynciterations` on (finetune::mcmc, src/finetune.cpp:762-769).
    Returns the final blockmodel and the search's trace: merge phases, merge proposals (ten per block) and
    finetune sweeps.  max_outer bounds the outer iterations (a mistaken search step can stall the search); a
    search stopped by it has trace["stopped"] set."""
    st = State(g, g.N, range(g.N), prm.nonparametric)
    tri = Triplet(mutate)
    trace = {"merge_phases": 0, "merge_proposals": 0, "sweeps": 0, "block_counts": [], "stopped": False}
    epoch = 0
    iteration = 0
    while not tri.optimal_num_blocks_found:
        if max_outer is not None and iteration >= max_outer:
            trace["stopped"] = True
            break
        trace["merge_phases"] += 1
        trace["merge_proposals"] += NUM_AGG_PROPOSALS_PER_BLOCK * st.B
        merge_blocks(st, og, prm.seed, epoch, prm.approximate, mutate)
        trace["block_counts"].append(st.B)
        epoch += 1
        alg = prm.algorithm
        if alg == 0 and iteration >= prm.asynciterations:
            alg = 1
        q = O.params(["async_gibbs", "metropolis_hastings", "hybrid_mcmc"][alg], batches=prm.batches,
                     greedy=bool(prm.greedy), ordered=bool(prm.ordered), mh_percent=prm.mh_percent, seed=prm.seed,
                     max_iterations=prm.max_iterations, nonparametric=bool(prm.nonparametric))
        epoch = mcmc(st, og, q, tri.golden_ratio_not_reached(), epoch, mutate, trace)
        iteration += 1
        st = tri.get_next_blockmodel(st)
    return st, trace
