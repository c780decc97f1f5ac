"""Directed graphs in CSR form, the reference's TSV format, and a DCSBM generator.

The CSR keeps each vertex's neighbours in file (insertion) order, as the
reference's ``Graph`` does with ``utils::insert`` (src/graph.cpp:185-207):
proposals pick a neighbour by its position in that order.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class CSRGraph:
    N: int
    out_off: np.ndarray  # int64 [N+1]
    out_adj: np.ndarray  # int32 [E], file order per source
    in_off: np.ndarray   # int64 [N+1]
    in_adj: np.ndarray   # int32 [E], file order per target
    truth: np.ndarray = None  # int32 [N] planted partition, if known

    @property
    def E(self):
        return int(self.out_off[-1])

    @classmethod
    def from_edges(cls, src, dst, N=None, truth=None):
        """Builds out/in CSR from an edge list in file order (Graph::load_text semantics)."""
        src = np.asarray(src, dtype=np.int64)
        dst = np.asarray(dst, dtype=np.int64)
        if N is None:
            N = int(max(src.max(initial=-1), dst.max(initial=-1)) + 1)
        order_out = np.argsort(src, kind="stable")
        order_in = np.argsort(dst, kind="stable")
        out_off = np.zeros(N + 1, np.int64)
        in_off = np.zeros(N + 1, np.int64)
        np.cumsum(np.bincount(src, minlength=N), out=out_off[1:])
        np.cumsum(np.bincount(dst, minlength=N), out=in_off[1:])
        return cls(N, out_off, dst[order_out].astype(np.int32), in_off, src[order_in].astype(np.int32),
                   None if truth is None else np.asarray(truth, np.int32))

    def edges(self):
        src = np.repeat(np.arange(self.N, dtype=np.int64), np.diff(self.out_off))
        return src, self.out_adj.astype(np.int64)

    def degrees(self):
        """Graph::degrees (src/graph.cpp:24-31): out + in - self-loop flag."""
        src, dst = self.edges()
        self_flag = np.zeros(self.N, np.int64)
        self_flag[src[src == dst]] = 1
        return np.diff(self.out_off) + np.diff(self.in_off) - self_flag


def load_tsv(base_path):
    """Reads ``<base>.tsv`` (1-indexed "from to" lines) and ``<base>_truePartition.tsv`` if present."""
    data = np.loadtxt(base_path + ".tsv", dtype=np.int64, ndmin=2)
    src, dst = data[:, 0] - 1, data[:, 1] - 1
    truth = None
    try:
        t = np.loadtxt(base_path + "_truePartition.tsv", dtype=np.int64, ndmin=2)
        N = int(max(src.max(), dst.max(), t[:, 0].max() - 1) + 1)
        truth = np.full(N, -1, np.int32)
        truth[t[:, 0] - 1] = t[:, 1] - 1
    except OSError:
        N = None
    return CSRGraph.from_edges(src, dst, N, truth)


def save_tsv(g, base_path):
    src, dst = g.edges()
    np.savetxt(base_path + ".tsv", np.stack([src + 1, dst + 1], 1), fmt="%d", delimiter="\t")
    if g.truth is not None:
        np.savetxt(base_path + "_truePartition.tsv",
                   np.stack([np.arange(g.N) + 1, g.truth.astype(np.int64) + 1], 1), fmt="%d", delimiter="\t")


def _power_law_pmf(kmin, kmax, gamma, shift):
    k = np.arange(kmin, kmax + 1, dtype=np.float64)
    w = (k + shift) ** (-gamma)
    return k, w / w.sum()


def degree_distribution(kmin, kmax, gamma, mean):
    """Discrete power law p(k) ~ (k + c)^-gamma on [kmin, kmax], c solved so the mean matches.

    When the target mean is below what c = 0 gives, c is negative (down to -kmin + 1e-6);
    when it cannot be reached, the closest achievable mean is used.
    """
    lo, hi = -kmin + 1e-6, float(kmax) * 50
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        k, p = _power_law_pmf(kmin, kmax, gamma, mid)
        if (k * p).sum() < mean:
            lo = mid
        else:
            hi = mid
    return _power_law_pmf(kmin, kmax, gamma, 0.5 * (lo + hi))


def dcsbm(N, B, *, gamma=2.5, kmin=1, kmax=100, mean_degree=10.0, p_in=0.75, sigma=None, seed=0):
    """Directed, unweighted degree-corrected SBM with planted blocks.

    Block sizes are equal (sigma None) or lognormal(0, sigma) scaled to sum to N
    with every block non-empty.  Out-degrees follow the clipped power law above;
    each out-edge stays in the source's block with probability p_in, otherwise it
    goes to a uniformly chosen other block; the target vertex inside the block is
    drawn proportionally to an independent power-law in-propensity.  Duplicate
    edges are redrawn; self-loops are allowed.  Returns a CSRGraph whose edges are
    in generation order (grouped by source) and whose ``truth`` is the planting.
    """
    rng = np.random.Generator(np.random.Philox(seed))
    if sigma is None:
        sizes = np.full(B, N // B, np.int64)
        sizes[: N % B] += 1
    else:
        w = rng.lognormal(0.0, sigma, B)
        sizes = np.maximum(1, np.floor(w / w.sum() * (N - B)).astype(np.int64) + 1)
        diff = N - sizes.sum()
        order = np.argsort(-w)
        i = 0
        while diff != 0:
            b = order[i % B]
            if diff > 0:
                sizes[b] += 1
                diff -= 1
            elif sizes[b] > 1:
                sizes[b] -= 1
                diff += 1
            i += 1
    truth = np.repeat(np.arange(B, dtype=np.int32), sizes)
    block_start = np.concatenate([[0], np.cumsum(sizes)])
    kvals, kp = degree_distribution(kmin, kmax, gamma, mean_degree)
    deg = kvals[np.searchsorted(np.cumsum(kp), rng.random(N), side="right").clip(0, len(kp) - 1)].astype(np.int64)
    deg = np.minimum(deg, N)  # a vertex cannot have more distinct targets than vertices
    theta = kvals[np.searchsorted(np.cumsum(kp), rng.random(N), side="right").clip(0, len(kp) - 1)]
    # per-block CDF of theta for within-block target choice
    theta_cdf = np.empty(N)
    for b in range(B):
        s, e = block_start[b], block_start[b + 1]
        c = np.cumsum(theta[s:e])
        theta_cdf[s:e] = c / c[-1]

    def draw_targets(src):
        own = truth[src]
        stay = rng.random(len(src)) < p_in
        other = rng.integers(0, max(B - 1, 1), len(src))
        other = other + (other >= own)
        tb = np.where(stay | (B == 1), own, other)
        u = rng.random(len(src))
        # within the target block: the first vertex whose CDF value is >= u (the CDF is non-decreasing),
        # clipped to the block's last vertex; one search per block over the edges grouped by block
        order = np.argsort(tb.astype(np.int16) if B < 32768 else tb, kind="stable")
        cnt = np.bincount(tb, minlength=B)
        ends = np.cumsum(cnt)
        out = np.empty(len(src), np.int64)
        for b in np.flatnonzero(cnt):
            sel = order[ends[b] - cnt[b]:ends[b]]
            s, e = block_start[b], block_start[b + 1]
            out[sel] = s + np.minimum(np.searchsorted(theta_cdf[s:e], u[sel], side="left"), e - s - 1)
        return out

    src = np.repeat(np.arange(N, dtype=np.int64), deg)
    src_off = np.concatenate([[0], np.cumsum(deg)])
    slot = np.arange(len(src), dtype=np.int64)
    dst = draw_targets(src)
    # Each pass redraws every edge that repeats an earlier (lower-index) edge of the same source.  A
    # duplicate can only involve edges of one source, so after the first pass only the edges of sources
    # that had a redraw are checked again; the redrawn set, and so the draws, are those of a check over
    # all edges.
    cand = slot
    for _ in range(64):
        key = src[cand] * N + dst[cand]
        _, first = np.unique(key, return_index=True)
        dup = np.ones(len(key), bool)
        dup[first] = False
        if not dup.any():
            break
        redraw = cand[dup]
        dst[redraw] = draw_targets(src[redraw])
        touched = np.unique(src[redraw])
        lens = deg[touched]
        starts = src_off[touched]
        cand = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens) + np.arange(lens.sum())
    else:
        key = src * N + dst
        _, first = np.unique(key, return_index=True)
        keep = np.zeros(len(key), bool)
        keep[first] = True
        src, dst, slot = src[keep], dst[keep], slot[keep]
    order = np.argsort(slot, kind="stable")
    return CSRGraph.from_edges(src[order], dst[order], N, truth)
