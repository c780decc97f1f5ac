"""ctypes binding of the CPU oracle (test infrastructure only).

Only the tests, the smoke check and the CPU-baseline leg of the benchmark load
this module; the product library never does.  The shared library is built by
``oracle/Makefile`` into ``oracle/_build/libsbp_oracle.so``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libsbp_oracle.so")

ALG = {"async_gibbs": 0, "metropolis_hastings": 1, "hybrid_mcmc": 2}


class Graph(C.Structure):
    _fields_ = [("N", C.c_int32), ("E", C.c_int64),
                ("out_off", C.c_void_p), ("out_adj", C.c_void_p),
                ("in_off", C.c_void_p), ("in_adj", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [("algorithm", C.c_int32), ("batches", C.c_int32), ("greedy", C.c_int32),
                ("approximate", C.c_int32), ("ordered", C.c_int32), ("asynciterations", C.c_int32),
                ("mh_percent", C.c_double), ("seed", C.c_uint64), ("max_iterations", C.c_int32),
                ("threads", C.c_int32), ("nonparametric", C.c_int32), ("degreecorrected", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("proposals", C.c_int64), ("moves", C.c_int64), ("refused", C.c_int64),
                ("sweeps", C.c_int64), ("merge_proposals", C.c_int64), ("merge_phases", C.c_int64),
                ("mdl", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u32, u64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double
        G, P, S = C.POINTER(Graph), C.POINTER(Params), C.POINTER(Stats)
        sig = {
            "orc_philox": (None, [u64, u32, u32, u32, u32, vp]),
            "orc_uniform_below": (u32, [u32, u32]),
            "orc_unit_double": (dbl, [u32, u32]),
            "orc_logf_range": (None, [u32, u32, vp]),
            "orc_fx_from_double": (None, [dbl, vp]),
            "orc_fx_to_double": (dbl, [i64, u64]),
            "orc_bm_new": (vp, [G, i32, vp]),
            "orc_bm_copy": (vp, [vp]),
            "orc_bm_free": (None, [vp]),
            "orc_bm_num_blocks": (i32, [vp]),
            "orc_bm_nnz": (i64, [vp]),
            "orc_bm_export": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "orc_bm_validate": (C.c_int, [vp, G]),
            "orc_bm_set_model": (None, [vp, vp, i32]),
            "orc_spence": (C.c_double, [C.c_double]),
            "orc_log_q": (C.c_double, [C.c_uint64, C.c_uint64]),
            "orc_np_deg_dl": (C.c_double, [vp]),
            "orc_lgamma": (dbl, [i64]),
            "orc_lbinom": (dbl, [i64, i64]),
            "orc_np_sparse_entropy": (dbl, [vp]),
            "orc_np_partition_dl": (dbl, [vp]),
            "orc_np_edges_dl": (dbl, [i64, i64]),
            "orc_np_mdl": (dbl, [vp, G]),
            "orc_log_posterior": (dbl, [vp]),
            "orc_mdl": (dbl, [vp, G]),
            "orc_null_mdl_v1": (dbl, [i64]),
            "orc_vertex_delta_mdl": (dbl, [vp, G, i32, i32]),
            "orc_vertex_hastings": (dbl, [vp, G, i32, i32]),
            "orc_merge_delta_mdl": (dbl, [vp, i32, i32]),
            "orc_vertex_delta_mdl_cells": (dbl, [vp, G, i32, i32, vp]),
            "orc_merge_delta_mdl_cells": (dbl, [vp, i32, i32, vp]),
            "orc_ft_propose": (C.c_int, [vp, G, u64, u32, i32, i32, vp, vp, vp]),
            "orc_async_gibbs_sweep": (C.c_int, [vp, G, P, u32, vp, S]),
            "orc_mh_sweep": (C.c_int, [vp, G, P, u32, i32, vp, S]),
            "orc_mcmc": (C.c_int, [vp, G, P, i32, vp, S]),
            "orc_finetune_assignment": (C.c_int, [vp, G, P, vp, S]),
            "orc_merge_propose": (C.c_int, [vp, u64, u32, vp, vp]),
            "orc_merge_apply": (vp, [vp, G, vp, vp, i64, i32]),
            "orc_merge_blocks": (vp, [vp, G, P, u32, S]),
            "orc_sbp": (C.c_int, [G, P, vp, vp, S]),
            "orc_sample": (i64, [G, i32, dbl, u64, vp, vp]),
            "orc_degree_weight_key": (C.c_float, [u32, i64]),
            "orc_subgraph": (i64, [G, vp, vp, i32, vp, vp, vp, vp]),
            "orc_extend": (C.c_int, [G, vp, vp, i32, u64, vp]),
            "orc_detach": (i64, [G, vp, vp]),
            "orc_reattach": (C.c_int, [G, vp, vp, i32, u64, vp]),
            "orc_split_communities": (i32, [G, vp, i32, i32, u64, u32, i32, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class OGraph:
    """Keeps the CSR arrays alive alongside the ctypes struct."""

    def __init__(self, out_off, out_adj, in_off, in_adj):
        self.out_off = np.ascontiguousarray(out_off, dtype=np.int64)
        self.out_adj = np.ascontiguousarray(out_adj, dtype=np.int32)
        self.in_off = np.ascontiguousarray(in_off, dtype=np.int64)
        self.in_adj = np.ascontiguousarray(in_adj, dtype=np.int32)
        self.N = len(self.out_off) - 1
        self.E = int(self.out_off[-1])
        self.s = Graph(self.N, self.E, _p(self.out_off).value, _p(self.out_adj).value,
                       _p(self.in_off).value, _p(self.in_adj).value)

    @classmethod
    def from_csr(cls, g):
        return cls(g.out_off, g.out_adj, g.in_off, g.in_adj)

    def ref(self):
        return C.byref(self.s)


def params(algorithm="async_gibbs", batches=1, greedy=True, approximate=False, ordered=False,
           asynciterations=2 ** 31 - 1, mh_percent=0.075, seed=0, max_iterations=100, threads=1,
           nonparametric=False, degreecorrected=False):
    return Params(ALG[algorithm] if isinstance(algorithm, str) else algorithm, batches, int(greedy),
                  int(approximate), int(ordered), asynciterations, mh_percent, seed, max_iterations, threads,
                  int(nonparametric), int(degreecorrected))


class OBlockmodel:
    def __init__(self, handle, graph):
        self.h = C.c_void_p(handle)
        self.g = graph

    @classmethod
    def build(cls, graph, B, assignment, nonparametric=False, degreecorrected=False):
        a = np.ascontiguousarray(assignment, dtype=np.int32)
        bm = cls(lib().orc_bm_new(graph.ref(), B, _p(a)), graph)
        if nonparametric:
            bm.set_model(True, degreecorrected)
        return bm

    def set_model(self, nonparametric, degreecorrected=False):
        """Entropy model used by mdl / delta-MDL: parametric (entropy::mdl), nonparametric, or nonparametric
        with degree correction (degree histograms built from the graph)."""
        lib().orc_bm_set_model(self.h, self.g.ref(), (2 if degreecorrected else 1) if nonparametric else 0)

    def np_parts(self):
        """(sparse_entropy, partition_dl, edges_dl) of the nonparametric description length."""
        L = lib()
        x = self.export()
        nonempty = int((x["sizes"] > 0).sum())
        return (L.orc_np_sparse_entropy(self.h), L.orc_np_partition_dl(self.h),
                L.orc_np_edges_dl(nonempty, self.g.E))

    def np_deg_dl(self):
        """get_deg_dl_dist: the degree description length (0 without degree correction)."""
        return lib().orc_np_deg_dl(self.h)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.orc_bm_free(self.h)
            self.h = None

    def copy(self):
        return OBlockmodel(lib().orc_bm_copy(self.h), self.g)

    @property
    def B(self):
        return lib().orc_bm_num_blocks(self.h)

    def export(self):
        L = lib()
        B = self.B
        nnz = L.orc_bm_nnz(self.h)
        row_ptr = np.zeros(B + 1, np.int32)
        cols = np.zeros(max(nnz, 1), np.int32)
        cnts = np.zeros(max(nnz, 1), np.int32)
        d_out = np.zeros(B, np.int64)
        d_in = np.zeros(B, np.int64)
        d_tot = np.zeros(B, np.int64)
        sizes = np.zeros(B, np.int32)
        assign = np.zeros(self.g.N, np.int32)
        L.orc_bm_export(self.h, _p(row_ptr), _p(cols), _p(cnts), _p(d_out), _p(d_in), _p(d_tot),
                        _p(sizes), _p(assign))
        return dict(B=B, row_ptr=row_ptr, cols=cols[:nnz], cnts=cnts[:nnz], d_out=d_out, d_in=d_in,
                    d_tot=d_tot, sizes=sizes, assignment=assign)

    def dense(self):
        x = self.export()
        M = np.zeros((x["B"], x["B"]), np.int64)
        for r in range(x["B"]):
            for i in range(x["row_ptr"][r], x["row_ptr"][r + 1]):
                M[r, x["cols"][i]] = x["cnts"][i]
        return M

    def validate(self):
        return bool(lib().orc_bm_validate(self.h, self.g.ref()))

    def log_posterior(self):
        return lib().orc_log_posterior(self.h)

    def mdl(self):
        return lib().orc_mdl(self.h, self.g.ref())

    def vertex_delta_mdl(self, v, s):
        return lib().orc_vertex_delta_mdl(self.h, self.g.ref(), v, s)

    def vertex_hastings(self, v, s):
        return lib().orc_vertex_hastings(self.h, self.g.ref(), v, s)

    def merge_delta_mdl(self, r, s):
        return lib().orc_merge_delta_mdl(self.h, r, s)

    def vertex_delta_mdl_cells(self, v, s):
        """Per-cell (reference order) delta-MDL and the sum of |cell terms|."""
        a = C.c_double()
        d = lib().orc_vertex_delta_mdl_cells(self.h, self.g.ref(), v, s, C.byref(a))
        return d, a.value

    def merge_delta_mdl_cells(self, r, s):
        a = C.c_double()
        d = lib().orc_merge_delta_mdl_cells(self.h, r, s, C.byref(a))
        return d, a.value

    def ft_propose(self, seed, epoch, v, greedy=True):
        s, dS, acc = C.c_int32(), C.c_double(), C.c_int32()
        lib().orc_ft_propose(self.h, self.g.ref(), seed, epoch, v, int(greedy), C.byref(s), C.byref(dS),
                             C.byref(acc))
        return s.value, dS.value, bool(acc.value)

    def async_gibbs_sweep(self, prm, epoch):
        proposed = np.full(self.g.N, -1, np.int32)
        st = Stats()
        lib().orc_async_gibbs_sweep(self.h, self.g.ref(), C.byref(prm), epoch, _p(proposed), C.byref(st))
        return proposed, st

    def mh_sweep(self, prm, epoch, ordered=False):
        tot = C.c_double()
        st = Stats()
        lib().orc_mh_sweep(self.h, self.g.ref(), C.byref(prm), epoch, int(ordered), C.byref(tot), C.byref(st))
        return tot.value, st

    def mcmc(self, prm, golden_ratio_not_reached, epoch):
        ep = C.c_uint32(epoch)
        st = Stats()
        lib().orc_mcmc(self.h, self.g.ref(), C.byref(prm), int(golden_ratio_not_reached), C.byref(ep),
                       C.byref(st))
        return ep.value, st

    def finetune_assignment(self, prm, epoch):
        ep = C.c_uint32(epoch)
        st = Stats()
        lib().orc_finetune_assignment(self.h, self.g.ref(), C.byref(prm), C.byref(ep), C.byref(st))
        return ep.value, st

    def merge_propose(self, seed, epoch):
        B = self.B
        best_s = np.zeros(B, np.int32)
        best_dS = np.zeros(B, np.float64)
        lib().orc_merge_propose(self.h, seed, epoch, _p(best_s), _p(best_dS))
        return best_s, best_dS

    def merge_apply(self, best_s, best_dS, num_to_merge, approximate=False):
        bs = np.ascontiguousarray(best_s, np.int32)
        bd = np.ascontiguousarray(best_dS, np.float64)
        h = lib().orc_merge_apply(self.h, self.g.ref(), _p(bs), _p(bd), int(num_to_merge), int(approximate))
        return OBlockmodel(h, self.g)


def philox(seed, c0, c1, c2, c3):
    out = np.zeros(4, np.uint32)
    lib().orc_philox(seed, c0, c1, c2, c3, _p(out))
    return out


def logf_range(first_bits, count):
    out = np.zeros(count, np.float32)
    lib().orc_logf_range(first_bits, count, _p(out))
    return out


def null_mdl_v1(E):
    return lib().orc_null_mdl_v1(E)


def sbp(graph, prm):
    assign = np.zeros(graph.N, np.int32)
    nb = C.c_int32()
    st = Stats()
    lib().orc_sbp(graph.ref(), C.byref(prm), _p(assign), C.byref(nb), C.byref(st))
    return assign, nb.value, st


SAMPLING = {"max_degree": 0, "random": 1, "degree_product": 2, "degree_weighted": 3, "expansion_snowball": 4,
            "degree_product_snowball": 5}


def sample(graph, alg, fraction, seed=0):
    """orc_sample: (mapping[N] full -> sample id or -1, vertices[n] sample -> full id)."""
    a = SAMPLING[alg] if isinstance(alg, str) else int(alg)
    mapping = np.zeros(graph.N, np.int32)
    verts = np.zeros(graph.N, np.int32)
    n = lib().orc_sample(graph.ref(), a, fraction, seed, _p(mapping), _p(verts))
    return mapping, verts[:n]


def degree_weight_key(w0, degree):
    return lib().orc_degree_weight_key(int(w0), int(degree))


def subgraph(graph, mapping, vertices):
    """orc_subgraph (sample::from_vertices): the induced subgraph as an OGraph."""
    n = len(vertices)
    m = np.ascontiguousarray(mapping, np.int32)
    vs = np.ascontiguousarray(vertices, np.int32)
    out_off = np.zeros(n + 1, np.int64)
    in_off = np.zeros(n + 1, np.int64)
    out_adj = np.zeros(max(graph.E, 1), np.int32)
    in_adj = np.zeros(max(graph.E, 1), np.int32)
    E = lib().orc_subgraph(graph.ref(), _p(m), _p(vs), n, _p(out_off), _p(out_adj), _p(in_off), _p(in_adj))
    return OGraph(out_off, out_adj[:E], in_off, in_adj[:E])


def detach(graph):
    """orc_detach (sample::detach): (mapping[N] -> detached id or -1, vertices[n])."""
    mapping = np.zeros(graph.N, np.int32)
    verts = np.zeros(graph.N, np.int32)
    n = lib().orc_detach(graph.ref(), _p(mapping), _p(verts))
    return mapping, verts[:n]


def reattach(graph, mapping, detached_assignment, detached_blocks, seed=0):
    """orc_reattach (sample::reattach): the full graph's assignment."""
    out = np.zeros(graph.N, np.int32)
    m = np.ascontiguousarray(mapping, np.int32)
    da = np.ascontiguousarray(detached_assignment, np.int32)
    lib().orc_reattach(graph.ref(), _p(m), _p(da), detached_blocks, seed, _p(out))
    return out


def extend(graph, mapping, sample_assignment, sample_blocks, seed=0):
    out = np.zeros(graph.N, np.int32)
    m = np.ascontiguousarray(mapping, np.int32)
    sa = np.ascontiguousarray(sample_assignment, np.int32)
    lib().orc_extend(graph.ref(), _p(m), _p(sa), sample_blocks, seed, _p(out))
    return out


def split_communities(graph, assignment, B, target, seed=0, epoch=0, nonparametric=False):
    """orc_split_communities (divisive::split_communities): (new assignment, new block count, delta[B])."""
    a = np.array(assignment, np.int32, copy=True)
    delta = np.zeros(max(B, 1), np.float64)
    nb = lib().orc_split_communities(graph.ref(), _p(a), B, target, seed, epoch, int(bool(nonparametric)),
                                     _p(delta))
    return a, nb, delta[:B]
