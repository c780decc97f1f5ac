"""ctypes binding of ``libsbp_amd.so`` (the C-ABI declared in ``include/sbp_amd.h``).

The library must exist (built by ``sbp_amd/build.py``); there is no fallback path: every call runs the
HIP kernels, and a missing library or device raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsbp_amd.so")

ALGORITHMS = {"async_gibbs": 0, "metropolis_hastings": 1, "hybrid_mcmc": 2}

# every symbol the header declares
EXPORTS = [
    "sbp_create", "sbp_destroy", "sbp_last_error", "sbp_set_seed", "sbp_synchronize",
    "sbp_graph_upload", "sbp_graph_free", "sbp_graph_info", "sbp_graph_download",
    "sbp_sample", "sbp_extend", "sbp_detach", "sbp_reattach",
    "sbp_bm_build", "sbp_bm_copy", "sbp_bm_free", "sbp_bm_info", "sbp_bm_download", "sbp_bm_validate",
    "sbp_bm_rebuild", "sbp_bm_move_vertex", "sbp_bm_merge_block", "sbp_ft_evaluate", "sbp_merge_evaluate",
    "sbp_merge_apply",
    "sbp_mdl", "sbp_mdl_noalloc", "sbp_bm_set_model", "sbp_np_description_length",
    "sbp_ft_propose_all", "sbp_async_gibbs_sweep", "sbp_async_gibbs_sweep_recorded", "sbp_mh_sweep", "sbp_mcmc", "sbp_finetune_assignment",
    "sbp_merge_propose", "sbp_merge_blocks", "sbp_split_communities",
    "sbp_partition",
    "sbp_ctx_memory", "sbp_ctx_trim", "sbp_set_carry_threads", "sbp_set_finetune_shape", "sbp_get_finetune_shape",
    "sbp_set_recorded_sweeps",
    "sbp_graph_bytes", "sbp_footprint_bytes",
    "sbp_partition_batch",
    "sbp_server_create", "sbp_server_last_error", "sbp_server_submit", "sbp_server_wait", "sbp_server_info",
    "sbp_server_destroy",
    "sbp_evaluate_table", "sbp_evaluate",
    "sbp_debug_logf", "sbp_debug_sort", "sbp_debug_scan", "sbp_debug_rle",
]


SBP_E_ARG = -3


class SbpError(RuntimeError):
    pass


class Params(C.Structure):
    """sbp_params: the hot-path subset of the reference's Args (include/args.hpp:61-212)."""
    _fields_ = [("algorithm", C.c_int32), ("batches", C.c_int32), ("greedy", C.c_int32),
                ("approximate", C.c_int32), ("ordered", C.c_int32), ("asynciterations", C.c_int32),
                ("mh_percent", C.c_double), ("seed", C.c_uint64), ("max_iterations", C.c_int32),
                ("threads", C.c_int32), ("nonparametric", C.c_int32), ("degreecorrected", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("proposals", C.c_int64), ("moves", C.c_int64), ("refused", C.c_int64), ("sweeps", C.c_int64),
                ("merge_proposals", C.c_int64), ("merge_phases", C.c_int64), ("mdl", C.c_double),
                ("finetune_ms", C.c_double), ("merge_ms", C.c_double), ("carry_ms", C.c_double),
                ("build_ms", C.c_double), ("mdl_ms", C.c_double), ("total_ms", C.c_double),
                ("ft_bytes", C.c_uint64), ("merge_bytes", C.c_uint64), ("ft_launches", C.c_int64),
                ("ft_kernel_ms", C.c_double), ("carry_evaluations", C.c_int64), ("carry_speculative", C.c_int64),
                ("carry_reused", C.c_int64), ("carry_requeues", C.c_int64), ("carry_load_ms", C.c_double),
                ("ft_high_bytes", C.c_uint64), ("ft_high_launches", C.c_int64), ("ft_high_kernel_ms", C.c_double),
                ("mh_bytes", C.c_uint64), ("mh_launches", C.c_int64), ("mh_kernel_ms", C.c_double),
                ("ft_class_bytes", C.c_uint64 * 6), ("ft_class_launches", C.c_int64 * 6), ("ft_class_ms", C.c_double * 6),
                ("recorded_sweeps", C.c_int64), ("cell_updates", C.c_int64)]

    def as_dict(self):
        return {k: (list(getattr(self, k)) if k.startswith("ft_class_") else getattr(self, k)) for k, _ in self._fields_}


class Evaluation(C.Structure):
    """sbp_evaluation: scores of a partition against a true one (src/evaluate.cpp:9-96)."""
    _fields_ = [("f1_score", C.c_double), ("nmi", C.c_double), ("accuracy", C.c_double),
                ("rand_index", C.c_double), ("precision", C.c_double), ("recall", C.c_double),
                ("found_blocks", C.c_int32), ("true_blocks", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def params(algorithm="hybrid_mcmc", batches=1, greedy=True, approximate=False, ordered=False,
           asynciterations=2 ** 31 - 1, mh_percent=0.075, seed=0, max_iterations=100, threads=1,
           nonparametric=False, degreecorrected=False):
    """Defaults follow include/args.hpp (algorithm hybrid_mcmc, greedy on, one batch, mh_percent 0.075)."""
    alg = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
    return Params(alg, batches, int(greedy), int(approximate), int(ordered), asynciterations, mh_percent, seed,
                  max_iterations, threads, int(nonparametric), int(degreecorrected))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SbpError(f"{LIB_PATH} is missing: build it with `python sbp_amd/build.py` "
                           "(or __graft_entry__.build()) before using the GPU path")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u32, u64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double
        P, S = C.POINTER(Params), C.POINTER(Stats)
        sig = {
            "sbp_create": [vp, C.c_int, u64],
            "sbp_destroy": [vp],
            "sbp_set_seed": [vp, u64],
            "sbp_synchronize": [vp],
            "sbp_graph_upload": [vp, i32, i64, vp, vp, vp, vp, vp],
            "sbp_graph_free": [vp],
            "sbp_graph_info": [vp, vp, vp],
            "sbp_graph_download": [vp, vp, vp, vp, vp, vp],
            "sbp_sample": [vp, vp, i32, dbl, u64, vp, vp, vp],
            "sbp_extend": [vp, vp, vp, vp, u64, vp, vp],
            "sbp_detach": [vp, vp, vp, vp, vp],
            "sbp_reattach": [i32, vp, vp, vp, vp, vp, vp, i32, u64, vp],
            "sbp_bm_build": [vp, vp, i32, vp, vp],
            "sbp_bm_copy": [vp, vp, vp],
            "sbp_bm_free": [vp],
            "sbp_bm_info": [vp, vp, vp],
            "sbp_bm_download": [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
            "sbp_bm_validate": [vp, vp, vp, vp],
            "sbp_bm_rebuild": [vp, vp, vp],
            "sbp_bm_move_vertex": [vp, vp, vp, i32, i32, vp],
            "sbp_bm_merge_block": [vp, vp, vp, i32, i32],
            "sbp_ft_evaluate": [vp, vp, vp, i64, vp, vp, vp, vp],
            "sbp_merge_evaluate": [vp, vp, i64, vp, vp, vp],
            "sbp_merge_apply": [vp, vp, vp, vp, vp, i64, i32, vp],
            "sbp_mdl": [vp, vp, vp, vp, vp],
            "sbp_mdl_noalloc": [vp, vp, vp, vp, vp],
            "sbp_bm_set_model": [vp, vp, vp, i32],
            "sbp_np_description_length": [vp, vp, vp, vp, vp, vp, vp, vp],
            "sbp_ft_propose_all": [vp, vp, vp, u32, i32, vp, vp, vp],
            "sbp_async_gibbs_sweep": [vp, vp, vp, P, u32, S],
            "sbp_async_gibbs_sweep_recorded": [vp, vp, vp, P, u32, vp, S],
            "sbp_mh_sweep": [vp, vp, vp, P, u32, i32, vp, S],
            "sbp_mcmc": [vp, vp, vp, P, i32, vp, vp, S],
            "sbp_finetune_assignment": [vp, vp, vp, P, vp, vp, S],
            "sbp_merge_propose": [vp, vp, u32, vp, vp],
            "sbp_merge_blocks": [vp, vp, vp, u32, i64, i32, vp, S],
            "sbp_split_communities": [vp, vp, vp, i32, u32, vp, vp],
            "sbp_partition": [vp, vp, P, vp, vp, S],
            "sbp_debug_logf": [vp, u32, u32, vp],
            "sbp_debug_sort": [vp, i64, i32, vp, vp, i32, i32, vp, vp],
            "sbp_debug_scan": [vp, i64, vp, i32, vp],
            "sbp_debug_rle": [vp, i64, vp, vp, vp, vp],
            "sbp_ctx_memory": [vp, vp, vp],
            "sbp_ctx_trim": [vp],
            "sbp_set_carry_threads": [vp, i32],
            "sbp_set_finetune_shape": [vp, vp, vp],
            "sbp_get_finetune_shape": [vp, vp, vp],
            "sbp_set_recorded_sweeps": [vp, i32],
            "sbp_partition_batch": [vp, vp, i32, vp, i32, vp, vp, vp, vp],
            "sbp_server_create": [C.c_int, i32, u64, vp],
            "sbp_server_submit": [vp, i32, i64, vp, vp, vp, vp, P, vp],
            "sbp_server_wait": [vp, i64, vp, vp, S, vp, vp],
            "sbp_server_info": [vp, vp, vp, vp, vp, vp, vp],
            "sbp_server_destroy": [vp],
            "sbp_evaluate_table": [i64, i32, i32, vp, C.POINTER(Evaluation)],
            "sbp_evaluate": [i32, vp, vp, C.POINTER(Evaluation), vp, vp, vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = args
        for name in ("sbp_last_error", "sbp_server_last_error"):
            getattr(L, name).restype = C.c_char_p
            getattr(L, name).argtypes = [vp]
        L.sbp_graph_bytes.restype = u64
        L.sbp_graph_bytes.argtypes = [i32, i64]
        L.sbp_footprint_bytes.restype = u64
        L.sbp_footprint_bytes.argtypes = [i32, i64, i64, i64, i32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One HIP stream on one device plus the replay seed (sbp_create)."""

    def __init__(self, device=0, seed=0):
        self.h = C.c_void_p()
        rc = lib().sbp_create(C.byref(self.h), device, seed)
        if rc != 0:
            raise SbpError(f"sbp_create failed with status {rc} (no HIP device?)")
        from .autotune import tuned_shape
        shape = tuned_shape()  # the proposal launch shape measured on this device model, if any
        if shape is not None:
            self.set_finetune_shape(*shape)

    def check(self, rc, what):
        if rc != 0:
            msg = lib().sbp_last_error(self.h)
            raise SbpError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            lib().sbp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_seed(self, seed):
        self.check(lib().sbp_set_seed(self.h, seed), "sbp_set_seed")

    def synchronize(self):
        self.check(lib().sbp_synchronize(self.h), "sbp_synchronize")

    def upload(self, g):
        return DeviceGraph(self, g)

    def debug_logf(self, first_bits, count):
        out = np.zeros(count, np.float32)
        self.check(lib().sbp_debug_logf(self.h, first_bits, count, _p(out)), "sbp_debug_logf")
        return out

    def debug_sort(self, keys, vals=None, begin_bit=0, end_bit=64, key_bits=64):
        """The library's stable radix sort of keys (uint64 array) by bits [begin_bit, end_bit), carrying vals."""
        keys = np.ascontiguousarray(keys, np.uint64)
        n = len(keys)
        ko = np.zeros(n, np.uint64)
        v = None if vals is None else np.ascontiguousarray(vals, np.int32)
        vo = None if vals is None else np.zeros(n, np.int32)
        self.check(lib().sbp_debug_sort(self.h, n, key_bits, _p(keys), None if v is None else _p(v), begin_bit, end_bit,
                                        _p(ko), None if vo is None else _p(vo)), "sbp_debug_sort")
        return ko, vo

    def debug_scan(self, values, inclusive=False):
        x = np.ascontiguousarray(values, np.int64)
        out = np.zeros(len(x), np.int64)
        self.check(lib().sbp_debug_scan(self.h, len(x), _p(x), int(inclusive), _p(out)), "sbp_debug_scan")
        return out

    def debug_rle(self, keys):
        k = np.ascontiguousarray(keys, np.uint64)
        u = np.zeros(len(k), np.uint64)
        c = np.zeros(len(k), np.int32)
        m = C.c_int64()
        self.check(lib().sbp_debug_rle(self.h, len(k), _p(k), _p(u), _p(c), C.byref(m)), "sbp_debug_rle")
        return u[:m.value], c[:m.value]

    def memory(self):
        """(current, peak) device bytes of the context (sbp_ctx_memory)."""
        cur, peak = C.c_uint64(), C.c_uint64()
        self.check(lib().sbp_ctx_memory(self.h, C.byref(cur), C.byref(peak)), "sbp_ctx_memory")
        return cur.value, peak.value

    def trim(self):
        self.check(lib().sbp_ctx_trim(self.h), "sbp_ctx_trim")

    def set_carry_threads(self, threads):
        self.check(lib().sbp_set_carry_threads(self.h, threads), "sbp_set_carry_threads")

    def set_finetune_shape(self, lanes, waves_per_simd=(4, 4, 4, 4)):
        """Lanes (threads) per vertex of the finetune degree classes k <= 8, 32, 128, 256, 1024 and wavefronts per
        SIMD of the first four (sbp_set_finetune_shape)."""
        la = np.ascontiguousarray(lanes, np.int32)
        wa = np.ascontiguousarray(waves_per_simd, np.int32)
        assert la.shape == (5,) and wa.shape == (4,)
        self.check(lib().sbp_set_finetune_shape(self.h, _p(la), _p(wa)), "sbp_set_finetune_shape")

    def set_recorded_sweeps(self, mode):
        """-1: graph replays for graphs of >= 100,000 vertices, 0 never, 1 always (sbp_set_recorded_sweeps)."""
        self.check(lib().sbp_set_recorded_sweeps(self.h, mode), "sbp_set_recorded_sweeps")

    def finetune_shape(self):
        la = np.zeros(5, np.int32)
        wa = np.zeros(4, np.int32)
        self.check(lib().sbp_get_finetune_shape(self.h, _p(la), _p(wa)), "sbp_get_finetune_shape")
        return la.tolist(), wa.tolist()

    def partition_batch(self, graphs, prms, slots=None):
        """sbp_partition_batch over DeviceGraphs with one Params each (slots None: the autotuned count for
        this host, else the library's default).  Returns (assignments, num_blocks, stats, finish_ms) per
        graph, in input order."""
        if slots is None:
            from .autotune import tuned_slots
            slots = tuned_slots(0)
        n = len(graphs)
        assert len(prms) == n
        hs = (C.c_void_p * n)(*[g.h for g in graphs])
        ps = (Params * n)(*prms)
        outs = [np.zeros(g.N, np.int32) for g in graphs]
        ptrs = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
        nb = np.zeros(n, np.int32)
        st = (Stats * n)()
        fin = np.zeros(n, np.float64)
        self.check(lib().sbp_partition_batch(self.h, hs, n, ps, slots, ptrs, _p(nb), st, _p(fin)),
                   "sbp_partition_batch")
        return outs, nb, list(st), fin


def evaluate_table(num_vertices, table):
    """sbp_evaluate_table on a contingency table (rows = the algorithm's communities)."""
    t = np.ascontiguousarray(table, np.int32)
    e = Evaluation()
    rc = lib().sbp_evaluate_table(num_vertices, t.shape[0], t.shape[1], _p(t), C.byref(e))
    if rc != 0:
        raise SbpError(f"sbp_evaluate_table failed ({rc})")
    return e


def evaluate(true_assignment, assignment):
    """sbp_evaluate: (Evaluation, matched contingency table with rows = the algorithm's communities)."""
    t = np.ascontiguousarray(true_assignment, np.int32)
    a = np.ascontiguousarray(assignment, np.int32)
    labels = len(np.unique(t[t >= 0]))
    table = np.zeros(max(1, (int(a.max(initial=0)) + 1) * labels), np.int32)
    e, r, c = Evaluation(), C.c_int32(), C.c_int32()
    rc = lib().sbp_evaluate(len(t), _p(t), _p(a), C.byref(e), _p(table), C.byref(r), C.byref(c))
    if rc != 0:
        raise SbpError(f"sbp_evaluate failed ({rc})")
    return e, table[: r.value * c.value].reshape(r.value, c.value)


def reattach(g, mapping, detached_assignment, detached_blocks, seed=0):
    """sbp_reattach (host) on a CSRGraph: the full graph's assignment."""
    oo, oa = np.ascontiguousarray(g.out_off, np.int64), np.ascontiguousarray(g.out_adj, np.int32)
    io, ia = np.ascontiguousarray(g.in_off, np.int64), np.ascontiguousarray(g.in_adj, np.int32)
    m = np.ascontiguousarray(mapping, np.int32)
    da = np.ascontiguousarray(detached_assignment, np.int32)
    out = np.zeros(g.N, np.int32)
    rc = lib().sbp_reattach(g.N, _p(oo), _p(oa), _p(io), _p(ia), _p(m), _p(da), detached_blocks,
                            C.c_uint64(seed), _p(out))
    if rc != 0:
        raise SbpError(f"sbp_reattach failed ({rc})")
    return out


def graph_bytes(N, E):
    return int(lib().sbp_graph_bytes(N, E))


def footprint_bytes(g, algorithm="async_gibbs"):
    """sbp_footprint_bytes of a CSRGraph for `algorithm` (name or sbp_params.algorithm value)."""
    k = np.diff(g.out_off) + np.diff(g.in_off)
    alg = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
    return int(lib().sbp_footprint_bytes(g.N, g.E, int(k.max(initial=0)), int((k > 256).sum()), alg))


class Server:
    """Request server (sbp_server_*): workers take partition requests in arrival order as a worker frees up
    and the request's footprint fits in the HBM budget."""

    def __init__(self, device=0, slots=0, hbm_budget_bytes=200 * 10 ** 9):
        self.h = C.c_void_p()
        rc = lib().sbp_server_create(device, slots, hbm_budget_bytes, C.byref(self.h))
        if rc != 0:
            raise SbpError(f"sbp_server_create failed with status {rc}")
        self._held = {}  # ticket -> the host arrays the request reads

    def check(self, rc, what):
        if rc != 0:
            msg = lib().sbp_server_last_error(self.h)
            raise SbpError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def submit(self, g, prm):
        arrs = (np.ascontiguousarray(g.out_off, np.int64), np.ascontiguousarray(g.out_adj, np.int32),
                np.ascontiguousarray(g.in_off, np.int64), np.ascontiguousarray(g.in_adj, np.int32))
        t = C.c_int64()
        self.check(lib().sbp_server_submit(self.h, g.N, g.E, _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), _p(arrs[3]),
                                           C.byref(prm), C.byref(t)), "sbp_server_submit")
        self._held[t.value] = (g.N, arrs)
        return t.value

    def wait(self, ticket):
        """Returns (assignment, num_blocks, stats, latency_ms, queue_ms)."""
        held = self._held.get(ticket)
        if held is None:
            raise SbpError(f"sbp_server_wait({ticket}): unknown ticket")
        out = np.zeros(held[0], np.int32)
        nb, st, lat, q = C.c_int32(), Stats(), C.c_double(), C.c_double()
        rc = lib().sbp_server_wait(self.h, ticket, _p(out), C.byref(nb), C.byref(st), C.byref(lat), C.byref(q))
        if rc != SBP_E_ARG:  # SBP_E_ARG: another call is waiting for this ticket and still needs its arrays
            self._held.pop(ticket, None)
        self.check(rc, f"sbp_server_wait({ticket})")
        return out, nb.value, st, lat.value, q.value

    def info(self):
        r, q, m = C.c_int32(), C.c_int32(), C.c_int32()
        b, u, pk = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.check(lib().sbp_server_info(self.h, C.byref(r), C.byref(q), C.byref(m), C.byref(b), C.byref(u),
                                         C.byref(pk)), "sbp_server_info")
        return dict(running=r.value, queued=q.value, max_running=m.value, budget=b.value, in_use=u.value,
                    peak_in_use=pk.value)

    def close(self):
        if self.h:
            lib().sbp_server_destroy(self.h)
            self.h = C.c_void_p()
            self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SAMPLING = {"max_degree": 0, "random": 1, "degree_product": 2, "degree_weighted": 3, "expansion_snowball": 4,
            "degree_product_snowball": 5}


class DeviceGraph:
    def __init__(self, ctx, g=None, handle=None, parent=None):
        self.ctx = ctx
        self.parent = parent  # a sample keeps the graph it was drawn from alive
        if handle is not None:
            self.h = handle
            N, E = C.c_int32(), C.c_int64()
            ctx.check(lib().sbp_graph_info(self.h, C.byref(N), C.byref(E)), "sbp_graph_info")
            self.N, self.E = N.value, E.value
            return
        self.N, self.E = g.N, g.E
        self.h = C.c_void_p()
        arrs = [np.ascontiguousarray(g.out_off, np.int64), np.ascontiguousarray(g.out_adj, np.int32),
                np.ascontiguousarray(g.in_off, np.int64), np.ascontiguousarray(g.in_adj, np.int32)]
        ctx.check(lib().sbp_graph_upload(ctx.h, g.N, g.E, *[_p(a) for a in arrs], C.byref(self.h)),
                  "sbp_graph_upload")

    def download(self):
        """(out_off, out_adj, in_off, in_adj) of the device CSR."""
        out_off = np.zeros(self.N + 1, np.int64)
        in_off = np.zeros(self.N + 1, np.int64)
        out_adj = np.zeros(max(self.E, 1), np.int32)
        in_adj = np.zeros(max(self.E, 1), np.int32)
        self.ctx.check(lib().sbp_graph_download(self.ctx.h, self.h, _p(out_off), _p(out_adj), _p(in_off),
                                                _p(in_adj)), "sbp_graph_download")
        return out_off, out_adj[:self.E], in_off, in_adj[:self.E]

    def sample(self, alg, fraction, seed=0):
        """sample::sample + from_vertices on the device: (sample DeviceGraph, mapping[N])."""
        a = SAMPLING[alg] if isinstance(alg, str) else int(alg)
        h = C.c_void_p()
        n = C.c_int32()
        mapping = np.zeros(self.N, np.int32)
        self.ctx.check(lib().sbp_sample(self.ctx.h, self.h, a, float(fraction), C.c_uint64(seed), C.byref(h),
                                        C.byref(n), _p(mapping)), "sbp_sample")
        return DeviceGraph(self.ctx, handle=h, parent=self), mapping

    def detach(self):
        """sample::detach on the device: (detached DeviceGraph, mapping[N])."""
        h = C.c_void_p()
        n = C.c_int32()
        mapping = np.zeros(self.N, np.int32)
        self.ctx.check(lib().sbp_detach(self.ctx.h, self.h, C.byref(h), C.byref(n), _p(mapping)), "sbp_detach")
        return DeviceGraph(self.ctx, handle=h, parent=self), mapping

    def __del__(self):
        if getattr(self, "h", None):
            lib().sbp_graph_free(self.h)
            self.h = None

    def partition(self, prm):
        assign = np.zeros(self.N, np.int32)
        nb = C.c_int32()
        st = Stats()
        self.ctx.check(lib().sbp_partition(self.ctx.h, self.h, C.byref(prm), _p(assign), C.byref(nb), C.byref(st)),
                       "sbp_partition")
        return assign, nb.value, st


class DeviceBlockmodel:
    def __init__(self, graph, handle):
        self.g = graph
        self.ctx = graph.ctx
        self.h = handle

    @classmethod
    def build(cls, graph, B, assignment, nonparametric=False, degreecorrected=False):
        a = np.ascontiguousarray(assignment, np.int32)
        h = C.c_void_p()
        graph.ctx.check(lib().sbp_bm_build(graph.ctx.h, graph.h, int(B), _p(a), C.byref(h)), "sbp_bm_build")
        bm = cls(graph, h)
        if nonparametric:
            bm.set_model(True, degreecorrected)
        return bm

    def set_model(self, nonparametric, degreecorrected=False):
        model = (2 if degreecorrected else 1) if nonparametric else 0
        self.ctx.check(lib().sbp_bm_set_model(self.ctx.h, self.h, self.g.h, model), "sbp_bm_set_model")

    def np_parts(self):
        """(mdl, sparse_entropy, partition_dl, edges_dl) of the nonparametric description length."""
        return self.np_parts_dc()[:4]

    def np_parts_dc(self):
        """(mdl, sparse_entropy, partition_dl, edges_dl, degree_dl) of the nonparametric description length."""
        out = [C.c_double() for _ in range(5)]
        self.ctx.check(lib().sbp_np_description_length(self.ctx.h, self.h, self.g.h, *[C.byref(x) for x in out]),
                       "sbp_np_description_length")
        return tuple(x.value for x in out)

    def __del__(self):
        if getattr(self, "h", None):
            lib().sbp_bm_free(self.h)
            self.h = None

    def copy(self):
        h = C.c_void_p()
        self.ctx.check(lib().sbp_bm_copy(self.ctx.h, self.h, C.byref(h)), "sbp_bm_copy")
        return DeviceBlockmodel(self.g, h)

    def extend(self, full_graph, seed=0):
        """sample::extend of this blockmodel (of a sample graph) to full_graph: (blockmodel, assignment)."""
        h = C.c_void_p()
        a = np.zeros(full_graph.N, np.int32)
        self.ctx.check(lib().sbp_extend(self.ctx.h, full_graph.h, self.g.h, self.h, C.c_uint64(seed), C.byref(h),
                                        _p(a)), "sbp_extend")
        return DeviceBlockmodel(full_graph, h), a

    def info(self):
        B, nnz = C.c_int32(), C.c_int64()
        self.ctx.check(lib().sbp_bm_info(self.h, C.byref(B), C.byref(nnz)), "sbp_bm_info")
        return B.value, nnz.value

    @property
    def B(self):
        return self.info()[0]

    def export(self):
        B, nnz = self.info()
        out = dict(B=B, assignment=np.zeros(self.g.N, np.int32), sizes=np.zeros(B, np.int32),
                   d_out=np.zeros(B, np.int32), d_in=np.zeros(B, np.int32), d_tot=np.zeros(B, np.int32),
                   row_ptr=np.zeros(B + 1, np.int32), cols=np.zeros(max(nnz, 1), np.int32),
                   cnts=np.zeros(max(nnz, 1), np.int32))
        self.ctx.check(lib().sbp_bm_download(self.ctx.h, self.h, _p(out["assignment"]), _p(out["sizes"]),
                                             _p(out["d_out"]), _p(out["d_in"]), _p(out["d_tot"]),
                                             _p(out["row_ptr"]), _p(out["cols"]), _p(out["cnts"])),
                       "sbp_bm_download")
        out["cols"] = out["cols"][:nnz]
        out["cnts"] = out["cnts"][:nnz]
        return out

    def validate(self):
        ok = C.c_int32()
        self.ctx.check(lib().sbp_bm_validate(self.ctx.h, self.h, self.g.h, C.byref(ok)), "sbp_bm_validate")
        return bool(ok.value)

    def mdl_noalloc(self):
        """(mdl, log_posterior) through sbp_mdl_noalloc: the same values, no device allocation."""
        m, lp = C.c_double(), C.c_double()
        self.ctx.check(lib().sbp_mdl_noalloc(self.ctx.h, self.h, self.g.h, C.byref(m), C.byref(lp)), "sbp_mdl_noalloc")
        return m.value, lp.value

    def mdl(self):
        m, lp = C.c_double(), C.c_double()
        self.ctx.check(lib().sbp_mdl(self.ctx.h, self.h, self.g.h, C.byref(m), C.byref(lp)), "sbp_mdl")
        return m.value, lp.value

    def propose_all(self, epoch, greedy=True):
        N = self.g.N
        s = np.zeros(N, np.int32)
        dS = np.zeros(N, np.float64)
        acc = np.zeros(N, np.int32)
        self.ctx.check(lib().sbp_ft_propose_all(self.ctx.h, self.h, self.g.h, epoch, int(greedy), _p(s), _p(dS),
                                                _p(acc)), "sbp_ft_propose_all")
        return s, dS, acc.astype(bool)

    def async_gibbs_sweep(self, prm, epoch):
        st = Stats()
        self.ctx.check(lib().sbp_async_gibbs_sweep(self.ctx.h, self.h, self.g.h, C.byref(prm), epoch, C.byref(st)),
                       "sbp_async_gibbs_sweep")
        return st

    def async_gibbs_sweep_recorded(self, prm, epoch):
        """One sweep plus the MDL as a replay of the recorded graph: (stats, mdl)."""
        st = Stats()
        m = C.c_double()
        self.ctx.check(lib().sbp_async_gibbs_sweep_recorded(self.ctx.h, self.h, self.g.h, C.byref(prm), epoch,
                                                            C.byref(m), C.byref(st)), "sbp_async_gibbs_sweep_recorded")
        return st, m.value

    def mh_sweep(self, prm, epoch, ordered=False):
        st = Stats()
        tot = C.c_double()
        self.ctx.check(lib().sbp_mh_sweep(self.ctx.h, self.h, self.g.h, C.byref(prm), epoch, int(ordered),
                                          C.byref(tot), C.byref(st)), "sbp_mh_sweep")
        return tot.value, st

    def mcmc(self, prm, golden_ratio_not_reached, epoch):
        ep = C.c_uint32(epoch)
        mdl = C.c_double()
        st = Stats()
        self.ctx.check(lib().sbp_mcmc(self.ctx.h, self.h, self.g.h, C.byref(prm), int(golden_ratio_not_reached),
                                      C.byref(ep), C.byref(mdl), C.byref(st)), "sbp_mcmc")
        return ep.value, mdl.value, st

    def finetune_assignment(self, prm, epoch):
        ep = C.c_uint32(epoch)
        mdl = C.c_double()
        st = Stats()
        self.ctx.check(lib().sbp_finetune_assignment(self.ctx.h, self.h, self.g.h, C.byref(prm), C.byref(ep),
                                                     C.byref(mdl), C.byref(st)), "sbp_finetune_assignment")
        return ep.value, mdl.value, st

    def rebuild(self):
        self.ctx.check(lib().sbp_bm_rebuild(self.ctx.h, self.h, self.g.h), "sbp_bm_rebuild")

    def move_vertex(self, vertex, block):
        """Blockmodel::move_vertex(const VertexMove_v3&): True when moved, False when refused."""
        moved = C.c_int32()
        self.ctx.check(lib().sbp_bm_move_vertex(self.ctx.h, self.h, self.g.h, int(vertex), int(block),
                                                C.byref(moved)), "sbp_bm_move_vertex")
        return bool(moved.value)

    def merge_block(self, frm, to):
        self.ctx.check(lib().sbp_bm_merge_block(self.ctx.h, self.h, self.g.h, int(frm), int(to)), "sbp_bm_merge_block")

    def evaluate_moves(self, vertices, targets):
        """(delta-MDL, Hastings correction) of moving vertices[i] to targets[i], each against the current state."""
        v = np.ascontiguousarray(vertices, np.int32)
        t = np.ascontiguousarray(targets, np.int32)
        dS, H = np.zeros(len(v)), np.zeros(len(v))
        self.ctx.check(lib().sbp_ft_evaluate(self.ctx.h, self.h, self.g.h, len(v), _p(v), _p(t), _p(dS), _p(H)),
                       "sbp_ft_evaluate")
        return dS, H

    def evaluate_merges(self, blocks, targets):
        b = np.ascontiguousarray(blocks, np.int32)
        t = np.ascontiguousarray(targets, np.int32)
        dS = np.zeros(len(b))
        self.ctx.check(lib().sbp_merge_evaluate(self.ctx.h, self.h, len(b), _p(b), _p(t), _p(dS)),
                       "sbp_merge_evaluate")
        return dS

    def merge_apply(self, best_s, best_dS, num_to_merge, approximate=False):
        s = np.ascontiguousarray(best_s, np.int32)
        d = np.ascontiguousarray(best_dS, np.float64)
        h = C.c_void_p()
        self.ctx.check(lib().sbp_merge_apply(self.ctx.h, self.h, self.g.h, _p(s), _p(d), int(num_to_merge),
                                             int(approximate), C.byref(h)), "sbp_merge_apply")
        return DeviceBlockmodel(self.g, h)

    def merge_propose(self, epoch):
        B = self.B
        s = np.zeros(B, np.int32)
        dS = np.zeros(B, np.float64)
        self.ctx.check(lib().sbp_merge_propose(self.ctx.h, self.h, epoch, _p(s), _p(dS)), "sbp_merge_propose")
        return s, dS

    def split_communities(self, target, epoch):
        """divisive::split_communities in place: (new block count, each old block's kept split value)."""
        B = self.B
        nb = C.c_int32()
        delta = np.zeros(max(B, 1), np.float64)
        self.ctx.check(lib().sbp_split_communities(self.ctx.h, self.h, self.g.h, int(target), int(epoch),
                                                   C.byref(nb), _p(delta)), "sbp_split_communities")
        return nb.value, delta[:B]

    def merge_blocks(self, epoch, num_to_merge, approximate=False):
        h = C.c_void_p()
        st = Stats()
        self.ctx.check(lib().sbp_merge_blocks(self.ctx.h, self.h, self.g.h, epoch, int(num_to_merge),
                                              int(approximate), C.byref(h), C.byref(st)), "sbp_merge_blocks")
        return DeviceBlockmodel(self.g, h), st
