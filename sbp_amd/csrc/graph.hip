// Graph-side device work: per-vertex degrees and the degree order (Graph::degrees, src/graph.cpp:24-31;
// Graph::sort_vertices, :235-268), vertex sampling (sample::max_degree / random / degree_product,
// src/sample.cpp:28-49,219-247, plus a degree-weighted sampler), the induced subgraph of a sample
// (sample::from_vertices, :202-217) and the extension of a sample partition to the full graph
// (sample::extend, :157-200).
//
// Orders are produced by stable radix sorts of 64-bit keys (key << 32 | vertex), so ties fall to the lower
// vertex id, as the reference's stable sorts do.
//
// extend is sequential in the reference: an unsampled vertex votes with the blocks of its sampled
// neighbours and of the unsampled neighbours with a lower id, which it extended before.  Here every
// unsampled vertex counts those lower-id unsampled neighbours (adjacency entries, self-loops excluded); the
// ones with none are ready.  A ready vertex takes the plurality block of its assigned neighbours (ties to
// the smallest block, a Philox-keyed random block when none is assigned) and releases its higher-id
// unsampled neighbours; those reaching zero form the next frontier.  Each vertex is computed exactly once,
// from exactly the neighbour blocks the sequential loop sees: its lower-id unsampled neighbours are final
// and its higher-id ones are still unassigned, because they wait for it.
#include <math.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "device_common.hpp"
#include "device_table.hpp"
#include "internal.hpp"
#include "primitives.hpp"
#include "snowball.hpp"

namespace sbpimpl {
using namespace sbpdev;

namespace {

constexpr int GB = 256;

int grid_for(int64_t n) {
    int64_t g = (n + GB - 1) / GB;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65535));
}

template <typename T>
struct DevBuf {  // temporary device buffer for one call
    T *p = nullptr;
    explicit DevBuf(size_t n) { SBP_HIP(dev_malloc(&p, sizeof(T) * std::max<size_t>(n, 1))); }
    ~DevBuf() {
        if (p) (void)dev_free(p);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

__global__ void k_edge_src(int32_t N, const int64_t *__restrict__ out_off, int32_t *__restrict__ edge_src, int64_t E) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        for (int64_t e = out_off[v]; e < out_off[v + 1]; ++e) edge_src[e] = v;
        edge_src[E + v] = v;  // identity vector after the edge sources
    }
}

// self[v] = 1 when v has a self-loop (Graph::_self_edges)
__global__ void k_self_flags(int64_t E, const int32_t *__restrict__ edge_src, const int32_t *__restrict__ out_adj,
                             int32_t *__restrict__ self) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x)
        if (out_adj[e] == edge_src[e]) self[edge_src[e]] = 1;
}

// degree = |out| + |in| - self (Graph::degrees); also max(|out| + |in|) and the count above the LDS limit
__global__ void k_degrees(int32_t N, const int64_t *__restrict__ out_off, const int64_t *__restrict__ in_off,
                          const int32_t *__restrict__ self, int32_t *__restrict__ deg, unsigned long long *stats) {
    unsigned long long mx = 0, high = 0;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const int64_t kall = (out_off[v + 1] - out_off[v]) + (in_off[v + 1] - in_off[v]);
        deg[v] = (int32_t)(kall - self[v]);
        mx = mx > (unsigned long long)kall ? mx : (unsigned long long)kall;
        high += kall > FT_SMALL_DEGREE;
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        unsigned long long o = __shfl_down(mx, off, WAVE);
        mx = mx > o ? mx : o;
        high += __shfl_down(high, off, WAVE);
    }
    if (lane_id() == 0) {
        atomicMax(&stats[0], mx);
        atomicAdd(&stats[1], high);
    }
}

__global__ void k_degree_keys(int32_t N, const int32_t *__restrict__ deg, uint64_t *__restrict__ keys) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        keys[v] = ((uint64_t)(0xFFFFFFFFu - (uint32_t)deg[v]) << 32) | (uint32_t)v;  // descending degree
}

__global__ void k_random_keys(int32_t N, uint64_t seed, uint64_t *__restrict__ keys) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const U4 w = philox(seed, (uint32_t)v, 0, TAG_SAMPLE, 0);
        keys[v] = ((uint64_t)w.x << 32) | (uint32_t)v;
    }
}

// Efraimidis-Spirakis: key = -logf(u) / degree (float, correctly rounded), ascending; degree 0 never wins
__global__ void k_weighted_keys(int32_t N, uint64_t seed, const int32_t *__restrict__ deg, uint64_t *__restrict__ keys) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const U4 w = philox(seed, (uint32_t)v, 0, TAG_SAMPLE, 0);
        uint32_t bits = 0x7F800000u;  // +inf
        if (deg[v] > 0) {
            const float u = (float)(((w.x >> 9) << 1) | 1u) * 0x1p-24f;
            const float key = __fdiv_rn(-logf_glibc(u), (float)deg[v]);
            bits = __float_as_uint(key);
        }
        keys[v] = ((uint64_t)bits << 32) | (uint32_t)v;
    }
}

// degree_product: per out-edge, descending degree(src) * degree(dst); the edge index breaks ties (stable)
__global__ void k_product_keys(int64_t E, const int32_t *__restrict__ edge_src, const int32_t *__restrict__ out_adj,
                               const int32_t *__restrict__ deg, uint64_t *__restrict__ keys, int32_t *__restrict__ idx) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        keys[e] = ~((uint64_t)deg[edge_src[e]] * (uint64_t)deg[out_adj[e]]);
        idx[e] = (int32_t)e;
    }
}

// first position of each vertex in the sequence (src0, dst0, src1, dst1, ...) of the sorted edges
__global__ void k_first_position(int64_t E, const int32_t *__restrict__ sorted_idx, const int32_t *__restrict__ edge_src,
                                 const int32_t *__restrict__ out_adj, int32_t *__restrict__ first) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < E; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e = sorted_idx[i];
        atomicMin(&first[edge_src[e]], (int32_t)(2 * i));
        atomicMin(&first[out_adj[e]], (int32_t)(2 * i + 1));
    }
}

__global__ void k_first_keys(int32_t N, const int32_t *__restrict__ first, uint64_t *__restrict__ keys) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        keys[v] = ((uint64_t)(uint32_t)first[v] << 32) | (uint32_t)v;
}

__global__ void k_low_words(int64_t n, const uint64_t *__restrict__ keys, int32_t *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (int32_t)(uint32_t)keys[i];
}

__global__ void k_fill(int64_t n, int32_t *__restrict__ p, int32_t value) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_map_sample(int64_t n, const int32_t *__restrict__ order, int32_t *__restrict__ mapping) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        mapping[order[i]] = (int32_t)i;
}

// induced subgraph: keep[e] for every full out-edge with both ends sampled
__global__ void k_keep_edges(int64_t E, const int32_t *__restrict__ edge_src, const int32_t *__restrict__ out_adj,
                             const int32_t *__restrict__ mapping, int32_t *__restrict__ keep) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x)
        keep[e] = (mapping[edge_src[e]] >= 0 && mapping[out_adj[e]] >= 0) ? 1 : 0;
}

__global__ void k_compact_edges(int64_t E, const int32_t *__restrict__ edge_src, const int32_t *__restrict__ out_adj,
                                const int32_t *__restrict__ mapping, const int32_t *__restrict__ keep,
                                const int64_t *__restrict__ pos, int32_t *__restrict__ s_src,
                                int32_t *__restrict__ s_dst, int64_t *__restrict__ out_cnt, int64_t *__restrict__ in_cnt) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        if (!keep[e]) continue;
        const int32_t a = mapping[edge_src[e]], b = mapping[out_adj[e]];
        s_src[pos[e]] = a;
        s_dst[pos[e]] = b;
        atomicAdd((unsigned long long *)&out_cnt[a + 1], 1ull);
        atomicAdd((unsigned long long *)&in_cnt[b + 1], 1ull);
    }
}

// ---- extend

__global__ void k_extend_init(int32_t N, const int32_t *__restrict__ mapping, const int32_t *__restrict__ sample_assign,
                              int32_t *__restrict__ assign) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        assign[v] = mapping[v] >= 0 ? sample_assign[mapping[v]] : -1;
}

// pending[v] = adjacency entries of unsampled v to unsampled lower-id neighbours (self-loops excluded)
__global__ void k_extend_pending(int64_t E, const int32_t *__restrict__ edge_src, const int32_t *__restrict__ out_adj,
                                 const int32_t *__restrict__ mapping, int32_t *__restrict__ pending) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = edge_src[e], y = out_adj[e];
        if (x == y || mapping[x] >= 0 || mapping[y] >= 0) continue;
        if (y < x) atomicAdd(&pending[x], 1);  // x's out-list entry y
        else atomicAdd(&pending[y], 1);        // y's in-list entry x
    }
}

__global__ void k_extend_frontier0(int32_t N, const int32_t *__restrict__ mapping, const int32_t *__restrict__ pending,
                                   int32_t *__restrict__ frontier, int32_t *__restrict__ count) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        if (mapping[v] < 0 && pending[v] == 0) frontier[atomicAdd(count, 1)] = v;
}

constexpr int EX_WAVES = 4;
constexpr int EX_LDS_LOG2 = 9;
constexpr int EX_SMALL = 256;  // max |out| + |in| for the LDS table

struct ExArgs {
    int32_t N, B;
    uint64_t seed;
    const int64_t *out_off, *in_off;
    const int32_t *out_adj, *in_adj;
    const int32_t *mapping;
    int32_t *assign, *pending;
    const int32_t *frontier;
    const int32_t *count;
    int32_t *next, *next_count;
    int *gtable;  // per-wave global tables for high-degree vertices
    int gcap_log2;
    int64_t gwaves;
};

// One wavefront per frontier vertex: plurality of the assigned neighbour blocks, ties to the smallest.
template <bool BIG>
__global__ void __launch_bounds__(EX_WAVES * WAVE) k_extend_level(ExArgs a) {
    __shared__ int lds[BIG ? 1 : EX_WAVES * table_ints(EX_LDS_LOG2)];
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    const int64_t gwave = (int64_t)blockIdx.x * EX_WAVES + wid;
    const int64_t nwaves = BIG ? a.gwaves : (int64_t)gridDim.x * EX_WAVES;  // waves that own a table
    if (gwave >= nwaves) return;
    const Table t = BIG ? table_at(a.gtable + gwave * table_ints(a.gcap_log2), a.gcap_log2)
                        : table_at(lds + wid * table_ints(EX_LDS_LOG2), EX_LDS_LOG2);
    table_clear_all(t);
    const int32_t n = *a.count;
    for (int64_t item = gwave; item < n; item += nwaves) {
        const int32_t v = a.frontier[item];
        const int64_t o0 = a.out_off[v], o1 = a.out_off[v + 1], i0 = a.in_off[v], i1 = a.in_off[v + 1];
        if (((o1 - o0) + (i1 - i0) > EX_SMALL) != BIG) continue;  // the other launch takes it
        for (int64_t e = o0 + lane; e < o1; e += WAVE) {
            const int32_t c = a.assign[a.out_adj[e]];
            if (c >= 0) table_add(t, c, 1, 0);
        }
        for (int64_t e = i0 + lane; e < i1; e += WAVE) {
            const int32_t u = a.in_adj[e];
            const int32_t c = u == v ? -1 : a.assign[u];
            if (c >= 0) table_add(t, c, 1, 0);
        }
        wave_sync();
        const int nb = *t.n_used;
        // max count, then min block: maximise (count << 32) | (INT32_MAX - block)
        unsigned long long best = 0;
        for (int i = lane; i < nb; i += WAVE) {
            const int slot = t.used[i];
            const unsigned long long k =
                ((unsigned long long)(uint32_t)t.outc[slot] << 32) | (uint32_t)(0x7FFFFFFF - t.keys[slot]);
            best = best > k ? best : k;
        }
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(best, off, WAVE);
            best = best > o ? best : o;
        }
        int32_t block;
        if (nb == 0) {  // no assigned neighbour: a random block (common::random_integer(0, B - 1))
            const U4 w = philox(a.seed, (uint32_t)v, 0, TAG_EXTEND, 0);
            block = (int32_t)uniform_below(w.x, (uint32_t)a.B);
        } else {
            block = 0x7FFFFFFF - (int32_t)(uint32_t)(best & 0xFFFFFFFFull);
        }
        table_clear_used(t);
        if (lane == 0) a.assign[v] = block;
        // release the higher-id unsampled neighbours waiting for v
        for (int64_t e = o0 + lane; e < o1; e += WAVE) {
            const int32_t x = a.out_adj[e];
            if (x > v && a.mapping[x] < 0 && atomicSub(&a.pending[x], 1) == 1) a.next[atomicAdd(a.next_count, 1)] = x;
        }
        for (int64_t e = i0 + lane; e < i1; e += WAVE) {
            const int32_t x = a.in_adj[e];
            if (x > v && a.mapping[x] < 0 && atomicSub(&a.pending[x], 1) == 1) a.next[atomicAdd(a.next_count, 1)] = x;
        }
    }
}

// Sorts n 64-bit keys ascending into `out` (stable; keys are unique here).
void sort_keys(hipStream_t st, const uint64_t *in, uint64_t *out, int64_t n) {
    size_t tmp = 0;
    const int32_t *no_vals = nullptr;
    SBP_HIP(sbpprim::radix_sort(nullptr, tmp, in, out, no_vals, (int32_t *)nullptr, n, 0, 64, st));
    DevBuf<char> t(tmp);
    SBP_HIP(sbpprim::radix_sort((void *)t.p, tmp, in, out, no_vals, (int32_t *)nullptr, n, 0, 64, st));
    SBP_HIP(hipStreamSynchronize(st));
}

}  // namespace

// Device-side derived arrays of a graph whose CSR is on the device: edge sources (+ identity vector),
// degrees, max |out| + |in| and the count of high-degree vertices.
void graph_finish(sbp_ctx ctx, sbp_graph g) {
    hipStream_t st = ctx->stream;
    const int32_t N = g->N;
    const int64_t E = g->E;
    if (!g->edge_src) SBP_HIP(dev_malloc(&g->edge_src, sizeof(int32_t) * (E + N)));
    if (!g->degree) SBP_HIP(dev_malloc(&g->degree, sizeof(int32_t) * (N + 1)));
    k_edge_src<<<grid_for(N), GB, 0, st>>>(N, g->out_off, g->edge_src, E);
    DevBuf<int32_t> self(N);
    DevBuf<unsigned long long> stats(2);
    SBP_HIP(hipMemsetAsync(self.p, 0, sizeof(int32_t) * N, st));
    SBP_HIP(hipMemsetAsync(stats.p, 0, 2 * sizeof(unsigned long long), st));
    if (E > 0) k_self_flags<<<grid_for(E), GB, 0, st>>>(E, g->edge_src, g->out_adj, self.p);
    k_degrees<<<grid_for(N), GB, 0, st>>>(N, g->out_off, g->in_off, self.p, g->degree, stats.p);
    SBP_HIP(hipGetLastError());
    unsigned long long h[2];
    SBP_HIP(hipMemcpyAsync(h, stats.p, sizeof(h), hipMemcpyDeviceToHost, st));
    SBP_HIP(hipStreamSynchronize(st));
    g->max_degree = (int64_t)h[0];
    g->n_high_degree = (int64_t)h[1];
}

// Graph::sort_vertices order: stable descending degree (computed once per graph).  Worker contexts of a batch may
// share one graph from their own streams: the order is computed under the graph's lock and published only after
// its stream has finished writing it, so no caller reads a partly written order.
const int32_t *degree_order(sbp_ctx ctx, sbp_graph g) {
    std::lock_guard<std::mutex> lk(g->deg_mu);
    if (g->deg_order) return g->deg_order;
    hipStream_t st = ctx->stream;
    const int32_t N = g->N;
    DevBuf<uint64_t> keys(N), sorted(N);
    k_degree_keys<<<grid_for(N), GB, 0, st>>>(N, g->degree, keys.p);
    sort_keys(st, keys.p, sorted.p, N);
    int32_t *order = nullptr;
    SBP_HIP(dev_malloc(&order, sizeof(int32_t) * N));
    k_low_words<<<grid_for(N), GB, 0, st>>>(N, sorted.p, order);
    SBP_HIP(hipGetLastError());
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)dev_free(order);
        SBP_HIP(e);
    }
    g->deg_order = order;
    return order;
}

// sample::sample + from_vertices.  Returns the sample as a new device graph that keeps its mapping.
sbp_graph sample_graph(sbp_ctx ctx, sbp_graph g, int32_t alg, double fraction, uint64_t seed) {
    hipStream_t st = ctx->stream;
    const int32_t N = g->N;
    const int64_t E = g->E;
    const int64_t target = (int64_t)(fraction * (double)N);
    SBP_CHECK(target >= 0 && target <= N, SBP_E_ARG, "sample fraction must be in [0, 1]");
    SBP_CHECK(alg == 2 || alg >= 4 || target > 0, SBP_E_ARG, "sample would be empty");
    SBP_CHECK(alg >= 0 && alg <= SAMPLE_DETACH, SBP_E_ARG, "unknown sampling algorithm");
    DevBuf<uint64_t> keys(N), sorted(N);
    DevBuf<int32_t> order(N);
    int64_t n = target;
    std::vector<int32_t> host_order, host_mapping;  // the snowball samplers run on the host
    if (alg == SAMPLE_DETACH) {  // sample::detach (src/sample.cpp:12-26): the vertices of degree > 1, in id order
        std::vector<int32_t> deg(N + 1);
        SBP_HIP(hipStreamSynchronize(st));
        SBP_HIP(dev_copy(deg.data(), g->degree, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
        host_mapping.assign(N, -1);
        for (int32_t v = 0; v < N; ++v)
            if (deg[v] > 1) {
                host_mapping[v] = (int32_t)host_order.size();
                host_order.push_back(v);
            }
        n = (int64_t)host_order.size();
        SBP_CHECK(n > 0, SBP_E_ARG, "detach: no vertex has degree > 1");
    } else if (alg == 4 || alg == 5) {
        std::vector<int64_t> oo(N + 1), io(N + 1);
        std::vector<int32_t> oa(E + 1), ia(E + 1), deg(N + 1);
        SBP_HIP(hipStreamSynchronize(st));
        SBP_HIP(dev_copy(oo.data(), g->out_off, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost));
        SBP_HIP(dev_copy(io.data(), g->in_off, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost));
        if (E > 0) {
            SBP_HIP(dev_copy(oa.data(), g->out_adj, sizeof(int32_t) * E, hipMemcpyDeviceToHost));
            SBP_HIP(dev_copy(ia.data(), g->in_adj, sizeof(int32_t) * E, hipMemcpyDeviceToHost));
        }
        SBP_HIP(dev_copy(deg.data(), g->degree, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
        const sbpsnow::Csr c{N, oo.data(), io.data(), oa.data(), ia.data()};
        host_order = alg == 4 ? sbpsnow::expansion_snowball(c, target, seed, host_mapping)
                              : sbpsnow::degree_product_snowball(c, deg.data(), target, seed, host_mapping);
        n = (int64_t)host_order.size();
    } else if (alg == 0) {
        k_degree_keys<<<grid_for(N), GB, 0, st>>>(N, g->degree, keys.p);
    } else if (alg == 1) {
        k_random_keys<<<grid_for(N), GB, 0, st>>>(N, seed, keys.p);
    } else if (alg == 3) {
        k_weighted_keys<<<grid_for(N), GB, 0, st>>>(N, seed, g->degree, keys.p);
    } else {
        SBP_CHECK(alg == 2, SBP_E_ARG, "unknown sampling algorithm (0..5)");
        SBP_CHECK(2 * E < INT32_MAX, SBP_E_ARG, "degree_product sampling needs 2E < 2^31");
        n = 0;
        DevBuf<int32_t> first(N);
        k_fill<<<grid_for(N), GB, 0, st>>>(N, first.p, INT32_MAX);
        if (E > 0) {
            DevBuf<uint64_t> ekeys(E), esorted(E);
            DevBuf<int32_t> eidx(E), eidx_sorted(E);
            k_product_keys<<<grid_for(E), GB, 0, st>>>(E, g->edge_src, g->out_adj, g->degree, ekeys.p, eidx.p);
            size_t tmp = 0;
            SBP_HIP(sbpprim::radix_sort(nullptr, tmp, (const uint64_t *)ekeys.p, esorted.p, (const int32_t *)eidx.p,
                                        eidx_sorted.p, E, 0, 64, st));
            DevBuf<char> t(tmp);
            SBP_HIP(sbpprim::radix_sort((void *)t.p, tmp, (const uint64_t *)ekeys.p, esorted.p, (const int32_t *)eidx.p,
                                        eidx_sorted.p, E, 0, 64, st));
            k_first_position<<<grid_for(E), GB, 0, st>>>(E, eidx_sorted.p, g->edge_src, g->out_adj, first.p);
        }
        k_first_keys<<<grid_for(N), GB, 0, st>>>(N, first.p, keys.p);
        sort_keys(st, keys.p, sorted.p, N);
        // vertices that occur, in order of first occurrence; the loop stops after the edge at which the count
        // reaches the target, so it takes every vertex first seen at or before that edge's destination slot
        std::vector<uint64_t> h(N);
        SBP_HIP(dev_copy(h.data(), sorted.p, sizeof(uint64_t) * N, hipMemcpyDeviceToHost));
        int64_t occurring = 0;
        while (occurring < N && (uint32_t)(h[occurring] >> 32) != (uint32_t)INT32_MAX) ++occurring;
        if (occurring > 0) {
            const int64_t k = std::max<int64_t>(target, 1);
            if (k >= occurring) {
                n = occurring;
            } else {
                const uint32_t pos = (uint32_t)(h[k - 1] >> 32);
                const uint32_t limit = (pos / 2) * 2 + 1;
                n = k;
                while (n < occurring && (uint32_t)(h[n] >> 32) <= limit) ++n;
            }
        }
    }
    if (host_order.empty()) {
        if (alg != 2) sort_keys(st, keys.p, sorted.p, N);
        k_low_words<<<grid_for(N), GB, 0, st>>>(N, sorted.p, order.p);
    }

    SBP_CHECK(n > 0, SBP_E_ARG, "sample would be empty");
    std::unique_ptr<sbp_graph_s, int (*)(sbp_graph)> s(new sbp_graph_s(), sbp_graph_free);
    s->ctx = ctx;
    s->N = (int32_t)n;
    s->parent = g;
    s->parent_N = N;
    SBP_HIP(dev_malloc(&s->parent_map, sizeof(int32_t) * N));
    SBP_HIP(dev_malloc(&s->sample_vertices, sizeof(int32_t) * std::max<int64_t>(n, 1)));
    if (!host_order.empty()) {  // as sampled: a vertex appended twice keeps an isolated earlier slot
        SBP_HIP(hipMemcpyAsync(s->parent_map, host_mapping.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st));
        SBP_HIP(hipMemcpyAsync(s->sample_vertices, host_order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    } else {
        k_fill<<<grid_for(N), GB, 0, st>>>(N, s->parent_map, -1);
        if (n > 0) {
            k_map_sample<<<grid_for(n), GB, 0, st>>>(n, order.p, s->parent_map);
            SBP_HIP(hipMemcpyAsync(s->sample_vertices, order.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
        }
    }
    // induced subgraph (from_vertices): the kept out-edges in full CSR order (source id ascending, list order),
    // stably sorted by sample source (out-lists) and by sample destination (in-lists)
    DevBuf<int32_t> keep(E);
    DevBuf<int64_t> pos(E + 1);
    int64_t Es = 0;
    if (E > 0) {
        k_keep_edges<<<grid_for(E), GB, 0, st>>>(E, g->edge_src, g->out_adj, s->parent_map, keep.p);
        size_t tmp = 0;
        SBP_HIP(sbpprim::scan(nullptr, tmp, (const int32_t *)keep.p, pos.p, E, false, st));
        DevBuf<char> t(tmp);
        SBP_HIP(sbpprim::scan((void *)t.p, tmp, (const int32_t *)keep.p, pos.p, E, false, st));
        int64_t last_pos = 0;
        int32_t last_keep = 0;
        SBP_HIP(hipMemcpyAsync(&last_pos, pos.p + E - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        SBP_HIP(hipMemcpyAsync(&last_keep, keep.p + E - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        SBP_HIP(hipStreamSynchronize(st));
        Es = last_pos + last_keep;
    }
    s->E = Es;
    SBP_HIP(dev_malloc(&s->out_off, sizeof(int64_t) * (n + 1)));
    SBP_HIP(dev_malloc(&s->in_off, sizeof(int64_t) * (n + 1)));
    SBP_HIP(dev_malloc(&s->out_adj, sizeof(int32_t) * (Es + 1)));
    SBP_HIP(dev_malloc(&s->in_adj, sizeof(int32_t) * (Es + 1)));
    SBP_HIP(hipMemsetAsync(s->out_off, 0, sizeof(int64_t) * (n + 1), st));
    SBP_HIP(hipMemsetAsync(s->in_off, 0, sizeof(int64_t) * (n + 1), st));
    if (Es > 0) {
        DevBuf<int32_t> ssrc(Es), sdst(Es), tmpk(Es);
        DevBuf<int64_t> ocnt(n + 1), icnt(n + 1);
        SBP_HIP(hipMemsetAsync(ocnt.p, 0, sizeof(int64_t) * (n + 1), st));
        SBP_HIP(hipMemsetAsync(icnt.p, 0, sizeof(int64_t) * (n + 1), st));
        k_compact_edges<<<grid_for(E), GB, 0, st>>>(E, g->edge_src, g->out_adj, s->parent_map, keep.p, pos.p, ssrc.p,
                                                     sdst.p, ocnt.p, icnt.p);
        // vertex ids are non-negative, so their int32 bits sort as unsigned
        auto *usrc = reinterpret_cast<const uint32_t *>(ssrc.p);
        auto *udst = reinterpret_cast<const uint32_t *>(sdst.p);
        auto *ukey = reinterpret_cast<uint32_t *>(tmpk.p);
        size_t tmp = 0, tmp2 = 0;
        SBP_HIP(sbpprim::radix_sort(nullptr, tmp, usrc, ukey, (const int32_t *)sdst.p, s->out_adj, Es, 0, 32, st));
        SBP_HIP(sbpprim::scan(nullptr, tmp2, (const int64_t *)ocnt.p, s->out_off, n + 1, true, st));
        DevBuf<char> t(std::max(tmp, tmp2));
        SBP_HIP(sbpprim::radix_sort((void *)t.p, tmp, usrc, ukey, (const int32_t *)sdst.p, s->out_adj, Es, 0, 32, st));
        SBP_HIP(sbpprim::radix_sort((void *)t.p, tmp, udst, ukey, (const int32_t *)ssrc.p, s->in_adj, Es, 0, 32, st));
        SBP_HIP(sbpprim::scan((void *)t.p, tmp2, (const int64_t *)ocnt.p, s->out_off, n + 1, true, st));
        SBP_HIP(sbpprim::scan((void *)t.p, tmp2, (const int64_t *)icnt.p, s->in_off, n + 1, true, st));
        SBP_HIP(hipStreamSynchronize(st));
    }
    graph_finish(ctx, s.get());
    SBP_HIP(hipStreamSynchronize(st));
    return s.release();
}

// sample::extend: full-graph assignment (device, N entries) from the sample's blockmodel.
void extend_assignment(sbp_ctx ctx, sbp_graph g, sbp_graph s, sbp_bm sample_bm, uint64_t seed, int32_t *d_assign) {
    hipStream_t st = ctx->stream;
    const int32_t N = g->N;
    const int64_t E = g->E;
    SBP_CHECK(s->parent == g && s->parent_map, SBP_E_ARG, "extend: the sample was not drawn from this graph");
    SBP_CHECK(sample_bm->N == s->N, SBP_E_ARG, "extend: blockmodel does not belong to the sample");
    DevBuf<int32_t> pending(N), fa(N + 1), fb(N + 1), counts(2);
    k_extend_init<<<grid_for(N), GB, 0, st>>>(N, s->parent_map, sample_bm->assign, d_assign);
    SBP_HIP(hipMemsetAsync(pending.p, 0, sizeof(int32_t) * N, st));
    SBP_HIP(hipMemsetAsync(counts.p, 0, 2 * sizeof(int32_t), st));
    if (E > 0) k_extend_pending<<<grid_for(E), GB, 0, st>>>(E, g->edge_src, g->out_adj, s->parent_map, pending.p);
    k_extend_frontier0<<<grid_for(N), GB, 0, st>>>(N, s->parent_map, pending.p, fa.p, counts.p);
    SBP_HIP(hipGetLastError());
    ExArgs a;
    a.N = N;
    a.B = sample_bm->B;
    a.seed = seed;
    a.out_off = g->out_off;
    a.in_off = g->in_off;
    a.out_adj = g->out_adj;
    a.in_adj = g->in_adj;
    a.mapping = s->parent_map;
    a.assign = d_assign;
    a.pending = pending.p;
    // high-degree vertices: one global table of 2 * min(max degree, B) slots per wave
    a.gcap_log2 = 10;
    while ((int64_t(1) << a.gcap_log2) < 2 * std::min<int64_t>(g->max_degree, a.B) + 2) ++a.gcap_log2;
    a.gwaves = std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>(g->n_high_degree, 1),
                                                      (int64_t(1) << 28) / table_ints(a.gcap_log2)));
    DevBuf<int> gtable(g->max_degree > EX_SMALL ? a.gwaves * table_ints(a.gcap_log2) : 1);
    a.gtable = gtable.p;
    int32_t *cur = fa.p, *nxt = fb.p;
    int32_t *cur_count = counts.p, *nxt_count = counts.p + 1;
    int32_t n_cur = 0;
    SBP_HIP(hipMemcpyAsync(&n_cur, cur_count, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SBP_HIP(hipStreamSynchronize(st));
    int64_t done = 0, levels = 0;
    while (n_cur > 0) {
        SBP_CHECK(levels++ <= N, SBP_E_STATE, "extend: dependency levels exceed the vertex count");
        done += n_cur;
        SBP_HIP(hipMemsetAsync(nxt_count, 0, sizeof(int32_t), st));
        a.frontier = cur;
        a.count = cur_count;
        a.next = nxt;
        a.next_count = nxt_count;
        const int blocks = (int)std::min<int64_t>((n_cur + EX_WAVES - 1) / EX_WAVES, 8192);
        k_extend_level<false><<<blocks, EX_WAVES * WAVE, 0, st>>>(a);
        if (g->max_degree > EX_SMALL) {
            const int big_blocks = (int)((a.gwaves + EX_WAVES - 1) / EX_WAVES);
            k_extend_level<true><<<big_blocks, EX_WAVES * WAVE, 0, st>>>(a);
        }
        SBP_HIP(hipGetLastError());
        SBP_HIP(hipMemcpyAsync(&n_cur, nxt_count, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        SBP_HIP(hipStreamSynchronize(st));
        std::swap(cur, nxt);
        std::swap(cur_count, nxt_count);
    }
    SBP_CHECK(done == (int64_t)N - s->N, SBP_E_STATE, "extend: not every unsampled vertex was reached");
}

}  // namespace sbpimpl
