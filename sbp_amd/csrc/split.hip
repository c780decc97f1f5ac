// Divisive step: split_communities (src/top_down_sbp.cpp:488-549) with random splits (propose_random_split,
// :228-240, the reference's default --split), on the device for every block and proposal at once.
//
// For each block b of n_b vertices (blocks under 0.001 N are not split, :504-511) the reference extracts the
// induced subgraph (its vertices ascending, out-edges in list order, :98-132) and draws NUM_AGG_PROPOSALS_PER_BLOCK
// random splits: a size k = random_integer(0.2 n_b, 0.8 n_b), then each vertex to side 0 with weight k and side 1
// with weight n_b - k.  A split's value is the nonparametric description length of the two-block blockmodel of
// the subgraph minus null_mdl_v1 of the subgraph (:514-521); the block keeps its smallest (the first smallest in
// proposal order).  apply_best_splits (:30-66) then walks the blocks from the smallest value up (argsort is a
// stable descending sort, walked from its end), moving each split's side-1 vertices to a new block until the
// block count reaches the target, and the blockmodel is rebuilt (initialize_edge_counts).
//
// The subgraphs are never materialised.  A vertex's sides in the ten proposals are one 10-bit mask, drawn in one
// pass over the vertices; one pass over the edges then counts, for every (block, proposal), the 2 x 2 block
// matrix of the two-block model, and the vertex pass its side sizes.  The description lengths of those small
// blockmodels are evaluated on the host from the counts.  Counting is wave-aggregated (lanes that hold the same
// block add by ballot and popcount) and, for few blocks, privatised in LDS per workgroup.
//
// Replay contract: with w = Philox(seed; b, epoch, TAG_SPLIT_SIZE, p / 4) word p % 4, proposal p of block b has
// size k = lo + uniform_below(w, hi - lo + 1), lo = (long)(0.2 n_b), hi = (long)(0.8 n_b); with
// w = Philox(seed; v, epoch, TAG_SPLIT_SIDE, p / 4) word p % 4, vertex v is on side 0 of proposal p when
// uniform_below(w, n_b) < k.  A block of fewer than two vertices has the one-block model (:191-197): every vertex on
// side 0 and, as its subgraph holds no edges (:114-116), E = 0.
#include <math.h>

#include <algorithm>
#include <limits>
#include <numeric>
#include <vector>

#include "carry.hpp"
#include "device_common.hpp"
#include "internal.hpp"
#include "primitives.hpp"

namespace sbpimpl {
using namespace sbpdev;

constexpr int SPLIT_PROPOSALS = 10;  // NUM_AGG_PROPOSALS_PER_BLOCK (include/top_down_sbp.hpp:28)
constexpr int SPLIT_THREADS = 256;
constexpr int SPLIT_LDS_BLOCKS = 256;  // blocks up to which the counters are privatised in LDS

static int split_grid(int64_t n, int cap) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + SPLIT_THREADS - 1) / SPLIT_THREADS, cap));
}

__device__ __forceinline__ uint32_t u4_word(const U4 &w, int i) {
    return i == 0 ? w.x : i == 1 ? w.y : i == 2 ? w.z : w.w;
}

// Split sizes k[b][p] of the blocks that are split (active[b]).
__global__ void k_split_sizes(int32_t B, const int32_t *__restrict__ sizes, const int32_t *__restrict__ active,
                              uint64_t seed, uint32_t epoch, int32_t *__restrict__ k) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)B * SPLIT_PROPOSALS;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = (int32_t)(i / SPLIT_PROPOSALS), p = (int32_t)(i % SPLIT_PROPOSALS);
        const long long n = sizes[b];
        if (!active[b] || n < 2) {
            k[i] = 0;
            continue;
        }
        const long long lo = (long long)((double)n * 0.2), hi = (long long)((double)n * 0.8);
        const U4 w = philox(seed, (uint32_t)b, epoch, TAG_SPLIT_SIZE, (uint32_t)(p >> 2));
        k[i] = (int32_t)(lo + (long long)uniform_below(u4_word(w, p & 3), (uint32_t)(hi - lo + 1)));
    }
}

// Adds x to counter i: privatised (LDS) or global.
template <bool LDS>
__device__ __forceinline__ void split_add(int32_t *lds, unsigned long long *glob, int64_t i, uint32_t x) {
    if (LDS) atomicAdd(&lds[i], (int32_t)x);
    else atomicAdd(&glob[i], (unsigned long long)x);
}

// Side masks of every vertex, and the side-1 count of every (block, proposal).  Lanes of a wave that hold the
// wave's first block add by popcount; the others one by one.
template <bool LDS>
__global__ __launch_bounds__(SPLIT_THREADS) void k_split_masks(int32_t N, int32_t B, const int32_t *__restrict__ assign,
                                                               const int32_t *__restrict__ sizes,
                                                               const int32_t *__restrict__ active,
                                                               const int32_t *__restrict__ k, uint64_t seed,
                                                               uint32_t epoch, uint16_t *__restrict__ mask,
                                                               unsigned long long *__restrict__ ones) {
    extern __shared__ int32_t lds[];  // LDS: ones [B * SPLIT_PROPOSALS]
    if (LDS) {
        for (int i = threadIdx.x; i < B * SPLIT_PROPOSALS; i += blockDim.x) lds[i] = 0;
        __syncthreads();
    }
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x - lane); base < N; base += stride) {
        const int64_t v = base + lane;
        int32_t b = -1;
        uint32_t m = 0;
        if (v < N) {
            const int32_t bv = assign[v];
            const int32_t n = sizes[bv];
            if (active[bv] && n >= 2) {
                b = bv;
                for (int q = 0; q < (SPLIT_PROPOSALS + 3) / 4; ++q) {
                    const U4 w = philox(seed, (uint32_t)v, epoch, TAG_SPLIT_SIDE, (uint32_t)q);
                    for (int j = 0; j < 4 && 4 * q + j < SPLIT_PROPOSALS; ++j) {
                        const int p = 4 * q + j;
                        if ((int32_t)uniform_below(u4_word(w, j), (uint32_t)n) >= k[bv * SPLIT_PROPOSALS + p])
                            m |= 1u << p;
                    }
                }
            }
            mask[v] = (uint16_t)m;
        }
        const unsigned long long in = __ballot(b >= 0);
        if (in == 0) continue;
        const int32_t b0 = __shfl(b, __ffsll((long long)in) - 1, WAVE);
        const bool grp = b == b0;
        for (int p = 0; p < SPLIT_PROPOSALS; ++p) {
            const uint32_t c = (uint32_t)__popcll(__ballot(grp && ((m >> p) & 1)));
            if (lane == 0 && c) split_add<LDS>(lds, ones, (int64_t)b0 * SPLIT_PROPOSALS + p, c);
        }
        if (b >= 0 && !grp)
            for (int p = 0; p < SPLIT_PROPOSALS; ++p)
                if ((m >> p) & 1) split_add<LDS>(lds, ones, (int64_t)b * SPLIT_PROPOSALS + p, 1);
    }
    if (LDS) {
        __syncthreads();
        for (int i = threadIdx.x; i < B * SPLIT_PROPOSALS; i += blockDim.x)
            if (lds[i]) atomicAdd(&ones[i], (unsigned long long)lds[i]);
    }
}

// The 2 x 2 block matrix (cell su * 2 + sv) of every (block, proposal) over the block's induced edges, and the
// induced edge count per block.
template <bool LDS>
__global__ __launch_bounds__(SPLIT_THREADS) void k_split_edges(int64_t E, int32_t B, const int32_t *__restrict__ src,
                                                               const int32_t *__restrict__ dst,
                                                               const int32_t *__restrict__ assign,
                                                               const int32_t *__restrict__ sizes,
                                                               const int32_t *__restrict__ active,
                                                               const uint16_t *__restrict__ mask,
                                                               unsigned long long *__restrict__ cells,
                                                               unsigned long long *__restrict__ edges) {
    extern __shared__ int32_t lds[];  // LDS: cells [B * SPLIT_PROPOSALS * 4], then edge counts [B]
    constexpr int NC = SPLIT_PROPOSALS * 4;
    if (LDS) {
        for (int i = threadIdx.x; i < B * (NC + 1); i += blockDim.x) lds[i] = 0;
        __syncthreads();
    }
    int32_t *lds_edges = lds + (LDS ? B * NC : 0);
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x - lane); base < E; base += stride) {
        const int64_t e = base + lane;
        int32_t b = -1;
        uint32_t mu = 0, mv = 0;
        if (e < E) {
            const int32_t u = src[e], v = dst[e];
            const int32_t bu = assign[u];
            if (assign[v] == bu && active[bu] && sizes[bu] >= 2) {
                b = bu;
                mu = mask[u];
                mv = mask[v];
            }
        }
        const unsigned long long in = __ballot(b >= 0);
        if (in == 0) continue;
        const int32_t b0 = __shfl(b, __ffsll((long long)in) - 1, WAVE);
        const bool grp = b == b0;
        const uint32_t ng = (uint32_t)__popcll(__ballot(grp));
        if (lane == 0) {
            if (LDS) atomicAdd(&lds_edges[b0], (int32_t)ng);
            else atomicAdd(&edges[b0], (unsigned long long)ng);
        }
        for (int p = 0; p < SPLIT_PROPOSALS; ++p) {
            const uint32_t su = (mu >> p) & 1, sv = (mv >> p) & 1;
            const uint32_t c01 = (uint32_t)__popcll(__ballot(grp && !su && sv));
            const uint32_t c10 = (uint32_t)__popcll(__ballot(grp && su && !sv));
            const uint32_t c11 = (uint32_t)__popcll(__ballot(grp && su && sv));
            if (lane == 0) {
                const int64_t at = ((int64_t)b0 * SPLIT_PROPOSALS + p) * 4;
                const uint32_t c00 = ng - c01 - c10 - c11;
                if (c00) split_add<LDS>(lds, cells, at + 0, c00);
                if (c01) split_add<LDS>(lds, cells, at + 1, c01);
                if (c10) split_add<LDS>(lds, cells, at + 2, c10);
                if (c11) split_add<LDS>(lds, cells, at + 3, c11);
            }
        }
        if (b >= 0 && !grp) {
            if (LDS) atomicAdd(&lds_edges[b], 1);
            else atomicAdd(&edges[b], 1ull);
            for (int p = 0; p < SPLIT_PROPOSALS; ++p) {
                const uint32_t su = (mu >> p) & 1, sv = (mv >> p) & 1;
                split_add<LDS>(lds, cells, ((int64_t)b * SPLIT_PROPOSALS + p) * 4 + su * 2 + sv, 1);
            }
        }
    }
    if (LDS) {
        __syncthreads();
        for (int i = threadIdx.x; i < B * NC; i += blockDim.x)
            if (lds[i]) atomicAdd(&cells[i], (unsigned long long)lds[i]);
        for (int i = threadIdx.x; i < B; i += blockDim.x)
            if (lds_edges[i]) atomicAdd(&edges[i], (unsigned long long)lds_edges[i]);
    }
}

// Side-1 vertices of the chosen proposal of each applied block move to its new block.
__global__ void k_split_apply(int32_t N, int32_t *__restrict__ assign, const uint16_t *__restrict__ mask,
                              const int32_t *__restrict__ choice, const int32_t *__restrict__ new_block) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const int32_t b = assign[v];
        const int32_t nb = new_block[b];
        if (nb >= 0 && ((mask[v] >> choice[b]) & 1)) assign[v] = nb;
    }
}

// ---------------------------------------------------------------- description lengths on the host

// entropy::nonparametric::mdl (src/entropy.cpp:740-767) of a blockmodel of nb blocks over a graph of n vertices
// and e edges, from its cells (row-major) and block sizes; directed, not degree-corrected.  sparse_entropy's cell
// and block terms are summed in 128-bit fixed point, as everywhere on this path.
static double np_mdl_small(int nb, const int64_t *cells, const int64_t *size, int64_t n, int64_t e) {
    sbpcarry::Fx128 acc;
    int64_t d_out[2] = {0, 0}, d_in[2] = {0, 0};
    for (int r = 0; r < nb; ++r)
        for (int c = 0; c < nb; ++c) {
            const int64_t w = cells[r * nb + c];
            if (w == 0) continue;
            acc.add(sbpcarry::np_eterm(w));
            d_out[r] += w;
            d_in[c] += w;
        }
    for (int b = 0; b < nb; ++b) acc.add(sbpcarry::np_vterm(d_out[b], d_in[b], size[b]));
    const double S = acc.value();
    int64_t ne = 0;
    for (int b = 0; b < nb; ++b) ne += size[b] > 0;
    double part = 0;  // get_partition_dl (src/entropy.cpp:657-669)
    part += host_lbinom(n - 1, ne - 1);
    part += host_lgamma(n + 1);
    for (int b = 0; b < nb; ++b) part -= host_lgamma(size[b] + 1);
    part += sbpcarry::fastlog(n);
    double S_dl = 0;
    S_dl += part;
    S_dl += 0.0;  // get_deg_dl_dist: 0 without degree correction
    S_dl += host_edges_dl(ne, e);
    return S + S_dl * 1.0;
}

// entropy::null_mdl_v1 (src/entropy.cpp:571-585) of a graph of n vertices and e edges: the one-block
// nonparametric model, or the closed form (NaN for e = 0, which then loses every comparison).
static double null_mdl_small(bool nonparametric, int64_t n, int64_t e) {
    if (nonparametric) return np_mdl_small(1, &e, &n, n, e);
    const double log_posterior_p = (double)e * log(1.0 / (double)e);
    const double x = 1.0 / (double)e;
    const double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return ((double)e * h) - log_posterior_p;
}

int32_t split_communities(sbp_ctx ctx, sbp_graph g, sbp_bm bm, int32_t target, uint32_t epoch,
                          std::vector<double> *delta_out) {
    SBP_CHECK(!bm->dc, SBP_E_ARG, "split_communities: degree-corrected description lengths are not supported");
    SBP_CHECK(bm->N == g->N, SBP_E_ARG, "split_communities: blockmodel and graph differ in vertex count");
    hipStream_t st = ctx->stream;
    const int32_t N = g->N, B = bm->B;
    const bool nonparametric = bm->np;
    SBP_CHECK((int64_t)B * SPLIT_PROPOSALS < (int64_t)1 << 31, SBP_E_ARG, "split_communities: too many blocks");
    std::vector<int32_t> sizes(B), active(B);
    SBP_HIP(hipMemcpyAsync(sizes.data(), bm->sizes, sizeof(int32_t) * B, hipMemcpyDeviceToHost, st));
    SBP_HIP(hipStreamSynchronize(st));
    for (int32_t b = 0; b < B; ++b) active[b] = !((double)sizes[b] < 0.001 * (double)N);  // :506
    const int64_t BP = (int64_t)B * SPLIT_PROPOSALS;
    // device work areas: counters (ones, cells, edges) in one 64-bit array; k, active, choice, new block in one
    // 32-bit array; the side masks
    const int64_t n_cnt = BP + 4 * BP + B;
    auto *cnt = ctx->scratch[S_COUNTS].as<unsigned long long>((size_t)n_cnt);
    unsigned long long *d_ones = cnt, *d_cells = cnt + BP, *d_edges = cnt + 5 * BP;
    int32_t *iv = ctx->scratch[S_MISC].as<int32_t>((size_t)(BP + 3 * B + 8));
    int32_t *d_k = iv, *d_active = iv + BP, *d_choice = d_active + B, *d_new = d_choice + B;
    auto *d_mask = ctx->scratch[S_ASSIGN_TMP].as<uint16_t>((size_t)std::max<int32_t>(N, 1));
    SBP_HIP(hipMemcpyAsync(d_active, active.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
    SBP_HIP(sbpprim::fill<unsigned long long>(cnt, n_cnt, 0ull, st));
    k_split_sizes<<<split_grid(BP, 4096), SPLIT_THREADS, 0, st>>>(B, bm->sizes, d_active, ctx->seed, epoch, d_k);
    const bool lds = B <= SPLIT_LDS_BLOCKS;
    if (N > 0) {
        const size_t shm = lds ? sizeof(int32_t) * (size_t)B * SPLIT_PROPOSALS : 0;
        const int grid = split_grid(N, lds ? 1024 : 16384);
        if (lds)
            k_split_masks<true><<<grid, SPLIT_THREADS, shm, st>>>(N, B, bm->assign, bm->sizes, d_active, d_k,
                                                                   ctx->seed, epoch, d_mask, d_ones);
        else
            k_split_masks<false><<<grid, SPLIT_THREADS, 0, st>>>(N, B, bm->assign, bm->sizes, d_active, d_k,
                                                                  ctx->seed, epoch, d_mask, d_ones);
    }
    if (g->E > 0) {
        const size_t shm = lds ? sizeof(int32_t) * (size_t)B * (4 * SPLIT_PROPOSALS + 1) : 0;
        const int grid = split_grid(g->E, lds ? 1024 : 16384);
        if (lds)
            k_split_edges<true><<<grid, SPLIT_THREADS, shm, st>>>(g->E, B, g->edge_src, g->out_adj, bm->assign,
                                                                   bm->sizes, d_active, d_mask, d_cells, d_edges);
        else
            k_split_edges<false><<<grid, SPLIT_THREADS, 0, st>>>(g->E, B, g->edge_src, g->out_adj, bm->assign,
                                                                  bm->sizes, d_active, d_mask, d_cells, d_edges);
    }
    SBP_HIP(hipGetLastError());
    std::vector<unsigned long long> h((size_t)n_cnt);
    SBP_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(unsigned long long) * n_cnt, hipMemcpyDeviceToHost, st));
    SBP_HIP(hipStreamSynchronize(st));
    // the best split of each block (split_communities' loop, src/top_down_sbp.cpp:503-526)
    const double DMAX = std::numeric_limits<double>::max();
    std::vector<double> delta(B, DMAX);
    std::vector<int32_t> choice(B, -1);
    for (int32_t b = 0; b < B; ++b) {
        if (!active[b]) continue;
        int64_t n = sizes[b];
        int64_t e = n < 2 ? 0 : (int64_t)h[5 * BP + b];
        const double old_entropy = null_mdl_small(nonparametric, n, e);
        for (int p = 0; p < SPLIT_PROPOSALS; ++p) {
            double new_entropy;
            if (n < 2) {
                new_entropy = np_mdl_small(1, &e, &n, n, e);
            } else {
                int64_t c[4], sz[2];
                for (int i = 0; i < 4; ++i) c[i] = (int64_t)h[BP + ((size_t)b * SPLIT_PROPOSALS + p) * 4 + i];
                sz[1] = (int64_t)h[(size_t)b * SPLIT_PROPOSALS + p];
                sz[0] = n - sz[1];
                new_entropy = np_mdl_small(2, c, sz, n, e);
            }
            const double d = new_entropy - old_entropy;
            if (d < delta[b]) {
                delta[b] = d;
                choice[b] = p;
            }
        }
    }
    if (delta_out) *delta_out = delta;
    // apply_best_splits (:30-66): utils::argsort is a stable sort by descending value, walked from its end; a
    // block that kept no split has the default Split, whose vertex count (LONG_MAX) exceeds N, and is skipped
    std::vector<int32_t> order(B);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return delta[x] > delta[y]; });
    std::vector<int32_t> new_block(B, -1);
    int32_t num_blocks = B;
    for (int32_t j = B - 1; j >= 0; --j) {
        const int32_t b = order[j];
        if (choice[b] < 0) continue;
        new_block[b] = num_blocks;
        num_blocks++;
        if (num_blocks >= target) break;
    }
    for (int32_t b = 0; b < B; ++b)
        if (choice[b] < 0) choice[b] = 0;
    if (num_blocks != B) {
        SBP_HIP(hipMemcpyAsync(d_choice, choice.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
        SBP_HIP(hipMemcpyAsync(d_new, new_block.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
        if (N > 0)
            k_split_apply<<<split_grid(N, 16384), SPLIT_THREADS, 0, st>>>(N, bm->assign, d_mask, d_choice, d_new);
        SBP_HIP(hipGetLastError());
        SBP_HIP(hipStreamSynchronize(st));  // the host vectors above must outlive the copies
    }
    // Blockmodel::initialize_edge_counts with the new block count
    bm_alloc(bm, num_blocks, N);
    bm_rebuild(ctx, g, bm);
    SBP_HIP(hipStreamSynchronize(st));
    return num_blocks;
}

}  // namespace sbpimpl
