// Blockmodel build: M[r][c] = #edges from block r to block c, its transpose, degrees and sizes.
//
// Replaces Blockmodel::initialize_edge_counts (src/blockmodel/blockmodel.cpp:193-276), where every
// OpenMP thread scans all N vertices and inserts into robin_map rows.  Here the edge list is turned
// into packed (row, col) keys, radix-sorted and run-length encoded (the library's own primitives,
// primitives.hpp); row pointers come from key boundaries, and the transpose from a second sort of the
// nnz unique cells.  The result is exact integer state, identical to the reference's after any sequence
// of incremental updates.
#include <algorithm>

#include "device_common.hpp"
#include "internal.hpp"
#include "primitives.hpp"

namespace sbpimpl {

static int bits_for(int64_t x) {  // number of bits to represent values in [0, x)
    int b = 0;
    while (b < 62 && (int64_t(1) << b) < x) ++b;
    return b < 1 ? 1 : b;
}
int bits_for_blocks(int64_t B) { return bits_for(B); }

__global__ void k_make_keys(int64_t E, const int32_t *__restrict__ src, const int32_t *__restrict__ dst,
                            const int32_t *__restrict__ assign, int cbits, unsigned long long *__restrict__ keys) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long r = (unsigned)assign[src[e]], c = (unsigned)assign[dst[e]];
        keys[e] = (r << cbits) | c;
    }
}

// One thread per unique cell: write cells, row pointers at key boundaries, and degree sums.
// Cells and row pointers from sorted unique (major, minor) keys and their counts.  The degrees are the row
// totals, taken afterwards by k_cum (same-address atomics here would serialize whole waves when B is small).
// dnnz: when given, nnz is read there (a recorded rebuild, launched for the largest nnz).
__global__ void k_fill_major(int64_t nnz, const int64_t *__restrict__ dnnz, int32_t B, int cbits,
                             const unsigned long long *__restrict__ uniq, const int32_t *__restrict__ counts,
                             int32_t *__restrict__ ptr, int2 *__restrict__ cells) {
    if (dnnz) nnz = *dnnz;
    const unsigned long long mask = (1ull << cbits) - 1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t a = (int32_t)(uniq[i] >> cbits), b = (int32_t)(uniq[i] & mask), cnt = counts[i];
        cells[i] = make_int2(b, cnt);
        int32_t prev = i == 0 ? -1 : (int32_t)(uniq[i - 1] >> cbits);
        for (int32_t x = prev + 1; x <= a; ++x) ptr[x] = (int32_t)i;
        if (i == nnz - 1)
            for (int32_t x = a + 1; x <= B; ++x) ptr[x] = (int32_t)nnz;
    }
    if (nnz == 0 && blockIdx.x == 0)  // no cells: every row empty
        for (int32_t x = threadIdx.x; x <= B; x += blockDim.x) ptr[x] = 0;
}

__global__ void k_transpose_keys(int64_t nnz, const int64_t *__restrict__ dnnz, int cbits,
                                 const unsigned long long *__restrict__ uniq, unsigned long long *__restrict__ tkeys) {
    if (dnnz) nnz = *dnnz;
    const unsigned long long mask = (1ull << cbits) - 1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long r = uniq[i] >> cbits, c = uniq[i] & mask;
        tkeys[i] = (c << cbits) | r;
    }
}

__global__ void k_dtot_sizes(int32_t B, const int32_t *__restrict__ row_ptr, const int2 *__restrict__ row_cells,
                             const int32_t *__restrict__ d_out, const int32_t *__restrict__ d_in,
                             int32_t *__restrict__ d_tot) {
    for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        int self = sbpdev::cell_lookup(row_cells, row_ptr[b], row_ptr[b + 1], b);
        d_tot[b] = d_out[b] + d_in[b] - self;
    }
}

__global__ void k_sizes(int32_t N, const int32_t *__restrict__ assign, int32_t *__restrict__ sizes) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        atomicAdd(&sizes[assign[v]], 1);
}

// Block sizes through a histogram in LDS per workgroup (B <= SIZES_LDS_MAX): few blocks means heavy contention
// on the global counters.
constexpr int32_t SIZES_LDS_MAX = 16384;
__global__ void __launch_bounds__(256) k_sizes_lds(int32_t N, int32_t B, const int32_t *__restrict__ assign,
                                                   int32_t *__restrict__ sizes) {
    __shared__ int32_t hist[SIZES_LDS_MAX];
    for (int32_t b = threadIdx.x; b < B; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        atomicAdd(&hist[assign[v]], 1);
    __syncthreads();
    for (int32_t b = threadIdx.x; b < B; b += blockDim.x)
        if (hist[b]) atomicAdd(&sizes[b], hist[b]);
}

static int grid_for(int64_t n, int block = 256) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

// Blockmodel buffers come from the context's pool (recycled across the blockmodels of a run).
static DevicePool &pool_of(sbp_bm bm) {
    if (!bm->pool) bm->pool = bm->ctx->pool;
    return *bm->pool;
}

void bm_alloc(sbp_bm bm, int32_t B, int32_t N) {
    DevicePool &pool = pool_of(bm);
    if (bm->N != N || bm->assign == nullptr) {
        pool.release(bm->assign);
        bm->assign = static_cast<int32_t *>(pool.alloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1)));
        bm->N = N;
    }
    if (B + 1 > bm->bcap) {
        int32_t **arrs[] = {&bm->sizes, &bm->d_out, &bm->d_in, &bm->d_tot, &bm->row_ptr, &bm->col_ptr};
        for (auto a : arrs) {
            pool.release(*a);
            *a = static_cast<int32_t *>(pool.alloc(sizeof(int32_t) * (size_t)(B + 2)));
        }
        bm->bcap = B + 1;
    }
    bm->B = B;
}

void bm_release(sbp_bm bm) {
    void *ptrs[] = {bm->assign,    bm->sizes,     bm->d_out,   bm->d_in,    bm->d_tot,    bm->row_ptr,
                    bm->col_ptr,   bm->row_cells, bm->col_cells, bm->row_cum, bm->col_cum, bm->cell_hash,
                    bm->hist,      bm->row_keys,  bm->col_keys, bm->dense};
    if (bm->pool) {
        for (void *p : ptrs) bm->pool->release(p);
    } else {
        for (void *p : ptrs)
            if (p) (void)dev_free(p);
    }
    bm->cell_hash = nullptr;
    bm->hash_valid = bm->dense_valid = false;
    bm->dense = nullptr;
    bm->dense_B = 0;
    bm->hist = nullptr;
    bm->hist_log2 = 0;
    bm->hist_ready = false;
    bm->assign = bm->sizes = bm->d_out = bm->d_in = bm->d_tot = bm->row_ptr = bm->col_ptr = nullptr;
    bm->row_cells = bm->col_cells = nullptr;
    bm->row_cum = bm->col_cum = nullptr;
    bm->row_keys = bm->col_keys = nullptr;
    bm->keys_valid = false;
    bm->cap = 0;
    bm->bcap = 0;
}

void bm_ensure_cells(sbp_bm bm, int64_t nnz) {
    if (nnz > bm->cap) {
        DevicePool &pool = pool_of(bm);
        void *old[] = {bm->row_cells, bm->col_cells, bm->row_cum, bm->col_cum, bm->row_keys, bm->col_keys};
        for (void *p : old) pool.release(p);
        int64_t cap = nnz + nnz / 8 + 64;
        bm->row_cells = static_cast<int2 *>(pool.alloc(sizeof(int2) * (size_t)cap));
        bm->col_cells = static_cast<int2 *>(pool.alloc(sizeof(int2) * (size_t)cap));
        bm->row_cum = static_cast<int32_t *>(pool.alloc(sizeof(int32_t) * (size_t)cap));
        bm->col_cum = static_cast<int32_t *>(pool.alloc(sizeof(int32_t) * (size_t)cap));
        bm->row_keys = static_cast<unsigned long long *>(pool.alloc(sizeof(unsigned long long) * (size_t)cap));
        bm->col_keys = static_cast<unsigned long long *>(pool.alloc(sizeof(unsigned long long) * (size_t)cap));
        bm->keys_valid = false;
        bm->cap = cap;
    }
}
static void ensure_cells(sbp_bm bm, int64_t nnz) { bm_ensure_cells(bm, nnz); }

// Inclusive cumulative counts per row (and per column without the diagonal cell): the multinomial over
// neighbors_weights(t) then selects by binary search instead of a scan.  One wavefront per row.
// Inclusive cumulative counts per row (the diagonal cell counted as 0 when skip_diag) and, when deg is given,
// each row's total including the diagonal (0 for an empty row).
__global__ void k_cum(int32_t B, const int32_t *__restrict__ ptr, const int2 *__restrict__ cells, int skip_diag,
                      int32_t *__restrict__ cum, int32_t *__restrict__ deg) {
    const int lane = sbpdev::lane_id();
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / sbpdev::WAVE);
    for (int64_t b = (int64_t)blockIdx.x * (blockDim.x / sbpdev::WAVE) + threadIdx.x / sbpdev::WAVE; b < B;
         b += waves) {
        int base = 0, diag = 0;
        for (int c0 = ptr[b]; c0 < ptr[b + 1]; c0 += sbpdev::WAVE) {
            int i = c0 + lane;
            int x = 0;
            if (i < ptr[b + 1]) {
                x = cells[i].y;
                if (skip_diag && cells[i].x == (int)b) {
                    diag += x;
                    x = 0;
                }
            }
            int incl = sbpdev::wave_inclusive_scan(x) + base;
            if (i < ptr[b + 1]) cum[i] = incl;
            base = __shfl(incl, sbpdev::WAVE - 1, sbpdev::WAVE);
        }
        if (deg) {
            diag = sbpdev::wave_sum(diag);
            if (lane == 0) deg[b] = base + diag;
        }
    }
}

static void launch_cum(sbp_ctx ctx, sbp_bm bm, int32_t *d_out, int32_t *d_in) {
    int64_t blocks = (bm->B + 3) / 4;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
    k_cum<<<(int)blocks, 256, 0, ctx->stream>>>(bm->B, bm->row_ptr, bm->row_cells, 0, bm->row_cum, d_out);
    k_cum<<<(int)blocks, 256, 0, ctx->stream>>>(bm->B, bm->col_ptr, bm->col_cells, 1, bm->col_cum, d_in);
    SBP_HIP(hipGetLastError());
}

void bm_compute_cum(sbp_ctx ctx, sbp_bm bm) {
    if (bm->nnz == 0) return;
    launch_cum(ctx, bm, nullptr, nullptr);
}

// Cell hash: one wavefront per row inserts its cells; the table holds at least twice nnz slots.
__global__ void k_hash_cells(int32_t B, const int32_t *__restrict__ ptr, const int2 *__restrict__ cells,
                             sbpdev::CellSlot *__restrict__ table, int shift, unsigned mask) {
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / sbpdev::WAVE);
    for (int64_t b = (int64_t)blockIdx.x * (blockDim.x / sbpdev::WAVE) + threadIdx.x / sbpdev::WAVE; b < B;
         b += waves) {
        for (int i = ptr[b] + sbpdev::lane_id(); i < ptr[b + 1]; i += sbpdev::WAVE) {
            const int2 cell = cells[i];
            const unsigned long long key = sbpdev::cell_key((int)b, cell.x);
            for (unsigned h = sbpdev::cell_slot(key, shift);; h = (h + 1) & mask) {
                if (atomicCAS(&table[h].key, sbpdev::CELL_EMPTY, key) == sbpdev::CELL_EMPTY) {
                    table[h].val = cell.y;
                    break;
                }
            }
        }
    }
}

// At least three slots per cell: the update after moves inserts new cells in place (and keeps emptied ones) until
// the load passes 0.7, so the room above the initial load of at most 1/3 sets how many sweeps run between rebuilds.
static int hash_log2_for(int64_t nnz) {
    int lg = 10;
    while ((int64_t(1) << lg) < 3 * nnz + 2) ++lg;
    return lg;
}

void bm_ensure_hash(sbp_ctx ctx, sbp_bm bm, bool capture) {
    if (bm->hash_valid && !capture) return;
    hipStream_t st = ctx->stream;
    const int lg_need = hash_log2_for(bm->nnz);
    if (!bm->cell_hash || bm->hash_log2 < lg_need) {
        SBP_CHECK(!capture, SBP_E_STATE, "cell hash not reserved for a recorded sweep");
        DevicePool &pool = pool_of(bm);
        pool.release(bm->cell_hash);
        bm->cell_hash = static_cast<sbpdev::CellSlot *>(pool.alloc(sizeof(sbpdev::CellSlot) << lg_need));
        bm->hash_log2 = lg_need;
    }
    const int lg = bm->hash_log2;
    // every key to CELL_EMPTY (all ones), with a kernel: no memset nodes in a recorded graph
    SBP_HIP(sbpprim::fill<unsigned long long>(reinterpret_cast<unsigned long long *>(bm->cell_hash),
                                              int64_t(2) << lg, ~0ull, st));
    if (bm->nnz > 0 || capture) {
        int64_t blocks = std::min<int64_t>((bm->B + 3) / 4, 16384);
        k_hash_cells<<<(int)std::max<int64_t>(blocks, 1), 256, 0, st>>>(bm->B, bm->row_ptr, bm->row_cells, bm->cell_hash,
                                                                       64 - lg, (1u << lg) - 1);
        SBP_HIP(hipGetLastError());
    }
    // a recorded hash is rebuilt by each replay; outside the graph the matrix it describes has changed
    bm->hash_valid = !capture;
    bm->hash_used = bm->nnz;
}

// Dense block matrix from the compact CSR: one wavefront per row writes the row's cells (the array is zeroed first).
__global__ void k_dense_cells(int32_t B, const int32_t *__restrict__ row_ptr, const int2 *__restrict__ row_cells,
                              int32_t *__restrict__ dense) {
    const int lane = threadIdx.x & 63;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; r < B;
         r += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        int32_t *row = dense + r * (int64_t)B;
        for (int32_t i = row_ptr[r] + lane; i < row_ptr[r + 1]; i += 64) {
            const int2 c = row_cells[i];
            row[c.x] = c.y;
        }
    }
}

// Largest block count with a dense copy (256 MB of counts), and the smallest graph for which it pays for its
// build on each rebuild.
constexpr int32_t DENSE_MAX_BLOCKS = 8192;
constexpr int64_t DENSE_MIN_EDGES = int64_t(1) << 22;

const int32_t *bm_ensure_dense(sbp_ctx ctx, sbp_bm bm, sbp_graph g, bool any_graph) {
    const char *e = getenv("SBP_AMD_DENSE");
    const bool off = e && e[0] == '0', forced = (e && e[0] == '1') || any_graph;
    if (off || bm->B > DENSE_MAX_BLOCKS || bm->B < 1 || (!forced && g->E < DENSE_MIN_EDGES)) return nullptr;
    if (bm->dense_valid && bm->dense_B == bm->B) return bm->dense;
    hipStream_t st = ctx->stream;
    const int64_t n = (int64_t)bm->B * bm->B;
    if (!bm->dense || (int64_t)bm->dense_B * bm->dense_B < n) {
        DevicePool &pool = pool_of(bm);
        pool.release(bm->dense);
        bm->dense = static_cast<int32_t *>(pool.alloc(sizeof(int32_t) * (size_t)n));
    }
    bm->dense_B = bm->B;
    SBP_HIP(sbpprim::fill<int32_t>(bm->dense, n, 0, st));
    if (bm->nnz > 0) {
        const int64_t blocks = std::min<int64_t>((bm->B + 3) / 4, 16384);
        k_dense_cells<<<(int)std::max<int64_t>(blocks, 1), 256, 0, st>>>(bm->B, bm->row_ptr, bm->row_cells, bm->dense);
        SBP_HIP(hipGetLastError());
    }
    bm->dense_valid = true;
    return bm->dense;
}

__global__ void k_hist_clear(int64_t n, sbpdev::HistSlot *__restrict__ t) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        t[i] = sbpdev::HistSlot{sbpdev::HIST_EMPTY, 0, 0};
}
__global__ void k_hist_build(int32_t N, const int32_t *__restrict__ assign, const int64_t *__restrict__ out_off,
                             const int64_t *__restrict__ in_off, sbpdev::HistSlot *__restrict__ t, int shift,
                             uint32_t mask) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const int b = assign[v];
        sbpdev::hist_add(t, shift, mask, sbpdev::hist_key(b, 0, out_off[v + 1] - out_off[v]), 1);
        sbpdev::hist_add(t, shift, mask, sbpdev::hist_key(b, 1, in_off[v + 1] - in_off[v]), 1);
    }
}

// Table size: at least 8N slots.  A table holds at most 4N live entries (non-zero counts: one per vertex and
// direction, plus bins left below zero) and one sweep opens at most 2N keys, so compacting when half of the
// slots are claimed keeps it under three quarters full.
static int hist_log2_for(int64_t N) {
    int lg = 10;
    while ((int64_t(1) << lg) < 8 * N + 16) ++lg;
    return lg;
}
static sbpdev::HistSlot *hist_alloc(sbp_bm bm, int lg) {
    return static_cast<sbpdev::HistSlot *>(pool_of(bm).alloc(sizeof(sbpdev::HistSlot) * ((size_t(1) << lg) + 1)));
}

__global__ void k_hist_compact(int64_t n, const sbpdev::HistSlot *__restrict__ from, sbpdev::HistSlot *__restrict__ to,
                               int shift, uint32_t mask) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const sbpdev::HistSlot e = from[i];
        if (e.key != sbpdev::HIST_EMPTY && e.val != 0) sbpdev::hist_add(to, shift, mask, e.key, e.val);
    }
}

void bm_hist_after_moves(sbp_ctx ctx, sbp_bm bm, int32_t claimed, int32_t full) {
    SBP_CHECK(full == 0, SBP_E_STATE, "degree histogram table full");
    const int64_t n = int64_t(1) << bm->hist_log2;
    if (claimed <= n / 2) return;
    // entries at zero count are dropped: a lookup of an absent bin gives 0
    hipStream_t st = ctx->stream;
    sbpdev::HistSlot *fresh = hist_alloc(bm, bm->hist_log2);
    k_hist_clear<<<grid_for(n + 1), 256, 0, st>>>(n + 1, fresh);
    k_hist_compact<<<grid_for(n), 256, 0, st>>>(n, bm->hist, fresh, 64 - bm->hist_log2, (uint32_t)(n - 1));
    SBP_HIP(hipGetLastError());
    pool_of(bm).release(bm->hist);
    bm->hist = fresh;
}

void bm_build_hist(sbp_ctx ctx, sbp_graph g, sbp_bm bm) {
    hipStream_t st = ctx->stream;
    const int lg = hist_log2_for(bm->N);
    if (!bm->hist || bm->hist_log2 != lg) {
        pool_of(bm).release(bm->hist);
        bm->hist = hist_alloc(bm, lg);
        bm->hist_log2 = lg;
    }
    const int64_t n = int64_t(1) << lg;
    k_hist_clear<<<grid_for(n + 1), 256, 0, st>>>(n + 1, bm->hist);
    k_hist_build<<<grid_for(bm->N), 256, 0, st>>>(bm->N, bm->assign, g->out_off, g->in_off, bm->hist, 64 - lg,
                                                  (uint32_t)(n - 1));
    SBP_HIP(hipGetLastError());
    bm->hist_ready = true;
}

void bm_set_model(sbp_ctx ctx, sbp_graph g, sbp_bm bm, int np, int dc) {
    bm->np = np ? 1 : 0;
    bm->dc = (np && dc) ? 1 : 0;
    if (bm->dc && !bm->hist_ready) bm_build_hist(ctx, g, bm);
}

// Largest block matrix of B blocks over E edges.
static int64_t max_cells(int64_t B, int64_t E) { return std::min<int64_t>(E, B * B); }

void bm_reserve_cells(sbp_ctx ctx, sbp_graph g, sbp_bm bm) {
    const int64_t bound = std::max<int64_t>(max_cells(bm->B, g->E), 1);
    if (bm->cap < bound) {  // grow, keeping the current cells (ensure_cells does not)
        DevicePool &pool = pool_of(bm);
        int2 *rc = static_cast<int2 *>(pool.alloc(sizeof(int2) * (size_t)bound));
        int2 *cc = static_cast<int2 *>(pool.alloc(sizeof(int2) * (size_t)bound));
        int32_t *rcum = static_cast<int32_t *>(pool.alloc(sizeof(int32_t) * (size_t)bound));
        int32_t *ccum = static_cast<int32_t *>(pool.alloc(sizeof(int32_t) * (size_t)bound));
        if (bm->nnz > 0) {
            hipStream_t st = ctx->stream;
            SBP_HIP(hipMemcpyAsync(rc, bm->row_cells, sizeof(int2) * bm->nnz, hipMemcpyDeviceToDevice, st));
            SBP_HIP(hipMemcpyAsync(cc, bm->col_cells, sizeof(int2) * bm->nnz, hipMemcpyDeviceToDevice, st));
            SBP_HIP(hipMemcpyAsync(rcum, bm->row_cum, sizeof(int32_t) * bm->nnz, hipMemcpyDeviceToDevice, st));
            SBP_HIP(hipMemcpyAsync(ccum, bm->col_cum, sizeof(int32_t) * bm->nnz, hipMemcpyDeviceToDevice, st));
        }
        void *old[] = {bm->row_cells, bm->col_cells, bm->row_cum, bm->col_cum, bm->row_keys, bm->col_keys};
        for (void *p : old) pool.release(p);  // reuse is ordered by the stream, after the copies
        bm->row_cells = rc;
        bm->col_cells = cc;
        bm->row_cum = rcum;
        bm->col_cum = ccum;
        bm->row_keys = static_cast<unsigned long long *>(pool.alloc(sizeof(unsigned long long) * (size_t)bound));
        bm->col_keys = static_cast<unsigned long long *>(pool.alloc(sizeof(unsigned long long) * (size_t)bound));
        bm->keys_valid = false;
        bm->cap = bound;
        bm->hash_valid = bm->dense_valid = false;
    }
    const int lg = hash_log2_for(bound);
    if (!bm->cell_hash || bm->hash_log2 < lg) {
        DevicePool &pool = pool_of(bm);
        pool.release(bm->cell_hash);
        bm->cell_hash = static_cast<sbpdev::CellSlot *>(pool.alloc(sizeof(sbpdev::CellSlot) << lg));
        bm->hash_log2 = lg;
        bm->hash_valid = bm->dense_valid = false;
    }
    // the rebuild's scratch at its largest: keys of every edge, the transposed sort of every cell
    const int64_t E = std::max<int64_t>(g->E, 1);
    ctx->scratch[S_KEYS_A].as<unsigned long long>(E);
    ctx->scratch[S_KEYS_B].as<unsigned long long>(E);
    ctx->scratch[S_UNIQ].as<unsigned long long>(E);
    ctx->scratch[S_COUNTS].as<int32_t>(E);
    ctx->scratch[S_VALS_B].as<int32_t>(bound);
    ctx->scratch[S_MISC].as<int64_t>(4);
    const int cbits = bits_for(bm->B);
    size_t a = 0, b = 0, c = 0;
    const int32_t *no_vals = nullptr;
    unsigned long long *k = nullptr;
    SBP_HIP(sbpprim::radix_sort(nullptr, a, (const unsigned long long *)k, k, no_vals, (int32_t *)nullptr, E, 0,
                                2 * cbits, ctx->stream));
    SBP_HIP(sbpprim::run_length_encode(nullptr, b, (const unsigned long long *)k, k, (int32_t *)nullptr,
                                       (int64_t *)nullptr, E, ctx->stream));
    SBP_HIP(sbpprim::radix_sort(nullptr, c, (const unsigned long long *)k, k, no_vals, (int32_t *)nullptr, bound, 0,
                                2 * cbits, ctx->stream));
    ctx->scratch[S_SORT_TMP].get(std::max(a, std::max(b, c)));
}

void bm_sync_nnz(sbp_ctx ctx, sbp_bm bm) {
    int32_t nnz = 0;
    read_small(ctx, {{&nnz, bm->row_ptr + bm->B, sizeof(int32_t)}});
    bm->nnz = nnz;
}

void bm_rebuild(sbp_ctx ctx, sbp_graph g, sbp_bm bm, bool keep_hist, bool capture) {
    hipStream_t st = ctx->stream;
    bm->hash_valid = bm->dense_valid = false;
    SBP_CHECK(!capture || !bm->dc, SBP_E_STATE, "recorded rebuild: degree-corrected histograms need host bookkeeping");
    if (bm->dc && !keep_hist) bm_build_hist(ctx, g, bm);
    if (!bm->dc) bm->hist_ready = false;
    const int32_t B = bm->B;
    const int64_t E = g->E;
    bm->E = E;
    const int cbits = bits_for(B);
    const int end_bit = 2 * cbits;
    SBP_CHECK(end_bit <= 64, SBP_E_ARG, "block count too large for packed keys");
    bm->keys_valid = false;
    bm_count_sizes(ctx, bm);
    if (E == 0) {
        SBP_HIP(sbpprim::fill<int32_t>(bm->d_out, B, 0, st));
        SBP_HIP(sbpprim::fill<int32_t>(bm->d_in, B, 0, st));
        SBP_HIP(sbpprim::fill<int32_t>(bm->row_ptr, B + 1, 0, st));
        SBP_HIP(sbpprim::fill<int32_t>(bm->col_ptr, B + 1, 0, st));
        SBP_HIP(sbpprim::fill<int32_t>(bm->d_tot, B, 0, st));
        bm->nnz = 0;
        return;
    }
    auto *keys = ctx->scratch[S_KEYS_A].as<unsigned long long>(E);
    auto *sorted = ctx->scratch[S_KEYS_B].as<unsigned long long>(E);
    auto *uniq = ctx->scratch[S_UNIQ].as<unsigned long long>(E);
    auto *counts = ctx->scratch[S_COUNTS].as<int32_t>(E);
    auto *num_runs = ctx->scratch[S_MISC].as<int64_t>(4);
    k_make_keys<<<grid_for(E), 256, 0, st>>>(E, g->edge_src, g->out_adj, bm->assign, cbits, keys);
    SBP_HIP(hipGetLastError());
    size_t tmp_a = 0, tmp_b = 0;
    const int32_t *no_vals = nullptr;
    SBP_HIP(sbpprim::radix_sort(nullptr, tmp_a, keys, sorted, no_vals, (int32_t *)nullptr, E, 0, end_bit, st));
    SBP_HIP(sbpprim::run_length_encode(nullptr, tmp_b, sorted, uniq, counts, num_runs, E, st));
    size_t tmp = tmp_a > tmp_b ? tmp_a : tmp_b;
    void *tmp_ptr = ctx->scratch[S_SORT_TMP].get(tmp);
    SBP_HIP(sbpprim::radix_sort(tmp_ptr, tmp_a, keys, sorted, no_vals, (int32_t *)nullptr, E, 0, end_bit, st));
    SBP_HIP(sbpprim::run_length_encode(tmp_ptr, tmp_b, sorted, uniq, counts, num_runs, E, st));
    // nnz: read back (eager), or left on the device with the launches sized for the largest block matrix (recorded)
    int64_t nnz = 0;
    const int64_t *dnnz = nullptr;
    if (capture) {
        nnz = max_cells(B, E);
        SBP_CHECK(bm->cap >= nnz, SBP_E_STATE, "cells not reserved for a recorded rebuild");
        dnnz = num_runs;
    } else {
        read_small(ctx, {{&nnz, num_runs, sizeof(int64_t)}});
        ensure_cells(bm, nnz);
    }
    bm->nnz = nnz;
    k_fill_major<<<grid_for(nnz), 256, 0, st>>>(nnz, dnnz, B, cbits, uniq, counts, bm->row_ptr, bm->row_cells);
    SBP_HIP(hipGetLastError());
    // transpose: sort (col, row) keys carrying the counts
    auto *tkeys = keys;  // reuse
    auto *tkeys_sorted = sorted;
    auto *tvals_sorted = ctx->scratch[S_VALS_B].as<int32_t>(nnz);
    k_transpose_keys<<<grid_for(nnz), 256, 0, st>>>(nnz, dnnz, cbits, uniq, tkeys);
    size_t tmp_c = 0;
    SBP_HIP(sbpprim::radix_sort(nullptr, tmp_c, tkeys, tkeys_sorted, (const int32_t *)counts, tvals_sorted, nnz, 0,
                                end_bit, st, dnnz));
    tmp_ptr = ctx->scratch[S_SORT_TMP].get(tmp_c);
    SBP_HIP(sbpprim::radix_sort(tmp_ptr, tmp_c, tkeys, tkeys_sorted, (const int32_t *)counts, tvals_sorted, nnz, 0,
                                end_bit, st, dnnz));
    k_fill_major<<<grid_for(nnz), 256, 0, st>>>(nnz, dnnz, B, cbits, tkeys_sorted, tvals_sorted, bm->col_ptr,
                                                 bm->col_cells);
    launch_cum(ctx, bm, bm->d_out, bm->d_in);  // cumulative counts, and the degrees as row / column totals
    k_dtot_sizes<<<grid_for(B), 256, 0, st>>>(B, bm->row_ptr, bm->row_cells, bm->d_out, bm->d_in, bm->d_tot);
    SBP_HIP(hipGetLastError());
    if (!capture && nnz > 0) {  // the sorted keys, for the update after moves
        SBP_HIP(hipMemcpyAsync(bm->row_keys, uniq, sizeof(unsigned long long) * nnz, hipMemcpyDeviceToDevice, st));
        SBP_HIP(hipMemcpyAsync(bm->col_keys, tkeys_sorted, sizeof(unsigned long long) * nnz, hipMemcpyDeviceToDevice,
                               st));
        bm->keys_cbits = cbits;
        bm->keys_valid = true;
    }
}

void bm_count_sizes(sbp_ctx ctx, sbp_bm bm) {
    hipStream_t st = ctx->stream;
    const int32_t B = bm->B;
    SBP_HIP(sbpprim::fill<int32_t>(bm->sizes, B, 0, st));
    if (B <= SIZES_LDS_MAX) k_sizes_lds<<<std::min(grid_for(bm->N), 512), 256, 0, st>>>(bm->N, B, bm->assign, bm->sizes);
    else k_sizes<<<grid_for(bm->N), 256, 0, st>>>(bm->N, bm->assign, bm->sizes);
    SBP_HIP(hipGetLastError());
}

// The CSR and CSC, cumulative counts and degrees from sorted row-major and column-major keys with their counts
// (bm->nnz of each; capacity already ensured); the keys are kept for the next update.
void bm_fill_cells(sbp_ctx ctx, sbp_bm bm, const unsigned long long *row_keys, const int32_t *row_counts,
                   const unsigned long long *col_keys, const int32_t *col_counts) {
    hipStream_t st = ctx->stream;
    const int32_t B = bm->B;
    const int64_t nnz = bm->nnz;
    const int cbits = bm->keys_cbits;
    k_fill_major<<<grid_for(nnz), 256, 0, st>>>(nnz, nullptr, B, cbits, row_keys, row_counts, bm->row_ptr,
                                                 bm->row_cells);
    k_fill_major<<<grid_for(nnz), 256, 0, st>>>(nnz, nullptr, B, cbits, col_keys, col_counts, bm->col_ptr,
                                                 bm->col_cells);
    if (nnz > 0) {
        SBP_HIP(hipMemcpyAsync(bm->row_keys, row_keys, sizeof(unsigned long long) * nnz, hipMemcpyDeviceToDevice, st));
        SBP_HIP(hipMemcpyAsync(bm->col_keys, col_keys, sizeof(unsigned long long) * nnz, hipMemcpyDeviceToDevice, st));
    }
    launch_cum(ctx, bm, bm->d_out, bm->d_in);
    k_dtot_sizes<<<grid_for(B), 256, 0, st>>>(B, bm->row_ptr, bm->row_cells, bm->d_out, bm->d_in, bm->d_tot);
    SBP_HIP(hipGetLastError());
    bm->keys_valid = true;
}

void bm_copy_into(sbp_ctx ctx, sbp_bm src, sbp_bm dst) {
    hipStream_t st = ctx->stream;
    bm_alloc(dst, src->B, src->N);
    ensure_cells(dst, src->nnz > 0 ? src->nnz : 1);
    dst->nnz = src->nnz;
    const int32_t B = src->B;
    SBP_HIP(hipMemcpyAsync(dst->assign, src->assign, sizeof(int32_t) * src->N, hipMemcpyDeviceToDevice, st));
    SBP_HIP(hipMemcpyAsync(dst->sizes, src->sizes, sizeof(int32_t) * B, hipMemcpyDeviceToDevice, st));
    SBP_HIP(hipMemcpyAsync(dst->d_out, src->d_out, sizeof(int32_t) * B, hipMemcpyDeviceToDevice, st));
    SBP_HIP(hipMemcpyAsync(dst->d_in, src->d_in, sizeof(int32_t) * B, hipMemcpyDeviceToDevice, st));
    SBP_HIP(hipMemcpyAsync(dst->d_tot, src->d_tot, sizeof(int32_t) * B, hipMemcpyDeviceToDevice, st));
    SBP_HIP(hipMemcpyAsync(dst->row_ptr, src->row_ptr, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToDevice, st));
    SBP_HIP(hipMemcpyAsync(dst->col_ptr, src->col_ptr, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToDevice, st));
    if (src->nnz > 0) {
        SBP_HIP(hipMemcpyAsync(dst->row_cells, src->row_cells, sizeof(int2) * src->nnz, hipMemcpyDeviceToDevice, st));
        SBP_HIP(hipMemcpyAsync(dst->col_cells, src->col_cells, sizeof(int2) * src->nnz, hipMemcpyDeviceToDevice, st));
        SBP_HIP(hipMemcpyAsync(dst->row_cum, src->row_cum, sizeof(int32_t) * src->nnz, hipMemcpyDeviceToDevice, st));
        SBP_HIP(hipMemcpyAsync(dst->col_cum, src->col_cum, sizeof(int32_t) * src->nnz, hipMemcpyDeviceToDevice, st));
    }
    dst->overall_entropy = src->overall_entropy;
    dst->num_to_merge = 0;
    dst->hash_valid = dst->dense_valid = false;
    dst->keys_valid = false;
    dst->np = src->np;
    dst->dc = src->dc;
    dst->hist_ready = false;
    if (src->dc && src->hist_ready) {
        if (!dst->hist || dst->hist_log2 != src->hist_log2) {
            pool_of(dst).release(dst->hist);
            dst->hist = hist_alloc(dst, src->hist_log2);
            dst->hist_log2 = src->hist_log2;
        }
        SBP_HIP(hipMemcpyAsync(dst->hist, src->hist, sizeof(sbpdev::HistSlot) * ((size_t(1) << src->hist_log2) + 1),
                               hipMemcpyDeviceToDevice, st));
        dst->hist_ready = true;
    }
    dst->E = src->E;
}

}  // namespace sbpimpl
