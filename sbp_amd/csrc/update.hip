// Block matrix after committed vertex moves, updated in place of a rebuild.
//
// Blockmodel::move_vertex (src/blockmodel/blockmodel.cpp:408-494) changes the cells of the moved vertices' edges
// only; the rebuild recounts every edge (a sort of E keys).  After an asynchronous-Gibbs commit that moved m
// vertices, the cells change by the edges of those vertices alone: for an out-edge (v, u) of a moved v, cell
// (old(v), old(u)) loses one and (new(v), new(u)) gains one; for an in-edge (u, v) from a vertex u that did not
// move, (b(u), old(v)) loses one and (b(u), new(v)) gains one (an edge between two moved vertices is one of the
// out-edges).  The signed deltas are sorted and summed per cell, and the sorted row-major and column-major cell
// lists are merged with them: a cell whose count reaches zero leaves the list, a cell that was absent enters it
// at its rank.  The result is the same CSR and CSC the rebuild produces (both list each row's / column's
// non-zero cells in ascending order), so everything downstream is unchanged.  The cell hash, when valid, is
// updated in place (a cell that reaches zero keeps its key with count 0, which reads as absent).
//
// The update keeps, per blockmodel, the sorted cell keys (row << cbits | col and col << cbits | row) that the
// rebuild produces anyway.  It is used by eager sweeps of graphs of at least 2^22 edges; when the moved vertices
// carry more than a quarter of the edges, or the graph is smaller, the rebuild is cheaper and runs instead.
#include <stdlib.h>

#include <algorithm>

#include "device_common.hpp"
#include "internal.hpp"
#include "primitives.hpp"

namespace sbpimpl {
using namespace sbpdev;

namespace {

constexpr int UPD_THREADS = 256;

int upd_grid(int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + UPD_THREADS - 1) / UPD_THREADS, 65536));
}

__device__ __forceinline__ unsigned long long lanes_below() {
    const int lane = lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// The vertices whose block changed, ascending, without atomics: wave w owns vertices [w C, (w + 1) C); a count
// pass, a scan of the per-wave counts, then a write pass.
constexpr int MOVED_WAVES = 4096;

__global__ void k_moved_count(int32_t N, int64_t chunk, const int32_t *__restrict__ prev,
                              const int32_t *__restrict__ assign, int32_t *__restrict__ wave_count) {
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (w >= MOVED_WAVES) return;
    const int64_t lo = w * chunk, hi = std::min<int64_t>(lo + chunk, N);
    int32_t n = 0;
    for (int64_t base = lo; base < hi; base += WAVE) {
        const int64_t v = base + lane_id();
        n += __popcll(__ballot(v < hi && prev[v] != assign[v]));
    }
    if (lane_id() == 0) wave_count[w] = n;
}

// Writes the moved vertices at their ranks, with their degrees (out + in); wave_base: exclusive scan of the counts.
__global__ void k_moved_write(int32_t N, int64_t chunk, const int32_t *__restrict__ prev,
                              const int32_t *__restrict__ assign, const int64_t *__restrict__ out_off,
                              const int64_t *__restrict__ in_off, const int32_t *__restrict__ wave_base,
                              int32_t *__restrict__ moved, int32_t *__restrict__ degree) {
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (w >= MOVED_WAVES) return;
    const int64_t lo = w * chunk, hi = std::min<int64_t>(lo + chunk, N);
    int64_t at = wave_base[w];
    for (int64_t base = lo; base < hi; base += WAVE) {
        const int64_t v = base + lane_id();
        const bool mv = v < hi && prev[v] != assign[v];
        const unsigned long long b = __ballot(mv);
        if (mv) {
            const int64_t r = at + __popcll(b & lanes_below());
            moved[r] = (int32_t)v;
            degree[r] = (int32_t)((out_off[v + 1] - out_off[v]) + (in_off[v + 1] - in_off[v]));
        }
        at += __popcll(b);
    }
}

// Signed cell deltas of the moved vertices' edges: key (cell << 1) | (1 for +1, 0 for -1), cell = row << cbits | col.
// One wavefront per moved vertex; vertex i writes two keys per edge at 2 off[i] + 2 j (off: exclusive scan of the
// degrees), out-edges first; an in-edge from another moved vertex (one of that vertex's out-edges) writes `filler`,
// which sorts after every delta.
__global__ void k_move_deltas(int64_t m, const int32_t *__restrict__ moved, const int64_t *__restrict__ off,
                              const int64_t *__restrict__ out_off, const int32_t *__restrict__ out_adj,
                              const int64_t *__restrict__ in_off, const int32_t *__restrict__ in_adj,
                              const int32_t *__restrict__ prev, const int32_t *__restrict__ assign, int cbits,
                              unsigned long long filler, unsigned long long *__restrict__ keys) {
    const int lane = lane_id();
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / WAVE);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; i < m; i += waves) {
        const int32_t v = moved[i];
        const unsigned long long rv = (unsigned)prev[v], sv = (unsigned)assign[v];
        unsigned long long *out = keys + 2 * off[i];
        for (int pass = 0; pass < 2; ++pass) {
            const int64_t e0 = pass == 0 ? out_off[v] : in_off[v], e1 = pass == 0 ? out_off[v + 1] : in_off[v + 1];
            for (int64_t e = e0 + lane; e < e1; e += WAVE) {
                const int32_t u = pass == 0 ? out_adj[e] : in_adj[e];
                const unsigned long long ru = (unsigned)prev[u], su = (unsigned)assign[u];
                unsigned long long minus, plus;
                if (pass == 0) {
                    minus = ((rv << cbits) | ru) << 1;
                    plus = (((sv << cbits) | su) << 1) | 1ull;
                } else if (ru == su) {
                    minus = ((ru << cbits) | rv) << 1;
                    plus = (((su << cbits) | sv) << 1) | 1ull;
                } else {
                    minus = plus = filler;
                }
                out[2 * (e - e0)] = minus;
                out[2 * (e - e0) + 1] = plus;
            }
            out += 2 * (e1 - e0);
        }
    }
}

// Net delta per cell from the sorted signed keys' runs: the first run of a cell (its -1 run when it has one)
// carries the net; a cell with a net of zero, and every second run, get the invalid key `none`.
__global__ void k_delta_net(const int64_t *__restrict__ n_runs, const unsigned long long *__restrict__ uniq,
                            const int32_t *__restrict__ cnt, unsigned long long none,
                            unsigned long long *__restrict__ dkey, int32_t *__restrict__ dnet) {
    const int64_t n = *n_runs;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = uniq[i], cell = k >> 1;
        const bool plus = k & 1ull;
        if (cell >= none || (plus && i > 0 && (uniq[i - 1] >> 1) == cell)) {
            dkey[i] = none;
            dnet[i] = 0;
            continue;
        }
        int32_t net = plus ? cnt[i] : -cnt[i];
        if (!plus && i + 1 < n && (uniq[i + 1] >> 1) == cell) net += cnt[i + 1];
        dkey[i] = net ? cell : none;
        dnet[i] = net;
    }
}

// The same deltas keyed column-major (col << cbits | row), for the column-major merge.
__global__ void k_delta_transpose(const int64_t *__restrict__ n_runs, const unsigned long long *__restrict__ dkey,
                                  int cbits, unsigned long long none, unsigned long long *__restrict__ tkey) {
    const int64_t n = *n_runs;
    const unsigned long long mask = (1ull << cbits) - 1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = dkey[i];
        tkey[i] = k == none ? none : (((k & mask) << cbits) | (k >> cbits));
    }
}

// Inserts key or sets its count in the cell hash.  The host keeps the table's load under 0.9 with every insertion
// counted; the probe is bounded by the table size all the same and reports a full table in *err.
__device__ __forceinline__ void hash_upsert(CellSlot *table, int shift, unsigned mask, unsigned long long key, int val,
                                            unsigned long long *inserted, int32_t *err) {
    unsigned h = cell_slot(key, shift);
    for (unsigned n = 0; n <= mask; ++n, h = (h + 1) & mask) {
        unsigned long long k = table[h].key;
        if (k == CELL_EMPTY) {
            k = atomicCAS(&table[h].key, CELL_EMPTY, key);
            if (k == CELL_EMPTY) {
                table[h].val = val;
                atomicAdd(inserted, 1ull);
                return;
            }
        }
        if (k == key) {
            table[h].val = val;
            return;
        }
    }
    atomicOr(err, 2);
}

// For each delta: its cell's rank in the old sorted list.  A cell already present takes the delta into its count
// in place; an absent one is an insertion before that rank (counted per rank in at[]).  err[0] is set when a
// count would become negative or an absent cell would get a non-positive count (deltas inconsistent with the
// cells).  The hash (row-major pass only) gets every changed count.
__global__ void k_locate(const int64_t *__restrict__ n_runs, const unsigned long long *__restrict__ dkey,
                         const int32_t *__restrict__ dnet, unsigned long long none, int64_t nnz,
                         const unsigned long long *__restrict__ keys, int2 *__restrict__ cells,
                         int32_t *__restrict__ pos, int32_t *__restrict__ ins, int32_t *__restrict__ at,
                         CellSlot *__restrict__ hash, int shift, unsigned hmask, int cbits,
                         unsigned long long *__restrict__ inserted, int32_t *__restrict__ err,
                         int32_t *__restrict__ dense, int64_t dstride) {
    const int64_t n = *n_runs;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = dkey[i];
        if (k == none) {
            ins[i] = 0;
            continue;
        }
        int64_t lo = 0, hi = nnz;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        int32_t c;
        pos[i] = (int32_t)lo;
        if (lo < nnz && keys[lo] == k) {
            c = cells[lo].y + dnet[i];
            cells[lo].y = c;
            ins[i] = 0;
            if (c < 0) atomicOr(err, 1);
        } else {
            c = dnet[i];
            ins[i] = 1;
            atomicAdd(&at[lo], 1);
            if (c <= 0) atomicOr(err, 1);
        }
        if (hash) {
            const unsigned long long mask = (1ull << cbits) - 1;
            hash_upsert(hash, shift, hmask, cell_key((int)(k >> cbits), (int)(k & mask)), c, inserted, err);
        }
        if (dense) dense[(int64_t)(k >> cbits) * dstride + (int64_t)(k & ((1ull << cbits) - 1))] = c;
    }
}

__global__ void k_removed(int64_t nnz, const int2 *__restrict__ cells, int32_t *__restrict__ rem) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= nnz; i += (int64_t)gridDim.x * blockDim.x)
        rem[i] = i < nnz && cells[i].y == 0;
}

// Kept old cells to their new rank: i - (removed before i) + (insertions at ranks <= i).
__global__ void k_scatter_old(int64_t nnz, const unsigned long long *__restrict__ keys, const int2 *__restrict__ cells,
                              const int32_t *__restrict__ remb, const int32_t *__restrict__ insb,
                              unsigned long long *__restrict__ nkeys, int32_t *__restrict__ ncnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = cells[i].y;
        if (c > 0) {
            const int64_t p = i - remb[i] + insb[i];
            nkeys[p] = keys[i];
            ncnt[p] = c;
        }
    }
}

// Inserted cells to their new rank: (insertions before it) + (kept old cells before its old rank).
__global__ void k_scatter_ins(const int64_t *__restrict__ n_runs, const unsigned long long *__restrict__ dkey,
                              const int32_t *__restrict__ dnet, const int32_t *__restrict__ pos,
                              const int32_t *__restrict__ ins, const int32_t *__restrict__ ins_rank,
                              const int32_t *__restrict__ remb, unsigned long long *__restrict__ nkeys,
                              int32_t *__restrict__ ncnt) {
    const int64_t n = *n_runs;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!ins[i]) continue;
        const int64_t p = (int64_t)ins_rank[i] + pos[i] - remb[pos[i]];
        nkeys[p] = dkey[i];
        ncnt[p] = dnet[i];
    }
}

__global__ void k_new_nnz(int64_t nnz, const int64_t *__restrict__ n_runs, const int32_t *__restrict__ remb,
                          const int32_t *__restrict__ ins_rank, int64_t *__restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *out = nnz - remb[nnz] + ins_rank[*n_runs];
}

bool incremental_allowed() {
    const char *e = getenv("SBP_AMD_INCREMENTAL");
    return !(e && e[0] == '0');
}
bool incremental_forced() {
    const char *e = getenv("SBP_AMD_INCREMENTAL");
    return e && e[0] == '1';
}

template <typename TI, typename T = TI>
void scan_into(sbp_ctx ctx, const TI *in, T *out, int64_t n, bool inclusive) {
    size_t bytes = 0;
    SBP_HIP((sbpprim::scan<TI, T>(nullptr, bytes, in, out, n, inclusive, ctx->stream)));
    SBP_HIP((sbpprim::scan<TI, T>(ctx->scratch[S_SORT_TMP].get(bytes), bytes, in, out, n, inclusive, ctx->stream)));
}

// One merge of the delta list (keys dkey, nets dnet, n_runs of them on the device, n_max bound) into a sorted cell
// list (keys, cells): the new sorted keys and counts in nkeys / ncnt, their number in *new_nnz (device).
void merge_major(sbp_ctx ctx, sbp_bm bm, int64_t n_max, const int64_t *n_runs, const unsigned long long *dkey,
                 const int32_t *dnet, unsigned long long none, const unsigned long long *keys, int2 *cells,
                 bool with_hash, unsigned long long *inserted, int32_t *err, unsigned long long *nkeys, int32_t *ncnt,
                 int64_t *new_nnz, ScratchSlot s_pos, ScratchSlot s_at, ScratchSlot s_rem, int32_t *dense = nullptr) {
    hipStream_t st = ctx->stream;
    const int64_t nnz = bm->nnz;
    int32_t *pos = ctx->scratch[s_pos].as<int32_t>(3 * (size_t)n_max + 4);
    int32_t *ins = pos + n_max + 1, *ins_rank = ins + n_max + 1;
    int32_t *at = ctx->scratch[s_at].as<int32_t>((size_t)nnz + 2);
    int32_t *rem = ctx->scratch[s_rem].as<int32_t>((size_t)nnz + 2);
    SBP_HIP(sbpprim::fill<int32_t>(at, nnz + 1, 0, st));
    SBP_HIP(sbpprim::fill<int32_t>(ins, n_max + 1, 0, st));
    CellSlot *hash = with_hash ? bm->cell_hash : nullptr;
    const int lg = bm->hash_log2;
    k_locate<<<upd_grid(n_max), UPD_THREADS, 0, st>>>(n_runs, dkey, dnet, none, nnz, keys, cells, pos, ins, at, hash,
                                                       64 - lg, (unsigned)((1ull << lg) - 1), bm->keys_cbits, inserted,
                                                       err, dense, (int64_t)bm->B);
    k_removed<<<upd_grid(nnz + 1), UPD_THREADS, 0, st>>>(nnz, cells, rem);
    scan_into<int32_t>(ctx, rem, rem, nnz + 1, false);  // removed before i (rem[nnz]: all removed)
    scan_into<int32_t>(ctx, at, at, nnz + 1, true);     // insertions at ranks <= i
    scan_into<int32_t>(ctx, ins, ins_rank, n_max + 1, false);
    k_scatter_old<<<upd_grid(nnz), UPD_THREADS, 0, st>>>(nnz, keys, cells, rem, at, nkeys, ncnt);
    k_scatter_ins<<<upd_grid(n_max), UPD_THREADS, 0, st>>>(n_runs, dkey, dnet, pos, ins, ins_rank, rem, nkeys, ncnt);
    k_new_nnz<<<1, 64, 0, st>>>(nnz, n_runs, rem, ins_rank, new_nnz);
    SBP_HIP(hipGetLastError());
}

}  // namespace

// Graphs below this many edges rebuild after every commit: their rebuild is a few short launches and one read,
// the update two dozen launches and five reads (measured: the batched workload, graphs of up to 50,000 vertices,
// 41.3 graphs/s with rebuilds against 36.0 with updates).
constexpr int64_t UPDATE_MIN_EDGES = int64_t(1) << 22;

bool bm_update_wanted(sbp_graph g) {
    if (!incremental_allowed()) return false;
    return incremental_forced() || g->E >= UPDATE_MIN_EDGES;
}

bool bm_update_moves(sbp_ctx ctx, sbp_graph g, sbp_bm bm, const int32_t *prev, int64_t moved_hint) {
    if (!bm->keys_valid || bm->dc || !bm_update_wanted(g)) return false;
    hipStream_t st = ctx->stream;
    const int32_t N = bm->N, B = bm->B;
    const int cbits = bm->keys_cbits;
    SBP_CHECK(cbits == bits_for_blocks(B), SBP_E_STATE, "cell keys of another block count");
    if (moved_hint == 0) return true;  // nothing moved: the cells, sizes and the hash stand
    auto *counters = ctx->scratch[S_UPD_CNT].as<unsigned long long>(8);
    SBP_HIP(sbpprim::fill<unsigned long long>(counters, 8, 0ull, st));
    int32_t *err = reinterpret_cast<int32_t *>(counters + 6);
    int32_t *moved = ctx->scratch[S_UPD_MOVED].as<int32_t>(2 * (size_t)N + 2);
    int32_t *degree = moved + N + 1;
    int32_t *wave_count = ctx->scratch[S_UPD_POS].as<int32_t>(2 * (size_t)MOVED_WAVES + 2);
    int32_t *wave_base = wave_count + MOVED_WAVES + 1;
    const int64_t chunk = ((int64_t)N + MOVED_WAVES - 1) / MOVED_WAVES;
    k_moved_count<<<MOVED_WAVES / 4, 256, 0, st>>>(N, chunk, prev, bm->assign, wave_count);
    SBP_HIP(hipGetLastError());
    SBP_HIP(sbpprim::fill<int32_t>(wave_count + MOVED_WAVES, 1, 0, st));
    scan_into<int32_t, int32_t>(ctx, wave_count, wave_base, MOVED_WAVES + 1, false);
    k_moved_write<<<MOVED_WAVES / 4, 256, 0, st>>>(N, chunk, prev, bm->assign, g->out_off, g->in_off, wave_base,
                                                   moved, degree);
    SBP_HIP(hipGetLastError());
    int32_t mcount = 0;
    read_small(ctx, {{&mcount, wave_base + MOVED_WAVES, sizeof(int32_t)}});
    const int64_t m = mcount;
    if (m == 0) return true;
    // per-vertex key offsets (exclusive scan of the degrees, in 64 bits) and the key count
    int64_t *off = reinterpret_cast<int64_t *>(ctx->scratch[S_UPD_REM].as<int64_t>((size_t)m + 2));
    SBP_HIP(sbpprim::fill<int32_t>(degree + m, 1, 0, st));
    scan_into<int32_t, int64_t>(ctx, degree, off, m + 1, false);
    int64_t sumdeg = 0;
    read_small(ctx, {{&sumdeg, off + m, sizeof(int64_t)}});
    const int64_t nk = 2 * sumdeg;
    if (!incremental_forced() && nk > g->E / 2) return false;  // the rebuild is cheaper
    // signed deltas, sorted, summed per cell
    auto *keys = ctx->scratch[S_UPD_KEYS].as<unsigned long long>((size_t)nk + 1);
    auto *sorted = ctx->scratch[S_UPD_SORTED].as<unsigned long long>((size_t)nk + 1);
    const unsigned long long none = 1ull << (2 * cbits);      // above every cell key
    const unsigned long long filler = none << 1;              // above every signed delta key
    const int end_bit = 2 * cbits + 2;
    {
        const int64_t waves = std::min<int64_t>(m, 65536 * 4);
        k_move_deltas<<<(int)((waves + 3) / 4), 256, 0, st>>>(m, moved, off, g->out_off, g->out_adj, g->in_off,
                                                               g->in_adj, prev, bm->assign, cbits, filler, keys);
        SBP_HIP(hipGetLastError());
    }
    auto *uniq = ctx->scratch[S_UPD_UNIQ].as<unsigned long long>((size_t)nk + 1);
    int32_t *cnt = ctx->scratch[S_UPD_CNTS].as<int32_t>((size_t)nk + 1);
    int64_t *n_runs = reinterpret_cast<int64_t *>(counters + 3);
    {
        size_t a = 0, b = 0;
        const int32_t *no_vals = nullptr;
        SBP_HIP(sbpprim::radix_sort(nullptr, a, keys, sorted, no_vals, (int32_t *)nullptr, nk, 0, end_bit, st));
        SBP_HIP(sbpprim::run_length_encode(nullptr, b, sorted, uniq, cnt, n_runs, nk, st));
        void *tmp = ctx->scratch[S_SORT_TMP].get(std::max(a, b));
        SBP_HIP(sbpprim::radix_sort(tmp, a, keys, sorted, no_vals, (int32_t *)nullptr, nk, 0, end_bit, st));
        SBP_HIP(sbpprim::run_length_encode(tmp, b, sorted, uniq, cnt, n_runs, nk, st));
    }
    // deltas per cell, row-major (in keys / sorted, which are free again) and column-major
    unsigned long long *dkey = keys, *tkey = sorted;
    int32_t *dnet = ctx->scratch[S_UPD_NET].as<int32_t>(2 * (size_t)nk + 2);
    int32_t *tnet = dnet + nk + 1;
    k_delta_net<<<upd_grid(nk), UPD_THREADS, 0, st>>>(n_runs, uniq, cnt, none, dkey, dnet);
    k_delta_transpose<<<upd_grid(nk), UPD_THREADS, 0, st>>>(n_runs, dkey, cbits, none, uniq);
    int64_t runs = 0;
    {
        // sort the transposed keys with their nets; `none` entries sort last
        read_small(ctx, {{&runs, n_runs, sizeof(runs)}});
        size_t a = 0;
        SBP_HIP(sbpprim::radix_sort(nullptr, a, uniq, tkey, (const int32_t *)dnet, tnet, runs, 0, end_bit, st));
        SBP_HIP(sbpprim::radix_sort(ctx->scratch[S_SORT_TMP].get(a), a, uniq, tkey, (const int32_t *)dnet, tnet, runs,
                                    0, end_bit, st));
    }
    // merges into new row-major and column-major lists
    const int64_t cap_new = bm->nnz + nk + 1;
    auto *rk = ctx->scratch[S_UPD_RKEYS].as<unsigned long long>((size_t)cap_new);
    auto *ck = ctx->scratch[S_UPD_CKEYS].as<unsigned long long>((size_t)cap_new);
    int32_t *rc = ctx->scratch[S_UPD_RCNT].as<int32_t>((size_t)cap_new);
    int32_t *cc = ctx->scratch[S_UPD_CCNT].as<int32_t>((size_t)cap_new);
    int64_t *new_nnz = reinterpret_cast<int64_t *>(counters + 4);
    // the hash takes the update when even every cell being new keeps its load under 0.9
    const bool with_hash = bm->hash_valid && bm->cell_hash && (bm->hash_used + runs) * 10 < (int64_t(9) << bm->hash_log2);
    // the dense copy, when kept, takes every changed count in the same pass
    int32_t *dense = bm->dense_valid && bm->dense_B == bm->B ? bm->dense : nullptr;
    merge_major(ctx, bm, nk, n_runs, dkey, dnet, none, bm->row_keys, bm->row_cells, with_hash, counters + 5, err, rk,
                rc, new_nnz, S_UPD_POS, S_UPD_AT, S_UPD_REM, dense);
    merge_major(ctx, bm, nk, n_runs, tkey, tnet, none, bm->col_keys, bm->col_cells, false, counters + 5, err, ck, cc,
                new_nnz + 1, S_UPD_POS, S_UPD_AT, S_UPD_REM);
    int64_t nn[2] = {0, 0};
    int32_t bad = 0;
    unsigned long long hash_ins = 0;
    read_small(ctx, {{nn, new_nnz, 2 * sizeof(int64_t)}, {&bad, err, sizeof(int32_t)},
                     {&hash_ins, counters + 5, sizeof(hash_ins)}});
    SBP_CHECK(!(bad & 1) && nn[0] == nn[1], SBP_E_STATE, "cell update inconsistent with the block matrix");
    SBP_CHECK(!(bad & 2), SBP_E_STATE, "cell hash full during the update");
    const int64_t nnz = nn[0];
    bm_ensure_cells(bm, nnz > 0 ? nnz : 1);
    bm->nnz = nnz;
    bm_fill_cells(ctx, bm, rk, rc, ck, cc);
    bm_count_sizes(ctx, bm);
    if (with_hash) {
        bm->hash_used += (int64_t)hash_ins;
        // keep the load under 0.7; past it the next lookup rebuilds the hash
        bm->hash_valid = bm->hash_used * 10 < (int64_t(7) << bm->hash_log2);
    } else {
        bm->hash_valid = false;
    }
    ctx->cell_updates++;
    return true;
}

}  // namespace sbpimpl
