// Block-merge proposals: one wavefront per block, NUM_AGG_PROPOSALS_PER_BLOCK = 10 proposals each.
//
// Replaces the OpenMP loop of block_merge::merge_blocks (src/block_merge.cpp:189-217) with
// propose_merge_sparse (:251-285), merge-mode common::propose_new_block (src/common.cpp:138-198) and
// entropy::block_merge_delta_mdl (src/entropy.cpp:129-182).  The merge Delta of
// block_merge::blockmodel_delta (src/block_merge.cpp:14-38) moves all of row r and column r into s; its
// cells are enumerated straight from the row/column cells of r, with hash lookups of the cells of row s /
// column s, so no per-proposal map is built.  Repeated proposals of the same s are skipped as
// past_proposals does.  Terms are summed in 128-bit fixed point.
#include <float.h>

#include <algorithm>

#include "device_common.hpp"
#include "internal.hpp"
#include "np_entropy.hpp"

namespace sbpimpl {
using namespace sbpdev;

constexpr int MG_BLOCK = 256;
constexpr int MG_WAVES = MG_BLOCK / WAVE;

struct MgArgs {
    int32_t B;
    uint64_t seed;
    uint32_t epoch;
    const int32_t *d_out, *d_in;
    const int32_t *row_ptr, *col_ptr;
    const int2 *row_cells, *col_cells;
    int32_t *best_s;
    double *best_dS;
    unsigned long long *bytes;
    int *err;
    const int32_t *sizes;
    NpConst np;  // nonparametric runs only
    const CellSlot *hash;
    int hash_shift;
    unsigned hash_mask;
};

__device__ __forceinline__ long long mg_get(const MgArgs &a, int row, int col) {
    return cell_hash_get(a.hash, a.hash_shift, a.hash_mask, row, col);
}

// Wave-cooperative selection: first cell (in list order, skipping excluded ids) whose inclusive prefix of
// counts exceeds target.  Returns -1 when the list is exhausted; base carries the running prefix.
__device__ __forceinline__ int wave_select(const int2 *cells, int begin, int end, int ex1, int ex2,
                                           long long target, long long &base) {
    const int lane = lane_id();
    for (int c0 = begin; c0 < end; c0 += WAVE) {
        int ci = c0 + lane;
        int2 cell = ci < end ? cells[ci] : make_int2(-1, 0);
        bool ok = ci < end && cell.x != ex1 && cell.x != ex2;
        long long incl = wave_inclusive_scan(ok ? (long long)cell.y : 0ll) + base;
        unsigned long long hit = ballot(ok && target < incl);
        if (hit) return __shfl(cell.x, __ffsll((long long)hit) - 1, WAVE);
        base = __shfl(incl, WAVE - 1, WAVE);
    }
    return -1;
}

// delta-MDL of merging block r into block s (entropy::block_merge_delta_mdl, src/entropy.cpp:129-182):
// the Delta entries term by term (row r and column r move into s), cell (s, s) when it is unchanged, and
// the other unchanged cells of row s / column s in bulk -- such a cell changes only through d_out[s] or
// d_in[s], contributing M * (fastlog(d') - fastlog(d)), and M summed over them is the row/column total minus
// the cells the merge touches.  Work is O(|row r| + |col r|) lookups; row s and column s are not scanned.
__device__ double merge_delta_mdl(const MgArgs &a, int r, int s) {
    const int lane = lane_id();
    const long long k_out = a.d_out[r], k_in = a.d_in[r];
    const int rr0 = a.row_ptr[r], rr1 = a.row_ptr[r + 1];
    const int cr0 = a.col_ptr[r], cr1 = a.col_ptr[r + 1];
    const long long din_s_new = (long long)a.d_in[s] + k_in, dout_s_new = (long long)a.d_out[s] + k_out;
    Fx acc = fx_zero();
    long long sub_row = 0, sub_col = 0;
    for (int i = rr0 + lane; i < rr1; i += WAVE) {  // (r, c) old terms; (s, c) for c outside {r, s}
        int2 cell = a.row_cells[i];
        int c = cell.x;
        fx_acc(acc, cell_entropy(cell.y, a.d_in[c], a.d_out[r]));
        if (c != r && c != s) {
            long long msc = mg_get(a, s, c);
            fx_acc(acc, cell_entropy(msc, a.d_in[c], a.d_out[s]));
            fx_sub_acc(acc, cell_entropy(msc + cell.y, a.d_in[c], dout_s_new));
            sub_row += msc;
        }
    }
    for (int i = cr0 + lane; i < cr1; i += WAVE) {  // (x, r) old terms; (x, s) for x outside {r, s}
        int2 cell = a.col_cells[i];
        int x = cell.x;
        if (x == r) continue;
        fx_acc(acc, cell_entropy(cell.y, a.d_in[r], a.d_out[x]));
        if (x != s) {
            long long mxs = mg_get(a, x, s);
            fx_acc(acc, cell_entropy(mxs, a.d_in[s], a.d_out[x]));
            fx_sub_acc(acc, cell_entropy(mxs + cell.y, din_s_new, a.d_out[x]));
            sub_col += mxs;
        }
    }
    const long long mrr = mg_get(a, r, r), mrs = mg_get(a, r, s), msr = mg_get(a, s, r), mss = mg_get(a, s, s);
    const long long dss = mrr + mrs + msr;  // what (s, s) receives
    acc = wave_sum_fx(acc);
    sub_row = wave_sum(sub_row);
    sub_col = wave_sum(sub_col);
    if (dss > 0) {
        fx_acc(acc, cell_entropy(mss, a.d_in[s], a.d_out[s]));
        fx_sub_acc(acc, cell_entropy(mss + dss, din_s_new, dout_s_new));
    } else if (mss > 0) {
        fx_acc(acc, cell_entropy(mss, a.d_in[s], a.d_out[s]));
        fx_sub_acc(acc, cell_entropy(mss, din_s_new, dout_s_new));
    }
    const long long S_row = (long long)a.d_out[s] - mss - msr - sub_row;
    const long long S_col = (long long)a.d_in[s] - mss - mrs - sub_col;
    fx_acc(acc, (double)S_row * (fastlog(dout_s_new) - fastlog(a.d_out[s])));
    fx_acc(acc, (double)S_col * (fastlog(din_s_new) - fastlog(a.d_in[s])));
    (void)lane;
    return fx_to(acc);
}

// Nonparametric delta-MDL of merging r into s (entropy::nonparametric::block_merge_delta_mdl,
// src/entropy.cpp:1183-1206): eterm differences of the merge Delta's entries -- row r and column r empty,
// (s, c) and (x, s) receive them, (s, s) receives (r, r) + (r, s) + (s, r) -- then the block terms and the
// description-length deltas with weight = size(r).
__device__ double merge_delta_mdl_np(const MgArgs &a, int r, int s) {
    const int lane = lane_id();
    if (a.sizes[r] == 0 || a.sizes[s] == 0) return 0.0;
    const int rr0 = a.row_ptr[r], rr1 = a.row_ptr[r + 1];
    const int cr0 = a.col_ptr[r], cr1 = a.col_ptr[r + 1];
    Fx acc = fx_zero();
    for (int i = rr0 + lane; i < rr1; i += WAVE) {
        const int2 cell = a.row_cells[i];
        const int c = cell.x;
        fx_acc(acc, np_entry(a.np, cell.y, -(long long)cell.y));
        if (c != r && c != s) fx_acc(acc, np_entry(a.np, mg_get(a, s, c), cell.y));
    }
    for (int i = cr0 + lane; i < cr1; i += WAVE) {
        const int2 cell = a.col_cells[i];
        const int x = cell.x;
        if (x == r) continue;
        fx_acc(acc, np_entry(a.np, cell.y, -(long long)cell.y));
        if (x != s) fx_acc(acc, np_entry(a.np, mg_get(a, x, s), cell.y));
    }
    const long long mrr = mg_get(a, r, r), mrs = mg_get(a, r, s), msr = mg_get(a, s, r), mss = mg_get(a, s, s);
    const long long dss = mrr + mrs + msr;
    acc = wave_sum_fx(acc);
    if (dss > 0) fx_acc(acc, np_entry(a.np, mss, dss));
    return np_finish(a.np, fx_to(acc), r, s, a.d_out[r], a.d_in[r], a.sizes[r], a.d_out[s], a.d_in[s], a.sizes[s],
                     a.d_out[r], a.d_in[r], a.sizes[r], a.d_in[r], a.d_out[r]);
}

template <bool NP>
__global__ void __launch_bounds__(MG_BLOCK) __attribute__((amdgpu_waves_per_eu(NP ? 1 : 4))) k_merge_propose(MgArgs a) {
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    const int64_t nwaves = (int64_t)gridDim.x * MG_WAVES;
    unsigned long long bytes = 0;
    for (int64_t rb = (int64_t)blockIdx.x * MG_WAVES + wid; rb < a.B; rb += nwaves) {
        const int r = (int)rb;
        const long long k = (long long)a.d_out[r] + a.d_in[r];
        const int rr0 = a.row_ptr[r], rr1 = a.row_ptr[r + 1], cr0 = a.col_ptr[r], cr1 = a.col_ptr[r + 1];
        int past[10];
        int npast = 0;
        double best = DBL_MAX;
        int best_s = -1;
        for (uint32_t i = 0; i < 10; ++i) {
            const U4 w = philox(a.seed, (uint32_t)r, a.epoch, TAG_MERGE, i);
            int s = -1, t = -1;
            if (k == 0) {
                s = random_block(r, a.B, w.y);
            } else {
                long long target = (long long)uniform_below(w.x, (uint32_t)k), base = 0;
                t = wave_select(a.row_cells, rr0, rr1, -2, -2, target, base);
                if (t < 0) t = wave_select(a.col_cells, cr0, cr1, -2, -2, target, base);
                if (t < 0) {  // degrees disagree with the cells: report and stop proposing for r
                    if (lane == 0) atomicExch(a.err, 1);
                    break;
                }
                // neighbors_weights(t) with edges[r] = 0 (src/common.cpp:169-178)
                long long mtr = mg_get(a, t, r), mtt = mg_get(a, t, t), mrt = (r != t) ? mg_get(a, r, t) : 0;
                long long total = (long long)a.d_out[t] - mtr + (long long)a.d_in[t] - mtt - mrt;
                if (total == 0) {
                    s = random_block(r, a.B, w.y);
                } else {
                    long long tg = (long long)uniform_below(w.y, (uint32_t)total), b2 = 0;
                    s = wave_select(a.row_cells, a.row_ptr[t], a.row_ptr[t + 1], r, -2, tg, b2);
                    if (s < 0) s = wave_select(a.col_cells, a.col_ptr[t], a.col_ptr[t + 1], t, r, tg, b2);
                    if (s < 0) {
                        if (lane == 0) atomicExch(a.err, 1);
                        break;
                    }
                }
            }
            bool seen = false;
            for (int j = 0; j < npast; ++j) seen |= (past[j] == s);
            double dS = DBL_MAX;
            if (!seen) {
                dS = NP ? merge_delta_mdl_np(a, r, s) : merge_delta_mdl(a, r, s);
                past[npast++] = s;
                long long rc_r = (rr1 - rr0) + (cr1 - cr0);
                long long rc_s = (a.row_ptr[s + 1] - a.row_ptr[s]) + (a.col_ptr[s + 1] - a.col_ptr[s]);
                long long rc_t = t >= 0 ? (a.row_ptr[t + 1] - a.row_ptr[t]) + (a.col_ptr[t + 1] - a.col_ptr[t]) : 0;
                bytes += 16ull * rc_r + 8ull * rc_t + 12ull * (rc_r + rc_s);  // per-proposal model of the reference
            }
            if (dS < best) {
                best = dS;
                best_s = s;
            }
        }
        if (lane == 0) {
            a.best_s[r] = best_s;
            a.best_dS[r] = best;
        }
    }
    if (lane == 0 && bytes) atomicAdd(a.bytes, bytes);
}

static MgArgs mg_args(sbp_ctx ctx, sbp_bm bm, uint32_t epoch, unsigned long long *bytes) {
    MgArgs a;
    a.err = reinterpret_cast<int *>(bytes + 1);
    a.B = bm->B;
    a.seed = ctx->seed;
    a.epoch = epoch;
    a.d_out = bm->d_out;
    a.d_in = bm->d_in;
    a.row_ptr = bm->row_ptr;
    a.col_ptr = bm->col_ptr;
    a.row_cells = bm->row_cells;
    a.col_cells = bm->col_cells;
    a.best_s = nullptr;
    a.best_dS = nullptr;
    a.bytes = bytes;
    a.sizes = bm->sizes;
    bm_ensure_hash(ctx, bm);
    a.hash = bm->cell_hash;
    a.hash_shift = 64 - bm->hash_log2;
    a.hash_mask = (1u << bm->hash_log2) - 1;
    a.np = NpConst{};
    if (bm->np) a.np = np_const(ctx, bm);
    return a;
}

// entropy::block_merge_delta_mdl of given merges (block r into block s), one wavefront each, with the device
// function the proposals use.
template <bool NP>
__global__ void __launch_bounds__(MG_BLOCK) k_merge_eval(MgArgs a, const int32_t *__restrict__ blocks,
                                                         const int32_t *__restrict__ targets, int64_t n,
                                                         double *__restrict__ dS_out) {
    const int64_t gwave = (int64_t)blockIdx.x * MG_WAVES + threadIdx.x / WAVE;
    const int64_t nwaves = (int64_t)gridDim.x * MG_WAVES;
    for (int64_t i = gwave; i < n; i += nwaves) {
        const double dS = NP ? merge_delta_mdl_np(a, blocks[i], targets[i]) : merge_delta_mdl(a, blocks[i], targets[i]);
        if (lane_id() == 0) dS_out[i] = dS;
    }
}

void merge_evaluate(sbp_ctx ctx, sbp_bm bm, int64_t n, const int32_t *d_blocks, const int32_t *d_targets,
                    double *d_dS) {
    if (n == 0) return;
    auto *bytes = ctx->scratch[S_COUNTER].as<unsigned long long>(4);
    SBP_HIP(hipMemsetAsync(bytes, 0, 4 * sizeof(unsigned long long), ctx->stream));
    MgArgs a = mg_args(ctx, bm, 0, bytes);
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n + MG_WAVES - 1) / MG_WAVES, 8192));
    if (bm->np) k_merge_eval<true><<<blocks, MG_BLOCK, 0, ctx->stream>>>(a, d_blocks, d_targets, n, d_dS);
    else k_merge_eval<false><<<blocks, MG_BLOCK, 0, ctx->stream>>>(a, d_blocks, d_targets, n, d_dS);
    SBP_HIP(hipGetLastError());
}

uint64_t merge_propose(sbp_ctx ctx, sbp_bm bm, uint32_t epoch, int32_t *d_best_s, double *d_best_dS) {
    hipStream_t st = ctx->stream;
    auto *bytes = ctx->scratch[S_COUNTER].as<unsigned long long>(4);
    SBP_HIP(hipMemsetAsync(bytes, 0, 4 * sizeof(unsigned long long), st));
    MgArgs a = mg_args(ctx, bm, epoch, bytes);
    a.best_s = d_best_s;
    a.best_dS = d_best_dS;
    int64_t blocks = (bm->B + MG_WAVES - 1) / MG_WAVES;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    if (bm->np) k_merge_propose<true><<<(int)blocks, MG_BLOCK, 0, st>>>(a);
    else k_merge_propose<false><<<(int)blocks, MG_BLOCK, 0, st>>>(a);
    SBP_HIP(hipGetLastError());
    unsigned long long hb[2] = {0, 0};
    read_small(ctx, {{hb, bytes, sizeof(hb)}});
    SBP_CHECK(hb[1] == 0, SBP_E_STATE, "merge proposal: block degrees inconsistent with the block matrix");
    return hb[0];
}

__global__ void k_relabel(int32_t N, const int32_t *__restrict__ map, const int32_t *__restrict__ in,
                          int32_t *__restrict__ out) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) out[v] = map[in[v]];
}

void relabel_assignment(sbp_ctx ctx, sbp_bm bm, const int32_t *d_block_map, int32_t new_B, int32_t *d_out) {
    (void)new_B;
    int64_t g = (bm->N + 255) / 256;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    k_relabel<<<(int)g, 256, 0, ctx->stream>>>(bm->N, d_block_map, bm->assign, d_out);
    SBP_HIP(hipGetLastError());
}

}  // namespace sbpimpl
