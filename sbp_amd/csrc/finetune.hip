// Finetune: vertex-move proposals (one wavefront per vertex), the asynchronous-Gibbs commit, and
// sequential Metropolis-Hastings.
//
// Proposal (propose_gibbs_move_v3 + eval_vertex_move_v3, src/finetune.cpp:447-462,871-901):
//   1. a uniformly random neighbour position in concat(out(v), in(v) \ {v}) gives neighbour u, t = b(u)
//      (common::propose_new_block, src/common.cpp:138-198);
//   2. s is drawn from neighbors_weights(t) = row(t) + col(t) (dict_transpose_matrix.cpp:176-188), whose
//      total is d_tot[t]; the wave scans the cells with an inclusive prefix sum and a ballot;
//   3. delta-MDL (entropy::delta_mdl, src/entropy.cpp:298-403): the neighbour-block histogram of v is
//      built in a hash table (LDS, one per wave; global memory for high-degree vertices); the Delta cells
//      of finetune::blockmodel_delta (src/finetune.cpp:201-235) follow from it in closed form, including
//      cells whose adds and subs cancel (they stay in the reference's Delta maps and are counted twice);
//      rows r, s and columns r, s are scanned once for the unchanged cells;
//   4. Hastings correction (src/entropy.cpp:497-561) when not greedy; accept (src/finetune.cpp:22-32).
// Terms are summed in 128-bit fixed point.
//
// Async-Gibbs commit (Blockmodel::move_vertex(const VertexMove_v3&), src/blockmodel/blockmodel.cpp:437-494,
// in vertex-id order, src/finetune.cpp:80-85): the block matrix after the commits is a pure function of
// the final assignment, so only the refusals (a move out of a block whose size has reached 1) need the
// sequential order.  Blocks whose proposed out-moves reach their size are "at risk"; for them the
// refusal pattern is the unique fixed point of a per-block scan over their move events in vertex order,
// reached by monotone iteration from "all accepted".  The matrix is then rebuilt from the assignment.
//
// Metropolis-Hastings (src/finetune.cpp:773-869) is sequential by definition: one persistent wavefront
// walks the vertex order, evaluates each proposal exactly as above against the current state and applies
// accepted Deltas in place (Blockmodel::move_vertex(Vertex, const Delta&, ...), blockmodel.cpp:408-435) to
// a CSR whose rows and columns carry slack.  When a row or column runs out of slack the kernel stops
// before that vertex; the host re-lays the matrix out with more slack and resumes.
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "device_common.hpp"
#include "device_table.hpp"
#include "internal.hpp"
#include "np_entropy.hpp"
#include "primitives.hpp"

namespace sbpimpl {
using namespace sbpdev;

constexpr int FT_BLOCK = 256;
constexpr int FT_WAVES = FT_BLOCK / WAVE;
// Wavefronts of the Metropolis-Hastings workgroup: vertices evaluated per window of the sequential walk.
constexpr int MH_WAVES = 16;
static_assert(MH_WAVES > 4, "the Metropolis-Hastings commit gives four wavefronts a row or column of r and s each");
constexpr int LDS_CAP_LOG2 = 9;
// The LDS table serves vertices with |out| + |in| <= SMALL_K, whose distinct neighbour blocks are at most
// SMALL_K: its used-slot list needs no more entries, which keeps five 4-wave blocks within a CU's 160 KB.
constexpr int LDS_USED = FT_SMALL_DEGREE;
constexpr int64_t LDS_TABLE_INTS = table_ints(LDS_CAP_LOG2, LDS_USED);
constexpr int LDS_CAP = 1 << LDS_CAP_LOG2;  // hash slots per wave
constexpr int SMALL_K = FT_SMALL_DEGREE;  // max |out| + |in| for the LDS table
static_assert(2 * SMALL_K <= LDS_CAP, "LDS table load factor above 1/2");
// Degree classes of the proposal launches, by k = |out(v)| + |in(v)|: classes 0..FT_NB-3 hold k up to
// ft_bucket_max(b) (and above the previous bound) and run on lane groups; classes FT_NB-2 (k <= 1,024) and FT_NB-1
// (the rest) run one workgroup per vertex.  Each lane-group table has 2 * ft_bucket_max(b)
// slots (load factor <= 1/2; fewer lanes than fit that: ft_table_log2) and lists at most ft_bucket_max(b) used slots.
__host__ __device__ constexpr int ft_bucket_max(int b) {
    return b == 0 ? 8 : b == 1 ? 32 : b == 2 ? 128 : b == 3 ? SMALL_K : 1024;
}
__host__ __device__ constexpr int ft_bucket_log2(int b) { return b == 0 ? 4 : b == 1 ? 6 : b == 2 ? 8 : 9; }
// LDS ints per wavefront for the lane-group tables.  Four 4-wavefront workgroups per CU (the 4 wavefronts per SIMD
// the register budget allows) take just under 160 KB.
constexpr int LDS_WAVE_INTS = 2496;
// Table slots (log2) of a lane group of L lanes in class b: 2 * ft_bucket_max(b) (load factor <= 1/2) when WAVE / L
// such tables fit a wavefront's LDS, else ft_bucket_max(b) (the class's most distinct blocks fill it at worst;
// probes are bounded by the table size).
__host__ __device__ constexpr int ft_table_log2(int b, int L) {
    return (b >= 1 && (WAVE / L) * table_ints(ft_bucket_log2(b), ft_bucket_max(b)) > LDS_WAVE_INTS)
               ? ft_bucket_log2(b) - 1
               : ft_bucket_log2(b);
}
static_assert(FT_NB == 6 && ft_bucket_max(FT_NB - 3) == FT_SMALL_DEGREE, "degree classes");
// (WAVE / L) tables of class b must fit a wavefront's LDS, which allows L >= 2 for class 0, L >= 4 for class 1,
// L >= 16 for class 2 and L >= 32 for class 3.
__host__ __device__ constexpr bool ft_lanes_fit(int b, int L) {
    return (WAVE / L) * table_ints(ft_table_log2(b, L), ft_bucket_max(b)) <= LDS_WAVE_INTS;
}
static_assert(ft_lanes_fit(0, 2) && ft_lanes_fit(1, 4) && ft_lanes_fit(2, 16) && ft_lanes_fit(3, 32) &&
                  !ft_lanes_fit(1, 2) && !ft_lanes_fit(2, 8) && !ft_lanes_fit(3, 16),
              "LDS budget");
// Read-only view of the block matrix.  Compact CSR: row b = [ptr[b], ptr[b+1]), with per-row inclusive
// cumulative counts (columns: excluding the diagonal cell) for multinomial selection by binary search.
// Slack layout (Metropolis-Hastings): row b = [start[b], start[b] + len[b]), no cumulative counts.
struct BmView {
    const int32_t *rbeg, *rlen, *cbeg, *clen;  // rlen/clen null for compact CSR
    int2 *row_cells, *col_cells;
    const int32_t *row_cum, *col_cum;          // null for the slack layout
    const CellSlot *hash;                      // cell hash of the compact CSR (null for the slack layout)
    int hash_shift;
    unsigned hash_mask;
    const int32_t *dense;                      // B x B counts, row-major, when kept (looked up before the hash)
    int32_t *dense_w;                          // the same array, written by the Metropolis-Hastings commits
    int64_t dense_stride;
    __device__ __forceinline__ int row_begin(int b) const { return rbeg[b]; }
    __device__ __forceinline__ int row_end(int b) const { return rlen ? rbeg[b] + rlen[b] : rbeg[b + 1]; }
    __device__ __forceinline__ int col_begin(int b) const { return cbeg[b]; }
    __device__ __forceinline__ int col_end(int b) const { return clen ? cbeg[b] + clen[b] : cbeg[b + 1]; }
};

struct FtArgs {
    int32_t N, B;
    uint64_t seed;
    uint32_t epoch;
    const uint32_t *epoch_ptr;  // when set, the epoch is read here (a captured sweep replays with a new epoch)
    int greedy;
    const int64_t *out_off, *in_off;
    const int32_t *out_adj, *in_adj;
    int32_t *assign, *sizes, *d_out, *d_in, *d_tot;
    BmView m;
    const int32_t *items;  // vertices to evaluate
    int64_t count;
    int32_t *move_s;       // [N] accepted target or -1
    double *dS_out;        // optional [N]
    int32_t *s_out;        // optional [N]
    unsigned long long *bytes;
    int *gtable;           // high-degree tables: per wave 4 * 2^gcap_log2 + 1 ints
    int gcap_log2;
    int *err;              // set when a multinomial runs past its cells (inconsistent degrees)
    const int32_t *blist;  // the items grouped by degree class (class b: bmeta[b] items from blist + bmeta[FT_NB + b])
    const int32_t *bmeta;
    NpConst np;            // nonparametric description-length constants (nonparametric runs only)
    HistSlot *hist;        // degree-corrected runs: the histograms, updated in place by Metropolis-Hastings
};

__device__ __forceinline__ long long get_cell(const FtArgs &a, int row, int col) {
    if (a.m.dense) return a.m.dense[row * a.m.dense_stride + col];
    if (a.m.hash) return cell_hash_get(a.m.hash, a.m.hash_shift, a.m.hash_mask, row, col);
    int r0 = a.m.row_begin(row), r1 = a.m.row_end(row);
    int c0 = a.m.col_begin(col), c1 = a.m.col_end(col);
    if (r1 - r0 <= c1 - c0) return cell_lookup(a.m.row_cells, r0, r1, col);
    return cell_lookup(a.m.col_cells, c0, c1, row);
}

// get_cell of (row1, col1) and (row2, col2) with the two lower-bound searches advanced in lockstep, so each
// step issues both loads: the same values as two get_cell calls in half the dependent round trips.
__device__ __forceinline__ void get_cell_pair(const FtArgs &a, int row1, int col1, int row2, int col2, long long &v1,
                                              long long &v2) {
    if (a.m.dense) {
        v1 = a.m.dense[row1 * a.m.dense_stride + col1];
        v2 = a.m.dense[row2 * a.m.dense_stride + col2];
        return;
    }
    if (a.m.hash) {
        v1 = cell_hash_get(a.m.hash, a.m.hash_shift, a.m.hash_mask, row1, col1);
        v2 = cell_hash_get(a.m.hash, a.m.hash_shift, a.m.hash_mask, row2, col2);
        return;
    }
    const int r10 = a.m.row_begin(row1), r11 = a.m.row_end(row1), c10 = a.m.col_begin(col1), c11 = a.m.col_end(col1);
    const int r20 = a.m.row_begin(row2), r21 = a.m.row_end(row2), c20 = a.m.col_begin(col2), c21 = a.m.col_end(col2);
    const bool by_row1 = r11 - r10 <= c11 - c10, by_row2 = r21 - r20 <= c21 - c20;
    const int2 *x1 = by_row1 ? a.m.row_cells : a.m.col_cells, *x2 = by_row2 ? a.m.row_cells : a.m.col_cells;
    const int e1 = by_row1 ? r11 : c11, e2 = by_row2 ? r21 : c21;
    const int k1 = by_row1 ? col1 : row1, k2 = by_row2 ? col2 : row2;
    int lo1 = by_row1 ? r10 : c10, hi1 = e1, lo2 = by_row2 ? r20 : c20, hi2 = e2;
    while (lo1 < hi1 || lo2 < hi2) {
        const int m1 = (lo1 + hi1) >> 1, m2 = (lo2 + hi2) >> 1;
        const bool a1 = lo1 < hi1, a2 = lo2 < hi2;
        const int y1 = a1 ? x1[m1].x : 0, y2 = a2 ? x2[m2].x : 0;
        if (a1) {
            if (y1 < k1) lo1 = m1 + 1;
            else hi1 = m1;
        }
        if (a2) {
            if (y2 < k2) lo2 = m2 + 1;
            else hi2 = m2;
        }
    }
    const int2 f1 = lo1 < e1 ? x1[lo1] : make_int2(-1, 0), f2 = lo2 < e2 ? x2[lo2] : make_int2(-1, 0);
    v1 = (lo1 < e1 && f1.x == k1) ? f1.y : 0;
    v2 = (lo2 < e2 && f2.x == k2) ? f2.y : 0;
}

struct MoveCtx {
    int v, r, s, t;
    int selfc;  // self-loop occurrences in out(v)
    long long k_out, k_in, shift_in, k_all;
    int out_r, in_r, out_s, in_s;
    int64_t o0, o1, i0, i1;
};

__device__ __forceinline__ long long new_din(const FtArgs &a, const MoveCtx &m, int x) {
    long long v = a.d_in[x];
    if (x == m.r) v -= m.shift_in;
    else if (x == m.s) v += m.shift_in;
    return v;
}
__device__ __forceinline__ long long new_dout(const FtArgs &a, const MoveCtx &m, int x) {
    long long v = a.d_out[x];
    if (x == m.r) v -= m.k_out;
    else if (x == m.s) v += m.k_out;
    return v;
}

// old-minus-new entropy of one cell with current value M and change d

// Delta of row r / row s at column c from the histogram (0 when the cell is not in the Delta).
__device__ __forceinline__ long long delta_row_r(const MoveCtx &m, int c, int oc) {
    return -(long long)oc + (c == m.r ? -(long long)m.selfc - m.in_r : 0) + (c == m.s ? (long long)m.in_r : 0);
}
__device__ __forceinline__ long long delta_row_s(const MoveCtx &m, int c, int oc) {
    return (long long)oc + (c == m.s ? (long long)m.selfc + m.in_s : 0) + (c == m.r ? -(long long)m.in_s : 0);
}

__device__ __forceinline__ Table make_table(const FtArgs &a, bool big, int *lds, int64_t gwave) {
    if (big) return table_at(a.gtable + gwave * table_ints(a.gcap_log2), a.gcap_log2);
    return table_at(lds, LDS_CAP_LOG2, LDS_USED);
}

// The threads that evaluate one proposal together: L aligned lanes of a wavefront (LaneGroup<L>; a whole
// wavefront is LaneGroup<WAVE>), or a whole workgroup (BlockGroup, the high-degree launch).  Loops over edges
// and neighbour blocks stride by size(); sum() returns the group's total to every thread.  The fixed-point sums
// make the result independent of the decomposition.  A lane group's lanes take the same branches at every
// collective (their conditions are uniform over the group), so shuffles read only lanes that are active.
template <int L>
struct LaneGroup {
    static_assert(L >= 2 && L <= WAVE && (L & (L - 1)) == 0, "a lane group is a power of two within a wavefront");
    __device__ __forceinline__ int rank() const { return lane_id() & (L - 1); }
    __device__ __forceinline__ int size() const { return L; }
    __device__ __forceinline__ int first_lane() const { return lane_id() & ~(L - 1); }
    __device__ __forceinline__ void sync() const { wave_sync(); }
    __device__ __forceinline__ Fx sum(Fx v) const {
#pragma unroll
        for (int off = L / 2; off >= 1; off >>= 1) {
            Fx o;
            o.lo = __shfl_xor(v.lo, off, WAVE);
            o.hi = __shfl_xor(v.hi, off, WAVE);
            fx_add(v, o);
        }
        return v;
    }
    template <typename T>
    __device__ __forceinline__ T sum(T x) const {
#pragma unroll
        for (int off = L / 2; off >= 1; off >>= 1) x += __shfl_xor(x, off, WAVE);
        return x;
    }
    // bit i set when p holds on the group's lane of rank i
    __device__ __forceinline__ unsigned long long ballot(bool p) const {
        const unsigned long long b = __ballot(p);
        if constexpr (L == WAVE) return b;
        else return (b >> first_lane()) & ((1ull << L) - 1ull);
    }
    // the value held by the group's lane of rank r
    template <typename T>
    __device__ __forceinline__ T bcast(T v, int r) const {
        return __shfl(v, first_lane() + r, WAVE);
    }
    template <typename T>
    __device__ __forceinline__ T inclusive_scan(T v) const {
        const int k = rank();
#pragma unroll
        for (int off = 1; off < L; off <<= 1) {
            const T o = __shfl(v, lane_id() - (k >= off ? off : 0), WAVE);
            if (k >= off) v += o;
        }
        return v;
    }
};
using WaveGroup = LaneGroup<WAVE>;

// First index in [lo, hi) whose inclusive cumulative count exceeds target (hi when none), by an L-ary search:
// each round the group loads L evenly spaced counts and keeps the interval the first larger one closes.
template <class G>
__device__ __forceinline__ int group_cum_search(const int32_t *cum, int lo, int hi, long long target, const G &g) {
    const int L = g.size(), k = g.rank();
    while (hi - lo > L) {
        const int len = hi - lo;
        const int p = lo + (int)(((long long)(k + 1) * len) / L) - 1;
        const unsigned long long hit = g.ballot((long long)cum[p] > target);
        if (!hit) return hi;
        const int j = __ffsll((long long)hit) - 1;
        const int new_hi = g.bcast(p, j) + 1;
        const int prev = g.bcast(p, j > 0 ? j - 1 : 0) + 1;
        lo = j == 0 ? lo : prev;
        hi = new_hi;
    }
    const int i = lo + k;
    const unsigned long long hit = g.ballot(i < hi && (long long)cum[i] > target);
    return hit ? lo + __ffsll((long long)hit) - 1 : hi;
}

// Empties a lane group's table: every slot (table_clear_all) or the slots its last item claimed.
template <class G>
__device__ __forceinline__ void group_table_clear_all(const Table &t, const G &g) {
    for (int i = g.rank(); i < t.cap; i += g.size()) {
        t.keys[i] = EMPTY;
        t.outc[i] = 0;
        t.inc[i] = 0;
    }
    if (g.rank() == 0) *t.n_used = 0;
    g.sync();
}
template <class G>
__device__ __forceinline__ void group_table_clear_used(const Table &t, const G &g) {
    g.sync();
    const int n = *t.n_used;
    for (int i = g.rank(); i < n; i += g.size()) {
        const int slot = t.used[i];
        t.keys[slot] = EMPTY;
        t.outc[slot] = 0;
        t.inc[slot] = 0;
    }
    g.sync();
    if (g.rank() == 0) *t.n_used = 0;
    g.sync();
}

struct BlockGroup {
    Fx *fx;         // LDS, one per wavefront of the workgroup
    long long *ll;  // LDS, one per wavefront of the workgroup
    __device__ __forceinline__ int rank() const { return (int)threadIdx.x; }
    __device__ __forceinline__ int size() const { return (int)blockDim.x; }
    __device__ __forceinline__ void sync() const { __syncthreads(); }
    __device__ Fx sum(Fx x) const {
        x = wave_sum_fx(x);
        const int nw = (int)(blockDim.x / WAVE);
        if (lane_id() == 0) fx[threadIdx.x / WAVE] = x;
        __syncthreads();
        Fx t = fx[0];
        for (int w = 1; w < nw; ++w) fx_add(t, fx[w]);
        __syncthreads();
        return t;
    }
    __device__ long long sum(long long x) const {
        x = wave_sum(x);
        const int nw = (int)(blockDim.x / WAVE);
        if (lane_id() == 0) ll[threadIdx.x / WAVE] = x;
        __syncthreads();
        long long t = 0;
        for (int w = 0; w < nw; ++w) t += ll[w];
        __syncthreads();
        return t;
    }
};

// Proposal of vertex v against the current state by the lane group g (uniform v).  Returns s (== r for "no
// move").  Fills m.
template <class G = WaveGroup>
__device__ int propose_block(const FtArgs &a, int v, MoveCtx &m, U4 &w, const G &g = G{}) {
    const int k_rank = g.rank(), L = g.size();
    m.v = v;
    m.o0 = a.out_off[v];
    m.o1 = a.out_off[v + 1];
    m.i0 = a.in_off[v];
    m.i1 = a.in_off[v + 1];
    const long long ko = m.o1 - m.o0, kia = m.i1 - m.i0;
    m.r = a.assign[v];
    int self_out = 0, self_in = 0;
    for (int64_t e = m.o0 + k_rank; e < m.o1; e += L) self_out += (a.out_adj[e] == v);
    for (int64_t e = m.i0 + k_rank; e < m.i1; e += L) self_in += (a.in_adj[e] == v);
    self_out = g.sum(self_out);
    self_in = g.sum(self_in);
    m.selfc = self_out;
    m.k_out = ko;
    m.k_in = kia - self_in;
    m.k_all = ko + kia;
    m.shift_in = m.k_in + (self_out > 0 ? 1 : 0);
    const long long k = ko + m.k_in;
    w = philox(a.seed, (uint32_t)v, a.epoch_ptr ? *a.epoch_ptr : a.epoch, TAG_FT, 0);
    int s = -1;
    m.t = -1;
    if (k == 0) {
        s = random_block(m.r, a.B, w.y);
    } else {
        const uint32_t idx = uniform_below(w.x, (uint32_t)k);
        int u;
        if ((long long)idx < ko) {
            u = a.out_adj[m.o0 + idx];
        } else if (self_in == 0) {
            u = a.in_adj[m.i0 + (idx - ko)];
        } else {
            long long want = (long long)idx - ko;
            u = -1;
            for (int64_t base = m.i0; base < m.i1 && u < 0; base += L) {
                int64_t e = base + k_rank;
                int x = e < m.i1 ? a.in_adj[e] : v;
                bool ok = (e < m.i1) && (x != v);
                unsigned long long msk = g.ballot(ok);
                int cnt = __popcll(msk);
                if (want < cnt) {
                    int before = __popcll(msk & ((1ull << k_rank) - 1ull));
                    unsigned long long hit = g.ballot(ok && before == want);
                    u = g.bcast(x, __ffsll((long long)hit) - 1);
                } else {
                    want -= cnt;
                }
            }
        }
        const int t = a.assign[u];
        m.t = t;
        // neighbors_weights(t): row(t) then column(t) without the diagonal; total = d_tot[t]
        const long long target = (long long)uniform_below(w.y, (uint32_t)a.d_tot[t]);
        const int rp0 = a.m.row_begin(t), rp1 = a.m.row_end(t);
        const int cp0 = a.m.col_begin(t), cp1 = a.m.col_end(t);
        if (a.m.row_cum) {
            const long long dout_t = a.d_out[t];
            if (target < dout_t) {
                int i = group_cum_search(a.m.row_cum, rp0, rp1, target, g);
                if (i < rp1) s = a.m.row_cells[i].x;
            } else {
                int i = group_cum_search(a.m.col_cum, cp0, cp1, target - dout_t, g);
                if (i < cp1) s = a.m.col_cells[i].x;
            }
        } else {
            long long base_sum = 0;
            for (int c0 = rp0; c0 < rp1 && s < 0; c0 += L) {
                int ci = c0 + k_rank;
                int2 cell = ci < rp1 ? a.m.row_cells[ci] : make_int2(0, 0);
                long long incl = g.inclusive_scan((long long)cell.y) + base_sum;
                unsigned long long hit = g.ballot(ci < rp1 && target < incl);
                if (hit) s = g.bcast(cell.x, __ffsll((long long)hit) - 1);
                else base_sum = g.bcast(incl, L - 1);
            }
            for (int c0 = cp0; c0 < cp1 && s < 0; c0 += L) {
                int ci = c0 + k_rank;
                int2 cell = ci < cp1 ? a.m.col_cells[ci] : make_int2(0, 0);
                bool ok = ci < cp1 && cell.x != t;
                long long incl = g.inclusive_scan(ok ? (long long)cell.y : 0ll) + base_sum;
                unsigned long long hit = g.ballot(ok && target < incl);
                if (hit) s = g.bcast(cell.x, __ffsll((long long)hit) - 1);
                else base_sum = g.bcast(incl, L - 1);
            }
        }
        if (s < 0) {  // d_tot[t] disagrees with the cells of row/column t: report, propose no move
            if (k_rank == 0) atomicExch(a.err, 1);
            s = m.r;
        }
    }
    m.s = s;
    return s;
}

// Algorithmic bytes of one proposal, the work-unit figure the roofline is quoted on: adjacency and
// neighbour-block gathers 8 k(v), the multinomial over t 8 rc(t), the delta-MDL row/column scans
// 12 (rc(r) + rc(s)) and the Delta cells 16 nb(v), plus 16 + 8 for the vertex and move records; the
// Hastings correction adds 8 rc(s) + 8 nb(v).  rc(b) = cells in row b + cells in column b.  A proposal of
// the current block evaluates nothing further.
__device__ __forceinline__ unsigned long long proposal_bytes(const FtArgs &a, const MoveCtx &m, int nb, bool greedy) {
    auto rc = [&](int b) -> unsigned long long {
        return b < 0 ? 0ull : (unsigned long long)((a.m.row_end(b) - a.m.row_begin(b)) + (a.m.col_end(b) - a.m.col_begin(b)));
    };
    unsigned long long x = 8ull * (unsigned long long)m.k_all + 16 + 8ull * rc(m.t) + 8;
    if (m.s != m.r) {
        x += 12ull * (rc(m.r) + rc(m.s)) + 16ull * (unsigned long long)nb;
        if (!greedy) x += 8ull * rc(m.s) + 8ull * (unsigned long long)nb;
    }
    return x;
}

// Hastings correction over the neighbour blocks (src/entropy.cpp:497-561) and the accept test
// (src/finetune.cpp:22-32); out of line so the greedy path keeps its registers.
template <class G, class T>
__device__ __noinline__ double hastings_value(const FtArgs &a, const MoveCtx &m, const T &t, int nb, long long mrr,
                                              long long mss, const G &grp) {
    const int r = m.r, s = m.s;
    Fx fwd = fx_zero(), bwd = fx_zero();
    const long long Bn = a.B;
    const bool self_only_r = m.selfc > 0 && m.out_r == 0 && m.in_r == 0;
    for (int i = grp.rank(); i < nb + 1; i += grp.size()) {
        int x, cntx, oc, ic;
        if (i < nb) {
            table_entry(t, i, x, oc, ic);
            cntx = oc + ic + (x == r ? m.selfc : 0);
        } else {
            if (!self_only_r) continue;  // the self-loop is the only link to block r
            x = r;
            oc = 0;
            ic = 0;
            cntx = m.selfc;
        }
        double counts = (double)cntx;
        double pw = (double)(get_cell(a, s, x) + get_cell(a, x, s)) + 1.0;
        double bd = (double)((long long)a.d_tot[x] + Bn);
        long long d_rx = delta_row_r(m, x, oc);
        long long d_xr = (x == r) ? delta_row_r(m, r, m.out_r) : (x == s ? delta_row_s(m, r, m.out_r) : -(long long)ic);
        double bw = (double)(get_cell(a, r, x) + d_rx + get_cell(a, x, r) + d_xr) + 1.0;
        long long nbd = a.d_tot[x];
        if (x == r) {
            long long self = mrr + delta_row_r(m, r, m.out_r);
            nbd = (a.d_out[r] - m.k_out) + (a.d_in[r] - m.k_in) - self;
        } else if (x == s) {
            long long self = mss + delta_row_s(m, s, m.out_s);
            nbd = (a.d_out[s] + m.k_out) + (a.d_in[s] + m.k_in) - self;
        }
        double pd = (double)(nbd + Bn);
        fx_acc(fwd, counts * pw / bd);
        fx_acc(bwd, counts * bw / pd);
    }
    fwd = grp.sum(fwd);
    bwd = grp.sum(bwd);
    return (m.k_out + m.k_in == 0) ? 1.0 : fx_to(bwd) / fx_to(fwd);
}

template <class G, class T>
__device__ __forceinline__ bool hastings_accept(const FtArgs &a, const MoveCtx &m, const T &t, const U4 &w,
                                                double dS, int nb, long long mrr, long long mss, const G &grp) {
    const double H = hastings_value(a, m, t, nb, mrr, mss, grp);
    double u = unit_double(w.z, w.w);
    double p = exp(-3.0 * dS) * H;
    p = p >= 1.0 ? 1.0 : p;
    return u <= p;
}

// Builds the neighbour-block histogram and returns delta-MDL; sets accept.  The table stays populated
// for the caller (Metropolis-Hastings applies the Delta from it), which clears it.
//
// delta-MDL = the Delta entries term by term + the cells (r,r), (r,s), (s,r), (s,s) that are unchanged,
// term by term + the other unchanged cells of rows r, s and columns r, s in bulk: each changes only
// through one block degree d -> d', contributing M * (fastlog(d') - fastlog(d)), so the four bulk terms
// are (row or column total - Delta and special cells) * (fastlog(d') - fastlog(d)).
// Old-minus-new entropy of one cell: value M with log degrees (lin of its column, lout of its row) before the
// move, M + d with (lin2, lout2) after -- common::cell_entropy twice, with the degree logs computed once per
// proposal instead of per cell.
__device__ __forceinline__ void cell_change(Fx &acc, long long M, long long d, double lin, double lout, double lin2,
                                            double lout2) {
    if (M) fx_acc(acc, (double)M * (fastlog(M) - (lin + lout)));
    const long long M2 = M + d;
    if (M2) fx_sub_acc(acc, (double)M2 * (fastlog(M2) - (lin2 + lout2)));
}

// Counts the blocks of v's neighbours adj[e0, e1) (self-loops excluded) into the table: each lane takes its
// edges e0 + lane + 64 j four at a time and issues their adjacency and assignment loads before the table
// updates, so the loads of a high-degree vertex overlap instead of costing two round trips per 64 edges.
// The lanes' updates race as before; the table's counts do not depend on their order.
template <class G, class T>
__device__ __forceinline__ void add_neighbour_blocks(const FtArgs &a, const T &t, const int32_t *adj, int64_t e0,
                                                     int64_t e1, int v, int dout, int din, const G &grp) {
    constexpr int U = 4;
    const int n = grp.size();
    for (int64_t e = e0 + grp.rank(); e < e1; e += U * n) {
        int u[U], b[U];
#pragma unroll
        for (int j = 0; j < U; ++j) u[j] = e + j * n < e1 ? adj[e + j * n] : v;
#pragma unroll
        for (int j = 0; j < U; ++j) b[j] = u[j] != v ? a.assign[u[j]] : -1;
#pragma unroll
        for (int j = 0; j < U; ++j)
            if (b[j] >= 0) table_add(t, b[j], dout, din);
    }
}

template <bool GREEDY, class G, class T>
__device__ double evaluate_move(const FtArgs &a, MoveCtx &m, const T &t, const U4 &w, bool &accept, const G &grp) {
    const int v = m.v, r = m.r, s = m.s;
    add_neighbour_blocks(a, t, a.out_adj, m.o0, m.o1, v, 1, 0, grp);
    add_neighbour_blocks(a, t, a.in_adj, m.i0, m.i1, v, 0, 1, grp);
    grp.sync();
    table_get(t, r, m.out_r, m.in_r);
    table_get(t, s, m.out_s, m.in_s);
    long long mrr, mrs, msr, mss;
    get_cell_pair(a, r, r, r, s, mrr, mrs);
    get_cell_pair(a, s, r, s, s, msr, mss);
    // logs of the degrees of r and s before and after the move
    const double lo_r = fastlog(a.d_out[r]), lo_r2 = fastlog(new_dout(a, m, r));
    const double lo_s = fastlog(a.d_out[s]), lo_s2 = fastlog(new_dout(a, m, s));
    const double li_r = fastlog(a.d_in[r]), li_r2 = fastlog(new_din(a, m, r));
    const double li_s = fastlog(a.d_in[s]), li_s2 = fastlog(new_din(a, m, s));
    Fx acc = fx_zero();
    long long sub_rrow = 0, sub_srow = 0, sub_rcol = 0, sub_scol = 0;
    const int nb = table_size(t);
    for (int i = grp.rank(); i < nb; i += grp.size()) {
        int c, oc, ic;
        table_entry(t, i, c, oc, ic);
        const bool special = (c == r) || (c == s);
        if (oc > 0) {
            long long mrc, msc;
            if (c == r) mrc = mrr, msc = msr;
            else if (c == s) mrc = mrs, msc = mss;
            else get_cell_pair(a, r, c, s, c, mrc, msc);
            double lc, lc2;  // column c's in-degree log, before and after
            if (c == r) lc = li_r, lc2 = li_r2;
            else if (c == s) lc = li_s, lc2 = li_s2;
            else lc = lc2 = fastlog(a.d_in[c]);
            cell_change(acc, mrc, delta_row_r(m, c, oc), lc, lo_r, lc2, lo_r2);
            cell_change(acc, msc, delta_row_s(m, c, oc), lc, lo_s, lc2, lo_s2);
            if (!special) {
                sub_rrow += mrc;
                sub_srow += msc;
            }
        }
        if (ic > 0 && !special) {
            long long mcr, mcs;
            get_cell_pair(a, c, r, c, s, mcr, mcs);
            const double lr = fastlog(a.d_out[c]);  // row c's out-degree log (unchanged)
            cell_change(acc, mcr, -(long long)ic, li_r, lr, li_r2, lr);
            cell_change(acc, mcs, (long long)ic, li_s, lr, li_s2, lr);
            sub_rcol += mcr;
            sub_scol += mcs;
        }
    }
    if (grp.rank() == 0) {
        // Delta entries not produced by an out-neighbour block
        if (m.out_r == 0 && (m.selfc > 0 || m.in_r > 0))
            cell_change(acc, mrr, -(long long)m.selfc - m.in_r, li_r, lo_r, li_r2, lo_r2);
        if (m.out_s == 0 && m.in_r > 0) cell_change(acc, mrs, (long long)m.in_r, li_s, lo_r, li_s2, lo_r2);
        if (m.out_s == 0 && (m.selfc > 0 || m.in_s > 0))
            cell_change(acc, mss, (long long)m.selfc + m.in_s, li_s, lo_s, li_s2, lo_s2);
        if (m.out_r == 0 && m.in_s > 0) cell_change(acc, msr, -(long long)m.in_s, li_r, lo_s, li_r2, lo_s2);
        // unchanged cells at rows/columns {r, s} (a zero-net Delta entry counts here as well)
        if (mrr && delta_row_r(m, r, m.out_r) == 0) cell_change(acc, mrr, 0, li_r, lo_r, li_r2, lo_r2);
        if (mrs && delta_row_r(m, s, m.out_s) == 0) cell_change(acc, mrs, 0, li_s, lo_r, li_s2, lo_r2);
        if (msr && delta_row_s(m, r, m.out_r) == 0) cell_change(acc, msr, 0, li_r, lo_s, li_r2, lo_s2);
        if (mss && delta_row_s(m, s, m.out_s) == 0) cell_change(acc, mss, 0, li_s, lo_s, li_s2, lo_s2);
    }
    acc = grp.sum(acc);
    sub_rrow = grp.sum(sub_rrow);
    sub_srow = grp.sum(sub_srow);
    sub_rcol = grp.sum(sub_rcol);
    sub_scol = grp.sum(sub_scol);
    const long long S_rrow = (long long)a.d_out[r] - mrr - mrs - sub_rrow;
    const long long S_srow = (long long)a.d_out[s] - mss - msr - sub_srow;
    const long long S_rcol = (long long)a.d_in[r] - mrr - msr - sub_rcol;
    const long long S_scol = (long long)a.d_in[s] - mss - mrs - sub_scol;
    fx_acc(acc, (double)S_rrow * (lo_r2 - lo_r));
    fx_acc(acc, (double)S_srow * (lo_s2 - lo_s));
    fx_acc(acc, (double)S_rcol * (li_r2 - li_r));
    fx_acc(acc, (double)S_scol * (li_s2 - li_s));
    const double dS = fx_to(acc);
    if (GREEDY) {
        accept = dS < 0.0;
        return dS;
    }
    accept = hastings_accept(a, m, t, w, dS, nb, mrr, mss, grp);
    return dS;
}

// Nonparametric delta-MDL of the move (entropy::nonparametric::delta_mdl, src/entropy.cpp:1135-1164): the
// Delta entries' eterm differences (unchanged cells do not change eterm, so no row or column is scanned),
// the block terms of r and s and the description-length deltas.  No Hastings correction
// (src/entropy.cpp:499); accept as finetune::accept (src/finetune.cpp:22-32).
template <bool GREEDY, class G, class T>
__device__ double evaluate_move_np(const FtArgs &a, MoveCtx &m, const T &t, const U4 &w, bool &accept,
                                   const G &grp) {
    const int v = m.v, r = m.r, s = m.s;
    add_neighbour_blocks(a, t, a.out_adj, m.o0, m.o1, v, 1, 0, grp);
    add_neighbour_blocks(a, t, a.in_adj, m.i0, m.i1, v, 0, 1, grp);
    grp.sync();
    table_get(t, r, m.out_r, m.in_r);
    table_get(t, s, m.out_s, m.in_s);
    long long mrr, mrs, msr, mss;
    get_cell_pair(a, r, r, r, s, mrr, mrs);
    get_cell_pair(a, s, r, s, s, msr, mss);
    Fx acc = fx_zero();
    const int nb = table_size(t);
    for (int i = grp.rank(); i < nb; i += grp.size()) {
        int c, oc, ic;
        table_entry(t, i, c, oc, ic);
        if (oc > 0) {
            long long mrc, msc;
            if (c == r) mrc = mrr, msc = msr;
            else if (c == s) mrc = mrs, msc = mss;
            else get_cell_pair(a, r, c, s, c, mrc, msc);
            fx_acc(acc, np_entry(a.np, mrc, delta_row_r(m, c, oc)));
            fx_acc(acc, np_entry(a.np, msc, delta_row_s(m, c, oc)));
        }
        if (ic > 0 && c != r && c != s) {
            long long mcr, mcs;
            get_cell_pair(a, c, r, c, s, mcr, mcs);
            fx_acc(acc, np_entry(a.np, mcr, -(long long)ic));
            fx_acc(acc, np_entry(a.np, mcs, (long long)ic));
        }
    }
    if (grp.rank() == 0) {  // Delta entries not produced by an out-neighbour block
        if (m.out_r == 0 && (m.selfc > 0 || m.in_r > 0)) fx_acc(acc, np_entry(a.np, mrr, -(long long)m.selfc - m.in_r));
        if (m.out_s == 0 && m.in_r > 0) fx_acc(acc, np_entry(a.np, mrs, (long long)m.in_r));
        if (m.out_s == 0 && (m.selfc > 0 || m.in_s > 0)) fx_acc(acc, np_entry(a.np, mss, (long long)m.selfc + m.in_s));
        if (m.out_r == 0 && m.in_s > 0) fx_acc(acc, np_entry(a.np, msr, -(long long)m.in_s));
    }
    acc = grp.sum(acc);
    const double dS = np_finish(a.np, fx_to(acc), r, s, a.d_out[r], a.d_in[r], a.sizes[r], a.d_out[s], a.d_in[s],
                                a.sizes[s], m.k_out, m.k_in, 1, m.i1 - m.i0, m.o1 - m.o0);
    if (GREEDY) {
        accept = dS < 0.0;
    } else {
        const double u = unit_double(w.z, w.w);
        double p = exp(-3.0 * dS);
        p = p >= 1.0 ? 1.0 : p;
        accept = u <= p;
    }
    return dS;
}

template <bool GREEDY, bool NP, class G = WaveGroup, class T = Table>
__device__ __forceinline__ double evaluate(const FtArgs &a, MoveCtx &m, const T &t, const U4 &w, bool &accept,
                                           const G &grp = G{}) {
    if constexpr (NP) return evaluate_move_np<GREEDY>(a, m, t, w, accept, grp);
    else return evaluate_move<GREEDY>(a, m, t, w, accept, grp);
}

// Proposals of one degree class by lane groups: L lanes per vertex, WAVE / L vertices per wavefront, each group
// with its own neighbour-block table in LDS sized for the class's largest degree (distinct neighbour blocks <=
// |out| + |in|).  Small vertices keep few lanes busy, so packing several into a wavefront keeps more of their
// dependent load chains (offsets -> neighbour -> its block -> the multinomial -> the cell lookups) in flight.
// The result of each vertex does not depend on L: its sums are fixed point.
// OCC: wavefronts per SIMD the register budget is sized for (4: 128 VGPRs; 3: 168, which the greedy parametric
// evaluation needs to run without spills).
template <int L, bool GREEDY, bool NP, int OCC>
__global__ void __launch_bounds__(FT_BLOCK) __attribute__((amdgpu_waves_per_eu(OCC))) k_ft_lanes(FtArgs a, int bucket) {
    __shared__ int lds[FT_WAVES * LDS_WAVE_INTS];
    constexpr int GPW = WAVE / L;  // lane groups per wavefront
    const LaneGroup<L> grp;
    const int wid = threadIdx.x / WAVE, gi = lane_id() / L;
    const int dmax = ft_bucket_max(bucket), lg = ft_table_log2(bucket, L);
    const Table t = table_at(lds + wid * LDS_WAVE_INTS + gi * (int)table_ints(lg, dmax), lg, dmax);
    group_table_clear_all(t, grp);
    const int64_t count = a.bmeta[bucket];
    const int32_t *items = a.blist + a.bmeta[FT_NB + bucket];
    const int64_t stride = (int64_t)gridDim.x * FT_WAVES * GPW;
    unsigned long long bytes = 0;
    for (int64_t item = ((int64_t)blockIdx.x * FT_WAVES + wid) * GPW + gi; item < count; item += stride) {
        const int v = items[item];
        MoveCtx m;
        U4 w;
        const int s = propose_block(a, v, m, w, grp);
        if (a.s_out && grp.rank() == 0) a.s_out[v] = s;
        if (s == m.r) {
            if (grp.rank() == 0) {
                a.move_s[v] = -1;
                if (a.dS_out) a.dS_out[v] = 0.0;
            }
            bytes += proposal_bytes(a, m, 0, GREEDY);
            continue;
        }
        bool accept;
        const double dS = evaluate<GREEDY, NP>(a, m, t, w, accept, grp);
        if (grp.rank() == 0) {
            a.move_s[v] = accept ? s : -1;
            if (a.dS_out) a.dS_out[v] = dS;
        }
        bytes += proposal_bytes(a, m, *t.n_used, GREEDY);
        group_table_clear_used(t, grp);
    }
    // one count per workgroup
    __shared__ unsigned long long block_bytes;
    if (threadIdx.x == 0) block_bytes = 0;
    __syncthreads();
    bytes = wave_sum(grp.rank() == 0 ? bytes : 0ull);
    if (lane_id() == 0 && bytes) atomicAdd(&block_bytes, bytes);
    __syncthreads();
    if (threadIdx.x == 0 && block_bytes) atomicAdd(a.bytes, block_bytes);
}

// Vertices of total degree > SMALL_K, one workgroup each: THREADS threads split the vertex's edges and its distinct
// neighbour blocks, with the neighbour-block table in LDS (2^CAPLG slots).  A vertex that may have more distinct
// neighbour blocks than half the table's slots (min(k, B) > 2^(CAPLG-1)) uses its workgroup's table in global
// memory instead.  Wavefront 0 draws the proposal; the delta-MDL, the Hastings correction and the accept test are
// those of the one-wavefront evaluation, summed in fixed point over the workgroup (src/finetune.cpp:447-462,
// 871-901).  Degree class FT_NB-2 (k <= 1,024) runs HUB_MID_THREADS-thread workgroups with 2,048-slot tables
// (28 KB), the last class 512-thread workgroups with 4,096-slot tables (57 KB: two workgroups per CU).
constexpr int HUB_THREADS = 512;
constexpr int HUB_CAP_LOG2 = 12;
constexpr int HUB_CAP = 1 << HUB_CAP_LOG2;
constexpr int HUB_MID_CAP_LOG2 = 11;
// Block counts up to which the last class indexes its table by block id in LDS (74 KB: two workgroups per CU).
constexpr int HUB_DENSE_MAX = 6144;
static_assert((1 << (HUB_MID_CAP_LOG2 - 1)) >= ft_bucket_max(FT_NB - 2), "mid-hub tables hold every block");

// Given moves (vertex, target block) to evaluate instead of proposals to draw: entropy::delta_mdl and
// hastings_correction of each (src/entropy.cpp:298-403,497-561), as the proposal kernels compute them.
struct EvalIO {
    const int32_t *vertices, *targets;
    int64_t n;
    double *dS;  // [n]
    double *H;   // [n] (1 for the nonparametric description length, src/entropy.cpp:499)
};

// One high-degree item's evaluation after its proposal: delta-MDL (and, for EVAL, the Hastings correction) with
// table t, the outputs, and the table emptied again.
template <bool GREEDY, bool NP, bool EVAL, int THREADS, class T>
__device__ __forceinline__ void hub_finish(const FtArgs &a, MoveCtx &m, const T &t, const U4 &w, int s, int v,
                                           int64_t item, const EvalIO &io, const BlockGroup &grp,
                                           unsigned long long &bytes) {
    bool accept;
    const double dS = evaluate<EVAL ? true : GREEDY, NP>(a, m, t, w, accept, grp);
    const int nb = table_size(t);
    if (EVAL) {
        double H = 1.0;
        if (!NP) {
            long long mrr, mss;
            get_cell_pair(a, m.r, m.r, m.s, m.s, mrr, mss);
            H = hastings_value(a, m, t, nb, mrr, mss, grp);
        }
        if (threadIdx.x == 0) {
            io.dS[item] = dS;
            io.H[item] = H;
        }
    } else if (threadIdx.x == 0) {
        a.move_s[v] = accept ? s : -1;
        if (a.dS_out) a.dS_out[v] = dS;
        bytes += proposal_bytes(a, m, nb, GREEDY);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += THREADS) {
        if constexpr (std::is_same<T, DenseTable>::value) {
            t.cnt[t.used[i]] = 0ull;
        } else {
            const int slot = t.used[i];
            t.keys[slot] = EMPTY;
            t.outc[slot] = 0;
            t.inc[slot] = 0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *t.n_used = 0;
    __syncthreads();
}

// Where a hub's neighbour-block table lives: the LDS hash table (every vertex of the class fits it), the LDS table
// indexed by block id (B <= HUB_DENSE_MAX), or per vertex the LDS hash table or a global-memory one.  The first two
// let the compiler address the table as LDS (ds_ atomics) rather than through flat addressing.
enum HubTable { HUB_LDS = 0, HUB_DENSE = 1, HUB_LDS_OR_GLOBAL = 2 };

template <bool GREEDY, bool NP, bool EVAL = false, int THREADS = HUB_THREADS, int CAPLG = HUB_CAP_LOG2,
          int MODE = HUB_LDS_OR_GLOBAL>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4)))
k_ft_hub(FtArgs a, int bucket, EvalIO io = EvalIO{}) {
    constexpr int CAP = 1 << CAPLG;
    constexpr bool DENSE = MODE == HUB_DENSE;
    constexpr int64_t LDS_INTS = DENSE ? dense_table_ints(HUB_DENSE_MAX, HUB_DENSE_MAX) : table_ints(CAPLG, CAP / 2);
    __shared__ unsigned long long lds64[(LDS_INTS + 1) / 2];
    int *lds = reinterpret_cast<int *>(lds64);
    __shared__ Fx red_fx[THREADS / WAVE];
    __shared__ long long red_ll[THREADS / WAVE];
    __shared__ MoveCtx s_m;
    __shared__ U4 s_w;
    __shared__ int s_s;
    BlockGroup grp;
    grp.fx = red_fx;
    grp.ll = red_ll;
    const Table lt = table_at(lds, CAPLG, CAP / 2);
    const DenseTable dt = dense_table_at(lds, HUB_DENSE_MAX);
    if constexpr (DENSE) {
        for (int i = threadIdx.x; i < a.B; i += THREADS) dt.cnt[i] = 0ull;
        if (threadIdx.x == 0) *dt.n_used = 0;
    } else {
        for (int i = threadIdx.x; i < CAP; i += THREADS) {
            lt.keys[i] = EMPTY;
            lt.outc[i] = 0;
            lt.inc[i] = 0;
        }
        if (threadIdx.x == 0) *lt.n_used = 0;
    }
    __syncthreads();
    const Table gt = table_at(a.gtable + (int64_t)blockIdx.x * table_ints(a.gcap_log2), a.gcap_log2);
    const int64_t count = EVAL ? io.n : a.bmeta[bucket];
    const int32_t *hubs = EVAL ? io.vertices : a.blist + a.bmeta[FT_NB + bucket];
    unsigned long long bytes = 0;
    for (int64_t item = blockIdx.x; item < count; item += gridDim.x) {
        const int v = hubs[item];
        if (threadIdx.x < WAVE) {
            MoveCtx m;
            U4 w;
            int s = propose_block(a, v, m, w);
            if (EVAL) s = m.s = io.targets[item];
            if (threadIdx.x == 0) {
                s_m = m;
                s_w = w;
                s_s = s;
            }
        }
        __syncthreads();
        MoveCtx m = s_m;
        const U4 w = s_w;
        const int s = s_s;
        if (!EVAL && threadIdx.x == 0 && a.s_out) a.s_out[v] = s;
        if (s == m.r) {
            if (threadIdx.x == 0) {
                if (EVAL) {
                    io.dS[item] = 0.0;
                    io.H[item] = 1.0;
                } else {
                    a.move_s[v] = -1;
                    if (a.dS_out) a.dS_out[v] = 0.0;
                    bytes += proposal_bytes(a, m, 0, GREEDY);
                }
            }
            __syncthreads();  // s_m is written again for the next vertex
            continue;
        }
        if constexpr (MODE == HUB_DENSE) {
            hub_finish<GREEDY, NP, EVAL, THREADS>(a, m, dt, w, s, v, item, io, grp, bytes);
        } else if constexpr (MODE == HUB_LDS) {
            hub_finish<GREEDY, NP, EVAL, THREADS>(a, m, lt, w, s, v, item, io, grp, bytes);
        } else {
            const bool in_lds = 2 * (m.k_all < (long long)a.B ? m.k_all : (long long)a.B) <= CAP;
            if (in_lds) hub_finish<GREEDY, NP, EVAL, THREADS>(a, m, lt, w, s, v, item, io, grp, bytes);
            else hub_finish<GREEDY, NP, EVAL, THREADS>(a, m, gt, w, s, v, item, io, grp, bytes);
        }
    }
    if (threadIdx.x == 0 && bytes) atomicAdd(a.bytes, bytes);
}

template <int THREADS, int CAPLG, int MODE>
static void launch_hub_t(int blocks, hipStream_t st, const FtArgs &a, int bucket, bool greedy, bool np) {
    if (np) {
        if (greedy) k_ft_hub<true, true, false, THREADS, CAPLG, MODE><<<blocks, THREADS, 0, st>>>(a, bucket);
        else k_ft_hub<false, true, false, THREADS, CAPLG, MODE><<<blocks, THREADS, 0, st>>>(a, bucket);
    } else {
        if (greedy) k_ft_hub<true, false, false, THREADS, CAPLG, MODE><<<blocks, THREADS, 0, st>>>(a, bucket);
        else k_ft_hub<false, false, false, THREADS, CAPLG, MODE><<<blocks, THREADS, 0, st>>>(a, bucket);
    }
}

// The workgroup class FT_NB-2 (k <= 1,024) with `threads` (128 or 256) per vertex; the last class with 512 and,
// when the blocks fit (B <= HUB_DENSE_MAX), the table indexed by block id.
static void launch_hub(int blocks, hipStream_t st, const FtArgs &a, int bucket, int threads, bool greedy, bool np) {
    if (bucket == FT_NB - 1) {
        if (a.B <= HUB_DENSE_MAX) launch_hub_t<HUB_THREADS, HUB_CAP_LOG2, HUB_DENSE>(blocks, st, a, bucket, greedy, np);
        else launch_hub_t<HUB_THREADS, HUB_CAP_LOG2, HUB_LDS_OR_GLOBAL>(blocks, st, a, bucket, greedy, np);
    } else if (threads == 128) {
        launch_hub_t<128, HUB_MID_CAP_LOG2, HUB_LDS>(blocks, st, a, bucket, greedy, np);
    } else {
        launch_hub_t<256, HUB_MID_CAP_LOG2, HUB_LDS>(blocks, st, a, bucket, greedy, np);
    }
}

template <int L>
static void launch_lanes_t(int blocks, hipStream_t st, const FtArgs &a, int bucket, bool greedy, bool np, int occ) {
    if (np) {
        if (greedy) k_ft_lanes<L, true, true, 4><<<blocks, FT_BLOCK, 0, st>>>(a, bucket);
        else k_ft_lanes<L, false, true, 4><<<blocks, FT_BLOCK, 0, st>>>(a, bucket);
    } else if (greedy) {
        if (occ == 3) k_ft_lanes<L, true, false, 3><<<blocks, FT_BLOCK, 0, st>>>(a, bucket);
        else k_ft_lanes<L, true, false, 4><<<blocks, FT_BLOCK, 0, st>>>(a, bucket);
    } else {
        k_ft_lanes<L, false, false, 4><<<blocks, FT_BLOCK, 0, st>>>(a, bucket);
    }
}

// Lanes per vertex that a degree class may use: a power of two from 8 to 64 whose tables fit the LDS budget; and
// the wavefronts per SIMD of the greedy parametric launch (3 or 4).
static bool shape_allowed(int bucket, int L, int occ) {
    return (L == 2 || L == 4 || L == 8 || L == 16 || L == 32 || L == 64) && bucket >= 0 && bucket < FT_NB - 2 && ft_lanes_fit(bucket, L) &&
           (occ == 3 || occ == 4);
}

// One lane-group launch over degree class `bucket` (items counted on the device), sized for up to `upper` items.
static void launch_lanes(int64_t upper, hipStream_t st, const FtArgs &a, int bucket, int L, int occ, bool greedy,
                         bool np) {
    const int64_t per_block = (int64_t)FT_WAVES * (WAVE / L);
    int64_t blocks = (upper + per_block - 1) / per_block;
    blocks = std::max<int64_t>(1, std::min<int64_t>(blocks, 4096));
    switch (L) {
        case 2: launch_lanes_t<2>((int)blocks, st, a, bucket, greedy, np, occ); break;
        case 4: launch_lanes_t<4>((int)blocks, st, a, bucket, greedy, np, occ); break;
        case 8: launch_lanes_t<8>((int)blocks, st, a, bucket, greedy, np, occ); break;
        case 16: launch_lanes_t<16>((int)blocks, st, a, bucket, greedy, np, occ); break;
        case 32: launch_lanes_t<32>((int)blocks, st, a, bucket, greedy, np, occ); break;
        default: launch_lanes_t<64>((int)blocks, st, a, bucket, greedy, np, occ); break;
    }
}

static int log2_ceil(int64_t x) {
    int b = 0;
    while ((int64_t(1) << b) < x) ++b;
    return b;
}

static FtArgs base_args(sbp_ctx ctx, sbp_bm bm, sbp_graph g, uint32_t epoch, int greedy) {
    FtArgs a;
    a.N = g->N;
    a.B = bm->B;
    a.seed = ctx->seed;
    a.epoch = epoch;
    a.epoch_ptr = nullptr;
    a.greedy = greedy;
    a.out_off = g->out_off;
    a.in_off = g->in_off;
    a.out_adj = g->out_adj;
    a.in_adj = g->in_adj;
    a.assign = bm->assign;
    a.sizes = bm->sizes;
    a.d_out = bm->d_out;
    a.d_in = bm->d_in;
    a.d_tot = bm->d_tot;
    a.m.rbeg = bm->row_ptr;
    a.m.rlen = nullptr;
    a.m.cbeg = bm->col_ptr;
    a.m.clen = nullptr;
    a.m.row_cells = bm->row_cells;
    a.m.col_cells = bm->col_cells;
    a.m.row_cum = bm->row_cum;
    a.m.col_cum = bm->col_cum;
    a.m.hash = nullptr;
    a.m.hash_shift = 0;
    a.m.hash_mask = 0;
    a.m.dense = nullptr;
    a.m.dense_w = nullptr;
    a.m.dense_stride = 0;
    a.items = nullptr;
    a.count = 0;
    a.move_s = nullptr;
    a.dS_out = nullptr;
    a.s_out = nullptr;
    a.bytes = nullptr;
    a.gtable = nullptr;
    a.gcap_log2 = 0;
    a.err = nullptr;
    a.blist = nullptr;
    a.bmeta = nullptr;
    a.np = NpConst{};
    a.hist = nullptr;
    if (bm->np) {
        bm_set_model(ctx, g, bm, bm->np, bm->dc);
        a.np = np_const(ctx, bm);
        if (bm->dc) a.hist = bm->hist;
    }
    return a;
}

// Empties `tables` global neighbour-block tables of capacity 2^lg at the start of slot S_HASH, unless they are
// already known to be empty: each use empties the slots it claims, so they stay empty between launches.
__global__ void k_table_init(int *base, int lg, int64_t tables) {
    const int64_t cap = (int64_t)1 << lg, ints = table_ints(lg);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tables * cap;
         i += (int64_t)gridDim.x * blockDim.x) {
        int *t = base + (i >> lg) * ints;
        const int64_t k = i & (cap - 1);
        t[k] = EMPTY;
        t[cap + k] = 0;
        t[2 * cap + k] = 0;
        if (k == 0) t[4 * cap] = 0;
    }
}

static int *prepare_tables(sbp_ctx ctx, int lg, int64_t tables) {
    int *base = ctx->scratch[S_HASH].as<int>((size_t)tables * (size_t)table_ints(lg));
    const uint64_t gen = ctx->scratch[S_HASH].generation;
    if (gen != ctx->hash_clean_gen || lg != ctx->hash_clean_lg || tables > ctx->hash_clean_tables) {
        const int64_t n = tables << lg;
        const int blocks = (int)std::min<int64_t>((n + 255) / 256, 16384);
        k_table_init<<<std::max(blocks, 1), 256, 0, ctx->stream>>>(base, lg, tables);
        SBP_HIP(hipGetLastError());
        ctx->hash_clean_gen = gen;
        ctx->hash_clean_lg = lg;
        ctx->hash_clean_tables = tables;
    }
    return base;
}

// ------------------------------------------------------------------ degree classes

__device__ __forceinline__ int ft_class_of(long long k) {
    return k <= ft_bucket_max(0)   ? 0
           : k <= ft_bucket_max(1) ? 1
           : k <= ft_bucket_max(2) ? 2
           : k <= ft_bucket_max(3) ? 3
           : k <= ft_bucket_max(4) ? 4
                                   : 5;
}

__global__ void k_vertex_class(int32_t N, const int64_t *__restrict__ out_off, const int64_t *__restrict__ in_off,
                               uint8_t *__restrict__ cls) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        cls[v] = (uint8_t)ft_class_of((out_off[v + 1] - out_off[v]) + (in_off[v + 1] - in_off[v]));
}

// meta: [0, FT_NB) counts, [FT_NB, 2 FT_NB) offsets, [2 FT_NB, 3 FT_NB) scatter cursors
__global__ void k_class_zero(int32_t *__restrict__ meta) {
    if (threadIdx.x < 3 * FT_NB) meta[threadIdx.x] = 0;
}

// Counts the items of each class with one atomic per wavefront and class.
__global__ void k_class_count(int64_t n, const int32_t *__restrict__ items, const int64_t *__restrict__ out_off,
                              const int64_t *__restrict__ in_off, int32_t *__restrict__ meta) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int v = items[i];
        const int c = ft_class_of((out_off[v + 1] - out_off[v]) + (in_off[v + 1] - in_off[v]));
#pragma unroll
        for (int b = 0; b < FT_NB; ++b) {
            const unsigned long long msk = __ballot(c == b);
            if (msk && lane_id() == __ffsll((long long)msk) - 1) atomicAdd(&meta[b], __popcll(msk));
        }
    }
}

// Writes each item to its class's range of list (order within a class is immaterial: every proposal reads only
// the frozen state).
__global__ void k_class_scatter(int64_t n, const int32_t *__restrict__ items, const int64_t *__restrict__ out_off,
                                const int64_t *__restrict__ in_off, int32_t *__restrict__ meta,
                                int32_t *__restrict__ list) {
    int off[FT_NB];
    int run = 0;
#pragma unroll
    for (int b = 0; b < FT_NB; ++b) {
        off[b] = run;
        run += meta[b];
    }
    if (blockIdx.x == 0 && threadIdx.x < FT_NB) meta[FT_NB + threadIdx.x] = off[threadIdx.x];
    const int lane = lane_id();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int v = items[i];
        const int c = ft_class_of((out_off[v + 1] - out_off[v]) + (in_off[v + 1] - in_off[v]));
#pragma unroll
        for (int b = 0; b < FT_NB; ++b) {
            const unsigned long long msk = __ballot(c == b);
            if (!msk) continue;
            const int leader = __ffsll((long long)msk) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(&meta[2 * FT_NB + b], __popcll(msk));
            base = __shfl(base, leader, WAVE);
            if (c == b) list[off[b] + base + __popcll(msk & ((1ull << lane) - 1ull))] = v;
        }
    }
}

// Every vertex of g grouped by degree class, ascending id within a class; built once per graph (degrees do not
// change) and published under deg_mu once complete, since workers on other streams may share the graph.
static void graph_classes(sbp_ctx ctx, sbp_graph g) {
    std::lock_guard<std::mutex> lk(g->deg_mu);
    if (g->ft_blist) return;
    hipStream_t st = ctx->stream;
    const int32_t N = g->N;
    uint8_t *d_cls = reinterpret_cast<uint8_t *>(ctx->scratch[S_BUCKETS].as<int32_t>((size_t)N / 4 + 16));
    std::vector<uint8_t> cls((size_t)N);
    if (N > 0) {
        k_vertex_class<<<(int)std::min<int64_t>((N + 255) / 256, 16384), 256, 0, st>>>(N, g->out_off, g->in_off, d_cls);
        SBP_HIP(hipGetLastError());
        SBP_HIP(hipMemcpyAsync(cls.data(), d_cls, (size_t)N, hipMemcpyDeviceToHost, st));
        SBP_HIP(hipStreamSynchronize(st));
    }
    int64_t cnt[FT_NB] = {};
    for (uint8_t c : cls) cnt[c]++;
    std::vector<int32_t> meta(2 * FT_NB);
    int64_t run = 0;
    for (int b = 0; b < FT_NB; ++b) {
        meta[b] = (int32_t)cnt[b];
        meta[FT_NB + b] = (int32_t)run;
        run += cnt[b];
    }
    std::vector<int32_t> list((size_t)std::max<int32_t>(N, 1));
    int64_t pos[FT_NB];
    for (int b = 0; b < FT_NB; ++b) pos[b] = meta[FT_NB + b];
    for (int32_t v = 0; v < N; ++v) list[(size_t)pos[cls[v]]++] = v;
    int32_t *d_list = nullptr, *d_meta = nullptr;
    SBP_HIP(dev_malloc(&d_list, sizeof(int32_t) * list.size()));
    if (dev_malloc(&d_meta, sizeof(int32_t) * meta.size()) != hipSuccess) {
        (void)dev_free(d_list);
        SBP_HIP(hipErrorOutOfMemory);
    }
    hipError_t e = dev_copy(d_list, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dev_copy(d_meta, meta.data(), sizeof(int32_t) * meta.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)dev_free(d_list);
        (void)dev_free(d_meta);
        SBP_HIP(e);
    }
    for (int b = 0; b < FT_NB; ++b) g->ft_bcount[b] = cnt[b];
    g->ft_bmeta = d_meta;
    g->ft_blist = d_list;
}

FtLaunchInfo ft_propose(sbp_ctx ctx, sbp_bm bm, sbp_graph g, uint32_t epoch, int greedy, const int32_t *d_vertices,
                        int64_t count, int32_t *d_move_s, double *d_dS, int32_t *d_s_out, const uint32_t *epoch_ptr) {
    hipStream_t st = ctx->stream;
    const bool capture = epoch_ptr != nullptr;  // recorded into a graph: no events, no host reads
    FtLaunchInfo info{};
    info.proposals = count;
    if (count == 0) return info;
    // [0, FT_NB): algorithmic bytes per degree class; [FT_NB]: the error flag
    auto *counters = ctx->scratch[S_COUNTER].as<unsigned long long>(8);
    SBP_HIP(sbpprim::fill<unsigned long long>(counters, 8, 0ull, st));
    FtArgs a = base_args(ctx, bm, g, epoch, greedy);
    a.epoch_ptr = epoch_ptr;
    // lookups read the dense copy when it is kept, else the cell hash (a recorded sweep keeps to the hash, which
    // each replay rebuilds); the hash is then neither built nor updated, and its other users build it on demand
    a.m.dense = capture ? nullptr : bm_ensure_dense(ctx, bm, g);
    a.m.dense_stride = bm->B;
    if (!a.m.dense) {
        bm_ensure_hash(ctx, bm, capture);
        a.m.hash = bm->cell_hash;
        a.m.hash_shift = 64 - bm->hash_log2;
        a.m.hash_mask = (1u << bm->hash_log2) - 1;
    }
    a.err = reinterpret_cast<int *>(counters + FT_NB);
    a.items = d_vertices;
    a.count = count;
    a.move_s = d_move_s;
    a.dS_out = d_dS;
    a.s_out = d_s_out;
    // the items grouped by degree class: every vertex (the common sweep) from the graph's grouping, any other list
    // grouped here on the device; upper[b] bounds class b's size
    int64_t upper[FT_NB];
    if (d_vertices == g->edge_src + g->E && count == g->N) {
        graph_classes(ctx, g);
        a.blist = g->ft_blist;
        a.bmeta = g->ft_bmeta;
        for (int b = 0; b < FT_NB; ++b) upper[b] = g->ft_bcount[b];
    } else {
        SBP_CHECK(!capture || count == g->N || ctx->scratch[S_BUCKETS].bytes >= sizeof(int32_t) * ((size_t)count + 4 * FT_NB) + 16,
                  SBP_E_STATE, "captured proposals: degree-class scratch not reserved");
        int32_t *meta = ctx->scratch[S_BUCKETS].as<int32_t>((size_t)count + 4 * FT_NB);
        int32_t *list = meta + 4 * FT_NB;
        const int grid = (int)std::min<int64_t>((count + 255) / 256, 4096);
        k_class_zero<<<1, WAVE, 0, st>>>(meta);
        k_class_count<<<grid, 256, 0, st>>>(count, d_vertices, g->out_off, g->in_off, meta);
        k_class_scatter<<<grid, 256, 0, st>>>(count, d_vertices, g->out_off, g->in_off, meta, list);
        SBP_HIP(hipGetLastError());
        a.blist = list;
        a.bmeta = meta;
        for (int b = 0; b < FT_NB; ++b) upper[b] = b < FT_NB - 2 ? count : std::min<int64_t>(count, g->n_high_degree);
    }
    auto mark = [&](int i) {
        if (!capture) SBP_HIP(hipEventRecord(ctx->ft_ev[i], st));
    };
    mark(0);
    for (int b = 0; b < FT_NB - 2; ++b) {
        info.class_launched[b] = upper[b] > 0;
        if (info.class_launched[b]) {
            a.bytes = counters + b;
            launch_lanes(upper[b], st, a, b, ctx->ft_lanes[b], ctx->ft_occ[b], greedy != 0, bm->np != 0);
        }
        mark(b + 1);
    }
    SBP_HIP(hipGetLastError());
    // k <= 1,024: one workgroup per vertex, LDS tables only (2 min(k, B) <= 2,048 slots)
    info.class_launched[FT_NB - 2] = upper[FT_NB - 2] > 0;
    if (info.class_launched[FT_NB - 2]) {
        a.bytes = counters + FT_NB - 2;
        launch_hub((int)std::min<int64_t>(upper[FT_NB - 2], 2048), st, a, FT_NB - 2, ctx->ft_lanes[FT_NB - 2],
                   greedy != 0, bm->np != 0);
    }
    mark(FT_NB - 1);
    info.class_launched[FT_NB - 1] = upper[FT_NB - 1] > 0;
    if (info.class_launched[FT_NB - 1]) {
        a.bytes = counters + FT_NB - 1;
        // one workgroup per vertex, up to 1,024 at once; global tables (one per workgroup, each holding the
        // distinct neighbour blocks of one vertex at load factor <= 1/2, within kMaxTableBytes) only when a vertex
        // can have more distinct neighbour blocks than the LDS table takes
        const int64_t kmax = std::min<int64_t>(g->max_degree, bm->B);
        const int lg = std::max(10, log2_ceil(2 * kmax + 2));
        const size_t tints = (size_t)table_ints(lg);
        constexpr size_t kMaxTableBytes = size_t(1) << 30;
        int64_t blocks_h = std::min<int64_t>(upper[FT_NB - 1], 1024);
        a.gtable = nullptr;
        if (2 * kmax > HUB_CAP && bm->B > HUB_DENSE_MAX) {
            blocks_h = std::min<int64_t>(blocks_h, (int64_t)(kMaxTableBytes / (tints * sizeof(int))));
            a.gtable = prepare_tables(ctx, lg, std::max<int64_t>(blocks_h, 1));
        }
        a.gcap_log2 = lg;
        launch_hub((int)std::max<int64_t>(blocks_h, 1), st, a, FT_NB - 1, HUB_THREADS, greedy != 0, bm->np != 0);
    }
    SBP_HIP(hipGetLastError());
    mark(FT_NB);
    if (capture) return info;  // counters stay on the device (ft_propose_counters)
    unsigned long long hb[8] = {};
    read_small(ctx, {{hb, counters, sizeof(hb)}});
    SBP_CHECK(hb[FT_NB] == 0, SBP_E_STATE, "finetune proposal: block degrees inconsistent with the block matrix");
    for (int b = 0; b < FT_NB; ++b) {
        info.class_bytes[b] = hb[b];
        info.class_ms[b] = info.class_launched[b] ? (float)elapsed_ms(ctx, ctx->ft_ev[b], ctx->ft_ev[b + 1]) : 0.f;
        if (b < FT_NB - 2) {
            info.bytes += hb[b];
            info.kernel_ms += info.class_ms[b];
        } else {
            info.high_bytes += hb[b];
            info.high_kernel_ms += info.class_ms[b];
        }
    }
    info.high_launched = info.class_launched[FT_NB - 2] || info.class_launched[FT_NB - 1];
    return info;
}

// Buffers a recorded sweep's proposals and commits need, at their largest, reserved before recording: the graph's
// degree classes, the per-batch degree-class lists (batch_max items), the commit's event arrays and sort, and the
// high-degree launch's global tables.
void ft_reserve(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int64_t batch_max) {
    const int64_t N = g->N, B = bm->B;
    graph_classes(ctx, g);
    ctx->scratch[S_COUNTER].as<unsigned long long>(8);
    ctx->scratch[S_BUCKETS].as<int32_t>((size_t)batch_max + 4 * FT_NB);
    ctx->scratch[S_MOVES].as<int32_t>(N);
    ctx->scratch[S_RISK].as<int32_t>(2 * (size_t)B + 8);
    ctx->scratch[S_FLAGS].as<int32_t>(16);
    ctx->scratch[S_EVENTS_A].as<unsigned long long>(2 * (size_t)N + 2);
    ctx->scratch[S_EVENTS_B].as<unsigned long long>(2 * (size_t)N + 2);
    ctx->scratch[S_ACCEPT].as<int32_t>(2 * (size_t)N + 2);
    ctx->scratch[S_MISC].as<int32_t>(3 * (size_t)B + 2);
    ctx->scratch[S_ORDER].as<int32_t>(N);
    ctx->scratch[S_KEYS_A].as<unsigned long long>(std::max<int64_t>(N, 1));
    ctx->scratch[S_KEYS_B].as<unsigned long long>(std::max<int64_t>(N, 1));
    size_t a = 0, b = 0;
    const int32_t *no_vals = nullptr;
    unsigned long long *k = nullptr;
    SBP_HIP(sbpprim::radix_sort(nullptr, a, (const unsigned long long *)k, k, no_vals, (int32_t *)nullptr, 2 * N, 0, 64,
                                ctx->stream));
    SBP_HIP(sbpprim::radix_sort(nullptr, b, (const unsigned long long *)k, k, no_vals, (int32_t *)nullptr, N, 0, 64,
                                ctx->stream));
    ctx->scratch[S_SORT_TMP].get(std::max(a, b));
    const int64_t kmax = std::min<int64_t>(g->max_degree, B);
    if (g->ft_bcount[FT_NB - 1] > 0 && 2 * kmax > HUB_CAP && B > HUB_DENSE_MAX) {
        const int lg = std::max(10, log2_ceil(2 * kmax + 2));
        int64_t blocks_h = std::min<int64_t>(g->ft_bcount[FT_NB - 1], 1024);
        blocks_h = std::min<int64_t>(blocks_h, (int64_t)((size_t(1) << 30) / ((size_t)table_ints(lg) * sizeof(int))));
        prepare_tables(ctx, lg, std::max<int64_t>(blocks_h, 1));
    }
}

// The counters the proposal launches leave on the device: algorithmic bytes per degree class, then the error flag.
const unsigned long long *ft_propose_counters(sbp_ctx ctx) {
    return ctx->scratch[S_COUNTER].as<unsigned long long>(8);
}

// Launch shape of the degree classes (sbp_set_finetune_shape).
bool set_finetune_shape(sbp_ctx ctx, const int32_t *lanes, const int32_t *waves_per_simd) {
    for (int b = 0; b < FT_NB - 2; ++b)
        if (!shape_allowed(b, lanes[b], waves_per_simd[b])) return false;
    if (lanes[FT_NB - 2] != 128 && lanes[FT_NB - 2] != 256) return false;
    for (int b = 0; b < FT_NB - 1; ++b) ctx->ft_lanes[b] = lanes[b];
    for (int b = 0; b < FT_NB - 2; ++b) ctx->ft_occ[b] = waves_per_simd[b];
    return true;
}

void ft_evaluate(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int64_t n, const int32_t *d_vertices, const int32_t *d_targets,
                 double *d_dS, double *d_H) {
    if (n == 0) return;
    hipStream_t st = ctx->stream;
    auto *counters = ctx->scratch[S_COUNTER].as<unsigned long long>(4);
    SBP_HIP(hipMemsetAsync(counters, 0, 4 * sizeof(unsigned long long), st));
    FtArgs a = base_args(ctx, bm, g, 0, 1);
    bm_ensure_hash(ctx, bm);
    a.m.hash = bm->cell_hash;
    a.m.hash_shift = 64 - bm->hash_log2;
    a.m.hash_mask = (1u << bm->hash_log2) - 1;
    a.bytes = counters;
    a.err = reinterpret_cast<int *>(counters + 1);
    const int64_t kmax = std::min<int64_t>(g->max_degree, bm->B);
    const int lg = std::max(10, log2_ceil(2 * kmax + 2));
    int64_t blocks = std::min<int64_t>(n, 1024);
    a.gtable = nullptr;
    if (2 * kmax > HUB_CAP) {
        blocks = std::min<int64_t>(blocks, (int64_t)((size_t(1) << 30) / ((size_t)table_ints(lg) * sizeof(int))));
        a.gtable = prepare_tables(ctx, lg, std::max<int64_t>(blocks, 1));
    }
    a.gcap_log2 = lg;
    const EvalIO io{d_vertices, d_targets, n, d_dS, d_H};
    if (bm->np) k_ft_hub<true, true, true><<<(int)blocks, HUB_THREADS, 0, st>>>(a, 0, io);
    else k_ft_hub<true, false, true><<<(int)blocks, HUB_THREADS, 0, st>>>(a, 0, io);
    SBP_HIP(hipGetLastError());
    unsigned long long hb[2] = {0, 0};
    read_small(ctx, {{hb, counters, sizeof(hb)}});
    SBP_CHECK(hb[1] == 0, SBP_E_STATE, "delta-MDL evaluation: block degrees inconsistent with the block matrix");
}

// ------------------------------------------------------------------ async-Gibbs commit

__global__ void k_count_out(int32_t N, const int32_t *__restrict__ move_s, const int32_t *__restrict__ assign,
                            int32_t *__restrict__ out_count) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        if (move_s[v] >= 0) atomicAdd(&out_count[assign[v]], 1);
}

__global__ void k_mark_risk(int32_t B, const int32_t *__restrict__ out_count, const int32_t *__restrict__ sizes,
                            int32_t *__restrict__ risk, int32_t *__restrict__ any,
                            unsigned long long *__restrict__ proposed) {
    unsigned long long local = 0;
    for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        int x = out_count[b] > 0 && out_count[b] >= sizes[b];
        risk[b] = x;
        local += (unsigned long long)out_count[b];
        if (x) atomicOr(any, 1);
    }
    local = wave_sum(local);
    if (lane_id() == 0 && local) atomicAdd(proposed, local);
}

// Events of moves touching at-risk blocks, keyed (block << 32 | vertex).
__global__ void k_make_events(int32_t N, const int32_t *__restrict__ move_s, const int32_t *__restrict__ assign,
                              const int32_t *__restrict__ risk, unsigned long long *__restrict__ events,
                              unsigned long long *__restrict__ n_events, int32_t *__restrict__ acc_a,
                              int32_t *__restrict__ acc_b) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        int s = move_s[v];
        acc_a[v] = 1;
        acc_b[v] = 1;
        if (s < 0) continue;
        int r = assign[v];
        if (risk[r]) events[atomicAdd(n_events, 1ull)] = ((unsigned long long)(unsigned)r << 32) | (unsigned)v;
        if (risk[s]) events[atomicAdd(n_events, 1ull)] = ((unsigned long long)(unsigned)s << 32) | (unsigned)v;
    }
}

// Event range of every block present in the sorted events, and the list of those blocks (n read on the device
// from dn when given).
__global__ void k_segments(int64_t n, const unsigned long long *__restrict__ dn, const unsigned long long *__restrict__ ev,
                           int32_t *__restrict__ seg_start, int32_t *__restrict__ seg_end,
                           int32_t *__restrict__ seg_list, int32_t *__restrict__ n_seg) {
    if (dn) n = (int64_t)*dn;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int b = (int)(ev[i] >> 32);
        if (i == 0 || (int)(ev[i - 1] >> 32) != b) {
            seg_start[b] = (int)i;
            seg_list[atomicAdd(n_seg, 1)] = b;
        }
        if (i == n - 1 || (int)(ev[i + 1] >> 32) != b) seg_end[b] = (int)i + 1;
    }
}

// The refusal pattern as the fixed point of per-block scans, in one workgroup: each round, every at-risk block with
// events walks them in vertex order with its running size, deciding its moves out (a move out is refused when the
// size has reached 1) from the previous round's decisions of the moves into it.  Moves out of blocks not at risk
// stay accepted (1 in both buffers, set by k_make_events), so a round writes exactly the decisions that can
// change; the rounds stop when one changes nothing, and both buffers then hold the fixed point.  The iteration is
// monotone and each round fixes at least the earliest undecided move, so it ends.
__global__ void __launch_bounds__(1024) k_resolve_all(const int32_t *__restrict__ seg_list,
                                                      const int32_t *__restrict__ n_seg,
                                                      const int32_t *__restrict__ seg_start,
                                                      const int32_t *__restrict__ seg_end,
                                                      const unsigned long long *__restrict__ ev,
                                                      const int32_t *__restrict__ assign,
                                                      const int32_t *__restrict__ sizes, int32_t *__restrict__ acc_a,
                                                      int32_t *__restrict__ acc_b, int32_t *__restrict__ rounds_out) {
    __shared__ int changed;
    const int m = *n_seg;
    int32_t *prev = acc_a, *next = acc_b;
    int rounds = 0;
    while (true) {
        if (threadIdx.x == 0) changed = 0;
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
            const int b = seg_list[k];
            int size = sizes[b];
            for (int i = seg_start[b]; i < seg_end[b]; ++i) {
                const int v = (int)(ev[i] & 0xffffffffull);
                if (assign[v] == b) {  // move out of b
                    const int ok = size > 1;
                    if (ok) size--;
                    next[v] = ok;
                    if (ok != prev[v]) changed = 1;
                } else if (prev[v]) {  // move into b
                    size++;
                }
            }
        }
        __threadfence();  // this round's decisions reach the other waves before they read them as prev
        __syncthreads();
        ++rounds;
        const int c = changed;
        __syncthreads();
        if (!c) break;
        int32_t *t = prev;
        prev = next;
        next = t;
    }
    if (threadIdx.x == 0) *rounds_out = rounds;
}

// Blockmodel::move_vertex(const VertexMove_v3&) (src/blockmodel/blockmodel.cpp:489-492): the committed moves'
// histogram entries, keyed by the vertex's out- and in-degree.  Runs before k_apply (reads the old blocks).
__global__ void k_hist_moves(int32_t N, const int32_t *__restrict__ move_s, const int32_t *__restrict__ acc,
                             const int32_t *__restrict__ assign, const int64_t *__restrict__ out_off,
                             const int64_t *__restrict__ in_off, HistSlot *__restrict__ t, int shift, uint32_t mask) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const int s = move_s[v];
        if (s < 0 || (acc != nullptr && !acc[v])) continue;
        const int r = assign[v];
        const long long ko = out_off[v + 1] - out_off[v], ki = in_off[v + 1] - in_off[v];
        hist_add(t, shift, mask, hist_key(r, 0, ko), -1);
        hist_add(t, shift, mask, hist_key(r, 1, ki), -1);
        hist_add(t, shift, mask, hist_key(s, 0, ko), 1);
        hist_add(t, shift, mask, hist_key(s, 1, ki), 1);
    }
}

// One count per workgroup: same-address atomics from every wavefront serialise at the L2.
__global__ void __launch_bounds__(256) k_apply(int32_t N, const int32_t *__restrict__ move_s,
                                               const int32_t *__restrict__ acc, int32_t *__restrict__ assign,
                                               unsigned long long *__restrict__ moved) {
    __shared__ unsigned long long block_total;
    if (threadIdx.x == 0) block_total = 0;
    __syncthreads();
    unsigned long long local = 0;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        int s = move_s[v];
        if (s >= 0 && (acc == nullptr || acc[v])) {
            assign[v] = s;
            local++;
        }
    }
    local = wave_sum(local);
    if (lane_id() == 0 && local) atomicAdd(&block_total, local);
    __syncthreads();
    if (threadIdx.x == 0 && block_total) atomicAdd(moved, block_total);
}

static int grid_of(int64_t n) {
    int64_t gsz = (n + 255) / 256;
    if (gsz < 1) gsz = 1;
    if (gsz > 16384) gsz = 16384;
    return (int)gsz;
}

// Commit of the batch's proposals (d_move_s), then the rebuild.  With capture, nothing is read on the host: the
// event count sizes nothing (the sort and the segment scan read it on the device, launched for the largest count),
// the moved and proposed counts stay on the device (ft_commit_counters), so the launches can be recorded into a
// graph; the moved count returned is then 0.
int64_t ft_commit(sbp_ctx ctx, sbp_bm bm, sbp_graph g, int32_t *d_move_s, int64_t *refused_out, bool capture) {
    hipStream_t st = ctx->stream;
    const int32_t N = bm->N, B = bm->B;
    int32_t *out_count = ctx->scratch[S_RISK].as<int32_t>(2 * (size_t)B + 8);
    int32_t *risk = out_count + B;
    int32_t *flags = ctx->scratch[S_FLAGS].as<int32_t>(16);
    // flags: [0] any block at risk, [1] blocks with events, [2] resolve rounds;
    // counters: [0] moved [1] events [2] proposed
    unsigned long long *counters = reinterpret_cast<unsigned long long *>(flags + 4);
    SBP_HIP(sbpprim::fill<int32_t>(out_count, B, 0, st));
    SBP_HIP(sbpprim::fill<int32_t>(flags, 16, 0, st));
    k_count_out<<<grid_of(N), 256, 0, st>>>(N, d_move_s, bm->assign, out_count);
    k_mark_risk<<<grid_of(B), 256, 0, st>>>(B, out_count, bm->sizes, risk, flags, counters + 2);
    int32_t any = 1;
    if (!capture) read_small(ctx, {{&any, flags, sizeof(int32_t)}});
    const int32_t *acc_final = nullptr;
    if (any) {
        const int64_t max_ev = 2 * (int64_t)N;
        auto *events = ctx->scratch[S_EVENTS_A].as<unsigned long long>((size_t)max_ev + 2);
        auto *events_sorted = ctx->scratch[S_EVENTS_B].as<unsigned long long>((size_t)max_ev + 2);
        int32_t *acc = ctx->scratch[S_ACCEPT].as<int32_t>(2 * (size_t)N + 2);
        int32_t *seg = ctx->scratch[S_MISC].as<int32_t>(3 * (size_t)B + 2);
        int32_t *seg_start = seg, *seg_end = seg + B, *seg_list = seg + 2 * (size_t)B;
        k_make_events<<<grid_of(N), 256, 0, st>>>(N, d_move_s, bm->assign, risk, events, counters + 1, acc, acc + N);
        // keys (block << 32) | vertex: the bits above 32 + log2(B) are zero
        int end_bit = 33;
        while (end_bit < 64 && (int64_t(1) << (end_bit - 32)) < B) ++end_bit;
        const int32_t *no_vals = nullptr;
        int64_t n_sort = max_ev;
        const int64_t *dn = reinterpret_cast<const int64_t *>(counters + 1);
        if (!capture) {  // the exact count sizes the launches
            unsigned long long n_ev = 0;
            read_small(ctx, {{&n_ev, counters + 1, sizeof(n_ev)}});
            n_sort = (int64_t)n_ev;
            dn = nullptr;
        }
        size_t tmp = 0;
        SBP_HIP(sbpprim::radix_sort(nullptr, tmp, events, events_sorted, no_vals, (int32_t *)nullptr, n_sort, 0, end_bit,
                                    st, dn));
        void *tp = ctx->scratch[S_SORT_TMP].get(tmp);
        SBP_HIP(sbpprim::radix_sort(tp, tmp, events, events_sorted, no_vals, (int32_t *)nullptr, n_sort, 0, end_bit, st,
                                    dn));
        if (n_sort > 0)
            k_segments<<<grid_of(n_sort), 256, 0, st>>>(n_sort, dn ? counters + 1 : nullptr, events_sorted, seg_start,
                                                        seg_end, seg_list, flags + 1);
        k_resolve_all<<<1, 1024, 0, st>>>(seg_list, flags + 1, seg_start, seg_end, events_sorted, bm->assign, bm->sizes,
                                          acc, acc + N, flags + 2);
        acc_final = acc;
    }
    if (bm->dc)
        k_hist_moves<<<grid_of(N), 256, 0, st>>>(N, d_move_s, acc_final, bm->assign, g->out_off, g->in_off, bm->hist,
                                                 64 - bm->hist_log2, (uint32_t)((1ull << bm->hist_log2) - 1));
    // eager: the assignment before the moves, for the update of the cells in place of a rebuild
    int32_t *prev = nullptr;
    if (!capture && bm->keys_valid && !bm->dc && N > 0 && bm_update_wanted(g)) {
        prev = ctx->scratch[S_PREV].as<int32_t>((size_t)N);
        SBP_HIP(hipMemcpyAsync(prev, bm->assign, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToDevice, st));
    }
    k_apply<<<std::min(grid_of(N), 1024), 256, 0, st>>>(N, d_move_s, acc_final, bm->assign, counters);
    SBP_HIP(hipGetLastError());
    if (capture) {
        SBP_CHECK(!bm->dc, SBP_E_STATE, "captured commit: degree-corrected histograms need host bookkeeping");
        bm_rebuild(ctx, g, bm, true, true);
        return 0;
    }
    unsigned long long c[2] = {0, 0};
    HistSlot book{0, 0, 0};
    if (bm->dc)
        read_small(ctx, {{&c[0], counters, sizeof(unsigned long long)}, {&c[1], counters + 2, sizeof(unsigned long long)},
                         {&book, bm->hist + (size_t(1) << bm->hist_log2), sizeof(HistSlot)}});
    else
        read_small(ctx, {{&c[0], counters, sizeof(unsigned long long)}, {&c[1], counters + 2, sizeof(unsigned long long)}});
    if (bm->dc) bm_hist_after_moves(ctx, bm, book.val, book.pad);
    if (refused_out) *refused_out = (int64_t)c[1] - (int64_t)c[0];
    if (!(prev && bm_update_moves(ctx, g, bm, prev, (int64_t)c[0]))) bm_rebuild(ctx, g, bm, true);
    return (int64_t)c[0];
}

// The counters a captured commit leaves on the device: moved, events, proposed (unsigned 64-bit each).
const unsigned long long *ft_commit_counters(sbp_ctx ctx) {
    return reinterpret_cast<const unsigned long long *>(ctx->scratch[S_FLAGS].as<int32_t>(16) + 4);
}

// ------------------------------------------------------------------ shuffle

__global__ void k_shuffle_keys(int64_t n, const int32_t *__restrict__ items, uint64_t seed, uint32_t epoch,
                               const uint32_t *__restrict__ epoch_ptr,
                               unsigned long long *__restrict__ keys) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int v = items[i];
        U4 w = philox(seed, (uint32_t)v, epoch_ptr ? *epoch_ptr : epoch, TAG_SHUFFLE, 0);
        keys[i] = ((unsigned long long)w.x << 32) | (unsigned)v;
    }
}

__global__ void k_low_word(int64_t n, const unsigned long long *__restrict__ keys, int32_t *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (int32_t)(keys[i] & 0xffffffffull);
}

// Ascending Philox sort key, ties by vertex id.
void shuffle_keys(sbp_ctx ctx, const int32_t *d_items, int64_t n, uint32_t epoch, int32_t *d_out,
                  const uint32_t *epoch_ptr) {
    if (n == 0) return;
    hipStream_t st = ctx->stream;
    auto *ka = ctx->scratch[S_KEYS_A].as<unsigned long long>(n);
    auto *kb = ctx->scratch[S_KEYS_B].as<unsigned long long>(n);
    k_shuffle_keys<<<grid_of(n), 256, 0, st>>>(n, d_items, ctx->seed, epoch, epoch_ptr, ka);
    size_t tmp = 0;
    const int32_t *no_vals = nullptr;
    SBP_HIP(sbpprim::radix_sort(nullptr, tmp, ka, kb, no_vals, (int32_t *)nullptr, n, 0, 64, st));
    void *tp = ctx->scratch[S_SORT_TMP].get(tmp);
    SBP_HIP(sbpprim::radix_sort(tp, tmp, ka, kb, no_vals, (int32_t *)nullptr, n, 0, 64, st));
    k_low_word<<<grid_of(n), 256, 0, st>>>(n, kb, d_out);
    SBP_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ Metropolis-Hastings

// State of the walk between launches.  The host queues several rounds of k_mh -> k_mh_scan -> k_mh_advance
// without reading anything back; each kernel acts only in the status it expects and is a no-op otherwise.
enum MhStatus : int32_t {
    MH_RUN = 0,     // k_mh walks on from stop_at
    MH_SLACK = 1,   // a row or column lacks room at stop_at: the host re-lays the matrix out
    MH_OCCUPY = 2,  // a move occupied an empty block (nonparametric): the host refreshes the constants
    MH_SCAN = 3,    // k_mh handed over after windows without an accepted move: the wide scan evaluates ahead
    MH_DONE = 4,    // the walk reached the end of the order
};

struct MhState {
    int32_t *rstart, *rlen, *rcap, *cstart, *clen, *ccap;  // [B]
    int32_t *rboost, *cboost;  // [B] multiplier of a row's / column's extra capacity at the next layout
    int32_t *stop_at;   // first order index not processed (== count when done)
    int32_t *status;    // MhStatus
    unsigned long long *moves;
    double *sum;        // running sum of accepted delta-MDL in visit order, carried across resumed launches
};

// The wide scan's work area.
struct MhScan {
    unsigned long long *first;       // lowest order index with an accepted proposal in the window (~0 when none)
    int32_t *width;                  // window length of the next scan (adapted by k_mh_advance)
    int32_t *n_queued;               // high-degree items of the window queued by the first scan launch
    int32_t *queue;                  // [MH_SCAN_MAX] their offsets in the window
    unsigned long long *item_bytes;  // [MH_SCAN_MAX] algorithmic bytes of each evaluated item of the window
};
// Windows without an accepted move after which k_mh hands the walk to the wide scan, and the scan's window
// lengths (first, least, most).
constexpr int MH_HANDOFF_WINDOWS = 2;
constexpr int32_t MH_SCAN_FIRST = 4096, MH_SCAN_MIN = 256, MH_SCAN_MAX = 1 << 16;

// Inserts/updates/erases one cell in a sorted slack row: cells[beg, beg+len), capacity cap.
__device__ void slack_add(int2 *cells, int beg, int32_t *len_p, int key, int delta) {
    const int lane = lane_id();
    const int len = *len_p;
    int lo = beg, hi = beg + len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cells[mid].x < key) lo = mid + 1;
        else hi = mid;
    }
    const int pos = lo;
    const bool found = pos < beg + len && cells[pos].x == key;
    if (found) {
        const int nv = cells[pos].y + delta;
        wave_sync();
        if (nv != 0) {
            if (lane == 0) cells[pos].y = nv;
        } else {
            // shift left by one: ascending chunks read before they write
            for (int i0 = pos; i0 < beg + len - 1; i0 += WAVE) {
                int i = i0 + lane;
                int2 x = (i < beg + len - 1) ? cells[i + 1] : make_int2(0, 0);
                wave_sync();
                if (i < beg + len - 1) cells[i] = x;
                wave_sync();
            }
            if (lane == 0) *len_p = len - 1;
        }
    } else if (delta != 0) {
        // shift right by one: descending chunks
        const int end = beg + len;
        for (int hi_i = end; hi_i > pos; hi_i -= WAVE) {
            int i = hi_i - 1 - lane;  // destination i+1 <- source i
            int2 x = (i >= pos) ? cells[i] : make_int2(0, 0);
            wave_sync();
            if (i >= pos) cells[i + 1] = x;
            wave_sync();
        }
        if (lane == 0) {
            cells[pos] = make_int2(key, delta);
            *len_p = len + 1;
        }
    }
    wave_sync();
}

// slack_add on row x, and the same change in the dense copy when the walk keeps one.  Each cell lives in one row
// list and each list is written by one wavefront, so the dense entry has one writer.
__device__ __forceinline__ void row_add(const FtArgs &a, const MhState &h, int x, int key, int d) {
    slack_add(a.m.row_cells, h.rstart[x], &h.rlen[x], key, d);
    if (a.m.dense && lane_id() == 0) a.m.dense_w[x * a.m.dense_stride + key] += d;
}

__device__ __forceinline__ void apply_cell(const FtArgs &a, const MhState &h, int row, int col, int d) {
    row_add(a, h, row, col, d);
    slack_add(a.m.col_cells, h.cstart[col], &h.clen[col], row, d);
}

// Asks for room for `more` new cells in row / column b at the next layout: its extra capacity (max(8, len / 2)
// times the boost) at least doubles and covers the request.  Lanes that ask for the same block store values
// that each double it, so any of them makes progress.
__device__ __forceinline__ void want_room(int32_t *boost, const int32_t *len, int b, int more) {
    const int l = len[b];
    const int extra = l / 2 > 8 ? l / 2 : 8;
    long long nb = 2ll * boost[b];
    const long long need = ((long long)more + extra - 1) / extra;
    if (need > nb) nb = need;
    boost[b] = (int32_t)(nb < (1 << 24) ? nb : (1 << 24));
}

// Would adding cell (row, col) need a new slot that row or column lacks?  A short row / column asks for room.
__device__ __forceinline__ bool lacks_room(const FtArgs &a, const MhState &h, int row, int col, long long d) {
    if (d <= 0) return false;
    if (get_cell(a, row, col) != 0) return false;
    const bool short_r = h.rlen[row] >= h.rcap[row], short_c = h.clen[col] >= h.ccap[col];
    if (short_r) want_room(h.rboost, h.rlen, row, 1);
    if (short_c) want_room(h.cboost, h.clen, col, 1);
    return short_r || short_c;
}

// One wavefront's share of applying an accepted move's Delta (finetune::blockmodel_delta,
// src/finetune.cpp:201-235) to the slack rows and columns: wavefront 0 writes row r, 1 row s, 2 column r, 3
// column s, and the others the rows and columns of the remaining neighbour blocks c (entries i = w - 4 mod
// MH_WAVES - 4).  A list is written by one wavefront only, with its updates in the order of a one-wavefront
// commit (table entries in used-slot order, then the cells (r, r), (r, s), (s, s), (s, r)), so each list
// passes through the same states and the room checked for that order suffices.
__device__ void mh_apply_share(const FtArgs &a, const MhState &h, const MoveCtx &m, const Table &t, int nb, int wid) {
    const int r = m.r, s = m.s;
    int2 *cols = a.m.col_cells;
    if (wid == 0 || wid == 1) {
        const int x = wid == 0 ? r : s;
        for (int i = 0; i < nb; ++i) {
            const int slot = t.used[i];
            const int c = t.keys[slot], oc = t.outc[slot];
            if (oc > 0) {
                const long long d = x == r ? delta_row_r(m, c, oc) : delta_row_s(m, c, oc);
                if (d) row_add(a, h, x, c, (int)d);
            }
        }
        if (x == r) {
            if (m.out_r == 0 && (m.selfc > 0 || m.in_r > 0)) row_add(a, h, r, r, -m.selfc - m.in_r);
            if (m.out_s == 0 && m.in_r > 0) row_add(a, h, r, s, m.in_r);
        } else {
            if (m.out_s == 0 && (m.selfc > 0 || m.in_s > 0)) row_add(a, h, s, s, m.selfc + m.in_s);
            if (m.out_r == 0 && m.in_s > 0) row_add(a, h, s, r, -m.in_s);
        }
    } else if (wid == 2 || wid == 3) {
        const int x = wid == 2 ? r : s;
        for (int i = 0; i < nb; ++i) {
            const int slot = t.used[i];
            const int c = t.keys[slot], oc = t.outc[slot], ic = t.inc[slot];
            if (oc > 0 && c == x) {
                const long long dr = delta_row_r(m, c, oc), ds = delta_row_s(m, c, oc);
                if (dr) slack_add(cols, h.cstart[x], &h.clen[x], r, (int)dr);
                if (ds) slack_add(cols, h.cstart[x], &h.clen[x], s, (int)ds);
            }
            if (ic > 0 && c != r && c != s) slack_add(cols, h.cstart[x], &h.clen[x], c, x == r ? -ic : ic);
        }
        if (x == r) {
            if (m.out_r == 0 && (m.selfc > 0 || m.in_r > 0)) slack_add(cols, h.cstart[r], &h.clen[r], r, -m.selfc - m.in_r);
            if (m.out_r == 0 && m.in_s > 0) slack_add(cols, h.cstart[r], &h.clen[r], s, -m.in_s);
        } else {
            if (m.out_s == 0 && m.in_r > 0) slack_add(cols, h.cstart[s], &h.clen[s], r, m.in_r);
            if (m.out_s == 0 && (m.selfc > 0 || m.in_s > 0)) slack_add(cols, h.cstart[s], &h.clen[s], s, m.selfc + m.in_s);
        }
    } else {
        for (int i = wid - 4; i < nb; i += MH_WAVES - 4) {
            const int slot = t.used[i];
            const int c = t.keys[slot], oc = t.outc[slot], ic = t.inc[slot];
            if (c == r || c == s) continue;
            if (oc > 0) {
                const long long dr = delta_row_r(m, c, oc), ds = delta_row_s(m, c, oc);
                if (dr) slack_add(cols, h.cstart[c], &h.clen[c], r, (int)dr);
                if (ds) slack_add(cols, h.cstart[c], &h.clen[c], s, (int)ds);
            }
            if (ic > 0) {
                row_add(a, h, c, r, -ic);
                row_add(a, h, c, s, ic);
            }
        }
    }
}

// Sequential Metropolis-Hastings over a.items[start, count) (src/finetune.cpp:773-869), evaluated a window of
// MH_WAVES vertices at a time: each wavefront of the one workgroup proposes and evaluates the next vertex of
// the window against the current blockmodel (propose_block and evaluate only read it).  The first accepted
// move of the window is committed by its wavefront, and the next window starts at the vertex after it.  Every
// evaluation up to and including that move read exactly the state a one-vertex-at-a-time walk reads (no move
// was applied between them), and the draws are keyed by vertex, so decisions, commits and the visit-order sum
// of accepted delta-MDL are those of the walk.  Evaluations after the committed move are discarded.
//
// The walk starts at *h.stop_at when *h.status is MH_RUN (else the launch does nothing).  With `handoff`, after
// MH_HANDOFF_WINDOWS windows in a row without an accepted move it stops with MH_SCAN at the next window, for the
// wide scan to evaluate far ahead against the same state.
template <bool GREEDY, bool NP>
__global__ void __launch_bounds__(MH_WAVES * WAVE) k_mh(FtArgs a, MhState h, int handoff) {
    if (*h.status != MH_RUN) return;
    __shared__ int lds[MH_WAVES * LDS_TABLE_INTS];
    __shared__ int s_accept[MH_WAVES];
    __shared__ long long s_next;      // first item of the next window
    __shared__ int s_status, s_stop;  // as *h.status / *h.stop_at
    __shared__ double s_sum;
    __shared__ unsigned long long s_moves;
    __shared__ int s_mv[10];  // the committed move: r, s, entries, selfc, in_r, in_s, out_r, out_s, big table, room
    const int lane = lane_id();
    const int wid = (int)(threadIdx.x / WAVE);
    unsigned long long bytes = 0;
    const Table small_t = make_table(a, false, lds + wid * LDS_TABLE_INTS, 0);
    const Table big_t = make_table(a, true, lds, wid);
    table_clear_all(small_t);  // the global tables arrive empty (prepare_tables) and are left empty
    if (threadIdx.x == 0) {
        s_next = *h.stop_at;
        s_status = MH_RUN;
        s_stop = (int32_t)a.count;
        s_sum = *h.sum;
        s_moves = 0;
    }
    __syncthreads();
    int empty_windows = 0;  // the same in every thread: `first` below is
    while (true) {
        const long long base = s_next;
        if (base >= a.count) break;
        const long long item = base + wid;
        bool accept = false;
        double dS = 0.0;
        unsigned long long pb = 0;
        long long kall = 0;
        MoveCtx m;
        U4 w;
        Table tsel = small_t;
        if (item < a.count) {
            const int v = a.items[item];
            if (a.sizes[a.assign[v]] != 1) {  // propose_move skips singleton blocks (src/finetune.cpp:845-847)
                kall = (a.out_off[v + 1] - a.out_off[v]) + (a.in_off[v + 1] - a.in_off[v]);
                tsel = kall > SMALL_K ? big_t : small_t;
                const int s = propose_block(a, v, m, w);
                if (s == m.r) {
                    pb = proposal_bytes(a, m, 0, GREEDY);
                } else {
                    dS = evaluate<GREEDY, NP>(a, m, tsel, w, accept);
                    pb = proposal_bytes(a, m, *tsel.n_used, GREEDY);
                    if (!accept) table_clear_used(tsel);
                }
            }
        }
        if (lane == 0) s_accept[wid] = accept ? 1 : 0;
        __syncthreads();
        int first = MH_WAVES;
        for (int i = 0; i < MH_WAVES; ++i)
            if (s_accept[i]) {
                first = i;
                break;
            }
        if (wid <= first) bytes += pb;  // later evaluations are repeated in the next window
        if (wid > first && accept) table_clear_used(tsel);
        if (first < MH_WAVES) {
            bool occupies = false;
            if (wid == first) {
                const Table &t = tsel;
                const int s = m.s;
                const int r = m.r;
                const int nb = *t.n_used;
                // occupying an empty block changes the non-empty block count the nonparametric constants are for
                occupies = NP && a.sizes[s] == 0;
                // ---- room check for every Delta cell that would create a new entry
                bool short_room = false;
                int new_row_s = 0, new_col_s = 0;
                for (int i = lane; i < nb; i += WAVE) {
                    const int slot = t.used[i];
                    const int c = t.keys[slot], oc = t.outc[slot], ic = t.inc[slot];
                    if (oc > 0) {
                        short_room |= lacks_room(a, h, s, c, delta_row_s(m, c, oc));
                        if (delta_row_s(m, c, oc) > 0 && get_cell(a, s, c) == 0) new_row_s++;
                    }
                    if (ic > 0 && c != r && c != s) {
                        short_room |= lacks_room(a, h, c, s, ic);
                        if (get_cell(a, c, s) == 0) new_col_s++;
                    }
                }
                if (lane == 0) {
                    if (m.out_s == 0 && m.in_r > 0) short_room |= lacks_room(a, h, r, s, m.in_r);
                    if (m.out_s == 0 && (m.selfc > 0 || m.in_s > 0))
                        short_room |= lacks_room(a, h, s, s, m.selfc + m.in_s);
                }
                // row s and column s may receive several new cells: check their totals
                new_row_s = wave_sum(new_row_s) + 2;
                new_col_s = wave_sum(new_col_s) + 2;
                const bool short_rs = h.rlen[s] + new_row_s > h.rcap[s], short_cs = h.clen[s] + new_col_s > h.ccap[s];
                if (lane == 0 && short_rs) want_room(h.rboost, h.rlen, s, new_row_s);
                if (lane == 0 && short_cs) want_room(h.cboost, h.clen, s, new_col_s);
                short_room |= short_rs || short_cs;
                const bool lacking = ballot(short_room) != 0;
                if (lacking) table_clear_used(t);
                if (lane == 0) {
                    s_mv[0] = r;
                    s_mv[1] = s;
                    s_mv[2] = nb;
                    s_mv[3] = m.selfc;
                    s_mv[4] = m.in_r;
                    s_mv[5] = m.in_s;
                    s_mv[6] = m.out_r;
                    s_mv[7] = m.out_s;
                    s_mv[8] = kall > SMALL_K ? 1 : 0;
                    s_mv[9] = lacking ? 0 : 1;
                    if (lacking) {
                        s_status = MH_SLACK;
                        s_stop = (int32_t)item;
                    }
                }
            }
            __syncthreads();
            if (s_mv[9]) {
                // ---- apply the Delta in place (update_edge_counts(delta), dict_transpose_matrix.cpp:430-446), each
                // wavefront its share of the lists
                MoveCtx mm;
                mm.r = s_mv[0];
                mm.s = s_mv[1];
                mm.selfc = s_mv[3];
                mm.in_r = s_mv[4];
                mm.in_s = s_mv[5];
                mm.out_r = s_mv[6];
                mm.out_s = s_mv[7];
                const Table t = make_table(a, s_mv[8] != 0, lds + first * LDS_TABLE_INTS, first);
                mh_apply_share(a, h, mm, t, s_mv[2], wid);
                __syncthreads();
                if (wid == first) {
                    const int r = m.r, s = m.s;
                    table_clear_used(tsel);
                    if (lane == 0) {
                        a.d_out[r] -= (int)m.k_out;
                        a.d_out[s] += (int)m.k_out;
                        a.d_in[r] -= (int)m.shift_in;
                        a.d_in[s] += (int)m.shift_in;
                        a.sizes[r] -= 1;
                        a.sizes[s] += 1;
                        a.assign[m.v] = s;
                        if (NP && a.hist) {  // keyed by the proposal's edge counts (src/blockmodel/blockmodel.cpp:431-434)
                            hist_add(a.hist, a.np.hist_shift, a.np.hist_mask, hist_key(r, 0, m.k_out), -1);
                            hist_add(a.hist, a.np.hist_shift, a.np.hist_mask, hist_key(r, 1, m.k_in), -1);
                            hist_add(a.hist, a.np.hist_shift, a.np.hist_mask, hist_key(s, 0, m.k_out), 1);
                            hist_add(a.hist, a.np.hist_shift, a.np.hist_mask, hist_key(s, 1, m.k_in), 1);
                        }
                    }
                    wave_sync();
                    if (lane == 0) {
                        a.d_tot[r] = a.d_out[r] + a.d_in[r] - (int)get_cell(a, r, r);
                        a.d_tot[s] = a.d_out[s] + a.d_in[s] - (int)get_cell(a, s, s);
                        s_sum += dS;
                        s_moves += 1;
                        if (occupies) {  // the host refreshes the constants and resumes with the next vertex
                            s_status = MH_OCCUPY;
                            s_stop = (int32_t)(item + 1);
                        }
                    }
                    bytes += 12ull * (unsigned long long)kall;
                }
            }
        }
        __syncthreads();
        empty_windows = first < MH_WAVES ? 0 : empty_windows + 1;
        if (threadIdx.x == 0) {
            long long nxt = base + (first < MH_WAVES ? first + 1 : MH_WAVES);
            if (s_status != MH_RUN) {
                nxt = a.count;
            } else if (handoff && empty_windows >= MH_HANDOFF_WINDOWS && nxt < a.count) {
                s_status = MH_SCAN;
                s_stop = (int32_t)nxt;
                nxt = a.count;
            }
            s_next = nxt;
        }
        __syncthreads();
    }
    if (lane == 0 && bytes) atomicAdd(a.bytes, bytes);
    if (threadIdx.x == 0) {
        *h.status = s_status == MH_RUN ? MH_DONE : s_status;
        *h.stop_at = s_stop;
        atomicAdd(h.moves, s_moves);
        *h.sum = s_sum;
    }
}

// Wide scan of the walk (MH_SCAN): every wavefront of the grid proposes and evaluates the items of
// [stop_at, stop_at + width) in turn against the current blockmodel, which nothing changes meanwhile, with the
// device functions k_mh uses, so its accept decisions are the walk's.  The lowest accepted item goes to
// *first; a wavefront stops once its items pass it (they would be evaluated again after that move).  Every
// item below the final *first is evaluated, so all of them are rejected (or skipped) proposals of the walk.
//
// Like the finetune proposal launch, the scan is two launches: the first (BIG false, neighbour-block tables in
// LDS) evaluates the items of total degree <= SMALL_K and queues the others; the second (BIG true) gives each
// queued item a wavefront with a global-memory table.
__device__ __forceinline__ unsigned long long uniform_first(const unsigned long long *first) {
    // the same value in every lane; a stale one only delays an exit, since *first only decreases
    const unsigned long long seen = __atomic_load_n(first, __ATOMIC_RELAXED);
    return ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(seen >> 32)) << 32) |
           (unsigned)__builtin_amdgcn_readfirstlane((unsigned)seen);
}

template <bool GREEDY, bool NP, bool BIG>
__global__ void __launch_bounds__(FT_BLOCK) k_mh_scan(FtArgs a, MhState h, MhScan sc) {
    if (*h.status != MH_SCAN) return;
    __shared__ int lds[BIG ? 1 : FT_WAVES * LDS_TABLE_INTS];
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    const int64_t gwave = (int64_t)blockIdx.x * FT_WAVES + wid;
    const int64_t nwaves = (int64_t)gridDim.x * FT_WAVES;
    const Table t = make_table(a, BIG, lds + (BIG ? 0 : wid * LDS_TABLE_INTS), gwave);
    if (!BIG) table_clear_all(t);  // the global tables arrive empty (prepare_tables) and are left empty
    const int64_t start = *h.stop_at;
    const int64_t end = start + *sc.width < a.count ? start + *sc.width : a.count;
    const int64_t n = BIG ? (int64_t)*sc.n_queued : end - start;
    for (int64_t j = gwave; j < n; j += nwaves) {
        const int64_t item = start + (BIG ? (int64_t)sc.queue[j] : j);
        if ((unsigned long long)item > uniform_first(sc.first)) {
            if (BIG) continue;  // the queue is in no particular order
            break;              // this wavefront's later items lie further still
        }
        const int v = a.items[item];
        bool accept = false;
        unsigned long long pb = 0;
        if (a.sizes[a.assign[v]] != 1) {  // propose_move skips singleton blocks (src/finetune.cpp:845-847)
            if (!BIG) {
                const long long kall = (a.out_off[v + 1] - a.out_off[v]) + (a.in_off[v + 1] - a.in_off[v]);
                if (kall > SMALL_K) {  // deferred to the global-memory-table launch
                    if (lane == 0) sc.queue[atomicAdd(sc.n_queued, 1)] = (int32_t)j;
                    continue;
                }
            }
            MoveCtx m;
            U4 w;
            const int s = propose_block(a, v, m, w);
            if (s == m.r) {
                pb = proposal_bytes(a, m, 0, GREEDY);
            } else {
                evaluate<GREEDY, NP>(a, m, t, w, accept);
                pb = proposal_bytes(a, m, *t.n_used, GREEDY);
                table_clear_used(t);
            }
        }
        if (lane == 0) {
            sc.item_bytes[item - start] = pb;
            if (accept) atomicMin(sc.first, (unsigned long long)item);
        }
    }
}

// After a scan: the walk skips the rejected items before the first accepted one (adding their bytes) and resumes
// there, or past the window when nothing was accepted.  The window doubles after a scan without an accepted
// move and halves when the first one came in its first quarter.  Resets *first for the next scan.
__global__ void __launch_bounds__(256) k_mh_advance(MhState h, MhScan sc, int64_t count,
                                                    unsigned long long *bytes_total) {
    if (*h.status != MH_SCAN) return;
    __shared__ unsigned long long part[256 / WAVE];
    const int64_t start = *h.stop_at;
    const int32_t width = *sc.width;
    const unsigned long long f = *sc.first;
    int64_t stop = start + width < count ? start + width : count;
    const bool found = f < (unsigned long long)stop;
    if (found) stop = (int64_t)f;
    unsigned long long b = 0;
    for (int64_t i = threadIdx.x; i < stop - start; i += blockDim.x) b += sc.item_bytes[i];
    b = wave_sum(b);
    if (lane_id() == 0) part[threadIdx.x / WAVE] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < 256 / WAVE; ++w) t += part[w];
        if (t) atomicAdd(bytes_total, t);
        int32_t w2 = width;
        if (!found) w2 = width < MH_SCAN_MAX / 2 ? 2 * width : MH_SCAN_MAX;
        else if ((int64_t)f - start < width / 4) w2 = width / 2 > MH_SCAN_MIN ? width / 2 : MH_SCAN_MIN;
        *sc.width = w2;
        *sc.first = ~0ull;
        *sc.n_queued = 0;
        *h.stop_at = (int32_t)stop;
        *h.status = stop >= count ? MH_DONE : MH_RUN;
    }
}

// Slack layout: capacity = len + max(8, len / 2) times the row's (column's) boost, at most B + 2: a row never
// holds more than B cells, and the room check for a destination row asks for its new cells plus 2.
__global__ void k_slack_caps(int32_t B, const int32_t *__restrict__ ptr, int32_t *__restrict__ len,
                             int32_t *__restrict__ cap, const int32_t *__restrict__ boost) {
    for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        const int l = ptr[b + 1] - ptr[b];
        len[b] = l;
        const long long extra = l / 2 > 8 ? l / 2 : 8;
        const long long c = (long long)l + extra * boost[b];
        cap[b] = (int32_t)(c < (long long)B + 2 ? c : (long long)B + 2);
    }
}

__global__ void k_fill_i32(int64_t n, int32_t *__restrict__ p, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_slack_copy(int32_t B, const int32_t *__restrict__ ptr, const int2 *__restrict__ src,
                             const int32_t *__restrict__ start, int2 *__restrict__ dst) {
    for (int32_t b = blockIdx.x; b < B; b += gridDim.x)
        for (int i = ptr[b] + threadIdx.x; i < ptr[b + 1]; i += blockDim.x) dst[start[b] + (i - ptr[b])] = src[i];
}

// Re-lays slack rows back out (used when growing): copies [start, start+len) to a compact CSR.
__global__ void k_compact_copy(int32_t B, const int32_t *__restrict__ start, const int32_t *__restrict__ len,
                               const int2 *__restrict__ src, const int32_t *__restrict__ ptr, int2 *__restrict__ dst) {
    for (int32_t b = blockIdx.x; b < B; b += gridDim.x)
        for (int i = threadIdx.x; i < len[b]; i += blockDim.x) dst[ptr[b] + i] = src[start[b] + i];
}

constexpr int MH_SCAN_BLOCKS = 1024;  // workgroups of the first scan launch: 4,096 wavefronts over the 256 CUs

template <bool GREEDY, bool NP>
static void launch_mh_t(const FtArgs &a, const MhState &h, const MhScan &sc, bool scan, int big_blocks,
                        hipStream_t st) {
    k_mh<GREEDY, NP><<<1, MH_WAVES * WAVE, 0, st>>>(a, h, scan ? 1 : 0);
    if (scan) {
        k_mh_scan<GREEDY, NP, false><<<MH_SCAN_BLOCKS, FT_BLOCK, 0, st>>>(a, h, sc);
        if (big_blocks > 0) k_mh_scan<GREEDY, NP, true><<<big_blocks, FT_BLOCK, 0, st>>>(a, h, sc);
        k_mh_advance<<<1, 256, 0, st>>>(h, sc, a.count, a.bytes);
    }
    SBP_HIP(hipGetLastError());
}

// One round of the walk: k_mh, then (with `scan`) the wide scan and the advance.  big_blocks: workgroups of the
// scan's high-degree launch, one global neighbour-block table per wavefront (0: the graph has no such vertex).
static void launch_mh(const FtArgs &a, const MhState &h, const MhScan &sc, bool np, bool greedy, bool scan,
                      int big_blocks, hipStream_t st) {
    if (np) {
        if (greedy) launch_mh_t<true, true>(a, h, sc, scan, big_blocks, st);
        else launch_mh_t<false, true>(a, h, sc, scan, big_blocks, st);
    } else {
        if (greedy) launch_mh_t<true, false>(a, h, sc, scan, big_blocks, st);
        else launch_mh_t<false, false>(a, h, sc, scan, big_blocks, st);
    }
}

void mh_run(sbp_ctx ctx, sbp_bm bm, sbp_graph g, const sbp_params &p, uint32_t epoch, const int32_t *d_order,
            int64_t count, double *sum_dS, sbp_stats *stats) {
    hipStream_t st = ctx->stream;
    if (count == 0) {
        if (sum_dS) *sum_dS = 0.0;
        return;
    }
    const int32_t B = bm->B;
    // per-block arrays: rstart rlen rcap cstart clen ccap + bookkeeping + row and column boosts
    int32_t *arr = ctx->scratch[S_EVENTS_A].as<int32_t>(8 * ((size_t)B + 1) + 64);
    MhState h;
    h.rstart = arr;
    h.rlen = arr + (B + 1);
    h.rcap = arr + 2 * (B + 1);
    h.cstart = arr + 3 * (B + 1);
    h.clen = arr + 4 * (B + 1);
    h.ccap = arr + 5 * (B + 1);
    int32_t *book = arr + 6 * (B + 1);
    h.rboost = book + 64;
    h.cboost = h.rboost + (B + 1);
    h.stop_at = book;
    h.status = book + 1;
    h.moves = reinterpret_cast<unsigned long long *>(book + 4);
    h.sum = reinterpret_cast<double *>(book + 8);
    auto *bytes = ctx->scratch[S_COUNTER].as<unsigned long long>(4);
    MhScan sc;
    sc.first = reinterpret_cast<unsigned long long *>(book + 12);
    sc.width = book + 14;
    sc.item_bytes = ctx->scratch[S_DS].as<unsigned long long>(MH_SCAN_MAX + MH_SCAN_MAX / 2);
    sc.queue = reinterpret_cast<int32_t *>(sc.item_bytes + MH_SCAN_MAX);
    sc.n_queued = book + 15;
    // book: stop_at 0, status MH_RUN, no accepted scan item, the first scan's width
    SBP_HIP(hipMemsetAsync(book, 0, 64 * sizeof(int32_t), st));
    SBP_HIP(hipMemsetAsync(bytes, 0, 4 * sizeof(unsigned long long), st));
    k_fill_i32<<<1, 64, 0, st>>>(2, book + 12, -1);
    k_fill_i32<<<1, 64, 0, st>>>(1, book + 14, MH_SCAN_FIRST);
    k_fill_i32<<<grid_of(2 * ((int64_t)B + 1)), 256, 0, st>>>(2 * ((int64_t)B + 1), h.rboost, 1);
    // The wide scan pays off once several windows pass without an accepted move; a walk shorter than a few
    // windows stays in k_mh.
    const bool scan = count > 4 * MH_WAVES;
    int32_t start = 0, retries = 0;
    float kernel_ms = 0.f;
    while (true) {
        // lay out rows and columns with slack
        k_slack_caps<<<grid_of(B), 256, 0, st>>>(B, bm->row_ptr, h.rlen, h.rcap, h.rboost);
        k_slack_caps<<<grid_of(B), 256, 0, st>>>(B, bm->col_ptr, h.clen, h.ccap, h.cboost);
        size_t tmp = 0;
        SBP_HIP(sbpprim::scan(nullptr, tmp, (const int32_t *)h.rcap, h.rstart, (int64_t)B + 1, false, st));
        void *tp = ctx->scratch[S_SORT_TMP].get(tmp);
        // capacities have B entries; the (B+1)-th scan input is the zeroed slot past the end
        SBP_HIP(hipMemsetAsync(h.rcap + B, 0, sizeof(int32_t), st));
        SBP_HIP(hipMemsetAsync(h.ccap + B, 0, sizeof(int32_t), st));
        SBP_HIP(sbpprim::scan(tp, tmp, (const int32_t *)h.rcap, h.rstart, (int64_t)B + 1, false, st));
        SBP_HIP(sbpprim::scan(tp, tmp, (const int32_t *)h.ccap, h.cstart, (int64_t)B + 1, false, st));
        int32_t tot[2] = {0, 0};
        read_small(ctx, {{&tot[0], h.rstart + B, sizeof(int32_t)}, {&tot[1], h.cstart + B, sizeof(int32_t)}});
        int2 *rc = ctx->scratch[S_KEYS_A].as<int2>((size_t)tot[0] + 1);
        int2 *cc = ctx->scratch[S_KEYS_B].as<int2>((size_t)tot[1] + 1);
        k_slack_copy<<<grid_of((int64_t)B * 64), 64, 0, st>>>(B, bm->row_ptr, bm->row_cells, h.rstart, rc);
        k_slack_copy<<<grid_of((int64_t)B * 64), 64, 0, st>>>(B, bm->col_ptr, bm->col_cells, h.cstart, cc);
        FtArgs a = base_args(ctx, bm, g, epoch, p.greedy);
        // the walk's lookups read a dense copy of the matrix at this layout's start, which its commits keep
        // exact (row_add); any B up to the dense bound, as a walk makes many dependent lookups per vertex
        int32_t *dn = const_cast<int32_t *>(bm_ensure_dense(ctx, bm, g, true));
        a.m.dense = dn;
        a.m.dense_w = dn;
        a.m.dense_stride = B;
        a.m.rbeg = h.rstart;
        a.m.rlen = h.rlen;
        a.m.cbeg = h.cstart;
        a.m.clen = h.clen;
        a.m.row_cells = rc;
        a.m.col_cells = cc;
        a.items = d_order;
        a.count = count;
        a.bytes = bytes;
        a.err = reinterpret_cast<int *>(bytes + 1);
        const int lg = std::max(10, log2_ceil(2 * std::min<int64_t>(g->max_degree, bm->B) + 2));
        // The scan's high-degree launch takes at most the tables the finetune proposal launch uses for this graph
        // (one per vertex of total degree > SMALL_K, within kMaxTableBytes), and at least one workgroup's.
        constexpr size_t kMaxTableBytes = size_t(1) << 30;
        int64_t big_waves = 0;
        if (g->max_degree > SMALL_K) {
            big_waves = std::min<int64_t>(g->n_high_degree, 4096);
            big_waves = std::min<int64_t>(big_waves, (int64_t)(kMaxTableBytes / (table_ints(lg) * sizeof(int))));
            big_waves = std::max<int64_t>(FT_WAVES, big_waves / FT_WAVES * FT_WAVES);
        }
        a.gtable = g->max_degree > SMALL_K ? prepare_tables(ctx, lg, std::max<int64_t>(MH_WAVES, big_waves)) : nullptr;
        a.gcap_log2 = lg;
        a.m.row_cum = nullptr;
        a.m.col_cum = nullptr;
        SBP_HIP(hipEventRecord(ctx->ev[4], st));
        // Rounds of walk -> scan -> advance, queued without reading back; each kernel is a no-op outside its
        // status.  The host reads the status after each batch of rounds.
        int32_t hs[2] = {MH_RUN, 0};
        while (true) {
            for (int round = 0; round < 4; ++round) {
                launch_mh(a, h, sc, bm->np != 0, p.greedy != 0, scan, (int)(big_waves / FT_WAVES), st);
                if (!scan) break;
            }
            read_small(ctx, {{hs, book, 2 * sizeof(int32_t)}});
            if (hs[1] != MH_RUN && hs[1] != MH_SCAN) break;
        }
        SBP_HIP(hipEventRecord(ctx->ev[5], st));
        // compact the slack layout back into the blockmodel's CSR (row and column counts changed)
        size_t tmp2 = 0;
        int32_t *rp = bm->row_ptr, *cp = bm->col_ptr;
        SBP_HIP(hipMemsetAsync(h.rlen + B, 0, sizeof(int32_t), st));
        SBP_HIP(hipMemsetAsync(h.clen + B, 0, sizeof(int32_t), st));
        SBP_HIP(sbpprim::scan(nullptr, tmp2, (const int32_t *)h.rlen, rp, (int64_t)B + 1, false, st));
        tp = ctx->scratch[S_SORT_TMP].get(tmp2);
        SBP_HIP(sbpprim::scan(tp, tmp2, (const int32_t *)h.rlen, rp, (int64_t)B + 1, false, st));
        SBP_HIP(sbpprim::scan(tp, tmp2, (const int32_t *)h.clen, cp, (int64_t)B + 1, false, st));
        int32_t nnz[2] = {0, 0};
        read_small(ctx, {{&nnz[0], rp + B, sizeof(int32_t)}, {&nnz[1], cp + B, sizeof(int32_t)}});
        kernel_ms += (float)elapsed_ms(ctx, ctx->ev[4], ctx->ev[5]);
        SBP_CHECK(nnz[0] == nnz[1], SBP_E_STATE, "Metropolis-Hastings: row and column cell counts diverged");
        bm_ensure_cells(bm, nnz[0]);  // more cells than the compact arrays hold: grow them
        k_compact_copy<<<grid_of((int64_t)B * 64), 64, 0, st>>>(B, h.rstart, h.rlen, rc, rp, bm->row_cells);
        k_compact_copy<<<grid_of((int64_t)B * 64), 64, 0, st>>>(B, h.cstart, h.clen, cc, cp, bm->col_cells);
        bm->nnz = nnz[0];
        bm->hash_valid = bm->dense_valid = false;
        bm->keys_valid = false;
        bm_compute_cum(ctx, bm);
        SBP_HIP(hipStreamSynchronize(st));
        if (hs[1] == MH_DONE) break;
        // MH_SLACK: resume at the vertex that ran out of room; the rows and columns it lacked room in asked for
        // more (their capacities at least double, up to B + 2, where no move lacks room)
        // MH_OCCUPY: resume after a move into an empty block, with the constants of the new block count
        SBP_CHECK(hs[1] == MH_SLACK || hs[1] == MH_OCCUPY, SBP_E_STATE, "Metropolis-Hastings: unexpected walk status");
        retries = (hs[1] == MH_SLACK && hs[0] == start) ? retries + 1 : 0;
        SBP_CHECK(retries <= 64, SBP_E_STATE, "Metropolis-Hastings: slack growth did not converge");
        start = hs[0];
        k_fill_i32<<<1, 64, 0, st>>>(1, h.status, MH_RUN);  // stop_at stays: the walk resumes there
    }
    struct {
        unsigned long long moves;
        unsigned long long pad;
        double sum;
    } res;
    unsigned long long hb[2] = {0, 0};
    HistSlot hbook{0, 0, 0};
    if (bm->dc)
        read_small(ctx, {{&res, book + 4, sizeof(res)}, {hb, bytes, sizeof(hb)},
                         {&hbook, bm->hist + (size_t(1) << bm->hist_log2), sizeof(HistSlot)}});
    else
        read_small(ctx, {{&res, book + 4, sizeof(res)}, {hb, bytes, sizeof(hb)}});
    if (bm->dc) bm_hist_after_moves(ctx, bm, hbook.val, hbook.pad);
    SBP_CHECK(hb[1] == 0, SBP_E_STATE, "Metropolis-Hastings: block degrees inconsistent with the block matrix");
    if (sum_dS) *sum_dS = res.sum;
    if (stats) {
        stats->proposals += count;
        stats->moves += (int64_t)res.moves;
        stats->mh_bytes += hb[0];
        stats->mh_kernel_ms += kernel_ms;
        stats->finetune_ms += kernel_ms;
        stats->mh_launches += 1;
    }
}

}  // namespace sbpimpl
