// Merge carry-out on the host (no device code; shared by the library and its CPU-side checks).
//
// carry_out_best_merges_advanced (src/block_merge.cpp:40-117) pops merges from a min-heap on delta-MDL,
// re-evaluates each against the blockmodel as already changed by the merges before it, and re-queues it
// when it got worse than the next candidate.  Every decision depends on the previous ones, so this is one
// sequential loop.  The reference relabels all B entries of block_map per merge and copies robin_map rows;
// here each block's row and column is an open-addressing hash map, a merge moves row/column r into s and
// renames r to s in the touched rows and columns in O(|row r| + |col r|), and block_map is a union-find
// whose roots equal the reference's (fully compressed) entries.  Heap ties are broken by ascending block id.
// delta-MDL terms use glibc logf and the same 128-bit fixed-point sum as the device kernels, so every
// re-evaluation is bit-identical to what the merge-proposal kernel computes for the same state.
//
// Blockmodel::carry_out_best_merges (src/blockmodel/blockmodel.cpp:55-122), used with --approximate, takes
// the merges in stable ascending order of delta-MDL without re-evaluation, following block_map chains.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <queue>
#include <vector>

namespace sbpcarry {

// Linear-probing map int32 -> int64 with backward-shift deletion.
struct FlatMap {
    std::vector<int32_t> keys;  // -1 = empty
    std::vector<int64_t> vals;
    uint32_t n = 0;

    static uint32_t h(int32_t k) { return (uint32_t)k * 0x9E3779B1u; }
    void init(uint32_t expected) {
        uint32_t cap = 4;
        while (cap < 2 * expected + 2) cap <<= 1;
        keys.assign(cap, -1);
        vals.assign(cap, 0);
        n = 0;
    }
    uint32_t mask() const { return (uint32_t)keys.size() - 1; }
    int64_t get(int32_t k) const {
        if (keys.empty()) return 0;
        uint32_t m = mask(), i = h(k) & m;
        while (keys[i] != -1) {
            if (keys[i] == k) return vals[i];
            i = (i + 1) & m;
        }
        return 0;
    }
    void grow() {
        std::vector<int32_t> ok;
        std::vector<int64_t> ov;
        ok.swap(keys);
        ov.swap(vals);
        init(n * 2 + 4);
        for (size_t i = 0; i < ok.size(); ++i)
            if (ok[i] != -1) add(ok[i], ov[i]);
    }
    void erase_at(uint32_t i) {
        uint32_t m = mask();
        keys[i] = -1;
        vals[i] = 0;
        n--;
        uint32_t j = (i + 1) & m;
        while (keys[j] != -1) {  // backward shift: re-place entries whose probe path crossed i
            uint32_t home = h(keys[j]) & m;
            bool move = (j > i) ? (home <= i || home > j) : (home <= i && home > j);
            if (move) {
                keys[i] = keys[j];
                vals[i] = vals[j];
                keys[j] = -1;
                vals[j] = 0;
                i = j;
            }
            j = (j + 1) & m;
        }
    }
    // adds d; erases the entry when it reaches 0
    void add(int32_t k, int64_t d) {
        if (d == 0) return;
        if (keys.empty()) init(4);
        uint32_t m = mask(), i = h(k) & m;
        while (keys[i] != -1) {
            if (keys[i] == k) {
                vals[i] += d;
                if (vals[i] == 0) erase_at(i);
                return;
            }
            i = (i + 1) & m;
        }
        keys[i] = k;
        vals[i] = d;
        n++;
        if (2 * n + 2 > keys.size()) grow();
    }
    void erase(int32_t k) {
        if (keys.empty()) return;
        uint32_t m = mask(), i = h(k) & m;
        while (keys[i] != -1) {
            if (keys[i] == k) {
                erase_at(i);
                return;
            }
            i = (i + 1) & m;
        }
    }
    template <typename F>
    void each(F f) const {
        for (size_t i = 0; i < keys.size(); ++i)
            if (keys[i] != -1) f(keys[i], vals[i]);
    }
    void clear() {
        keys.clear();
        vals.clear();
        n = 0;
    }
};

struct Fx128 {
    __int128 v = 0;
    static __int128 from(double d) {
        uint64_t bits;
        memcpy(&bits, &d, 8);
        int e = (int)((bits >> 52) & 0x7ff);
        if (e == 0) return 0;
        uint64_t m = (bits & 0xfffffffffffffull) | (1ull << 52);
        int shift = e - 1011;
        unsigned __int128 mag;
        if (shift >= 0) mag = (unsigned __int128)m << shift;
        else if (shift > -64) mag = (unsigned __int128)(m >> (-shift));
        else mag = 0;
        __int128 x = (__int128)mag;
        return (bits >> 63) ? -x : x;
    }
    void add(double d) { v += from(d); }
    void sub(double d) { v -= from(d); }
    double value() const { return (double)(int64_t)(v >> 64) + (double)(uint64_t)v * 0x1p-64; }
};

// fastlog (include/fastlog.hpp:21-30) with a table of logf(i) for small counts and degrees, like the
// reference's fastlog_cache (src/fastlog.cpp:13-23) -- the table holds logf's own values.
struct LogfTable {
    static constexpr int64_t SIZE = 1 << 22;
    std::vector<float> v;
    LogfTable() : v(SIZE) {
        v[0] = 0.0f;
        for (int64_t i = 1; i < SIZE; ++i) v[i] = logf((float)i);
    }
};
inline const LogfTable &logf_table() {
    static const LogfTable t;
    return t;
}
inline double fastlog(int64_t x) {
    if (x < LogfTable::SIZE) return x <= 0 ? 0.0 : (double)logf_table().v[x];
    return (double)logf((float)x);
}
// common::cell_entropy (include/common.hpp:39-49)
inline double cell(int64_t value, int64_t din, int64_t dout) {
    if (value == 0) return 0.0;
    return (double)value * (fastlog(value) - (fastlog(din) + fastlog(dout)));
}

struct Model {
    int32_t B = 0;
    std::vector<FlatMap> row, col;
    std::vector<int64_t> d_out, d_in;

    // row-major cells (col, count) per row and column-major cells (row, count) per column
    void load(int32_t nb, const int32_t *row_ptr, const int32_t *row_cols, const int32_t *row_cnts,
              const int32_t *col_ptr, const int32_t *col_rows, const int32_t *col_cnts, const int32_t *dout,
              const int32_t *din) {
        B = nb;
        row.assign(B, FlatMap());
        col.assign(B, FlatMap());
        d_out.assign(dout, dout + B);
        d_in.assign(din, din + B);
        for (int32_t b = 0; b < B; ++b) {
            row[b].init(row_ptr[b + 1] - row_ptr[b]);
            for (int32_t i = row_ptr[b]; i < row_ptr[b + 1]; ++i) row[b].add(row_cols[i], row_cnts[i]);
            col[b].init(col_ptr[b + 1] - col_ptr[b]);
            for (int32_t i = col_ptr[b]; i < col_ptr[b + 1]; ++i) col[b].add(col_rows[i], col_cnts[i]);
        }
    }
    int64_t get(int32_t r, int32_t c) const { return row[r].n <= col[c].n ? row[r].get(c) : col[c].get(r); }
};

// entropy::block_merge_delta_mdl (src/entropy.cpp:129-182) for merging r into s on the current state:
// Delta entries term by term, (s, s) when unchanged, the other unchanged cells of row s / column s in bulk
// (M * (fastlog(d') - fastlog(d)) each) -- the same terms as the merge-proposal kernel.
inline double merge_dS(const Model &m, int32_t r, int32_t s) {
    const int64_t k_out = m.d_out[r], k_in = m.d_in[r];
    const int64_t din_s_new = m.d_in[s] + k_in, dout_s_new = m.d_out[s] + k_out;
    Fx128 acc;
    int64_t sub_row = 0, sub_col = 0;
    m.row[r].each([&](int32_t c, int64_t v) {
        acc.add(cell(v, m.d_in[c], m.d_out[r]));
        if (c != r && c != s) {
            int64_t msc = m.row[s].get(c);
            acc.add(cell(msc, m.d_in[c], m.d_out[s]));
            acc.sub(cell(msc + v, m.d_in[c], dout_s_new));
            sub_row += msc;
        }
    });
    m.col[r].each([&](int32_t x, int64_t v) {
        if (x == r) return;
        acc.add(cell(v, m.d_in[r], m.d_out[x]));
        if (x != s) {
            int64_t mxs = m.col[s].get(x);
            acc.add(cell(mxs, m.d_in[s], m.d_out[x]));
            acc.sub(cell(mxs + v, din_s_new, m.d_out[x]));
            sub_col += mxs;
        }
    });
    const int64_t mrr = m.get(r, r), mrs = m.get(r, s), msr = m.get(s, r), mss = m.get(s, s);
    const int64_t dss = mrr + mrs + msr;
    if (dss > 0) {
        acc.add(cell(mss, m.d_in[s], m.d_out[s]));
        acc.sub(cell(mss + dss, din_s_new, dout_s_new));
    } else if (mss > 0) {
        acc.add(cell(mss, m.d_in[s], m.d_out[s]));
        acc.sub(cell(mss, din_s_new, dout_s_new));
    }
    const int64_t S_row = m.d_out[s] - mss - msr - sub_row, S_col = m.d_in[s] - mss - mrs - sub_col;
    acc.add((double)S_row * (fastlog(dout_s_new) - fastlog(m.d_out[s])));
    acc.add((double)S_col * (fastlog(din_s_new) - fastlog(m.d_in[s])));
    return acc.value();
}

// Blockmodel::merge_block (src/blockmodel/blockmodel.cpp:337-363): row r and column r move into s.
inline void merge_apply(Model &m, int32_t r, int32_t s) {
    std::vector<std::pair<int32_t, int64_t>> rr, rc;
    rr.reserve(m.row[r].n);
    rc.reserve(m.col[r].n);
    m.row[r].each([&](int32_t c, int64_t v) { rr.emplace_back(c, v); });
    m.col[r].each([&](int32_t x, int64_t v) { rc.emplace_back(x, v); });
    for (auto &cv : rr) {  // (r, c) -> (s, c'), c' = s for c in {r, s}
        int32_t c = cv.first, c2 = (c == r || c == s) ? s : c;
        m.row[s].add(c2, cv.second);
        m.col[c].erase(r);
        m.col[c2].add(s, cv.second);
    }
    for (auto &xv : rc) {  // (x, r) -> (x', s), x' = s for x = s
        int32_t x = xv.first;
        if (x == r) continue;
        int32_t x2 = (x == s) ? s : x;
        m.col[s].add(x2, xv.second);
        m.row[x].erase(r);
        m.row[x2].add(s, xv.second);
    }
    m.row[r].clear();
    m.col[r].clear();
    m.d_out[s] += m.d_out[r];
    m.d_in[s] += m.d_in[r];
    m.d_out[r] = 0;
    m.d_in[r] = 0;
}

struct CarryStats {
    int64_t pops = 0, evaluations = 0, requeues = 0, merges = 0;
};

// Returns the block count after the merges; block_map[b] = new label of old block b (labels compacted in
// ascending order of the surviving roots, as Blockmodel::build_mapping does).  `m` may be null when
// approximate.
inline int32_t carry_out(Model *m, int32_t B, const int32_t *best_s, const double *best_dS, const int32_t *sizes,
                         int64_t num_to_merge, bool approximate, std::vector<int32_t> &block_map,
                         CarryStats *cs = nullptr) {
    std::vector<int32_t> parent(B);
    for (int32_t b = 0; b < B; ++b) parent[b] = b;
    auto find = [&](int32_t x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    };
    int64_t merged = 0;
    CarryStats st;
    if (approximate) {
        std::vector<int32_t> order(B);
        for (int32_t b = 0; b < B; ++b) order[b] = b;
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return best_dS[a] < best_dS[b]; });
        int64_t counter = 0;
        while (merged < num_to_merge && counter < B) {
            int32_t from = order[counter++];
            int32_t to = find(best_s[from]);
            if (to != from) {
                parent[from] = to;
                merged++;
            }
        }
    } else {
        struct Item {
            double dS;
            int32_t r, s;
        };
        auto worse = [](const Item &a, const Item &b) {
            if (a.dS != b.dS) return a.dS > b.dS;
            return a.r > b.r;
        };
        std::vector<Item> init;
        init.reserve(B);
        for (int32_t b = 0; b < B; ++b) init.push_back(Item{best_dS[b], b, best_s[b]});
        std::priority_queue<Item, std::vector<Item>, decltype(worse)> heap(worse, std::move(init));
        while (merged < num_to_merge && !heap.empty()) {
            Item it = heap.top();
            heap.pop();
            st.pops++;
            const int32_t from = it.r;
            const int32_t to = find(it.s);
            if (from == to) continue;
            double actual = merge_dS(*m, from, to);
            st.evaluations++;
            if (!heap.empty() && actual > heap.top().dS) {
                it.dS = actual;
                heap.push(it);
                st.requeues++;
                continue;
            }
            parent[from] = to;
            merge_apply(*m, from, to);
            merged++;
        }
    }
    st.merges = merged;
    if (cs) *cs = st;
    std::vector<int32_t> used(B, 0);
    for (int32_t b = 0; b < B; ++b)
        if (sizes[b] > 0) used[find(b)] = 1;
    std::vector<int32_t> label(B, -1);
    int32_t next = 0;
    for (int32_t b = 0; b < B; ++b)
        if (used[b]) label[b] = next++;
    block_map.resize(B);
    for (int32_t b = 0; b < B; ++b) block_map[b] = label[find(b)];
    return next;
}

}  // namespace sbpcarry
