// Host-side mutable blockmodel for the merge carry-out.
//
// carry_out_best_merges_advanced (src/block_merge.cpp:40-117) pops merges from a min-heap on delta-MDL,
// re-evaluates each against the blockmodel as already modified by the previous merges and re-queues it
// when it got worse than the next candidate.  Every decision depends on the one before, so this is a
// sequential loop; it runs here on the host against a copy of the device CSR, with union-find roots in
// place of the reference's O(B) relabel per merge (the reference keeps block_map fully compressed, so a
// root lookup returns exactly its entry).  Heap ties are broken by ascending block id.  delta-MDL terms use
// glibc logf and the same 128-bit fixed-point summation as the device kernels, so re-evaluations agree bit
// for bit with the proposal kernel's values.
#pragma once
#include <math.h>
#include <string.h>

#include <algorithm>
#include <queue>
#include <vector>

#include "internal.hpp"

namespace sbpimpl {

struct HostBlockmodel {
    int32_t B = 0;
    std::vector<std::vector<std::pair<int32_t, int64_t>>> row, col;
    std::vector<int64_t> d_out, d_in, d_tot;
    std::vector<int32_t> sizes;

    void download(sbp_ctx ctx, sbp_bm bm) {
        SBP_HIP(hipStreamSynchronize(ctx->stream));
        B = bm->B;
        std::vector<int32_t> rp(B + 1), cp(B + 1), a(B), b(B), c(B), sz(B);
        std::vector<int2> rc(bm->nnz), cc(bm->nnz);
        SBP_HIP(hipMemcpy(rp.data(), bm->row_ptr, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost));
        SBP_HIP(hipMemcpy(cp.data(), bm->col_ptr, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost));
        if (bm->nnz) {
            SBP_HIP(hipMemcpy(rc.data(), bm->row_cells, sizeof(int2) * bm->nnz, hipMemcpyDeviceToHost));
            SBP_HIP(hipMemcpy(cc.data(), bm->col_cells, sizeof(int2) * bm->nnz, hipMemcpyDeviceToHost));
        }
        SBP_HIP(hipMemcpy(a.data(), bm->d_out, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
        SBP_HIP(hipMemcpy(b.data(), bm->d_in, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
        SBP_HIP(hipMemcpy(c.data(), bm->d_tot, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
        SBP_HIP(hipMemcpy(sz.data(), bm->sizes, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
        row.assign(B, {});
        col.assign(B, {});
        for (int32_t r = 0; r < B; ++r) {
            row[r].reserve(rp[r + 1] - rp[r]);
            for (int32_t i = rp[r]; i < rp[r + 1]; ++i) row[r].push_back({rc[i].x, rc[i].y});
            col[r].reserve(cp[r + 1] - cp[r]);
            for (int32_t i = cp[r]; i < cp[r + 1]; ++i) col[r].push_back({cc[i].x, cc[i].y});
        }
        d_out.assign(a.begin(), a.end());
        d_in.assign(b.begin(), b.end());
        d_tot.assign(c.begin(), c.end());
        sizes = sz;
    }

    bool equals(const HostBlockmodel &o) const {
        return B == o.B && row == o.row && col == o.col && d_out == o.d_out && d_in == o.d_in && d_tot == o.d_tot &&
               sizes == o.sizes;
    }

    static int64_t vget(const std::vector<std::pair<int32_t, int64_t>> &v, int32_t key) {
        auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(key, (int64_t)INT64_MIN));
        return (it != v.end() && it->first == key) ? it->second : 0;
    }
    static void vadd(std::vector<std::pair<int32_t, int64_t>> &v, int32_t key, int64_t d) {
        auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(key, (int64_t)INT64_MIN));
        if (it != v.end() && it->first == key) {
            it->second += d;
            if (it->second == 0) v.erase(it);
        } else if (d != 0) {
            v.insert(it, {key, d});
        }
    }
    int64_t get(int32_t r, int32_t c) const {
        return row[r].size() <= col[c].size() ? vget(row[r], c) : vget(col[c], r);
    }
    void add(int32_t r, int32_t c, int64_t d) {
        vadd(row[r], c, d);
        vadd(col[c], r, d);
    }
};

// ---- delta-MDL on the host (same terms and fixed-point sum as merge.hip)

struct HostFx {
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
    double value() const {
        int64_t hi = (int64_t)(v >> 64);
        uint64_t lo = (uint64_t)v;
        return (double)hi + (double)lo * 0x1p-64;
    }
};

inline double host_fastlog(int64_t x) { return x == 0 ? 0.0 : (double)logf((float)x); }
inline double host_cell(int64_t value, int64_t din, int64_t dout) {
    if (value == 0) return 0.0;
    return (double)value * (host_fastlog(value) - (host_fastlog(din) + host_fastlog(dout)));
}

// entropy::block_merge_delta_mdl for merging r into s on the current host state.
inline double host_merge_dS(const HostBlockmodel &m, int32_t r, int32_t s) {
    const int64_t k_out = m.d_out[r], k_in = m.d_in[r];
    const int64_t din_s_new = m.d_in[s] + k_in, dout_s_new = m.d_out[s] + k_out;
    HostFx acc;
    for (auto &cell : m.row[r]) {
        int32_t c = cell.first;
        acc.add(host_cell(cell.second, m.d_in[c], m.d_out[r]));
        if (c != r && c != s) {
            int64_t msc = HostBlockmodel::vget(m.row[s], c);
            acc.add(host_cell(msc, m.d_in[c], m.d_out[s]));
            acc.sub(host_cell(msc + cell.second, m.d_in[c], dout_s_new));
        }
    }
    for (auto &cell : m.col[r]) {
        int32_t x = cell.first;
        if (x == r) continue;
        acc.add(host_cell(cell.second, m.d_in[r], m.d_out[x]));
        if (x != s) {
            int64_t mxs = HostBlockmodel::vget(m.col[s], x);
            acc.add(host_cell(mxs, m.d_in[s], m.d_out[x]));
            acc.sub(host_cell(mxs + cell.second, din_s_new, m.d_out[x]));
        }
    }
    const int64_t dss = m.get(r, r) + m.get(r, s) + m.get(s, r), mss = m.get(s, s);
    if (dss > 0) {
        acc.add(host_cell(mss, m.d_in[s], m.d_out[s]));
        acc.sub(host_cell(mss + dss, din_s_new, dout_s_new));
    }
    for (auto &cell : m.row[s]) {
        int32_t c = cell.first;
        if (c == r) continue;
        if (c == s && dss > 0) continue;
        if (c != s && HostBlockmodel::vget(m.row[r], c) != 0) continue;
        int64_t din_new = (c == s) ? din_s_new : m.d_in[c];
        acc.add(host_cell(cell.second, m.d_in[c], m.d_out[s]));
        acc.sub(host_cell(cell.second, din_new, dout_s_new));
    }
    for (auto &cell : m.col[s]) {
        int32_t x = cell.first;
        if (x == r || x == s) continue;
        if (HostBlockmodel::vget(m.col[r], x) != 0) continue;
        acc.add(host_cell(cell.second, m.d_in[s], m.d_out[x]));
        acc.sub(host_cell(cell.second, din_s_new, m.d_out[x]));
    }
    return acc.value();
}

// Blockmodel::merge_block (src/blockmodel/blockmodel.cpp:337-363): row r and column r move into s.
inline void host_merge_apply(HostBlockmodel &m, int32_t r, int32_t s) {
    const int64_t k_out = m.d_out[r], k_in = m.d_in[r];
    std::vector<std::pair<int32_t, int64_t>> rrow = m.row[r], rcol = m.col[r];
    // Delta of block_merge::blockmodel_delta (src/block_merge.cpp:14-38)
    std::vector<std::tuple<int32_t, int32_t, int64_t>> cells;
    for (auto &cell : rrow) {
        int32_t c = cell.first;
        cells.emplace_back((c == r || c == s) ? s : s, (c == r || c == s) ? s : c, cell.second);
        cells.emplace_back(r, c, -cell.second);
    }
    for (auto &cell : rcol) {
        int32_t x = cell.first;
        if (x == r) continue;
        cells.emplace_back(x == s ? s : x, s, cell.second);
        cells.emplace_back(x, r, -cell.second);
    }
    for (auto &t : cells) m.add(std::get<0>(t), std::get<1>(t), std::get<2>(t));
    m.d_out[r] = 0;
    m.d_out[s] += k_out;
    m.d_in[r] = 0;
    m.d_in[s] += k_in;
    m.d_tot[r] = 0;
    m.d_tot[s] = m.d_out[s] + m.d_in[s] - m.get(s, s);
    m.sizes[s] += m.sizes[r];
    m.sizes[r] = 0;
}

// Returns the number of blocks after the merges; block_map[b] = new label of old block b.
inline int32_t carry_out_merges(sbp_ctx ctx, sbp_bm bm, const std::vector<int32_t> &best_s,
                                const std::vector<double> &best_dS, int64_t num_to_merge, bool approximate,
                                std::vector<int32_t> &block_map) {
    const int32_t B = bm->B;
    std::vector<int32_t> parent(B);
    for (int32_t b = 0; b < B; ++b) parent[b] = b;
    std::vector<int32_t> sizes(B);
    SBP_HIP(hipStreamSynchronize(ctx->stream));
    SBP_HIP(hipMemcpy(sizes.data(), bm->sizes, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
    int64_t merged = 0;
    auto find = [&](int32_t x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    };
    if (approximate) {
        // Blockmodel::carry_out_best_merges (src/blockmodel/blockmodel.cpp:55-122): stable ascending
        // order of delta-MDL, chains followed through block_map, no re-evaluation.
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
        auto worse = [](const Item &a, const Item &b) {  // priority_queue top = smallest (dS, r)
            if (a.dS != b.dS) return a.dS > b.dS;
            return a.r > b.r;
        };
        std::priority_queue<Item, std::vector<Item>, decltype(worse)> heap(worse);
        for (int32_t b = 0; b < B; ++b) heap.push(Item{best_dS[b], b, best_s[b]});
        HostBlockmodel m;
        m.download(ctx, bm);
        while (merged < num_to_merge && !heap.empty()) {
            Item it = heap.top();
            heap.pop();
            const int32_t from = it.r;
            const int32_t to = find(it.s);
            if (from == to) continue;
            double actual = host_merge_dS(m, from, to);
            if (!heap.empty() && actual > heap.top().dS) {
                it.dS = actual;
                heap.push(it);
                continue;
            }
            parent[from] = to;
            host_merge_apply(m, from, to);
            merged++;
        }
    }
    // Blockmodel::build_mapping (src/blockmodel/blockmodel.cpp:34-46): labels in use, ascending
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

}  // namespace sbpimpl
