// Snowball samplers on the host (no device code): sample::expansion_snowball (src/sample.cpp:94-155) and
// sample::degree_product_snowball (src/sample.cpp:51-92).  Both grow the sample one vertex at a time from a
// frontier, each step depending on the last, so they run sequentially over the graph's CSR; the induced
// subgraph is then built on the device like the other samplers' (from_vertices).
//
// Random draws: common::random_integer(lo, hi) is replaced by draw j of the replay stream,
// lo + floor(w0 * (hi - lo + 1) / 2^32) with (w0, ...) = Philox4x32-10(seed; j, 0, TAG_SNOWBALL, 0), j = 0, 1, ...
// in the order the reference makes the calls.
//
// expansion_snowball keeps the reference's data flow exactly (contributions are never reset when a vertex is
// sampled, argmax runs over all vertices and returns the first maximum, a chosen vertex that is already sampled
// is appended again) but replaces the O(N) argmax scan by a max segment tree and std::advance over the
// neighbour set by a Fenwick-tree order statistic.  degree_product_snowball uses std::priority_queue with the
// reference's comparator, so ties between equal degree products pop in the same order.  Where the reference
// calls frontier.top() on an empty queue (a restart vertex with no unsampled neighbours: undefined behaviour)
// the loop goes on to its next iteration.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <queue>
#include <tuple>
#include <vector>

namespace sbpsnow {

constexpr uint32_t TAG_SNOWBALL = 0x53410002u;

inline void philox(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4]) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * x0, p1 = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0, y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        x0 = y0;
        x1 = (uint32_t)p1;
        x2 = y2;
        x3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0, out[1] = x1, out[2] = x2, out[3] = x3;
}

// The replay stream of random_integer calls.
struct Draws {
    uint64_t seed;
    uint32_t next = 0;
    explicit Draws(uint64_t s) : seed(s) {}
    int64_t integer(int64_t lo, int64_t hi) {
        uint32_t w[4];
        philox(seed, next++, 0, TAG_SNOWBALL, 0, w);
        return lo + (int64_t)(((uint64_t)w[0] * (uint64_t)(hi - lo + 1)) >> 32);
    }
};

struct Csr {
    int32_t N;
    const int64_t *out_off, *in_off;
    const int32_t *out_adj, *in_adj;
};

// Leftmost maximum of a vector under point updates.
class MaxTree {
  public:
    explicit MaxTree(int32_t n) : n_(n), size_(1) {
        while (size_ < n_) size_ <<= 1;
        best_.assign(2 * (size_t)size_, 0);
        for (int32_t i = 0; i < size_; ++i) best_[size_ + i] = i;
        val_.assign(size_, 0);
        for (int32_t i = size_ - 1; i >= 1; --i) best_[i] = best_[2 * i];
    }
    void set(int32_t i, int64_t v) {
        val_[i] = v;
        for (int32_t p = (i + size_) >> 1; p >= 1; p >>= 1) {
            const int32_t a = best_[2 * p], b = best_[2 * p + 1];
            best_[p] = val_[b] > val_[a] ? b : a;  // ties keep the left (smaller) index
        }
    }
    int32_t argmax() const { return best_[1]; }

  private:
    int32_t n_, size_;
    std::vector<int32_t> best_;
    std::vector<int64_t> val_;
};

// Ordered set of vertex ids with k-th smallest.
class OrderedSet {
  public:
    explicit OrderedSet(int32_t n) : n_(n), tree_(n + 1, 0) {
        lg_ = 1;
        while ((1 << lg_) <= n_) ++lg_;
    }
    void add(int32_t v, int32_t d) {
        count_ += d;
        for (int32_t i = v + 1; i <= n_; i += i & -i) tree_[i] += d;
    }
    int64_t size() const { return count_; }
    int32_t kth(int64_t k) const {  // 0-based
        int32_t pos = 0;
        int64_t rem = k + 1;
        for (int b = lg_; b >= 0; --b) {
            const int32_t nxt = pos + (1 << b);
            if (nxt <= n_ && tree_[nxt] < rem) {
                pos = nxt;
                rem -= tree_[nxt];
            }
        }
        return pos;  // the (k+1)-th present id
    }

  private:
    int32_t n_, lg_;
    int64_t count_ = 0;
    std::vector<int32_t> tree_;
};

// Returns the sampled vertices in sample order (mapping[v] = the last index at which v was appended).
inline std::vector<int32_t> expansion_snowball(const Csr &g, int64_t target, uint64_t seed, std::vector<int32_t> &mapping) {
    const int32_t N = g.N;
    std::vector<int32_t> sampled;
    mapping.assign(N, -1);
    std::vector<int64_t> contribution(N, 0);
    std::vector<char> flag(N, 0);
    MaxTree maxc(N);
    OrderedSet neighbors(N);
    int64_t sum = 0;
    Draws rng(seed);
    auto update = [&](int32_t u) {  // es_update_contribution (sample.cpp:112-124)
        int64_t c = contribution[u];
        for (int64_t e = g.out_off[u]; e < g.out_off[u + 1]; ++e) {
            const int32_t nb = g.out_adj[e];
            if (flag[nb] || mapping[nb] >= 0) continue;
            c++;
            sum++;
        }
        if (c != contribution[u]) {
            contribution[u] = c;
            maxc.set(u, c);
        }
        for (int64_t e = g.in_off[u]; e < g.in_off[u + 1]; ++e) {
            const int32_t nb = g.in_adj[e];
            if (contribution[nb] > 0) {
                contribution[nb]--;
                maxc.set(nb, contribution[nb]);
                sum--;
            }
        }
    };
    auto add = [&](int32_t v) {  // es_add_vertex (sample.cpp:94-110)
        sampled.push_back(v);
        mapping[v] = (int32_t)sampled.size() - 1;
        for (int64_t e = g.out_off[v]; e < g.out_off[v + 1]; ++e) {
            const int32_t nb = g.out_adj[e];
            if (flag[nb] || mapping[nb] >= 0) continue;
            neighbors.add(nb, 1);
            flag[nb] = 1;
            if (contribution[nb] > 0) continue;
            update(nb);
        }
        if (flag[v]) neighbors.add(v, -1);
        flag[v] = 0;
    };
    add((int32_t)rng.integer(0, N - 1));
    while ((int64_t)sampled.size() < target) {
        if (neighbors.size() == 0) {
            int32_t v;
            do {
                v = (int32_t)rng.integer(0, N - 1);
            } while (mapping[v] >= 0);
            add(v);
        } else if (sum == 0) {
            const int64_t k = rng.integer(0, neighbors.size() - 1);
            add(neighbors.kth(k));
        } else {
            add(maxc.argmax());
        }
    }
    return sampled;
}

inline std::vector<int32_t> degree_product_snowball(const Csr &g, const int32_t *degree, int64_t target, uint64_t seed,
                                                    std::vector<int32_t> &mapping) {
    const int32_t N = g.N;
    std::vector<int32_t> sampled;
    mapping.assign(N, -1);
    typedef std::tuple<long, long, long> edge_t;
    auto cmp = [](const edge_t &l, const edge_t &r) { return std::get<2>(l) < std::get<2>(r); };
    std::priority_queue<edge_t, std::vector<edge_t>, decltype(cmp)> frontier(cmp);
    Draws rng(seed);
    auto add = [&](int32_t v) {
        mapping[v] = (int32_t)sampled.size();
        sampled.push_back(v);
        for (int64_t e = g.out_off[v]; e < g.out_off[v + 1]; ++e) {
            const int32_t nb = g.out_adj[e];
            if (mapping[nb] != -1) continue;
            frontier.push(edge_t(v, nb, (long)degree[v] * (long)degree[nb]));
        }
        for (int64_t e = g.in_off[v]; e < g.in_off[v + 1]; ++e) {
            const int32_t nb = g.in_adj[e];
            if (mapping[nb] != -1) continue;
            frontier.push(edge_t(nb, v, (long)degree[nb] * (long)degree[v]));
        }
    };
    add((int32_t)rng.integer(0, N - 1));
    while ((int64_t)sampled.size() < target) {
        if (frontier.empty()) {
            int32_t v;
            do {
                v = (int32_t)rng.integer(0, N - 1);
            } while (mapping[v] >= 0);
            add(v);
            if (frontier.empty()) continue;
        }
        const edge_t next = frontier.top();
        frontier.pop();
        const int32_t src = (int32_t)std::get<0>(next), dst = (int32_t)std::get<1>(next);
        if (mapping[src] == -1) add(src);
        if (mapping[dst] == -1) add(dst);
    }
    return sampled;
}

// sample::reattach (src/sample.cpp:367-405): the assignment of every vertex of the full graph from the
// partition of its detached graph.  Kept vertices take their block; the others, in id order, draw
// random_integer(0, B - 1) (draw j of the stream with tag TAG_REATTACH, one per unassigned vertex) and then
// scan their out- and then in-neighbours: each assigned neighbour's block overwrites theirs, and the first
// unassigned neighbour gets the drawn block together with the vertex and ends that list's scan.  A vertex
// with no neighbour keeps the drawn block.
constexpr uint32_t TAG_REATTACH = 0x52410001u;

inline void reattach(const Csr &g, const int32_t *mapping, const int32_t *sample_assignment, int32_t B, uint64_t seed,
                     int32_t *assignment) {
    const int32_t N = g.N;
    for (int32_t v = 0; v < N; ++v) assignment[v] = mapping[v] >= 0 ? sample_assignment[mapping[v]] : -1;
    uint32_t next = 0;
    auto draw = [&]() {
        uint32_t w[4];
        philox(seed, next++, 0, TAG_REATTACH, 0, w);
        return (int32_t)(((uint64_t)w[0] * (uint64_t)B) >> 32);
    };
    for (int32_t v = 0; v < N; ++v) {
        if (assignment[v] != -1) continue;
        const int32_t random_block = draw();
        for (int64_t e = g.out_off[v]; e < g.out_off[v + 1]; ++e) {
            const int32_t u = g.out_adj[e];
            if (assignment[u] == -1) {
                assignment[v] = assignment[u] = random_block;
                break;
            }
            assignment[v] = assignment[u];
        }
        for (int64_t e = g.in_off[v]; e < g.in_off[v + 1]; ++e) {
            const int32_t u = g.in_adj[e];
            if (assignment[u] == -1) {
                assignment[v] = assignment[u] = random_block;
                break;
            }
            assignment[v] = assignment[u];
        }
        if (assignment[v] < 0) assignment[v] = random_block;
    }
}

}  // namespace sbpsnow
