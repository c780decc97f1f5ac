// Host-only entries to the merge carry-out (sbp_amd/csrc/carry.hpp) and the snowball samplers
// (sbp_amd/csrc/snowball.hpp), built without HIP so the CPU test suite can check them against the oracle at
// full size.  Not part of the device C-ABI.
#include <chrono>
#include <memory>
#include <vector>

#include "../csrc/carry.hpp"
#include "../csrc/snowball.hpp"

extern "C" int sbp_host_carry_out_dc(int32_t B, const int32_t *row_ptr, const int32_t *row_cols,
                                     const int32_t *row_cnts, const int32_t *col_ptr, const int32_t *col_rows,
                                     const int32_t *col_cnts, const int32_t *d_out, const int32_t *d_in,
                                     const int32_t *sizes, const int32_t *best_s, const double *best_dS,
                                     int64_t num_to_merge, int32_t approximate, int32_t threads,
                                     int32_t seed_from_proposals, int32_t verify_proposals, int32_t nonparametric,
                                     int64_t num_vertices, const uint64_t *hist_keys, const int32_t *hist_counts,
                                     int64_t hist_n, int32_t *block_map_out, int64_t *stats_out);

// stats_out (10 entries): pops, sequential evaluations, speculative evaluations, reused evaluations,
// requeues, merges, model load microseconds, carry-out loop microseconds, threads, proposal values that
// differ from a host re-evaluation (checked only when verify_proposals != 0).
extern "C" int sbp_host_carry_out(int32_t B, const int32_t *row_ptr, const int32_t *row_cols, const int32_t *row_cnts,
                                  const int32_t *col_ptr, const int32_t *col_rows, const int32_t *col_cnts,
                                  const int32_t *d_out, const int32_t *d_in, const int32_t *sizes,
                                  const int32_t *best_s, const double *best_dS, int64_t num_to_merge,
                                  int32_t approximate, int32_t threads, int32_t seed_from_proposals,
                                  int32_t verify_proposals, int32_t nonparametric, int64_t num_vertices,
                                  int32_t *block_map_out, int64_t *stats_out) {
    return sbp_host_carry_out_dc(B, row_ptr, row_cols, row_cnts, col_ptr, col_rows, col_cnts, d_out, d_in, sizes,
                                 best_s, best_dS, num_to_merge, approximate, threads, seed_from_proposals,
                                 verify_proposals, nonparametric, num_vertices, nullptr, nullptr, 0, block_map_out,
                                 stats_out);
}

// As sbp_host_carry_out, degree-corrected when hist_keys is given: the vertex-degree histograms as (key,
// count) slots, key = block << 33 | direction << 32 | degree (direction 1 = in).
extern "C" int sbp_host_carry_out_dc(int32_t B, const int32_t *row_ptr, const int32_t *row_cols,
                                     const int32_t *row_cnts, const int32_t *col_ptr, const int32_t *col_rows,
                                     const int32_t *col_cnts, const int32_t *d_out, const int32_t *d_in,
                                     const int32_t *sizes, const int32_t *best_s, const double *best_dS,
                                     int64_t num_to_merge, int32_t approximate, int32_t threads,
                                     int32_t seed_from_proposals, int32_t verify_proposals, int32_t nonparametric,
                                     int64_t num_vertices, const uint64_t *hist_keys, const int32_t *hist_counts,
                                     int64_t hist_n, int32_t *block_map_out, int64_t *stats_out) {
    using clk = std::chrono::steady_clock;
    auto us = [](clk::time_point a, clk::time_point b) {
        return (int64_t)std::chrono::duration_cast<std::chrono::microseconds>(b - a).count();
    };
    const int nt = threads > 0 ? threads : sbpcarry::carry_threads();
    std::unique_ptr<sbpcarry::Pool> pool(nt > 1 ? new sbpcarry::Pool(nt) : nullptr);
    sbpcarry::Model m;
    auto t0 = clk::now();
    if (!approximate) {
        m.load(B, row_ptr, row_cols, row_cnts, col_ptr, col_rows, col_cnts, 1, d_out, d_in, pool.get());
        if (nonparametric) m.set_nonparametric(sizes, num_vertices);
        if (nonparametric && hist_keys) m.set_degree_corrected(hist_keys, hist_counts, hist_n);
    }
    auto t1 = clk::now();
    int64_t mismatches = 0;
    if (verify_proposals && !approximate)
        for (int32_t b = 0; b < B; ++b) {
            if (best_s[b] < 0 || best_s[b] == b) continue;
            const double v = nonparametric ? sbpcarry::np_merge_dS(m, b, best_s[b], sbpcarry::np_merge_entries(m, b, best_s[b]))
                                           : sbpcarry::merge_dS(m, b, best_s[b]);
            if (v != best_dS[b]) ++mismatches;
        }
    auto t2 = clk::now();
    std::vector<int32_t> map;
    sbpcarry::CarryStats cs;
    int32_t nb = sbpcarry::carry_out(approximate ? nullptr : &m, B, best_s, best_dS, sizes, num_to_merge,
                                     approximate != 0, map, &cs, pool.get(), seed_from_proposals != 0);
    auto t3 = clk::now();
    for (int32_t b = 0; b < B; ++b) block_map_out[b] = map[b];
    if (stats_out) {
        const int64_t v[10] = {cs.pops, cs.evaluations, cs.speculative, cs.reused, cs.requeues, cs.merges,
                               us(t0, t1), us(t2, t3), nt, mismatches};
        for (int i = 0; i < 10; ++i) stats_out[i] = v[i];
    }
    return nb;
}

// alg 4 expansion_snowball, 5 degree_product_snowball over a host CSR; degree[N] = |out| + |in| - self-loop
// flag.  Writes mapping[N] and vertices[<= N] (sample order); returns the number of sampled slots or -1.
extern "C" int64_t sbp_host_snowball(int32_t alg, int32_t N, const int64_t *out_off, const int32_t *out_adj,
                                     const int64_t *in_off, const int32_t *in_adj, const int32_t *degree,
                                     double fraction, uint64_t seed, int32_t *mapping, int32_t *vertices) {
    if (N <= 0 || (alg != 4 && alg != 5)) return -1;
    const sbpsnow::Csr c{N, out_off, in_off, out_adj, in_adj};
    const int64_t target = (int64_t)(fraction * (double)N);
    std::vector<int32_t> map;
    std::vector<int32_t> order = alg == 4 ? sbpsnow::expansion_snowball(c, target, seed, map)
                                          : sbpsnow::degree_product_snowball(c, degree, target, seed, map);
    for (int32_t v = 0; v < N; ++v) mapping[v] = map[v];
    for (size_t i = 0; i < order.size(); ++i) vertices[i] = order[i];
    return (int64_t)order.size();
}
