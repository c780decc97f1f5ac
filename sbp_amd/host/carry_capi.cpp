// Host-only entry to the merge carry-out (sbp_amd/csrc/carry.hpp), built without HIP so the CPU test
// suite can check it against the oracle at full size.  Not part of the device C-ABI.
#include <vector>

#include "../csrc/carry.hpp"

extern "C" int sbp_host_carry_out(int32_t B, const int32_t *row_ptr, const int32_t *row_cols, const int32_t *row_cnts,
                                  const int32_t *col_ptr, const int32_t *col_rows, const int32_t *col_cnts,
                                  const int32_t *d_out, const int32_t *d_in, const int32_t *sizes,
                                  const int32_t *best_s, const double *best_dS, int64_t num_to_merge,
                                  int32_t approximate, int32_t *block_map_out, int64_t *stats_out) {
    sbpcarry::Model m;
    if (!approximate) m.load(B, row_ptr, row_cols, row_cnts, col_ptr, col_rows, col_cnts, d_out, d_in);
    std::vector<int32_t> map;
    sbpcarry::CarryStats cs;
    int32_t nb = sbpcarry::carry_out(approximate ? nullptr : &m, B, best_s, best_dS, sizes, num_to_merge,
                                     approximate != 0, map, &cs);
    for (int32_t b = 0; b < B; ++b) block_map_out[b] = map[b];
    if (stats_out) {
        stats_out[0] = cs.pops;
        stats_out[1] = cs.evaluations;
        stats_out[2] = cs.requeues;
        stats_out[3] = cs.merges;
    }
    return nb;
}
