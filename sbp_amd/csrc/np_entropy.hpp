// Nonparametric description-length terms (entropy::nonparametric, src/entropy.cpp:617-1225;
// include/entropy.hpp:101-194), directed and not degree-corrected, as device functions.
//
// fastlgamma(x) is glibc lgamma(double(x)) (+inf at 0; include/fastlgamma.hpp:19-27, src/fastlgamma.cpp).
// The reference caches it for small x; here the cache covers every argument the path can produce (cell
// values and block sizes, so up to max(E, N) + 1) and is filled on the host with the library's lgamma, so
// device and host terms are the same numbers.  The description-length terms that depend only on the number
// of non-empty blocks (lbinom(N - 1, B - 1), edges_dl(B, E)) are constants of one launch, evaluated on the
// host for the block count before and after a block is vacated or occupied.
#pragma once
#include <stdint.h>

#include "device_common.hpp"

namespace sbpdev {

struct NpConst {
    const double *lg;  // lg[x] = fastlgamma(x), x < n
    int64_t n;
    double lb[3];   // lbinom(N - 1, nonempty - 2 + i): i = 0 (a block vacated), 1 (unchanged), 2 (one occupied)
    double edl[3];  // edges_dl(nonempty - 1 + i, E), same indexing
};

// eterm_exact (include/entropy.hpp:103-112), directed
__device__ __forceinline__ double np_eterm(const NpConst &c, long long w) { return -c.lg[w + 1]; }

// eterm(M + d) - eterm(M) for one Delta entry (entries_dS, src/entropy.cpp:771-798)
__device__ __forceinline__ double np_entry(const NpConst &c, long long M, long long d) {
    return np_eterm(c, M + d) - np_eterm(c, M);
}

// vterm_exact (include/entropy.hpp:114-135), directed, not degree-corrected
__device__ __forceinline__ double np_vterm(long long d_out, long long d_in, long long w) {
    if (w == 0) return 0.0;
    return (double)(d_out + d_in) * fastlog(w);
}

// get_delta_partition_dl + get_delta_deg_dl (0) + get_delta_edges_dl (src/entropy.cpp:865-1132) for moving
// `weight` vertices from a block of size size_r to one of size size_s, in the reference's operation order.
__device__ __forceinline__ double np_delta_dl(const NpConst &c, long long size_r, long long size_s, long long weight) {
    double S_b = 0, S_a = 0;
    S_b += -c.lg[size_r + 1];
    S_a += -c.lg[size_r - weight + 1];
    S_b += -c.lg[size_s + 1];
    S_a += -c.lg[size_s + weight + 1];
    int dB = 0;
    if (size_r == weight) dB--;
    if (size_s == 0) dB++;
    if (dB != 0) {
        S_b += c.lb[1];
        S_a += c.lb[1 + dB];
    }
    double dpart = 0;
    dpart += S_a - S_b;
    double eb = 0, ea = 0;
    if (dB != 0) {
        eb += c.edl[1];
        ea += c.edl[1 + dB];
    }
    const double dedges = ea - eb;
    double dS_dl = 0;
    dS_dl += dpart;
    dS_dl += 0.0;  // get_delta_deg_dl: 0 without degree correction
    dS_dl += dedges;
    return dS_dl;
}

// virtual_move_sparse's block terms (src/entropy.cpp:836-859) added to the entries sum, then the
// description-length deltas: nonparametric::delta_mdl (weight 1) / block_merge_delta_mdl (weight = size(r))
__device__ __forceinline__ double np_finish(const NpConst &c, double entries, long long dout_r, long long din_r,
                                            long long size_r, long long dout_s, long long din_s, long long size_s,
                                            long long k_out, long long k_in, long long weight) {
    double dS = entries;
    dS += np_vterm(dout_r - k_out, din_r - k_in, size_r - weight);
    dS -= np_vterm(dout_r, din_r, size_r);
    dS += np_vterm(dout_s + k_out, din_s + k_in, size_s + weight);
    dS -= np_vterm(dout_s, din_s, size_s);
    const double dS_dl = np_delta_dl(c, size_r, size_s, weight);
    return dS + 1.0 * dS_dl;  // BETA_DL = 1 (include/entropy.hpp:14)
}

}  // namespace sbpdev
