// Nonparametric description-length terms (entropy::nonparametric, src/entropy.cpp:617-1225;
// include/entropy.hpp:101-194), directed and not degree-corrected, as device functions.
//
// fastlgamma(x) is glibc lgamma(double(x)) (+inf at 0; include/fastlgamma.hpp:19-27, src/fastlgamma.cpp).
// The reference caches it for small x; here the cache covers every argument the path can produce (cell
// values and block sizes, so up to max(E, N) + 1) and is filled on the host with the library's lgamma, so
// device and host terms are the same numbers.  The description-length terms that depend only on the number
// of non-empty blocks (lbinom(N - 1, B - 1), edges_dl(B, E)) are constants of one launch, evaluated on the
// host for the block count before and after a block is vacated or occupied.
//
// Degree correction (args.degreecorrected): vterm becomes lgamma(d_out + 1) + lgamma(d_in + 1), and the
// degree description length adds log_q (src/entropy.cpp:688-717, with get_v and the Cephes dilogarithm
// spence, src/spence.cpp) and the per-block vertex-degree histograms.  log_q runs on the device with the
// device exp / log / log1p / sqrt, which may differ from glibc in the last place: these terms agree with the
// host to ~1e-15 relative, not bit for bit.  A histogram is a hash table keyed by (block, direction,
// degree) with int32 counts that may go below zero (see the Metropolis-Hastings update).
#pragma once
#include <stdint.h>

#include "device_common.hpp"

namespace sbpdev {

// Degree histograms: slot {key, count}; key = block << 33 | direction << 32 | degree (direction 1 = in).
struct __align__(16) HistSlot {
    unsigned long long key;
    int val;
    int pad;
};
constexpr unsigned long long HIST_EMPTY = ~0ull;
__host__ __device__ inline unsigned long long hist_key(long long block, int dir, long long degree) {
    return ((unsigned long long)block << 33) | ((unsigned long long)dir << 32) | (unsigned long long)(uint32_t)degree;
}
__device__ __forceinline__ uint32_t hist_home(unsigned long long key, int shift) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> shift);
}
// A table has mask + 1 slots followed by one bookkeeping slot: its val counts the claimed slots (entries are
// never removed; the host compacts the table when half of it is claimed) and its pad flags a full table.
// Probing is bounded by the table size, so a full table ends the probe and sets the flag instead of looping.
// Atomic loads: the Metropolis-Hastings kernel reads entries it has just updated, and a wave-uniform address
// must not be served from the scalar cache.
__device__ __forceinline__ int hist_get(const HistSlot *t, int shift, uint32_t mask, unsigned long long key) {
    uint32_t i = hist_home(key, shift);
    for (uint32_t n = 0; n <= mask; ++n, i = (i + 1) & mask) {
        const unsigned long long k = __hip_atomic_load(&t[i].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == key) return __hip_atomic_load(&t[i].val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == HIST_EMPTY) return 0;
    }
    return 0;  // not reachable while the table has an empty slot
}
__device__ __forceinline__ void hist_add(HistSlot *t, int shift, uint32_t mask, unsigned long long key, int delta) {
    uint32_t i = hist_home(key, shift);
    for (uint32_t n = 0; n <= mask; ++n, i = (i + 1) & mask) {
        unsigned long long k = t[i].key;  // a stale EMPTY is settled by the CAS; keys are never removed
        if (k == HIST_EMPTY) {
            k = atomicCAS(&t[i].key, HIST_EMPTY, key);
            if (k == HIST_EMPTY) {
                atomicAdd(&t[mask + 1].val, 1);
                atomicAdd(&t[i].val, delta);
                return;
            }
        }
        if (k == key) {
            atomicAdd(&t[i].val, delta);
            return;
        }
    }
    atomicOr(&t[mask + 1].pad, 1);
}

struct NpConst {
    const double *lg;  // lg[x] = fastlgamma(x), x < n
    int64_t n;
    double lb[3];   // lbinom(N - 1, nonempty - 2 + i): i = 0 (a block vacated), 1 (unchanged), 2 (one occupied)
    double edl[3];  // edges_dl(nonempty - 1 + i, E), same indexing
    int dc;         // degree-corrected
    const HistSlot *hist;  // dc: degree histograms
    int hist_shift;
    uint32_t hist_mask;
    double ln2, lnpi;  // log(2), log(M_PI) from the host library
};

// fastlgamma of a size_t argument: a negative int or long converts to 2^64 + x (the reference then takes
// lgamma of that); 0 gives +inf from the table.
__device__ __forceinline__ double np_flg(const NpConst &c, long long x) {
    if (x < 0) return lgamma((double)(unsigned long long)x);
    if (x < c.n) return c.lg[x];
    return lgamma((double)x);
}
// fastlbinom (include/entropy.hpp:145-149)
__device__ __forceinline__ double np_lbinom(const NpConst &c, long long n, long long k) {
    if (n == 0 || k == 0 || k > n) return 0;
    return ((np_flg(c, n + 1) - np_flg(c, k + 1)) - np_flg(c, n - k + 1));
}
// spence (src/spence.cpp:94-155): Cephes dilogarithm with its published rational coefficients
__device__ inline double dc_spence(double x) {
    const double A[8] = {4.65128586073990045278E-5, 7.31589045238094711071E-3, 1.33847639578309018650E-1,
                         8.79691311754530315341E-1, 2.71149851196553469920E0,  4.25697156008121755724E0,
                         3.29771340985225106936E0,  1.00000000000000000126E0};
    const double Bc[8] = {6.90990488912553276999E-4, 2.54043763932544379113E-2, 2.82974860602568089943E-1,
                          1.41172597751831069617E0,  3.63800533345137075418E0,  5.03278880143316990390E0,
                          3.54771340985225096217E0,  9.99999999999999998740E-1};
    const double PI2_6 = 3.14159265358979323846 * 3.14159265358979323846 / 6.0;
    if (x < 0.0) return __builtin_nan("");
    if (x == 1.0) return 0.0;
    if (x == 0.0) return PI2_6;
    int flag = 0;
    double w;
    if (x > 2.0) {
        x = 1.0 / x;
        flag |= 2;
    }
    if (x > 1.5) {
        w = (1.0 / x) - 1.0;
        flag |= 2;
    } else if (x < 0.5) {
        w = -x;
        flag |= 1;
    } else {
        w = x - 1.0;
    }
    double pa = A[0], pb = Bc[0];
    for (int i = 1; i < 8; ++i) {
        pa = pa * w + A[i];
        pb = pb * w + Bc[i];
    }
    double y = -w * pa / pb;
    if (flag & 1) y = PI2_6 - log(x) * log1p(-x) - y;
    if (flag & 2) {
        const double z = log(x);
        y = -0.5 * z * z - y;
    }
    return y;
}
// get_v (src/entropy.cpp:676-686).  The fixed point is reached in a few dozen steps; the bound only keeps a
// kernel finite.
__device__ inline double dc_get_v(double u) {
    double v = u, delta = 1;
    for (int it = 0; delta > 1e-8 && it < 100000; ++it) {
        const double n_v = u * sqrt(dc_spence(exp(-v)));
        delta = fabs(n_v - v);
        v = n_v;
    }
    return v;
}
// log_q (src/entropy.cpp:688-717) with size_t arguments.  k < pow(n, 1/4) is tested as k^4 < n (pow is exact
// at fourth powers, so the two agree for all integers).
__device__ inline double dc_log_q(const NpConst &c, unsigned long long n, unsigned long long k) {
    if (n == 0 || k < 1) return 0;
    if (k > n) k = n;
    if (k < 65536 && k * k * k * k < n) return np_lbinom(c, (long long)(n - 1), (long long)(k - 1)) - np_flg(c, (long long)(k + 1));
    const double u = (double)k / sqrt((double)n);
    const double v = dc_get_v(u);
    const double lf = log(v) - log1p(-exp(-v) * (1 + u * u / 2)) / 2 - c.ln2 * 3 / 2. - log(u) - c.lnpi;
    const double g = 2 * v / u - u * log1p(-exp(-v));
    return lf - log((double)n) + sqrt((double)n) * g;
}
// get_delta_deg_dl_dist_change (src/entropy.cpp:969-1060), directed: the degree description-length change of
// `block` when vweight vertices of in-/out-degree (vkin, vkout) leave it (diff -1) or join it (+1), with the
// reference's int and size_t conversions.
__device__ inline double dc_change(const NpConst &c, int block, long long size, long long dout, long long din,
                                   long long vkin, long long vkout, long long vweight, int diff) {
    const long long total_r = size, ep_r = dout, em_r = din;
    const int tkin = (int)vkin, tkout = (int)vkout, n = (int)vweight;
    const int nd_in = hist_get(c.hist, c.hist_shift, c.hist_mask, hist_key(block, 1, vkin));
    const int nd_out = hist_get(c.hist, c.hist_shift, c.hist_mask, hist_key(block, 0, vkout));
    auto sk = [&](int delta) {
        return (0.0 - np_flg(c, (long long)(nd_in + delta + 1))) - np_flg(c, (long long)(nd_out + delta + 1));
    };
    auto se = [&](long long delta, long long kin, long long kout) {
        return (0.0 + dc_log_q(c, (unsigned long long)(em_r + kin), (unsigned long long)(total_r + delta))) +
               dc_log_q(c, (unsigned long long)(ep_r + kout), (unsigned long long)(total_r + delta));
    };
    auto sr = [&](long long delta) { return 2 * np_flg(c, total_r + delta + 1); };
    double S_b = 0, S_a = 0;
    S_b += sk(0);
    S_a += sk((int)(diff * vweight));
    S_b += se(0, 0, 0);
    S_a += se(diff * n, diff * tkin, diff * tkout);
    S_b += sr(0);
    S_a += sr(diff * n);
    return S_a - S_b;
}

// eterm_exact (include/entropy.hpp:103-112), directed
__device__ __forceinline__ double np_eterm(const NpConst &c, long long w) { return -c.lg[w + 1]; }

// eterm(M + d) - eterm(M) for one Delta entry (entries_dS, src/entropy.cpp:771-798)
__device__ __forceinline__ double np_entry(const NpConst &c, long long M, long long d) {
    return np_eterm(c, M + d) - np_eterm(c, M);
}

// vterm_exact (include/entropy.hpp:114-135), directed
__device__ __forceinline__ double np_vterm(const NpConst &c, long long d_out, long long d_in, long long w) {
    if (c.dc) return np_flg(c, d_out + 1) + np_flg(c, d_in + 1);
    if (w == 0) return 0.0;
    return (double)(d_out + d_in) * fastlog(w);
}

// get_delta_partition_dl + get_delta_deg_dl (0) + get_delta_edges_dl (src/entropy.cpp:865-1132) for moving
// `weight` vertices from a block of size size_r to one of size size_s, in the reference's operation order.
__device__ __forceinline__ double np_delta_dl(const NpConst &c, long long size_r, long long size_s, long long weight,
                                              double ddeg) {
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
    dS_dl += ddeg;  // get_delta_deg_dl: 0 without degree correction
    dS_dl += dedges;
    return dS_dl;
}

// virtual_move_sparse's block terms (src/entropy.cpp:836-859) added to the entries sum, then the
// description-length deltas: nonparametric::delta_mdl (weight 1; vkin / vkout the vertex's in- and out-degree)
// and block_merge_delta_mdl (weight = size(r); vkin / vkout block r's degrees, src/entropy.cpp:1170-1173).
__device__ __forceinline__ double np_finish(const NpConst &c, double entries, int r, int s, long long dout_r,
                                            long long din_r, long long size_r, long long dout_s, long long din_s,
                                            long long size_s, long long k_out, long long k_in, long long weight,
                                            long long vkin, long long vkout) {
    double dS = entries;
    dS += np_vterm(c, dout_r - k_out, din_r - k_in, size_r - weight);
    dS -= np_vterm(c, dout_r, din_r, size_r);
    dS += np_vterm(c, dout_s + k_out, din_s + k_in, size_s + weight);
    dS -= np_vterm(c, dout_s, din_s, size_s);
    double ddeg = 0.0;
    if (c.dc && r != s) {
        ddeg += dc_change(c, r, size_r, dout_r, din_r, vkin, vkout, weight, -1);
        ddeg += dc_change(c, s, size_s, dout_s, din_s, vkin, vkout, weight, +1);
    }
    const double dS_dl = np_delta_dl(c, size_r, size_s, weight, ddeg);
    return dS + 1.0 * dS_dl;  // BETA_DL = 1 (include/entropy.hpp:14)
}

}  // namespace sbpdev
