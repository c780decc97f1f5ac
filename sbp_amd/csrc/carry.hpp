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
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <queue>
#include <thread>
#include <vector>

namespace sbpcarry {

// Linear-probing map int32 -> int32 (block id -> cell count) with backward-shift deletion.  Keys and
// counts share one slot array so a probe touches one cache line; the slot index is the top log2(cap) bits
// of a multiplicative hash.
struct FlatMap {
    struct Slot {
        int32_t key;  // -1 = empty
        int32_t val;
    };
    Slot *s = nullptr;
    uint32_t cap = 0, n = 0;
    int shift = 32;

    FlatMap() = default;
    FlatMap(const FlatMap &) = delete;
    FlatMap &operator=(const FlatMap &) = delete;
    FlatMap(FlatMap &&o) noexcept : s(o.s), cap(o.cap), n(o.n), shift(o.shift) { o.s = nullptr, o.cap = o.n = 0; }
    FlatMap &operator=(FlatMap &&o) noexcept {
        std::swap(s, o.s);
        std::swap(cap, o.cap);
        std::swap(n, o.n);
        std::swap(shift, o.shift);
        return *this;
    }
    ~FlatMap() { free(s); }

    uint32_t slot(int32_t k) const { return ((uint32_t)k * 0x9E3779B1u) >> shift; }
    void init(uint32_t expected) {
        uint32_t c = 4;
        int lg = 2;
        while (c < 2 * expected + 2) c <<= 1, ++lg;
        free(s);
        s = static_cast<Slot *>(malloc(sizeof(Slot) * c));
        for (uint32_t i = 0; i < c; ++i) s[i] = Slot{-1, 0};
        cap = c;
        shift = 32 - lg;
        n = 0;
    }
    int64_t get(int32_t k) const {
        if (!cap) return 0;
        const uint32_t m = cap - 1;
        for (uint32_t i = slot(k);; i = (i + 1) & m) {
            if (s[i].key == k) return s[i].val;
            if (s[i].key == -1) return 0;
        }
    }
    void grow() {
        Slot *old = s;
        const uint32_t oc = cap;
        s = nullptr;
        init(n * 2 + 4);
        for (uint32_t i = 0; i < oc; ++i)
            if (old[i].key != -1) add(old[i].key, old[i].val);
        free(old);
    }
    void erase_at(uint32_t i) {
        const uint32_t m = cap - 1;
        s[i] = Slot{-1, 0};
        n--;
        for (uint32_t j = (i + 1) & m; s[j].key != -1; j = (j + 1) & m) {
            // backward shift: re-place entries whose probe path crossed i
            const uint32_t home = slot(s[j].key);
            const bool move = (j > i) ? (home <= i || home > j) : (home <= i && home > j);
            if (move) {
                s[i] = s[j];
                s[j] = Slot{-1, 0};
                i = j;
            }
        }
    }
    // adds d; erases the entry when it reaches 0
    void add(int32_t k, int64_t d) {
        if (d == 0) return;
        if (!cap) init(4);
        const uint32_t m = cap - 1;
        uint32_t i = slot(k);
        for (; s[i].key != -1; i = (i + 1) & m) {
            if (s[i].key == k) {
                s[i].val += (int32_t)d;
                if (s[i].val == 0) erase_at(i);
                return;
            }
        }
        s[i] = Slot{k, (int32_t)d};
        n++;
        if (2 * n + 2 > cap) grow();
    }
    void erase(int32_t k) {
        if (!cap) return;
        const uint32_t m = cap - 1;
        for (uint32_t i = slot(k); s[i].key != -1; i = (i + 1) & m)
            if (s[i].key == k) {
                erase_at(i);
                return;
            }
    }
    template <typename F>
    void each(F f) const {
        for (uint32_t i = 0; i < cap; ++i)
            if (s[i].key != -1) f(s[i].key, (int64_t)s[i].val);
    }
    void clear() { free(detach()); }
    // empties the map and hands over its slot array
    Slot *detach() {
        Slot *p = s;
        s = nullptr;
        cap = n = 0;
        shift = 32;
        return p;
    }
    // room for `expected` keys without growing
    void reserve(uint32_t expected) {
        if (2 * expected + 2 <= cap) return;
        Slot *old = s;
        const uint32_t oc = cap;
        s = nullptr;
        init(expected);
        for (uint32_t i = 0; i < oc; ++i)
            if (old[i].key != -1) add(old[i].key, old[i].val);
        free(old);
    }
};

// Slot arrays of merged-away blocks, freed with the model instead of inside the merge loop.
struct Retired {
    std::vector<void *> p;
    Retired() = default;
    Retired(Retired &&) = default;
    Retired &operator=(Retired &&o) {
        std::swap(p, o.p);
        return *this;
    }
    ~Retired() {
        for (void *x : p) free(x);
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
    static constexpr int64_t SIZE = 1 << 16;
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

// fastlgamma (include/fastlgamma.hpp:19-27): glibc lgamma(double(x)), +inf at 0; tabulated for small x with
// the same function
struct LgammaTable {
    static constexpr int64_t SIZE = 1 << 20;
    std::vector<double> v;
    LgammaTable() : v(SIZE) {
        for (int64_t i = 0; i < SIZE; ++i) v[i] = raw(i);
    }
    static double raw(int64_t x) {
        if (x <= 0) return INFINITY;
        int sign;
        return lgamma_r((double)x, &sign);
    }
};
inline const LgammaTable &lgamma_table() {
    static const LgammaTable t;
    return t;
}
inline double flgamma(int64_t x) { return x < LgammaTable::SIZE ? lgamma_table().v[x] : LgammaTable::raw(x); }
inline double flbinom(int64_t n, int64_t k) {
    if (n == 0 || k == 0 || k > n) return 0;
    return ((flgamma(n + 1) - flgamma(k + 1)) - flgamma(n - k + 1));
}
inline double np_eterm(int64_t w) { return -flgamma(w + 1); }
inline double np_entry(int64_t M, int64_t d) { return np_eterm(M + d) - np_eterm(M); }
inline double np_vterm(int64_t d_out, int64_t d_in, int64_t w) {
    if (w == 0) return 0.0;
    return (double)(d_out + d_in) * fastlog(w);
}
inline double np_edges_dl(int64_t B, int64_t E) { return flbinom(B * B + E - 1, E); }

// Degree correction (args.degreecorrected), with the host library's math as the reference uses it.
// fastlgamma of a size_t argument: a negative int or long converts to 2^64 + x; 0 gives +inf.
inline double flg_size(int64_t x) {
    if (x < 0) {
        int sign;
        return lgamma_r((double)(uint64_t)x, &sign);
    }
    return flgamma(x);
}
inline double np_vterm_dc(int64_t d_out, int64_t d_in) { return flg_size(d_out + 1) + flg_size(d_in + 1); }
// spence (src/spence.cpp:94-155): the Cephes dilogarithm with its published rational coefficients
inline double dc_spence(double x) {
    static const double A[8] = {4.65128586073990045278E-5, 7.31589045238094711071E-3, 1.33847639578309018650E-1,
                                8.79691311754530315341E-1, 2.71149851196553469920E0,  4.25697156008121755724E0,
                                3.29771340985225106936E0,  1.00000000000000000126E0};
    static const double Bc[8] = {6.90990488912553276999E-4, 2.54043763932544379113E-2, 2.82974860602568089943E-1,
                                 1.41172597751831069617E0,  3.63800533345137075418E0,  5.03278880143316990390E0,
                                 3.54771340985225096217E0,  9.99999999999999998740E-1};
    if (x < 0.0) return NAN;
    if (x == 1.0) return 0.0;
    if (x == 0.0) return M_PI * M_PI / 6.0;
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
    if (flag & 1) y = (M_PI * M_PI) / 6.0 - log(x) * log1p(-x) - y;
    if (flag & 2) {
        const double z = log(x);
        y = -0.5 * z * z - y;
    }
    return y;
}
// log_q (src/entropy.cpp:676-717) with size_t arguments
inline double dc_log_q(uint64_t n, uint64_t k) {
    if (n <= 0 || k < 1) return 0;
    if (k > n) k = n;
    if ((double)k < pow((double)n, 1 / 4.)) return flbinom((int64_t)(n - 1), (int64_t)(k - 1)) - flg_size((int64_t)(k + 1));
    const double u = (double)k / sqrt((double)n);
    double v = u, delta = 1;
    while (delta > 1e-8) {
        const double n_v = u * sqrt(dc_spence(exp(-v)));
        delta = fabs(n_v - v);
        v = n_v;
    }
    const double lf = log(v) - log1p(-exp(-v) * (1 + u * u / 2)) / 2 - log(2) * 3 / 2. - log(u) - log(M_PI);
    const double g = 2 * v / u - u * log1p(-exp(-v));
    return lf - log((double)n) + sqrt((double)n) * g;
}
// get_delta_deg_dl_dist_change (src/entropy.cpp:969-1060), directed; nd_in / nd_out are block's histogram
// counts at degrees vkin / vkout.
inline double dc_change(int64_t size, int64_t dout, int64_t din, int nd_in, int nd_out, int64_t vkin, int64_t vkout,
                        int64_t vweight, int diff) {
    const int64_t total_r = size, ep_r = dout, em_r = din;
    const int tkin = (int)vkin, tkout = (int)vkout, n = (int)vweight;
    auto sk = [&](int delta) {
        return (0.0 - flg_size((int64_t)(nd_in + delta + 1))) - flg_size((int64_t)(nd_out + delta + 1));
    };
    auto se = [&](int64_t delta, int64_t kin, int64_t kout) {
        return (0.0 + dc_log_q((uint64_t)(em_r + kin), (uint64_t)(total_r + delta))) +
               dc_log_q((uint64_t)(ep_r + kout), (uint64_t)(total_r + delta));
    };
    auto sr = [&](int64_t delta) { return 2 * flg_size(total_r + delta + 1); };
    double S_b = 0, S_a = 0;
    S_b += sk(0);
    S_a += sk((int)(diff * vweight));
    S_b += se(0, 0, 0);
    S_a += se(diff * n, diff * tkin, diff * tkout);
    S_b += sr(0);
    S_a += sr(diff * n);
    return S_a - S_b;
}

// The block terms and description-length deltas of nonparametric::block_merge_delta_mdl
// (src/entropy.cpp:1183-1206; virtual_move_sparse :836-859, get_delta_partition_dl :865-967,
// get_delta_edges_dl :1098-1132) around the Delta entries' sum, for `nonempty` non-empty blocks -- the same
// operations in the same order as the device kernels.
inline double np_merge_finish(double entries, int64_t dout_r, int64_t din_r, int64_t size_r, int64_t dout_s,
                              int64_t din_s, int64_t size_s, int64_t N, int64_t E, int64_t nonempty, bool dc = false,
                              double ddeg = 0.0) {
    const int64_t k_out = dout_r, k_in = din_r, weight = size_r;
    auto vt = [dc](int64_t o, int64_t i, int64_t w) { return dc ? np_vterm_dc(o, i) : np_vterm(o, i, w); };
    double dS = entries;
    dS += vt(dout_r - k_out, din_r - k_in, size_r - weight);
    dS -= vt(dout_r, din_r, size_r);
    dS += vt(dout_s + k_out, din_s + k_in, size_s + weight);
    dS -= vt(dout_s, din_s, size_s);
    double S_b = 0, S_a = 0;
    S_b += -flgamma(size_r + 1);
    S_a += -flgamma(size_r - weight + 1);
    S_b += -flgamma(size_s + 1);
    S_a += -flgamma(size_s + weight + 1);
    int dB = 0;
    if (size_r == weight) dB--;
    if (size_s == 0) dB++;
    if (dB != 0) {
        S_b += flbinom(N - 1, nonempty - 1);
        S_a += flbinom(N - 1, nonempty + dB - 1);
    }
    double dpart = 0;
    dpart += S_a - S_b;
    double eb = 0, ea = 0;
    if (dB != 0) {
        eb += np_edges_dl(nonempty, E);
        ea += np_edges_dl(nonempty + dB, E);
    }
    const double dedges = ea - eb;
    double dS_dl = 0;
    dS_dl += dpart;
    dS_dl += ddeg;  // get_delta_deg_dl: 0 without degree correction
    dS_dl += dedges;
    return dS + 1.0 * dS_dl;
}

// Fork-join helper for the speculative evaluations: fn(i) for i in [0, n) on up to `threads` threads.  The
// carry-out forks every few hundred microseconds, so idle workers spin on the generation counter for about
// a millisecond before they sleep on the condition variable; a fork then costs a cache-line transfer, not
// a futex wake-up per worker.  A worker takes items of a generation only while its run is open, and the run
// closes only after the workers inside it have left, so a slow waker can never run a later generation's
// indices with an earlier task.
class Pool {
  public:
    explicit Pool(int threads, int spin_us = spin_default()) : nthreads_(std::max(1, threads)), spin_us_(spin_us) {
        for (int t = 1; t < nthreads_; ++t) workers_.emplace_back([this] { work(); });
    }
    ~Pool() {
        stop_.store(true);
        gen_.fetch_add(1);
        {
            std::lock_guard<std::mutex> lk(mu_);
            cv_.notify_all();
        }
        for (auto &w : workers_) w.join();
    }
    int threads() const { return nthreads_; }
    // SBP_AMD_CARRY_SPIN_US (0..100000) when set, else 1000
    static int spin_default() {
        if (const char *e = getenv("SBP_AMD_CARRY_SPIN_US")) {
            const int v = atoi(e);
            if (v >= 0 && v <= 100000) return v;
        }
        return 1000;
    }
    template <typename F>
    void run(int64_t n, F fn) {
        if (n <= 0) return;
        if (nthreads_ == 1 || n < 2) {
            for (int64_t i = 0; i < n; ++i) fn(i);
            return;
        }
        std::function<void(int64_t)> f = fn;
        task_ = &f;
        n_ = n;
        next_.store(0, std::memory_order_relaxed);
        done_.store(0, std::memory_order_relaxed);
        const uint64_t g = gen_.load(std::memory_order_relaxed) + 1;
        open_.store(g);
        gen_.store(g);  // publishes the task (sequentially consistent, paired with sleepers_ below)
        if (sleepers_.load() > 0) {
            std::lock_guard<std::mutex> lk(mu_);
            cv_.notify_all();
        }
        // The caller works too, and waits for the items rather than for every worker: a worker that is asleep
        // or descheduled does not hold up a short run.  Closing the run then waits only for the workers inside
        // it; one that enters after the close sees it closed and never touches the task.
        drain(f);
        while (done_.load(std::memory_order_acquire) != n) relax();
        open_.store(0);
        while (inside_.load() != 0) relax();
        task_ = nullptr;
    }

  private:
    static void relax() {
#if defined(__x86_64__)
        __builtin_ia32_pause();
#else
        std::this_thread::yield();
#endif
    }
    void drain(const std::function<void(int64_t)> &f) {
        for (int64_t i; (i = next_.fetch_add(1, std::memory_order_relaxed)) < n_;) {
            f(i);
            done_.fetch_add(1, std::memory_order_release);
        }
    }
    void work() {
        uint64_t seen = 0;
        while (true) {
            uint64_t g;
            auto t0 = std::chrono::steady_clock::now();
            for (uint32_t spins = 0; (g = gen_.load(std::memory_order_acquire)) == seen; ++spins) {
                relax();
                if ((spins & 255) != 255 || std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(spin_us_))
                    continue;
                std::unique_lock<std::mutex> lk(mu_);
                sleepers_.fetch_add(1);
                cv_.wait(lk, [&] { return gen_.load() != seen; });
                sleepers_.fetch_sub(1);
                t0 = std::chrono::steady_clock::now();
            }
            seen = g;
            if (stop_.load()) return;
            inside_.fetch_add(1);  // sequentially consistent with the caller's close of the run
            if (open_.load() == g) drain(*task_);
            inside_.fetch_sub(1, std::memory_order_release);
        }
    }
    int nthreads_, spin_us_;
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::atomic<uint64_t> gen_{0}, open_{0};  // open_: generation of the run that still takes workers, 0 if none
    std::atomic<int> sleepers_{0}, inside_{0};
    std::atomic<int64_t> done_{0};
    std::atomic<bool> stop_{false};
    const std::function<void(int64_t)> *task_ = nullptr;
    int64_t n_ = 0;
    std::atomic<int64_t> next_{0};
};

// Worker threads of the carry-out: SBP_AMD_CARRY_THREADS when set (1..64), else the hardware threads up to 16.
inline int carry_threads() {
    if (const char *e = getenv("SBP_AMD_CARRY_THREADS")) {
        const int t = atoi(e);
        if (t >= 1 && t <= 64) return t;
    }
    unsigned hw = std::thread::hardware_concurrency();
    return (int)std::min<unsigned>(16, hw ? hw : 1);
}

struct Model {
    int32_t B = 0;
    std::vector<FlatMap> row, col;
    Retired retired;
    std::vector<int64_t> d_out, d_in;
    // nonparametric entropy model: block sizes, vertices, edges and the non-empty block count
    bool np = false;
    std::vector<int64_t> size;
    int64_t N = 0, E = 0, nonempty = 0;
    // degree-corrected: per block, vertex in-/out-degree -> count (histograms), from (key, count) slots with
    // key = block << 33 | direction << 32 | degree (direction 1 = in; empty slots have key ~0)
    bool dc = false;
    std::vector<FlatMap> hin, hout;
    void set_degree_corrected(const uint64_t *keys, const int32_t *counts, int64_t n) {
        dc = true;
        hin.clear();
        hout.clear();
        hin.resize(B);
        hout.resize(B);
        for (int64_t i = 0; i < n; ++i) {
            if (keys[i] == ~0ull || counts[i] == 0) continue;
            const int32_t b = (int32_t)(keys[i] >> 33), deg = (int32_t)(uint32_t)keys[i];
            ((keys[i] >> 32) & 1 ? hin : hout)[b].add(deg, counts[i]);
        }
    }
    void set_nonparametric(const int32_t *sizes, int64_t num_vertices) {
        np = true;
        size.assign(sizes, sizes + B);
        N = num_vertices;
        E = 0;
        nonempty = 0;
        for (int32_t b = 0; b < B; ++b) {
            E += d_out[b];
            nonempty += size[b] > 0;
        }
    }

    // row-major cells (col, count) per row and column-major cells (row, count) per column, as separate
    // index/count arrays (stride 1) or interleaved (index, count) pairs (stride 2)
    void load(int32_t nb, const int32_t *row_ptr, const int32_t *row_idx, const int32_t *row_cnt,
              const int32_t *col_ptr, const int32_t *col_idx, const int32_t *col_cnt, int stride,
              const int32_t *dout, const int32_t *din, Pool *pool = nullptr) {
        B = nb;
        row.clear();
        col.clear();
        row.resize(B);
        col.resize(B);
        d_out.assign(dout, dout + B);
        d_in.assign(din, din + B);
        auto fill = [&](int64_t chunk) {
            const int32_t b0 = (int32_t)(chunk * 1024), b1 = std::min<int32_t>(B, b0 + 1024);
            for (int32_t b = b0; b < b1; ++b) {
                row[b].init(row_ptr[b + 1] - row_ptr[b]);
                for (int64_t i = row_ptr[b]; i < row_ptr[b + 1]; ++i) row[b].add(row_idx[stride * i], row_cnt[stride * i]);
                col[b].init(col_ptr[b + 1] - col_ptr[b]);
                for (int64_t i = col_ptr[b]; i < col_ptr[b + 1]; ++i) col[b].add(col_idx[stride * i], col_cnt[stride * i]);
            }
        };
        const int64_t chunks = ((int64_t)B + 1023) / 1024;
        if (pool) pool->run(chunks, fill);
        else for (int64_t c = 0; c < chunks; ++c) fill(c);
    }
    int64_t get(int32_t r, int32_t c) const { return row[r].n <= col[c].n ? row[r].get(c) : col[c].get(r); }
};

// Copies the occupied slots of `f` into `out` (no branch per slot); returns their number.
inline size_t gather(const FlatMap &f, std::vector<FlatMap::Slot> &out) {
    if (out.size() < f.cap) out.resize(f.cap);
    size_t k = 0;
    FlatMap::Slot *o = out.data();
    for (uint32_t i = 0; i < f.cap; ++i) {
        o[k] = f.s[i];
        k += f.s[i].key >= 0;
    }
    return k;
}

// Entries of row r + column r at or above which one evaluation, and renames at or above which one merge,
// is split over the pool, and the entries per piece.  SBP_AMD_CARRY_SPLIT="eval_min,eval_piece,merge_min,
// merge_piece" overrides them (all >= 1).
struct SplitSizes {
    size_t eval_min = 512, eval_piece = 128, merge_min = 128, merge_piece = 64;
    static SplitSizes from_env() {
        SplitSizes s;
        if (const char *e = getenv("SBP_AMD_CARRY_SPLIT")) {
            unsigned long a, b, c, d;
            if (sscanf(e, "%lu,%lu,%lu,%lu", &a, &b, &c, &d) == 4 && a && b && c && d) s = SplitSizes{a, b, c, d};
        }
        return s;
    }
};

// entropy::block_merge_delta_mdl (src/entropy.cpp:129-182) for merging r into s on the current state:
// Delta entries term by term, (s, s) when unchanged, the other unchanged cells of row s / column s in bulk
// (M * (fastlog(d') - fastlog(d)) each) -- the same terms as the merge-proposal kernel.  With a pool, the
// entries of a long row r / column r are split into pieces whose fixed-point partial sums add exactly, so
// the result does not depend on the split.  Only the thread that owns `pool` may pass it.
inline double merge_dS(const Model &m, int32_t r, int32_t s, Pool *pool = nullptr,
                       const SplitSizes &split = SplitSizes()) {
    const float *L = logf_table().v.data();
    auto lg = [L](int64_t x) -> double {
        if (x < LogfTable::SIZE) return x <= 0 ? 0.0 : (double)L[x];
        return (double)logf((float)x);
    };
    const int64_t k_out = m.d_out[r], k_in = m.d_in[r];
    const int64_t din_s_new = m.d_in[s] + k_in, dout_s_new = m.d_out[s] + k_out;
    const double l_dout_r = lg(m.d_out[r]), l_dout_s = lg(m.d_out[s]), l_dout_sn = lg(dout_s_new);
    const double l_din_r = lg(m.d_in[r]), l_din_s = lg(m.d_in[s]), l_din_sn = lg(din_s_new);
    const FlatMap &Rs = m.row[s], &Cs = m.col[s];
    thread_local std::vector<FlatMap::Slot> row_buf, col_buf;
    const size_t nr = gather(m.row[r], row_buf), nc = gather(m.col[r], col_buf);
    const FlatMap::Slot *er = row_buf.data(), *ec = col_buf.data();  // the caller's buffers, also in workers
    auto row_part = [&](size_t lo, size_t hi, Fx128 &acc, int64_t &sub) {
        for (size_t i = lo; i < hi; ++i) {
            if (i + 8 < hi) {
                __builtin_prefetch(&m.d_in[er[i + 8].key]);
                __builtin_prefetch(&Rs.s[Rs.slot(er[i + 8].key)]);
            }
            const int32_t c = er[i].key;
            const int64_t v = er[i].val;
            const double ld = lg(m.d_in[c]);
            acc.add((double)v * (lg(v) - (ld + l_dout_r)));
            if (c != r && c != s) {
                const int64_t msc = Rs.get(c);
                if (msc) acc.add((double)msc * (lg(msc) - (ld + l_dout_s)));
                acc.sub((double)(msc + v) * (lg(msc + v) - (ld + l_dout_sn)));
                sub += msc;
            }
        }
    };
    auto col_part = [&](size_t lo, size_t hi, Fx128 &acc, int64_t &sub) {
        for (size_t i = lo; i < hi; ++i) {
            if (i + 8 < hi) {
                __builtin_prefetch(&m.d_out[ec[i + 8].key]);
                __builtin_prefetch(&Cs.s[Cs.slot(ec[i + 8].key)]);
            }
            const int32_t x = ec[i].key;
            if (x == r) continue;
            const int64_t v = ec[i].val;
            const double ld = lg(m.d_out[x]);
            acc.add((double)v * (lg(v) - (l_din_r + ld)));
            if (x != s) {
                const int64_t mxs = Cs.get(x);
                if (mxs) acc.add((double)mxs * (lg(mxs) - (l_din_s + ld)));
                acc.sub((double)(mxs + v) * (lg(mxs + v) - (l_din_sn + ld)));
                sub += mxs;
            }
        }
    };
    Fx128 acc;
    int64_t sub_row = 0, sub_col = 0;
    if (pool && pool->threads() > 1 && nr + nc >= split.eval_min) {
        const size_t P = split.eval_piece;
        struct alignas(64) Part {
            Fx128 acc;
            int64_t sub;
        };
        const size_t kr = (nr + P - 1) / P, kc = (nc + P - 1) / P;
        thread_local std::vector<Part> part_buf;
        if (part_buf.size() < kr + kc) part_buf.resize(kr + kc);
        Part *parts = part_buf.data();
        pool->run((int64_t)(kr + kc), [&](int64_t i) {
            Part &p = parts[i];
            p.acc = Fx128();
            p.sub = 0;
            if ((size_t)i < kr) row_part(i * P, std::min(nr, (i + 1) * P), p.acc, p.sub);
            else {
                const size_t j = i - kr;
                col_part(j * P, std::min(nc, (j + 1) * P), p.acc, p.sub);
            }
        });
        for (size_t i = 0; i < kr + kc; ++i) {
            acc.v += parts[i].acc.v;
            (i < kr ? sub_row : sub_col) += parts[i].sub;
        }
    } else {
        row_part(0, nr, acc, sub_row);
        col_part(0, nc, acc, sub_col);
    }
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
    acc.add((double)S_row * (l_dout_sn - l_dout_s));
    acc.add((double)S_col * (l_din_sn - l_din_s));
    return acc.value();
}

// The Delta entries' eterm differences of nonparametric::block_merge_delta_mdl: row r and column r empty,
// (s, c) and (x, s) receive them, (s, s) receives (r, r) + (r, s) + (s, r); fixed-point sum.
inline double np_merge_entries(const Model &m, int32_t r, int32_t s) {
    Fx128 acc;
    m.row[r].each([&](int32_t c, int64_t v) {
        acc.add(np_entry(v, -v));
        if (c != r && c != s) acc.add(np_entry(m.row[s].get(c), v));
    });
    m.col[r].each([&](int32_t x, int64_t v) {
        if (x == r) return;
        acc.add(np_entry(v, -v));
        if (x != s) acc.add(np_entry(m.col[s].get(x), v));
    });
    const int64_t dss = m.get(r, r) + m.get(r, s) + m.get(s, r);
    if (dss > 0) acc.add(np_entry(m.get(s, s), dss));
    return acc.value();
}
// With degree correction the merge's degree terms take block r's degrees as the vertex degrees
// (src/entropy.cpp:1166-1181), looked up in the histograms of r and s.
inline double np_merge_dS(const Model &m, int32_t r, int32_t s, double entries) {
    if (m.size[r] == 0 || m.size[s] == 0) return 0.0;
    double ddeg = 0.0;
    if (m.dc && r != s) {
        const int64_t vkin = m.d_in[r], vkout = m.d_out[r], w = m.size[r];
        ddeg += dc_change(m.size[r], m.d_out[r], m.d_in[r], (int)m.hin[r].get((int32_t)vkin),
                          (int)m.hout[r].get((int32_t)vkout), vkin, vkout, w, -1);
        ddeg += dc_change(m.size[s], m.d_out[s], m.d_in[s], (int)m.hin[s].get((int32_t)vkin),
                          (int)m.hout[s].get((int32_t)vkout), vkin, vkout, w, +1);
    }
    return np_merge_finish(entries, m.d_out[r], m.d_in[r], m.size[r], m.d_out[s], m.d_in[s], m.size[s], m.N, m.E,
                           m.nonempty, m.dc, ddeg);
}

// Renames key `from` to `to` in the maps `maps[ent[i].key]`, i < n, which hold count ent[i].val under `from`,
// prefetching each map and then its two slot lines a few entries ahead: the maps are scattered over a model
// much larger than the caches.
inline void rename_in(std::vector<FlatMap> &maps, const FlatMap::Slot *ent, size_t n, int32_t from, int32_t to) {
    constexpr size_t D1 = 16, D2 = 6;
    for (size_t i = 0; i < std::min(n, D1); ++i) __builtin_prefetch(&maps[ent[i].key]);
    for (size_t i = 0; i < std::min(n, D2); ++i) {
        const FlatMap &f = maps[ent[i].key];
        __builtin_prefetch(&f.s[f.slot(from)]);
        __builtin_prefetch(&f.s[f.slot(to)]);
    }
    for (size_t i = 0; i < n; ++i) {
        if (i + D1 < n) __builtin_prefetch(&maps[ent[i + D1].key]);
        if (i + D2 < n) {
            const FlatMap &f = maps[ent[i + D2].key];
            __builtin_prefetch(&f.s[f.slot(from)]);
            __builtin_prefetch(&f.s[f.slot(to)]);
        }
        FlatMap &f = maps[ent[i].key];
        f.erase(from);
        f.add(to, ent[i].val);
    }
}

// Blockmodel::merge_block (src/blockmodel/blockmodel.cpp:337-363): row r and column r move into s.
// (r, c) -> (s, c') and (x, r) -> (x', s), with c' = s for c in {r, s} and x' = s for x = s.  Row s and
// column s take all of them on one thread; the renames r -> s in the columns of the other c and the rows of
// the other x touch distinct maps, so with a pool a long list of them is split into pieces.  Only the
// thread that owns `pool` may pass it.
// The block-matrix part of merging r into s: row and column s take r's cells, the rows and columns of r's
// neighbours rename key r to s, and r's maps are retired.  It writes the maps of r, s and of the keys of row r
// and column r, and nothing else.
inline void merge_cells(Model &m, int32_t r, int32_t s, Pool *pool = nullptr, const SplitSizes &split = SplitSizes()) {
    FlatMap &Rr = m.row[r], &Cr = m.col[r], &Rs = m.row[s], &Cs = m.col[s];
    thread_local std::vector<FlatMap::Slot> row_buf, col_buf, row_ren, col_ren;
    const size_t nr = gather(Rr, row_buf), nc = gather(Cr, col_buf);
    if (row_ren.size() < nr) row_ren.resize(nr);
    if (col_ren.size() < nc) col_ren.resize(nc);
    const FlatMap::Slot *er = row_buf.data(), *ec = col_buf.data();
    FlatMap::Slot *fr = row_ren.data(), *fc = col_ren.data();  // the caller's buffers, also in workers
    size_t kr = 0, kc = 0;
    for (size_t i = 0; i < nr; ++i) {
        fr[kr] = er[i];
        kr += er[i].key != r && er[i].key != s;
    }
    for (size_t i = 0; i < nc; ++i) {
        fc[kc] = ec[i];
        kc += ec[i].key != r && ec[i].key != s;
    }
    // start fetching the first neighbour maps to be renamed; row s / column s are updated meanwhile
    for (size_t i = 0; i < std::min<size_t>(kr, 16); ++i) __builtin_prefetch(&m.col[fr[i].key]);
    for (size_t i = 0; i < std::min<size_t>(kc, 16); ++i) __builtin_prefetch(&m.row[fc[i].key]);
    auto own = [&] {
        Rs.reserve(Rs.n + (uint32_t)nr);
        Cs.reserve(Cs.n + (uint32_t)nc);
        for (size_t i = 0; i < nr; ++i) {
            if (i + 8 < nr) __builtin_prefetch(&Rs.s[Rs.slot(er[i + 8].key)]);
            const int32_t c = er[i].key;
            const int64_t v = er[i].val;
            if (c == r || c == s) {
                Rs.add(s, v);
                m.col[c].erase(r);
                Cs.add(s, v);
            } else {
                Rs.add(c, v);
            }
        }
        for (size_t i = 0; i < nc; ++i) {
            if (i + 8 < nc) __builtin_prefetch(&Cs.s[Cs.slot(ec[i + 8].key)]);
            const int32_t x = ec[i].key;
            const int64_t v = ec[i].val;
            if (x == r) continue;
            if (x == s) {
                Cs.add(s, v);
                Rs.erase(r);
                Rs.add(s, v);
            } else {
                Cs.add(x, v);
            }
        }
    };
    if (pool && pool->threads() > 1 && kr + kc >= split.merge_min) {
        const size_t P = split.merge_piece;
        const size_t pr = (kr + P - 1) / P, pc = (kc + P - 1) / P;
        pool->run((int64_t)(1 + pr + pc), [&](int64_t i) {
            if (i == 0) {
                own();
            } else if ((size_t)i <= pr) {
                const size_t lo = (i - 1) * P;
                rename_in(m.col, fr + lo, std::min(kr - lo, P), r, s);
            } else {
                const size_t lo = (i - 1 - pr) * P;
                rename_in(m.row, fc + lo, std::min(kc - lo, P), r, s);
            }
        });
    } else {
        own();
        rename_in(m.col, fr, kr, r, s);
        rename_in(m.row, fc, kc, r, s);
    }
    m.retired.p.push_back(Rr.detach());
    m.retired.p.push_back(Cr.detach());
}

// The block degrees, sizes and degree histograms of merging r into s.
inline void merge_totals(Model &m, int32_t r, int32_t s) {
    m.d_out[s] += m.d_out[r];
    m.d_in[s] += m.d_in[r];
    m.d_out[r] = 0;
    m.d_in[r] = 0;
    if (m.np) {
        m.size[s] += m.size[r];
        m.size[r] = 0;
        m.nonempty--;
    }
    if (m.dc) {  // Blockmodel::merge_block adds r's histograms to s's (src/blockmodel/blockmodel.cpp:357-362)
        m.hin[r].each([&](int32_t deg, int64_t c) { m.hin[s].add(deg, c); });
        m.hout[r].each([&](int32_t deg, int64_t c) { m.hout[s].add(deg, c); });
    }
}

inline void merge_apply(Model &m, int32_t r, int32_t s, Pool *pool = nullptr,
                        const SplitSizes &split = SplitSizes()) {
    merge_cells(m, r, s, pool, split);
    merge_totals(m, r, s);
}

// Applies the block-matrix part of merges (merge_cells) on a thread of its own, in the order submitted, while
// the caller goes on deciding.  Ticket t is the t-th merge submitted; wait(t) returns once it is applied,
// after which its writes are visible to the caller.
class Applier {
  public:
    explicit Applier(Model &m) : m_(m), ring_(RING) { th_ = std::thread([this] { run(); }); }
    ~Applier() {
        stop_.store(true, std::memory_order_release);
        th_.join();
    }
    int64_t next_ticket() const { return submitted_ + 1; }
    int64_t submit(int32_t r, int32_t s) {
        const int64_t t = submitted_ + 1;
        while (t - applied_.load(std::memory_order_acquire) > RING) relax();
        ring_[(size_t)((t - 1) & (RING - 1))] = {r, s};
        submitted_ = t;
        published_.store(t, std::memory_order_release);
        return t;
    }
    void wait(int64_t ticket) const {
        while (applied_.load(std::memory_order_acquire) < ticket) relax();
    }
    void drain() const { wait(submitted_); }

  private:
    static constexpr int64_t RING = int64_t(1) << 14;
    static void relax() {
#if defined(__x86_64__)
        __builtin_ia32_pause();
#else
        std::this_thread::yield();
#endif
    }
    void run() {
        int64_t done = 0;
        uint32_t idle = 0;
        while (true) {
            const int64_t p = published_.load(std::memory_order_acquire);
            if (p == done) {
                if (stop_.load(std::memory_order_acquire) && published_.load(std::memory_order_acquire) == done) return;
                // idle while the caller evaluates in parallel on the pool: give the core to its workers
                if (++idle < 4096) relax();
                else std::this_thread::yield();
                continue;
            }
            idle = 0;
            for (; done < p; ++done) {
                const std::pair<int32_t, int32_t> job = ring_[(size_t)(done & (RING - 1))];
                merge_cells(m_, job.first, job.second);
                applied_.store(done + 1, std::memory_order_release);
            }
        }
    }
    Model &m_;
    std::vector<std::pair<int32_t, int32_t>> ring_;
    int64_t submitted_ = 0;  // caller only
    std::atomic<int64_t> published_{0}, applied_{0};
    std::atomic<bool> stop_{false};
    std::thread th_;
};

// Merges are applied on the Applier thread (SBP_AMD_CARRY_PIPE=0 turns this off) for the parametric model
// with a thread pool, at least this many blocks and rows of fewer than PIPE_MAX_ROW_CELLS cells on average.
// Fewer blocks share more neighbours, and longer rows make each merge's renames long: the one applier thread
// falls behind and the loop waits for it, where the in-loop path splits the renames over the pool.  On the
// 200,000-vertex workload the carry-out took 25% less time at B = 200,000, 14% less at 50,000 (about 75 cells
// per row) and 11% more at 25,000 (about 150).  On the 2,500,000-vertex sample of the 5,000,000-vertex
// workload it took 23%, 25% and 17% less at B = 2,500,000, 1,250,000 and 625,000 (up to about 95 cells per
// row), and 17% to 69% more from 312,500 (about 190) down to 39,063.
constexpr int32_t PIPE_MIN_BLOCKS = 32768;
constexpr int64_t PIPE_MAX_ROW_CELLS = 128;
inline bool pipe_enabled() {
    const char *e = getenv("SBP_AMD_CARRY_PIPE");
    return !(e && e[0] == '0');
}

struct CarryStats {
    int64_t pops = 0;          // items taken from the queue
    int64_t evaluations = 0;   // delta-MDL evaluations on the sequential path
    int64_t speculative = 0;   // delta-MDL evaluations run ahead in parallel
    int64_t reused = 0;        // pops answered from a still-valid earlier evaluation
    int64_t requeues = 0, merges = 0;
};

// Returns the block count after the merges; block_map[b] = new label of old block b (labels compacted in
// ascending order of the surviving roots, as Blockmodel::build_mapping does).  `m` may be null when
// approximate.
//
// The advanced carry-out takes exactly the decisions of the sequential loop (pop the smallest (dS, r),
// re-evaluate it on the current state, re-queue it if it got worse than the next item, else merge), but
// avoids most sequential evaluations:
//   * delta-MDL(r -> t) reads rows/columns r and t, the degrees of r, t and of N(r) = the blocks sharing a
//     cell with r, and the cells of t keyed by N(r) + {r, t}.  A merge a -> b changes rows/columns a and b,
//     the entries keyed a (which become b) of the rows/columns of N(a), and the degrees of a and b.  So
//     it leaves delta-MDL(r -> t) unchanged unless r or t is a or b, or r is in N(a) or N(b) (a in N(r)
//     or b in N(r)).  After the merge every block of N(b), which now holds N(a), is marked with the merge
//     count (its row and column were just written, so the scan is cache-hot), a and b are stamped with it,
//     and an evaluation made at time T is still exact while stamp[r], stamp[t] and mark[r] are <= T;
//   * the proposal values (computed on the same state with the same terms and fixed-point sum) are the
//     first evaluations;
//   * items are taken from the heap in batches and the ones without a valid evaluation are evaluated in
//     parallel against the current state before the sequential pass consumes them.  The batch size adapts
//     to the fraction of those evaluations that a merge inside the batch invalidates.
inline int32_t carry_out(Model *m, int32_t B, const int32_t *best_s, const double *best_dS, const int32_t *sizes,
                         int64_t num_to_merge, bool approximate, std::vector<int32_t> &block_map,
                         CarryStats *cs = nullptr, Pool *pool = nullptr, bool seed_from_proposals = true,
                         const SplitSizes &split = SplitSizes::from_env()) {
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
            if (best_s[from] < 0) continue;
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
        auto before = [](const Item &a, const Item &b) { return a.dS != b.dS ? a.dS < b.dS : a.r < b.r; };
        auto worse = [&](const Item &a, const Item &b) { return before(b, a); };
        std::vector<Item> init;
        init.reserve(B);
        for (int32_t b = 0; b < B; ++b)
            if (best_s[b] >= 0) init.push_back(Item{best_dS[b], b, best_s[b]});
        // The queue: the proposals sorted once, merged on the fly with a heap of the re-queued items (the front
        // is the smaller of the two).  Pops come in the order of one heap over all items, since (dS, r) orders
        // them totally (each block has one item); the sorted part is read in sequence instead of sifted.
        std::sort(init.begin(), init.end(), before);
        size_t q_head = 0;
        std::priority_queue<Item, std::vector<Item>, decltype(worse)> requeued(worse);
        auto q_from_requeued = [&] {
            return q_head == init.size() || (!requeued.empty() && before(requeued.top(), init[q_head]));
        };
        auto q_empty = [&] { return q_head == init.size() && requeued.empty(); };
        auto q_top = [&]() -> const Item & { return q_from_requeued() ? requeued.top() : init[q_head]; };
        auto q_pop = [&] {
            if (q_from_requeued()) requeued.pop();
            else ++q_head;
        };
        // evaluation cache per source block: target root, merge count when evaluated, value
        std::vector<int32_t> c_to(B, -1);
        std::vector<int64_t> c_time(B, -1), stamp(B, 0), mark(B + 1, 0);  // mark[B]: spare entry
        std::vector<double> c_val(B, 0.0);
        // nonparametric: the cache holds the Delta entries' sum; the rest depends on the non-empty block
        // count, which every merge changes, and is recomputed at use
        const bool np = m->np;
        auto evaluate = [&](int32_t r, int32_t t, Pool *p) {
            return np ? np_merge_entries(*m, r, t) : merge_dS(*m, r, t, p, split);
        };
        auto finalize = [&](int32_t r, int32_t t, double v) { return np ? np_merge_dS(*m, r, t, v) : v; };
        if (seed_from_proposals && !np)
            for (int32_t b = 0; b < B; ++b)
                if (best_s[b] >= 0 && best_s[b] != b) {
                    c_to[b] = best_s[b];
                    c_time[b] = 0;
                    c_val[b] = best_dS[b];
                }
        int64_t now = 0;
        auto valid = [&](int32_t r, int32_t t) {
            const int64_t T = c_time[r];
            return c_to[r] == t && T >= stamp[r] && T >= stamp[t] && T >= mark[r];
        };
        // marks the keys of f (no branch per slot: empty slots write the spare entry)
        int64_t *mk = mark.data();
        auto mark_keys = [&](const FlatMap &f, int64_t when) {
            for (uint32_t i = 0; i < f.cap; ++i) {
                const int32_t k = f.s[i].key;
                mk[k < 0 ? B : k] = when;
            }
        };
        auto root_ro = [&](int32_t x) {
            while (parent[x] != x) x = parent[x];
            return x;
        };
        // merges applied on a thread of their own: busy[x] is the ticket of the last queued merge that writes
        // the row or column of x (its r, s and the keys of row r and column r); the maps of x are read after
        // waiting for it
        std::unique_ptr<Applier> applier;
        std::vector<int64_t> busy;
        if (m && !m->np && !m->dc && pool && pool->threads() > 1 && B >= PIPE_MIN_BLOCKS && pipe_enabled()) {
            int64_t cells = 0;
            for (int32_t b = 0; b < B; ++b) cells += m->row[b].n;
            if (cells < PIPE_MAX_ROW_CELLS * (int64_t)B) {
                applier.reset(new Applier(*m));
                busy.assign(B, 0);
            }
        }
        auto ready = [&](int32_t x) {
            if (applier) applier->wait(busy[x]);
        };
        std::vector<Item> batch;
        std::atomic<int64_t> spec_count{0};
        size_t bi = 0;
        int64_t K = 256, batch_spec = 0, batch_lost = 0, spec_time = -1;
        while (merged < num_to_merge) {
            if (bi == batch.size()) {
                if (q_empty()) break;
                if (batch_spec >= 32) {  // adapt the batch size to the invalidated fraction of the last batch
                    if (4 * batch_lost > batch_spec) K = std::max<int64_t>(32, K / 2);
                    else if (10 * batch_lost < batch_spec) K = std::min<int64_t>(16384, K * 2);
                }
                batch.clear();
                bi = 0;
                batch_spec = batch_lost = 0;
                const int64_t want = std::min<int64_t>(K, 2 * (num_to_merge - merged) + 32);
                while ((int64_t)batch.size() < want && !q_empty()) {
                    batch.push_back(q_top());
                    q_pop();
                }
                if (pool && batch.size() >= 16) {
                    if (applier) applier->drain();
                    pool->run((int64_t)batch.size(), [&](int64_t i) {
                        const Item &it = batch[i];
                        const int32_t t = root_ro(it.s);
                        if (t == it.r || valid(it.r, t)) return;
                        c_val[it.r] = evaluate(it.r, t, nullptr);
                        c_to[it.r] = t;
                        c_time[it.r] = now;
                        spec_count.fetch_add(1, std::memory_order_relaxed);
                    });
                    batch_spec = spec_count.exchange(0);
                    st.speculative += batch_spec;
                    spec_time = now;
                }
            }
            Item it;
            if (!q_empty() && before(q_top(), batch[bi])) {
                it = q_top();
                q_pop();
            } else {
                it = batch[bi++];
                if (m && !applier) {  // the rows and columns of the next items are read when they are merged
                    if (bi + 4 < batch.size()) {
                        __builtin_prefetch(&m->row[batch[bi + 4].r]);
                        __builtin_prefetch(&m->col[batch[bi + 4].r]);
                    }
                    if (bi + 2 < batch.size()) {
                        __builtin_prefetch(m->row[batch[bi + 2].r].s);
                        __builtin_prefetch(m->col[batch[bi + 2].r].s);
                    }
                }
            }
            st.pops++;
            const int32_t from = it.r;
            const int32_t to = find(it.s);
            if (from == to) continue;
            double actual;
            if (valid(from, to)) {
                actual = finalize(from, to, c_val[from]);
                st.reused++;
            } else {
                if (c_to[from] == to && c_time[from] == spec_time) batch_lost++;  // speculated, then invalidated
                ready(from);
                ready(to);
                const double v = evaluate(from, to, pool);
                c_to[from] = to;
                c_time[from] = now;
                c_val[from] = v;
                actual = finalize(from, to, v);
                st.evaluations++;
            }
            const bool has_b = bi < batch.size();
            if (has_b || !q_empty()) {
                const double next = (has_b && (q_empty() || before(batch[bi], q_top()))) ? batch[bi].dS
                                                                                          : q_top().dS;
                if (actual > next) {
                    it.dS = actual;
                    requeued.push(it);
                    st.requeues++;
                    continue;
                }
            }
            ++now;
            stamp[from] = stamp[to] = now;
            parent[from] = to;
            if (applier) {
                // the merged row and column `to` hold the keys of rows and columns from and to (key from becoming
                // to, which needs no mark: from is merged and to is stamped)
                ready(from);
                ready(to);
                const int64_t ticket = applier->next_ticket();
                for (const FlatMap *f : {&m->row[from], &m->col[from]})
                    for (uint32_t i = 0; i < f->cap; ++i) {
                        const int32_t k = f->s[i].key;
                        if (k < 0) continue;
                        mk[k] = now;
                        busy[k] = ticket;
                    }
                mark_keys(m->row[to], now);
                mark_keys(m->col[to], now);
                busy[from] = busy[to] = ticket;
                merge_totals(*m, from, to);
                applier->submit(from, to);
            } else {
                merge_apply(*m, from, to, pool, split);
                mark_keys(m->row[to], now);
                mark_keys(m->col[to], now);
            }
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
