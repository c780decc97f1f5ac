// Device-side arithmetic shared by every SBP kernel (gfx950, wave64).
//
//  * logf_glibc: a restatement of the table-driven single-precision logarithm
//    published in ARM's optimized-routines (Szabolcs Nagy, 2017) that glibc
//    (>= 2.28) ships as logf.  The reference's fastlog and nat_log are logf
//    (include/fastlog.hpp:21-30, include/utils.hpp:228-238), so the device must
//    reproduce glibc's rounding exactly for FP64 delta-MDL sums to be bit-exact.
//    Checked against glibc on every positive float.
//  * Philox4x32-10 counter-based draws (Salmon et al., SC'11) keyed by
//    (seed) with counter {unit, epoch, tag, sub}.
//  * 128-bit fixed-point accumulation of FP64 terms (scale 2^64, truncation
//    toward zero), so sums are independent of reduction order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sbpdev {

constexpr uint32_t TAG_FT = 0x46540001u;
constexpr uint32_t TAG_MERGE = 0x4d470001u;
constexpr uint32_t TAG_SHUFFLE = 0x53480001u;
constexpr uint32_t TAG_EXTEND = 0x45580001u;
constexpr uint32_t TAG_SAMPLE = 0x53410001u;
constexpr uint32_t TAG_SPLIT_SIZE = 0x53500001u;  // split size: unit = block, sub = proposal / 4
constexpr uint32_t TAG_SPLIT_SIDE = 0x53500002u;  // split side: unit = vertex, sub = proposal / 4

constexpr int WAVE = 64;

// ---------------------------------------------------------------- logf

struct LogfEntry {
    double invc, logc;
};

__device__ __constant__ const LogfEntry kLogfTab[16] = {
    {0x1.661ec79f8f3bep+0, -0x1.57bf7808caadep-2}, {0x1.571ed4aaf883dp+0, -0x1.2bef0a7c06ddbp-2},
    {0x1.49539f0f010bp+0, -0x1.01eae7f513a67p-2},  {0x1.3c995b0b80385p+0, -0x1.b31d8a68224e9p-3},
    {0x1.30d190c8864a5p+0, -0x1.6574f0ac07758p-3}, {0x1.25e227b0b8eap+0, -0x1.1aa2bc79c81p-3},
    {0x1.1bb4a4a1a343fp+0, -0x1.a4e76ce8c0e5ep-4}, {0x1.12358f08ae5bap+0, -0x1.1973c5a611cccp-4},
    {0x1.0953f419900a7p+0, -0x1.252f438e10c1ep-5}, {0x1p+0, 0x0p+0},
    {0x1.e608cfd9a47acp-1, 0x1.aa5aa5df25984p-5},  {0x1.ca4b31f026aap-1, 0x1.c5e53aa362eb4p-4},
    {0x1.b2036576afce6p-1, 0x1.526e57720db08p-3},  {0x1.9c2d163a1aa2dp-1, 0x1.bc2860d22477p-3},
    {0x1.886e6037841edp-1, 0x1.1058bc8a07ee1p-2},  {0x1.767dcf5534862p-1, 0x1.4043057b6ee09p-2},
};

// log(x) = k*ln2 + log(c) + log1p(z/c - 1), z in [0.7, 1.4) after the exponent split,
// 16 subintervals indexed by the top mantissa bits, degree-3 polynomial for log1p.
__device__ __forceinline__ float logf_glibc(float x) {
    const double Ln2 = 0x1.62e42fefa39efp-1;
    const double A0 = -0x1.00ea348b88334p-2, A1 = 0x1.5575b0be00b6ap-2, A2 = -0x1.ffffef20a4123p-2;
    uint32_t ix = __float_as_uint(x);
    if (ix == 0x3f800000u) return 0.0f;
    if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
        if (ix * 2 == 0) return -__builtin_inff();
        if (ix == 0x7f800000u) return x;
        if ((ix & 0x80000000u) || ix * 2 >= 0xff000000u) return __builtin_nanf("");
        ix = __float_as_uint(x * 0x1p23f);  // subnormal: normalize
        ix -= 23u << 23;
    }
    uint32_t tmp = ix - 0x3f330000u;
    int i = (tmp >> (23 - 4)) % 16;
    int k = (int32_t)tmp >> 23;
    uint32_t iz = ix - (tmp & (0x1ffu << 23));
    double invc = kLogfTab[i].invc, logc = kLogfTab[i].logc;
    double z = (double)__uint_as_float(iz);
    double r = z * invc - 1.0;
    double y0 = logc + (double)k * Ln2;
    double r2 = r * r;
    double y = A1 * r + A2;
    y = A0 * r2 + y;
    y = y * r2 + (y0 + r);
    return (float)y;
}

// fastlog (include/fastlog.hpp:21-30): logf of an integer count, 0 for 0.
__device__ __forceinline__ double fastlog(long long x) {
    return x == 0 ? 0.0 : (double)logf_glibc((float)x);
}

// common::cell_entropy (include/common.hpp:39-49).
__device__ __forceinline__ double cell_entropy(long long value, long long d_in, long long d_out) {
    if (value == 0) return 0.0;
    return (double)value * (fastlog(value) - (fastlog(d_in) + fastlog(d_out)));
}

// ---------------------------------------------------------------- Philox4x32-10

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint32_t hi0 = __umulhi(0xD2511F53u, x0), lo0 = 0xD2511F53u * x0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, x2), lo1 = 0xCD9E8D57u * x2;
        uint32_t y0 = hi1 ^ x1 ^ k0, y2 = hi0 ^ x3 ^ k1;
        x0 = y0;
        x1 = lo1;
        x2 = y2;
        x3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{x0, x1, x2, x3};
}

__device__ __forceinline__ uint32_t uniform_below(uint32_t w, uint32_t n) {
    return (uint32_t)(((uint64_t)w * n) >> 32);
}

__device__ __forceinline__ double unit_double(uint32_t a, uint32_t b) {
    uint64_t u = (((uint64_t)a << 32) | b) >> 11;
    return (double)u * 0x1p-53;
}

// common::propose_random_block (src/common.cpp:200-209) with candidates = U[0, B-2].  With fewer than two blocks
// there is no other block (the reference's U[0, -1] is undefined there): propose the current one, no move.
__device__ __forceinline__ int random_block(int current, int B, uint32_t w) {
    if (B < 2) return current;
    int p = (int)uniform_below(w, (uint32_t)(B - 1));
    return p >= current ? p + 1 : p;
}

// ---------------------------------------------------------------- 128-bit fixed point

struct Fx {
    unsigned long long lo;
    long long hi;
};

__device__ __forceinline__ Fx fx_zero() { return Fx{0ull, 0ll}; }

__device__ __forceinline__ void fx_add(Fx &a, Fx b) {
    unsigned long long lo = a.lo + b.lo;
    a.hi = a.hi + b.hi + (lo < a.lo ? 1 : 0);
    a.lo = lo;
}

__device__ __forceinline__ Fx fx_neg(Fx a) {
    Fx r;
    r.lo = ~a.lo + 1ull;
    r.hi = (long long)(~(unsigned long long)a.hi + (r.lo == 0ull ? 1ull : 0ull));
    return r;
}

// d * 2^64 truncated toward zero; |d| < 2^63 by construction of the terms.
__device__ __forceinline__ Fx fx_from(double d) {
    unsigned long long bits = (unsigned long long)__double_as_longlong(d);
    int e = (int)((bits >> 52) & 0x7ff);
    if (e == 0) return fx_zero();
    unsigned long long m = (bits & 0xfffffffffffffull) | (1ull << 52);
    int shift = e - 1011;  // value * 2^64 = m * 2^shift
    Fx r;
    if (shift >= 64) {
        r.lo = 0;
        r.hi = (long long)(m << (shift - 64));
    } else if (shift > 0) {
        r.lo = m << shift;
        r.hi = (long long)(m >> (64 - shift));
    } else if (shift == 0) {
        r.lo = m;
        r.hi = 0;
    } else if (shift > -64) {
        r.lo = m >> (-shift);
        r.hi = 0;
    } else {
        return fx_zero();
    }
    return (bits >> 63) ? fx_neg(r) : r;
}

__device__ __forceinline__ double fx_to(Fx v) { return (double)v.hi + (double)v.lo * 0x1p-64; }

__device__ __forceinline__ void fx_acc(Fx &acc, double term) { fx_add(acc, fx_from(term)); }
__device__ __forceinline__ void fx_sub_acc(Fx &acc, double term) { fx_add(acc, fx_neg(fx_from(term))); }

__device__ __forceinline__ Fx wave_sum_fx(Fx v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        Fx o;
        o.lo = __shfl_xor(v.lo, off, WAVE);
        o.hi = __shfl_xor(v.hi, off, WAVE);
        fx_add(v, o);
    }
    return v;
}

// ---------------------------------------------------------------- wave helpers

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Inclusive prefix sum across the wave.
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        T o = __shfl_up(v, off, WAVE);
        if (lane >= off) v += o;
    }
    return v;
}

__device__ __forceinline__ unsigned long long ballot(bool p) { return __ballot(p); }

// First index in [lo, hi) of a sorted int array with value >= key (binary search).
__device__ __forceinline__ int lower_bound_i32(const int *a, int lo, int hi, int key) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Open-addressing hash of the block matrix: key (row << 32) | col, linear probing, empty key ~0.  One slot is
// 16 bytes, so a probe is one 128-bit load.
struct __align__(16) CellSlot {
    unsigned long long key;
    int val;
    int pad;
};
constexpr unsigned long long CELL_EMPTY = ~0ull;

__host__ __device__ __forceinline__ unsigned long long cell_key(int row, int col) {
    return ((unsigned long long)(unsigned)row << 32) | (unsigned)col;
}
__device__ __forceinline__ unsigned cell_slot(unsigned long long key, int shift) {
    return (unsigned)((key * 0x9E3779B97F4A7C15ull) >> shift);
}
__device__ __forceinline__ int cell_hash_get(const CellSlot *table, int shift, unsigned mask, int row, int col) {
    const unsigned long long key = cell_key(row, col);
    for (unsigned i = cell_slot(key, shift);; i = (i + 1) & mask) {
        const CellSlot c = table[i];
        if (c.key == key) return c.val;
        if (c.key == CELL_EMPTY) return 0;
    }
}

// Cells are stored as int2 {index, count}, sorted by index within a row/column.
__device__ __forceinline__ int cell_lookup(const int2 *cells, int begin, int end, int key) {
    int lo = begin, hi = end;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cells[mid].x < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < end && cells[lo].x == key) ? cells[lo].y : 0;
}

}  // namespace sbpdev
