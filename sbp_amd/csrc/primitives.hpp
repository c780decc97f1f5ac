// Device-wide primitives of the library, written for gfx950 (wave64): prefix sums, a stable LSD radix sort of
// 32- or 64-bit keys (optionally carrying int32 values) and run-length encoding of sorted keys.
//
// Calling convention (as the sorts of the blockmodel build, the commit's event sort, the shuffle, the samplers
// and the Metropolis-Hastings re-layout need them): a first call with temp == nullptr stores the temporary bytes
// in temp_bytes; a second call with temp of at least that size runs on `stream`.  Nothing is allocated.
//
// Radix sort, one pass per 8-bit digit of [begin_bit, end_bit):
//   1. k_digit_hist: each workgroup counts the digits of its tile of RS_TILE keys into an LDS histogram and
//      writes it digit-major (hist[digit * tiles + tile]);
//   2. an exclusive scan of the histogram gives, for every (digit, tile), the output position of its first key;
//   3. k_digit_scatter: each workgroup ranks its keys stably -- item j of wavefront w holds key
//      tile_base + (w * RS_ITEMS + j) * 64 + lane, so (w, j, lane) is input order -- with per-wavefront digit
//      counters in LDS, peers of a digit found by nine ballots (the digit's bits plus the tile-end sentinel), then
//      an exclusive scan of the counters over the wavefronts; each key lands at the (digit, tile) position plus
//      its rank.
// Keys equal in the sorted bits keep their input order, so the LSD passes compose into a stable sort.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sbpprim {

constexpr int PR_THREADS = 256;
constexpr int PR_WAVES = PR_THREADS / 64;
constexpr int RS_ITEMS = 16;                      // keys per thread in a radix tile
constexpr int RS_TILE = PR_THREADS * RS_ITEMS;    // 4,096 keys
constexpr int SC_ITEMS = 16;                      // values per thread in a scan chunk
constexpr int SC_CHUNK = PR_THREADS * SC_ITEMS;   // 4,096 values

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ int pr_lane() { return (int)(threadIdx.x & 63); }

// Inclusive scan of one value per thread over the workgroup (PR_THREADS threads); total in *total.
template <typename T>
__device__ __forceinline__ T block_inclusive_scan(T v, T *total) {
    __shared__ T wsum[PR_WAVES];
    const int lane = pr_lane(), w = (int)threadIdx.x / 64;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T o = __shfl(v, lane - (lane >= off ? off : 0), 64);
        if (lane >= off) v += o;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    T before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < PR_WAVES; ++i) {
        if (i < w) before += wsum[i];
        all += wsum[i];
    }
    __syncthreads();
    if (total) *total = all;
    return v + before;
}

// ---------------------------------------------------------------- prefix sums

// Sum of each chunk of SC_CHUNK values (summed in T).
template <typename TI, typename T>
__global__ void __launch_bounds__(PR_THREADS) k_chunk_sums(const TI *__restrict__ in, int64_t n, T *__restrict__ sums) {
    const int64_t base = (int64_t)blockIdx.x * SC_CHUNK;
    T s = 0;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
        const int64_t i = base + (int64_t)j * PR_THREADS + threadIdx.x;
        if (i < n) s += (T)in[i];
    }
    T total;
    block_inclusive_scan(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the chunk sums in place, by one workgroup walking them PR_THREADS at a time.
template <typename T>
__global__ void __launch_bounds__(PR_THREADS) k_scan_sums(T *__restrict__ sums, int64_t m) {
    T carry = 0;
    for (int64_t base = 0; base < m; base += PR_THREADS) {
        const int64_t i = base + threadIdx.x;
        const T v = i < m ? sums[i] : T(0);
        T total;
        const T incl = block_inclusive_scan(v, &total);
        if (i < m) sums[i] = carry + incl - v;
        carry += total;
    }
}

// Scan of each chunk from its offset: thread t owns values t*SC_ITEMS .. of the chunk (contiguous), so the
// per-thread partial is a sequential sum and one workgroup scan joins them.
template <typename TI, typename T>
__global__ void __launch_bounds__(PR_THREADS) k_chunk_scan(const TI *__restrict__ in, T *__restrict__ out, int64_t n,
                                                           const T *__restrict__ offs, int inclusive) {
    const int64_t base = (int64_t)blockIdx.x * SC_CHUNK + (int64_t)threadIdx.x * SC_ITEMS;
    T v[SC_ITEMS];
    T s = 0;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
        v[j] = base + j < n ? (T)in[base + j] : T(0);
        s += v[j];
    }
    T run = block_inclusive_scan(s, (T *)nullptr) - s + offs[blockIdx.x];
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
        if (inclusive) run += v[j];
        if (base + j < n) out[base + j] = run;
        if (!inclusive) run += v[j];
    }
}

// out[i] = sum of in[0..i) (exclusive) or in[0..i] (inclusive), summed in the output type; in and out may be the
// same array (same type).
template <typename TI, typename T = TI>
hipError_t scan(void *temp, size_t &temp_bytes, const TI *in, T *out, int64_t n, bool inclusive, hipStream_t st) {
    static_assert(std::is_same<T, int32_t>::value || std::is_same<T, int64_t>::value, "int32 or int64 sums");
    const int64_t chunks = ceil_div(n > 0 ? n : 1, SC_CHUNK);
    const size_t need = align256(sizeof(T) * (size_t)chunks);
    if (!temp) {
        temp_bytes = need;
        return hipSuccess;
    }
    if (temp_bytes < need) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    T *sums = static_cast<T *>(temp);
    k_chunk_sums<TI, T><<<(unsigned)chunks, PR_THREADS, 0, st>>>(in, n, sums);
    k_scan_sums<T><<<1, PR_THREADS, 0, st>>>(sums, chunks);
    k_chunk_scan<TI, T><<<(unsigned)chunks, PR_THREADS, 0, st>>>(in, out, n, sums, inclusive ? 1 : 0);
    return hipGetLastError();
}

// ---------------------------------------------------------------- radix sort

template <typename K>
__device__ __forceinline__ int key_digit(K k, int shift, int bits) {
    return (int)((k >> shift) & (K)((1u << bits) - 1u));
}

template <typename K>
__global__ void __launch_bounds__(PR_THREADS) k_digit_hist(const K *__restrict__ keys, int64_t n,
                                                           const int64_t *__restrict__ dn, int shift, int bits,
                                                           int64_t tiles, int32_t *__restrict__ hist) {
    if (dn) n = *dn;
    __shared__ int32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t i = base + (int64_t)j * PR_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[key_digit(keys[i], shift, bits)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Lanes of the wavefront holding the same value d (0..256; 256 marks no key), among the active lanes.
__device__ __forceinline__ unsigned long long peers_of(int d) {
    unsigned long long m = __ballot(true);
#pragma unroll
    for (int b = 0; b < 9; ++b) {
        const bool set = (d >> b) & 1;
        const unsigned long long bal = __ballot(set);
        m &= set ? bal : ~bal;
    }
    return m;
}

template <typename K, typename V, bool HAS_VALS>
__global__ void __launch_bounds__(PR_THREADS) k_digit_scatter(const K *__restrict__ keys, const V *__restrict__ vals,
                                                              K *__restrict__ keys_out, V *__restrict__ vals_out,
                                                              int64_t n, const int64_t *__restrict__ dn, int shift,
                                                              int bits, int64_t tiles,
                                                              const int32_t *__restrict__ offs) {
    if (dn) n = *dn;
    __shared__ int32_t wcnt[PR_WAVES][256];
    __shared__ int32_t goff[256];
    const int lane = pr_lane(), w = (int)threadIdx.x / 64;
    for (int i = lane; i < 256; i += 64) wcnt[w][i] = 0;
    goff[threadIdx.x] = offs[(int64_t)threadIdx.x * tiles + blockIdx.x];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE + (int64_t)w * RS_ITEMS * 64;
    K k[RS_ITEMS];
    int d[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t i = base + (int64_t)j * 64 + lane;
        k[j] = i < n ? keys[i] : K(0);
        d[j] = i < n ? key_digit(k[j], shift, bits) : 256;
    }
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const unsigned long long peers = peers_of(d[j]);
        const int before = __popcll(peers & ((1ull << lane) - 1ull));
        const int leader = __ffsll((long long)peers) - 1;
        const int dd = d[j] < 256 ? d[j] : 0;
        const int cur = d[j] < 256 ? wcnt[w][dd] : 0;
        rank[j] = cur + before;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if (lane == leader && d[j] < 256) wcnt[w][dd] = cur + __popcll(peers);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {  // exclusive scan over the wavefronts of each digit, added to the tile's global offset
        int run = goff[threadIdx.x];
#pragma unroll
        for (int x = 0; x < PR_WAVES; ++x) {
            const int c = wcnt[x][threadIdx.x];
            wcnt[x][threadIdx.x] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        if (d[j] < 256) {
            const int64_t pos = (int64_t)wcnt[w][d[j]] + rank[j];
            keys_out[pos] = k[j];
            if constexpr (HAS_VALS) vals_out[pos] = vals[base + (int64_t)j * 64 + lane];
        }
    }
}

// Stable ascending sort of keys by bits [begin_bit, end_bit) (keys agreeing there keep their order), values
// permuted alike when vals_in is given.  keys_in / vals_in are left unchanged; the results are in keys_out /
// vals_out.  Temporary: two key buffers (and value buffers), the histogram and its scan.  With dn, the key count is
// read on the device from *dn (at most n, which sizes the launches): no host read, so the sort can be captured in a
// graph; n = 0 then still launches nothing.
template <typename K, typename V = int32_t>
hipError_t radix_sort(void *temp, size_t &temp_bytes, const K *keys_in, K *keys_out, const V *vals_in, V *vals_out,
                      int64_t n, int begin_bit, int end_bit, hipStream_t st, const int64_t *dn = nullptr) {
    static_assert(std::is_unsigned<K>::value && (sizeof(K) == 4 || sizeof(K) == 8), "unsigned 32- or 64-bit keys");
    const int64_t tiles = ceil_div(n > 0 ? n : 1, RS_TILE);
    const int64_t hist_n = 256 * tiles;
    size_t scan_bytes = 0;
    scan<int32_t>(nullptr, scan_bytes, (const int32_t *)nullptr, (int32_t *)nullptr, hist_n, false, st);
    const size_t kb = align256(sizeof(K) * (size_t)(n > 0 ? n : 1));
    const size_t vb = vals_in ? align256(sizeof(V) * (size_t)(n > 0 ? n : 1)) : 0;
    const size_t hb = align256(sizeof(int32_t) * (size_t)hist_n);
    const size_t need = kb + vb + hb + scan_bytes;
    if (!temp) {
        temp_bytes = need;
        return hipSuccess;
    }
    if (temp_bytes < need) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    char *p = static_cast<char *>(temp);
    K *kt = reinterpret_cast<K *>(p);
    V *vt = vals_in ? reinterpret_cast<V *>(p + kb) : nullptr;
    int32_t *hist = reinterpret_cast<int32_t *>(p + kb + vb);
    void *scan_tmp = p + kb + vb + hb;
    const int passes = (int)ceil_div(end_bit > begin_bit ? end_bit - begin_bit : 0, 8);
    if (passes == 0) {
        if (dn) return hipErrorInvalidValue;  // callers sort at least one digit
        hipError_t e = hipMemcpyAsync(keys_out, keys_in, sizeof(K) * (size_t)n, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && vals_in)
            e = hipMemcpyAsync(vals_out, vals_in, sizeof(V) * (size_t)n, hipMemcpyDeviceToDevice, st);
        return e;
    }
    // ping-pong so that the last pass writes keys_out
    const K *src_k = keys_in;
    const V *src_v = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = begin_bit + 8 * pass;
        const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
        const bool to_out = ((passes - 1 - pass) % 2) == 0;
        K *dst_k = to_out ? keys_out : kt;
        V *dst_v = to_out ? vals_out : vt;
        k_digit_hist<K><<<(unsigned)tiles, PR_THREADS, 0, st>>>(src_k, n, dn, shift, bits, tiles, hist);
        size_t sb = scan_bytes;
        hipError_t e = scan<int32_t>(scan_tmp, sb, (const int32_t *)hist, hist, hist_n, false, st);
        if (e != hipSuccess) return e;
        if (vals_in)
            k_digit_scatter<K, V, true><<<(unsigned)tiles, PR_THREADS, 0, st>>>(src_k, src_v, dst_k, dst_v, n, dn,
                                                                              shift, bits, tiles, hist);
        else
            k_digit_scatter<K, V, false><<<(unsigned)tiles, PR_THREADS, 0, st>>>(src_k, nullptr, dst_k, nullptr, n,
                                                                               dn, shift, bits, tiles, hist);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        src_k = dst_k;
        src_v = dst_v;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------- fills

// Fills with kernels rather than hipMemsetAsync: memset nodes recorded into a captured graph are avoided.
template <typename T>
__global__ void k_fill(T *__restrict__ p, int64_t n, T v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T>
hipError_t fill(T *p, int64_t n, T v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t g = ceil_div(n, 256);
    k_fill<T><<<(unsigned)(g < 16384 ? g : 16384), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------- run-length encoding

template <typename K>
__global__ void k_run_heads(const K *__restrict__ keys, int64_t n, int32_t *__restrict__ head) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// run[i] = inclusive count of heads up to i; heads write their run's key and start.
template <typename K>
__global__ void k_run_write(const K *__restrict__ keys, int64_t n, const int32_t *__restrict__ run,
                            K *__restrict__ uniq, int32_t *__restrict__ start, int64_t *__restrict__ num_runs) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (i == 0 || keys[i] != keys[i - 1]) {
            uniq[run[i] - 1] = keys[i];
            start[run[i] - 1] = (int32_t)i;
        }
        if (i == n - 1) {
            *num_runs = run[i];
            start[run[i]] = (int32_t)n;
        }
    }
}

__global__ inline void k_run_counts(const int32_t *__restrict__ start, const int64_t *__restrict__ num_runs,
                                    int32_t *__restrict__ counts) {
    const int64_t m = *num_runs;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m; r += (int64_t)gridDim.x * blockDim.x)
        counts[r] = start[r + 1] - start[r];
}

// Runs of equal keys in sorted keys[0, n): unique keys in uniq, run lengths in counts, the number of runs in
// *num_runs (device memory).  n < 2^31.
template <typename K>
hipError_t run_length_encode(void *temp, size_t &temp_bytes, const K *keys, K *uniq, int32_t *counts,
                             int64_t *num_runs, int64_t n, hipStream_t st) {
    size_t scan_bytes = 0;
    scan<int32_t>(nullptr, scan_bytes, (const int32_t *)nullptr, (int32_t *)nullptr, n, true, st);
    const size_t ab = align256(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    const size_t sb = align256(sizeof(int32_t) * (size_t)(n + 1));
    const size_t need = ab + sb + scan_bytes;
    if (!temp) {
        temp_bytes = need;
        return hipSuccess;
    }
    if (temp_bytes < need) return hipErrorInvalidValue;
    if (n <= 0) return fill<int64_t>(num_runs, 1, 0, st);
    char *p = static_cast<char *>(temp);
    int32_t *run = reinterpret_cast<int32_t *>(p);
    int32_t *start = reinterpret_cast<int32_t *>(p + ab);
    const unsigned grid = (unsigned)(ceil_div(n, 256) < 65536 ? ceil_div(n, 256) : 65536);
    k_run_heads<K><<<grid, 256, 0, st>>>(keys, n, run);
    size_t s2 = scan_bytes;
    hipError_t e = scan<int32_t>(p + ab + sb, s2, (const int32_t *)run, run, n, true, st);
    if (e != hipSuccess) return e;
    k_run_write<K><<<grid, 256, 0, st>>>(keys, n, run, uniq, start, num_runs);
    k_run_counts<<<grid, 256, 0, st>>>(start, num_runs, counts);
    return hipGetLastError();
}

}  // namespace sbpprim
