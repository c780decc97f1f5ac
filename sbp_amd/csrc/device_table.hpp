// Per-wavefront open-addressing histogram (block id -> out count, in count) in LDS or global memory.
// The slots claimed for the current item are listed in `used`, so clearing costs O(distinct keys) rather
// than O(capacity).  Layout of one table of capacity C: keys[C], outc[C], inc[C], used[C], n_used[1].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace sbpdev {

constexpr int EMPTY = -1;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

struct Table {
    int *keys;
    int *outc;
    int *inc;
    int *used;
    int *n_used;
    unsigned mask;
    int shift;  // 32 - log2(cap)
    int cap;
};

// ints of one table of capacity 2^cap_log2 whose used-slot list holds used_cap entries (default: every slot)
__host__ __device__ constexpr int64_t table_ints(int cap_log2, int64_t used_cap = -1) {
    return 3 * ((int64_t)1 << cap_log2) + (used_cap < 0 ? ((int64_t)1 << cap_log2) : used_cap) + 1;
}

// A table takes at most used_cap distinct keys (default: its capacity).
__device__ __forceinline__ Table table_at(int *base, int cap_log2, int used_cap = -1) {
    Table t;
    t.cap = 1 << cap_log2;
    t.shift = 32 - cap_log2;
    t.keys = base;
    t.outc = t.keys + t.cap;
    t.inc = t.outc + t.cap;
    t.used = t.inc + t.cap;
    t.n_used = t.used + (used_cap < 0 ? t.cap : used_cap);
    t.mask = (unsigned)t.cap - 1;
    return t;
}

__device__ __forceinline__ unsigned hash_slot(const Table &t, int key) {
    return ((unsigned)key * 0x9E3779B1u) >> t.shift;
}

// Probes are bounded by the table size: a table may be filled to its last slot (class 1 on 4 lanes: at most 32
// distinct blocks in 32 slots), where a lookup of an absent key meets no empty slot.  An insertion always finds
// room, since a table is sized for the most distinct keys an item can have.
__device__ __forceinline__ void table_add(const Table &t, int key, int dout, int din) {
    unsigned slot = hash_slot(t, key);
    for (int n = 0; n < t.cap; ++n) {
        int prev = atomicCAS(&t.keys[slot], EMPTY, key);
        if (prev == EMPTY) t.used[atomicAdd(t.n_used, 1)] = (int)slot;
        if (prev == EMPTY || prev == key) {
            if (dout) atomicAdd(&t.outc[slot], dout);
            if (din) atomicAdd(&t.inc[slot], din);
            return;
        }
        slot = (slot + 1) & t.mask;
    }
}

__device__ __forceinline__ void table_get(const Table &t, int key, int &o, int &i) {
    unsigned slot = hash_slot(t, key);
    for (int n = 0; n < t.cap; ++n) {
        int k = t.keys[slot];
        if (k == key) {
            o = t.outc[slot];
            i = t.inc[slot];
            return;
        }
        if (k == EMPTY) break;
        slot = (slot + 1) & t.mask;
    }
    o = 0;  // absent (an empty slot, or a full table without the key)
    i = 0;
}

__device__ __forceinline__ int table_size(const Table &t) { return *t.n_used; }
// The i-th distinct key of the current item with its counts.
__device__ __forceinline__ void table_entry(const Table &t, int i, int &key, int &o, int &in) {
    const int slot = t.used[i];
    key = t.keys[slot];
    o = t.outc[slot];
    in = t.inc[slot];
}

// Histogram indexed by the key itself (keys in [0, capacity)): cnt[key] packs the out count (low 32 bits) and the
// in count (high 32 bits), so one 64-bit atomic adds both and its old value tells the first insertion.  No
// probing; the keys touched by the current item are listed in used.  Layout: cnt[capacity] (8-byte aligned),
// used[used_cap], n_used[1].
struct DenseTable {
    unsigned long long *cnt;
    int *used;
    int *n_used;
};

__host__ __device__ constexpr int64_t dense_table_ints(int64_t capacity, int64_t used_cap) {
    return 2 * capacity + used_cap + 1;
}

__device__ __forceinline__ DenseTable dense_table_at(int *base, int capacity) {
    DenseTable t;
    t.cnt = reinterpret_cast<unsigned long long *>(base);
    t.used = base + 2 * capacity;
    t.n_used = t.used + capacity;
    return t;
}

__device__ __forceinline__ void table_add(const DenseTable &t, int key, int dout, int din) {
    const unsigned long long add = (unsigned long long)(unsigned)dout | ((unsigned long long)(unsigned)din << 32);
    if (atomicAdd(&t.cnt[key], add) == 0ull) t.used[atomicAdd(t.n_used, 1)] = key;
}

__device__ __forceinline__ void table_get(const DenseTable &t, int key, int &o, int &i) {
    const unsigned long long v = t.cnt[key];
    o = (int)(unsigned)(v & 0xffffffffull);
    i = (int)(unsigned)(v >> 32);
}

__device__ __forceinline__ int table_size(const DenseTable &t) { return *t.n_used; }

__device__ __forceinline__ void table_entry(const DenseTable &t, int i, int &key, int &o, int &in) {
    key = t.used[i];
    table_get(t, key, o, in);
}

__device__ __forceinline__ void table_clear_all(const Table &t) {
    for (int i = lane_id(); i < t.cap; i += WAVE) {
        t.keys[i] = EMPTY;
        t.outc[i] = 0;
        t.inc[i] = 0;
    }
    if (lane_id() == 0) *t.n_used = 0;
    wave_sync();
}

__device__ __forceinline__ void table_clear_used(const Table &t) {
    wave_sync();
    const int n = *t.n_used;
    for (int i = lane_id(); i < n; i += WAVE) {
        int slot = t.used[i];
        t.keys[slot] = EMPTY;
        t.outc[slot] = 0;
        t.inc[slot] = 0;
    }
    wave_sync();
    if (lane_id() == 0) *t.n_used = 0;
    wave_sync();
}

}  // namespace sbpdev
