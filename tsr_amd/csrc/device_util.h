// Device-side helpers: HIP error checking, a 64-bit open-addressing hash map
// keyed by octree location codes, device buffers, and rocPRIM-backed radix
// sort / scan wrappers used by the tree build, occupancy and DMC stages.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <hipcub/hipcub.hpp>

#include <string>
#include <utility>

#include "common.h"
#include "tsr_capi_internal.h"

namespace tsr {

inline void hipCheck(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw tsr_error(TSR_ERR_OUT_OF_MEMORY, std::string(what) + ": out of device memory");
        throw tsr_error(TSR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}
#define HIPCHECK(x) ::tsr::hipCheck((x), #x)

// Owning device buffer (hipMallocAsync-free: plain hipMalloc / hipFree).
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) HIPCHECK(hipMalloc((void **)&p, sizeof(T) * count));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // grow-only: keeps the buffer when it already holds `count` elements
    T *ensure(size_t count) {
        if (count > n || !p) alloc(std::max<size_t>(count, 1));
        return p;
    }
    size_t bytes() const { return sizeof(T) * n; }
};

// ---------------------------------------------------------------------------
// Hash map u64 location code -> u32 value.  Empty slot key = 0 (location codes
// are >= 1).  Linear probing over a power-of-two table.
// ---------------------------------------------------------------------------
TSR_HD u64 hashMix(u64 k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

struct HashView {
    u64 *keys;
    u32 *vals;
    u64 mask;

    // returns true if inserted, false if already present; sets *full on overflow
    __device__ bool insert(u64 key, u32 val, u32 *full) const {
        u64 s = hashMix(key) & mask;
        for (u64 probe = 0; probe <= mask; ++probe) {
            const u64 prev = atomicCAS((unsigned long long *)&keys[s], 0ull, (unsigned long long)key);
            if (prev == 0ull) {
                vals[s] = val;
                return true;
            }
            if (prev == key) return false;
            s = (s + 1) & mask;
        }
        atomicOr(full, 1u);
        return false;
    }
    __device__ u32 find(u64 key) const {
        u64 s = hashMix(key) & mask;
        for (u64 probe = 0; probe <= mask; ++probe) {
            const u64 k = keys[s];
            if (k == key) return vals[s];
            if (k == 0ull) return INVALID;
            s = (s + 1) & mask;
        }
        return INVALID;
    }
    __device__ bool contains(u64 key) const {
        u64 s = hashMix(key) & mask;
        for (u64 probe = 0; probe <= mask; ++probe) {
            const u64 k = keys[s];
            if (k == key) return true;
            if (k == 0ull) return false;
            s = (s + 1) & mask;
        }
        return false;
    }
};

struct DHash {
    DBuf<u64> keys;
    DBuf<u32> vals;
    u64 capacity = 0;
    void alloc(u64 minEntries, hipStream_t st) {
        u64 cap = 1024;
        while (cap < 2 * minEntries) cap <<= 1;
        capacity = cap;
        keys.alloc(cap);
        vals.alloc(cap);
        HIPCHECK(hipMemsetAsync(keys.p, 0, keys.bytes(), st));
    }
    HashView view() const { return HashView{keys.p, vals.p, capacity - 1}; }
};

// ---------------------------------------------------------------------------
// Sort / scan wrappers (hipCUB on rocPRIM).  Temporary storage is kept in a
// growing scratch buffer owned by the caller.
// ---------------------------------------------------------------------------
struct Scratch {
    DBuf<unsigned char> buf;
    void *get(size_t bytes) {
        if (bytes > buf.n) buf.alloc(bytes + bytes / 4 + 256);
        return buf.p;
    }
};

template <typename K, typename V>
void sortPairs(Scratch &sc, const K *kin, K *kout, const V *vin, V *vout, size_t n, int beginBit, int endBit,
               hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, (int)n, beginBit, endBit, st));
    void *d = sc.get(tmp);
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(d, tmp, kin, kout, vin, vout, (int)n, beginBit, endBit, st));
}

template <typename K>
void sortKeys(Scratch &sc, const K *kin, K *kout, size_t n, int beginBit, int endBit, hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, kin, kout, (int)n, beginBit, endBit, st));
    void *d = sc.get(tmp);
    HIPCHECK(hipcub::DeviceRadixSort::SortKeys(d, tmp, kin, kout, (int)n, beginBit, endBit, st));
}

template <typename T>
void exclusiveScan(Scratch &sc, const T *in, T *out, size_t n, hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, st));
    void *d = sc.get(tmp);
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(d, tmp, in, out, (int)n, st));
}

template <typename T>
void inclusiveScan(Scratch &sc, const T *in, T *out, size_t n, hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, in, out, (int)n, st));
    void *d = sc.get(tmp);
    HIPCHECK(hipcub::DeviceScan::InclusiveSum(d, tmp, in, out, (int)n, st));
}

template <typename T>
void uniqueKeys(Scratch &sc, const T *in, T *out, u32 *numOut, size_t n, hipStream_t st) {
    size_t tmp = 0;
    HIPCHECK(hipcub::DeviceSelect::Unique(nullptr, tmp, in, out, numOut, (int)n, st));
    void *d = sc.get(tmp);
    HIPCHECK(hipcub::DeviceSelect::Unique(d, tmp, in, out, numOut, (int)n, st));
}

template <typename T>
T readScalar(const T *dptr, hipStream_t st) {
    T v;
    HIPCHECK(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return v;
}

inline unsigned gridFor(size_t n, unsigned block) {
    const size_t g = (n + block - 1) / block;
    return (unsigned)(g == 0 ? 1 : g);
}

}  // namespace tsr
