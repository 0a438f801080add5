// Device-side helpers: HIP error checking, a 64-bit open-addressing hash map
// keyed by octree location codes, device buffers, and radix sort / scan / select /
// reduce wrappers over rocPRIM used by the tree build, occupancy and DMC stages.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <cstring>
#include <limits>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

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

// Device bytes this process holds through the library's buffers (hipMalloc'd blocks,
// cached ones included) and their high-water mark since the last reset.
struct MemCounter {
    std::atomic<size_t> current{0}, peak{0};
    void add(size_t b) {
        const size_t now = current.fetch_add(b) + b;
        size_t p = peak.load();
        while (now > p && !peak.compare_exchange_weak(p, now)) {
        }
    }
    void sub(size_t b) { current.fetch_sub(b); }
};
inline MemCounter &memCounter() {
    static MemCounter m;
    return m;
}
inline hipError_t trackedMalloc(void **p, size_t bytes) {
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) memCounter().add(bytes);
    return e;
}
inline void trackedFree(void *p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    memCounter().sub(bytes);
}

// Device blocks kept for reuse by one context.  hipFree waits for the whole device and
// hipMalloc of large blocks maps memory, so the stages' many temporaries would stall the
// stream at every scope exit; instead a block released while its context's entry point
// runs goes back to the context's cache and the next allocation of a similar size takes it.
// All of a context's device work is ordered on its one stream, so a block handed out again
// is only touched by work queued after the work that used it.  Blocks released outside such
// a scope (handle destructors, other contexts) are freed as before.
class DevicePool {
  public:
    explicit DevicePool(int device) : device_(device) {}
    DevicePool(const DevicePool &) = delete;
    DevicePool &operator=(const DevicePool &) = delete;
    ~DevicePool() { trim(); }
    // block size for a request: 256-byte multiples below 2 MiB, 2 MiB multiples above
    static size_t roundUp(size_t bytes) {
        const size_t g = bytes >= (2u << 20) ? (2u << 20) : 256u;
        return (bytes + g - 1) / g * g;
    }
    // a cached block of at least `bytes` and at most 25% (+2 MiB) larger, or null
    void *take(size_t bytes, size_t &blockBytes) {
        std::lock_guard<std::mutex> g(m_);
        auto it = cached_.lower_bound(bytes);
        if (it == cached_.end() || it->first > bytes + bytes / 4 + (2u << 20)) return nullptr;
        void *q = it->second;
        blockBytes = it->first;
        cachedBytes_ -= it->first;
        cached_.erase(it);
        return q;
    }
    // keep the block, or free it when the cache would grow past its cap
    void give(void *q, size_t blockBytes) {
        {
            std::lock_guard<std::mutex> g(m_);
            if (cachedBytes_ + blockBytes <= kCapBytes) {
                cached_.insert({blockBytes, q});
                cachedBytes_ += blockBytes;
                return;
            }
        }
        trackedFree(q, blockBytes);
    }
    // free every cached block (device `device_` made current, then restored)
    void trim() {
        std::lock_guard<std::mutex> g(m_);
        if (cached_.empty()) return;
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(device_);
        for (auto &kv : cached_) trackedFree(kv.second, kv.first);
        cached_.clear();
        cachedBytes_ = 0;
        (void)hipSetDevice(prev);
    }

  private:
    static constexpr size_t kCapBytes = size_t(64) << 30;
    int device_;
    std::mutex m_;
    std::multimap<size_t, void *> cached_;
    size_t cachedBytes_ = 0;
};

// Every pool of the process, so that an out-of-memory retry can give back the cached blocks
// of all of them (a context's second pool caches the blocks of its second stream's work).
inline std::mutex &poolRegistryMutex() {
    static std::mutex m;
    return m;
}
inline std::vector<std::weak_ptr<DevicePool>> &poolRegistry() {
    static std::vector<std::weak_ptr<DevicePool>> r;
    return r;
}
inline std::shared_ptr<DevicePool> makePool(int device) {
    auto p = std::make_shared<DevicePool>(device);
    std::lock_guard<std::mutex> g(poolRegistryMutex());
    auto &r = poolRegistry();
    r.erase(std::remove_if(r.begin(), r.end(), [](const std::weak_ptr<DevicePool> &w) { return w.expired(); }),
            r.end());
    r.push_back(p);
    return p;
}
inline void trimAllPools() {
    std::vector<std::shared_ptr<DevicePool>> live;
    {
        std::lock_guard<std::mutex> g(poolRegistryMutex());
        for (const auto &w : poolRegistry())
            if (auto p = w.lock()) live.push_back(std::move(p));
    }
    for (auto &p : live) p->trim();
}

// the pool of the context whose entry point runs on this thread (null outside one)
inline std::shared_ptr<DevicePool> &currentPool() {
    thread_local std::shared_ptr<DevicePool> pool;
    return pool;
}
struct PoolScope {
    std::shared_ptr<DevicePool> prev;
    explicit PoolScope(std::shared_ptr<DevicePool> pool) : prev(currentPool()) { currentPool() = std::move(pool); }
    ~PoolScope() { currentPool() = std::move(prev); }
    PoolScope(const PoolScope &) = delete;
    PoolScope &operator=(const PoolScope &) = delete;
};

// Owning device buffer: hipMalloc / hipFree, through the current context's DevicePool
// inside an entry point.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    size_t blockBytes = 0;
    std::shared_ptr<DevicePool> owner;   // pool the block came from (null: plain hipMalloc)
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), blockBytes(o.blockBytes), owner(std::move(o.owner)) {
        o.p = nullptr;
        o.n = 0;
        o.blockBytes = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            blockBytes = o.blockBytes;
            owner = std::move(o.owner);
            o.p = nullptr;
            o.n = 0;
            o.blockBytes = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (!count) return;
        const std::shared_ptr<DevicePool> &pool = currentPool();
        if (!pool) {
            HIPCHECK(trackedMalloc((void **)&p, sizeof(T) * count));
            blockBytes = sizeof(T) * count;
            return;
        }
        const size_t want = DevicePool::roundUp(sizeof(T) * count);
        p = (T *)pool->take(want, blockBytes);
        if (!p) {
            // TSR_TRACE: fresh blocks of 1 GiB or more, and out-of-memory retries, on stderr
            static const bool trace = std::getenv("TSR_TRACE") != nullptr;
            const auto t0 = std::chrono::steady_clock::now();
            hipError_t e = trackedMalloc((void **)&p, want);
            bool retried = false;
            if (e == hipErrorOutOfMemory) {   // give every pool's cached blocks back and try once more
                (void)hipGetLastError();
                trimAllPools();
                e = trackedMalloc((void **)&p, want);
                retried = true;
            }
            if (trace && (want >= (size_t(1) << 30) || retried))
                std::fprintf(stderr, "[tsr memory] fresh block %.2f GiB%s: %.2f ms (held %.1f GiB)\n",
                             (double)want / (1u << 30), retried ? " after an out-of-memory trim" : "",
                             std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
                             (double)memCounter().current.load() / (1u << 30));
            if (e != hipSuccess) p = nullptr;
            HIPCHECK(e);
            blockBytes = want;
        }
        owner = pool;
    }
    void release() {
        if (p) {
            if (owner && owner == currentPool()) owner->give(p, blockBytes);
            else trackedFree(p, blockBytes);
        }
        p = nullptr;
        n = 0;
        blockBytes = 0;
        owner.reset();
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
    // the value slot of key, or null
    __device__ u32 *slot(u64 key) const {
        u64 s = hashMix(key) & mask;
        for (u64 probe = 0; probe <= mask; ++probe) {
            const u64 k = keys[s];
            if (k == key) return &vals[s];
            if (k == 0ull) return nullptr;
            s = (s + 1) & mask;
        }
        return nullptr;
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
// Sort / scan / select / reduce wrappers over rocPRIM's device primitives.  Temporary
// storage is kept in a growing scratch buffer owned by the caller.
// ---------------------------------------------------------------------------
struct Scratch {
    DBuf<unsigned char> buf;
    void *get(size_t bytes) {
        if (bytes > buf.n) buf.alloc(bytes + bytes / 4 + 256);
        return buf.p;
    }
};

// stable LSD radix sort of (key, value) pairs on bits [beginBit, endBit)
template <typename K, typename V>
void sortPairs(Scratch &sc, const K *kin, K *kout, const V *vin, V *vout, size_t n, int beginBit, int endBit,
               hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, n, (unsigned)beginBit, (unsigned)endBit,
                                       st));
    HIPCHECK(rocprim::radix_sort_pairs(sc.get(tmp), tmp, kin, kout, vin, vout, n, (unsigned)beginBit,
                                       (unsigned)endBit, st));
}

template <typename K>
void sortKeys(Scratch &sc, const K *kin, K *kout, size_t n, int beginBit, int endBit, hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(rocprim::radix_sort_keys(nullptr, tmp, kin, kout, n, (unsigned)beginBit, (unsigned)endBit, st));
    HIPCHECK(rocprim::radix_sort_keys(sc.get(tmp), tmp, kin, kout, n, (unsigned)beginBit, (unsigned)endBit, st));
}

template <typename T>
void exclusiveScan(Scratch &sc, const T *in, T *out, size_t n, hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(rocprim::exclusive_scan(nullptr, tmp, in, out, T(0), n, rocprim::plus<T>(), st));
    HIPCHECK(rocprim::exclusive_scan(sc.get(tmp), tmp, in, out, T(0), n, rocprim::plus<T>(), st));
}

template <typename T>
void inclusiveScan(Scratch &sc, const T *in, T *out, size_t n, hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(rocprim::inclusive_scan(nullptr, tmp, in, out, n, rocprim::plus<T>(), st));
    HIPCHECK(rocprim::inclusive_scan(sc.get(tmp), tmp, in, out, n, rocprim::plus<T>(), st));
}

// distinct consecutive values of in (n) into out, their number into *numOut (device)
template <typename T>
void uniqueKeys(Scratch &sc, const T *in, T *out, u32 *numOut, size_t n, hipStream_t st) {
    size_t tmp = 0;
    HIPCHECK(rocprim::unique(nullptr, tmp, in, out, numOut, n, rocprim::equal_to<T>(), st));
    HIPCHECK(rocprim::unique(sc.get(tmp), tmp, in, out, numOut, n, rocprim::equal_to<T>(), st));
}

// in[i] with flags[i] != 0, in order, into out; their number into *numOut (device)
template <typename T, typename F>
void selectFlagged(Scratch &sc, const T *in, const F *flags, T *out, u32 *numOut, size_t n, hipStream_t st) {
    size_t tmp = 0;
    HIPCHECK(rocprim::select(nullptr, tmp, in, flags, out, numOut, n, st));
    HIPCHECK(rocprim::select(sc.get(tmp), tmp, in, flags, out, numOut, n, st));
}

// in[i] with pred(in[i]), in order, into out; their number into *numOut (device)
template <typename T, typename Pred>
void selectIf(Scratch &sc, const T *in, T *out, u32 *numOut, size_t n, Pred pred, hipStream_t st) {
    size_t tmp = 0;
    HIPCHECK(rocprim::select(nullptr, tmp, in, out, numOut, n, pred, st));
    HIPCHECK(rocprim::select(sc.get(tmp), tmp, in, out, numOut, n, pred, st));
}

// sum (fixed reduction tree: deterministic) / maximum of in (n) into *out (device)
template <typename T, typename R>
void reduceSum(Scratch &sc, const T *in, R *out, size_t n, hipStream_t st) {
    size_t tmp = 0;
    HIPCHECK(rocprim::reduce(nullptr, tmp, in, out, R(0), n, rocprim::plus<R>(), st));
    HIPCHECK(rocprim::reduce(sc.get(tmp), tmp, in, out, R(0), n, rocprim::plus<R>(), st));
}
template <typename T>
void reduceMax(Scratch &sc, const T *in, T *out, size_t n, hipStream_t st) {
    size_t tmp = 0;
    HIPCHECK(rocprim::reduce(nullptr, tmp, in, out, std::numeric_limits<T>::lowest(), n, rocprim::maximum<T>(), st));
    HIPCHECK(rocprim::reduce(sc.get(tmp), tmp, in, out, std::numeric_limits<T>::lowest(), n, rocprim::maximum<T>(),
                             st));
}

// runs of equal consecutive values: their values, lengths and number (device)
template <typename T>
void runLengthEncode(Scratch &sc, const T *in, T *unique, u32 *counts, u32 *runs, size_t n, hipStream_t st) {
    size_t tmp = 0;
    HIPCHECK(rocprim::run_length_encode(nullptr, tmp, in, (unsigned)n, unique, counts, runs, st));
    HIPCHECK(rocprim::run_length_encode(sc.get(tmp), tmp, in, (unsigned)n, unique, counts, runs, st));
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
