// Process-wide caches behind the library's allocations: device blocks, pinned host
// blocks and stream pairs.
//
// hipFree and hipHostFree wait for the whole device, and a stream costs milliseconds
// to create.  A server that admits and retires scenarios continuously would pay both
// for every scenario, and the device-wide waits would stall the batch of scenarios
// running beside it.  A freed block goes back to a free list of its size class and is
// handed out again for the next request of that class; a block that would take the
// cache past its limit is released at once.  Whoever returns a block or a stream must
// make sure that no queued GPU work still uses it (a context synchronises its streams
// before it returns them).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

namespace thy {

class Pool {
  public:
    // the pool outlives every context; it is never destroyed, so that no HIP call runs
    // after the runtime has shut down at process exit
    static Pool &get() {
        static Pool *p = new Pool();
        return *p;
    }

    hipError_t dev_alloc(void **p, size_t bytes) { return alloc(dev_, p, bytes); }
    void dev_free(void *p) { give_back(dev_, p); }
    hipError_t host_alloc(void **p, size_t bytes) { return alloc(host_, p, bytes); }
    void host_free(void *p) { give_back(host_, p); }

    // a pair of streams (default priority), cached or new; non-blocking, so that work
    // other code puts on the null stream does not order against them
    hipError_t get_streams(hipStream_t &a, hipStream_t &b) {
        {
            std::lock_guard<std::mutex> g(m_);
            if (!streams_.empty()) {
                a = streams_.back().first;
                b = streams_.back().second;
                streams_.pop_back();
                return hipSuccess;
            }
        }
        return new_streams(a, b);
    }

    // a new pair, not from the cache
    hipError_t new_streams(hipStream_t &a, hipStream_t &b) {
        a = b = nullptr;
        // the runtime gives a new stream the hardware queue that serves the fewest
        // streams; the two of a pair are created back to back (no other thread's stream
        // in between), so that unless one queue serves two streams fewer than every
        // other, they go to two different queues and their launches overlap
        std::lock_guard<std::mutex> g(create_m_);
        hipError_t e = hipStreamCreateWithFlags(&a, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&b, hipStreamNonBlocking);
        if (e != hipSuccess) {
            if (a) (void)hipStreamDestroy(a);
            a = b = nullptr;
        }
        return e;
    }
    // streams with no queued work left
    void put_streams(hipStream_t a, hipStream_t b) {
        {
            std::lock_guard<std::mutex> g(m_);
            if (streams_.size() < kMaxStreamPairs) {
                streams_.emplace_back(a, b);
                return;
            }
        }
        (void)hipStreamDestroy(a);
        (void)hipStreamDestroy(b);
    }

    // release every cached block and stream
    void trim() {
        std::vector<void *> d, h;
        std::vector<std::pair<hipStream_t, hipStream_t>> s;
        {
            std::lock_guard<std::mutex> g(m_);
            take_all(dev_, d);
            take_all(host_, h);
            s.swap(streams_);
        }
        for (void *p : d) (void)hipFree(p);
        for (void *p : h) (void)hipHostFree(p);
        for (auto &x : s) {
            (void)hipStreamDestroy(x.first);
            (void)hipStreamDestroy(x.second);
        }
    }

    size_t cached_device_bytes() {
        std::lock_guard<std::mutex> g(m_);
        return dev_.cached;
    }

  private:
    static constexpr size_t kMaxStreamPairs = 64;
    struct Heap {
        bool host = false;
        size_t cached = 0, limit = 0;
        std::unordered_map<size_t, std::vector<void *>> free;
        std::unordered_map<void *, size_t> live;
    };

    Pool() {
        dev_.host = false;
        host_.host = true;
        // device cache: THYSEA_POOL_MB, else an eighth of the device memory
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) tot = size_t(8) << 30;
        dev_.limit = tot / 8;
        host_.limit = size_t(1) << 30;
        if (const char *v = std::getenv("THYSEA_POOL_MB")) dev_.limit = (size_t)std::strtoull(v, nullptr, 10) << 20;
    }

    // size classes: powers of two up to 1 MiB, multiples of 2 MiB beyond
    static size_t size_class(size_t b) {
        if (b <= (size_t(1) << 20)) {
            size_t c = 256;
            while (c < b) c <<= 1;
            return c;
        }
        const size_t g = size_t(2) << 20;
        return (b + g - 1) / g * g;
    }

    hipError_t raw_alloc(Heap &h, void **p, size_t bytes) {
        return h.host ? hipHostMalloc(p, bytes, hipHostMallocDefault) : hipMalloc(p, bytes);
    }
    void raw_free(Heap &h, void *p) {
        if (h.host) (void)hipHostFree(p);
        else (void)hipFree(p);
    }

    hipError_t alloc(Heap &h, void **p, size_t bytes) {
        const size_t c = size_class(bytes);
        {
            std::lock_guard<std::mutex> g(m_);
            auto it = h.free.find(c);
            if (it != h.free.end() && !it->second.empty()) {
                *p = it->second.back();
                it->second.pop_back();
                h.cached -= c;
                h.live[*p] = c;
                return hipSuccess;
            }
        }
        hipError_t e = raw_alloc(h, p, c);
        if (e != hipSuccess) {
            // out of memory with blocks of other sizes cached: release them and retry
            (void)hipGetLastError();
            std::vector<void *> all;
            {
                std::lock_guard<std::mutex> g(m_);
                take_all(h, all);
            }
            if (all.empty()) return e;
            for (void *q : all) raw_free(h, q);
            e = raw_alloc(h, p, c);
            if (e != hipSuccess) return e;
        }
        std::lock_guard<std::mutex> g(m_);
        h.live[*p] = c;
        return hipSuccess;
    }

    void give_back(Heap &h, void *p) {
        if (!p) return;
        {
            std::lock_guard<std::mutex> g(m_);
            auto it = h.live.find(p);
            if (it == h.live.end()) return;  // not a block of this pool
            const size_t c = it->second;
            h.live.erase(it);
            if (h.cached + c <= h.limit) {
                h.free[c].push_back(p);
                h.cached += c;
                return;
            }
        }
        raw_free(h, p);
    }

    // (with m_ held)
    static void take_all(Heap &h, std::vector<void *> &out) {
        for (auto &kv : h.free)
            for (void *q : kv.second) out.push_back(q);
        h.free.clear();
        h.cached = 0;
    }

    std::mutex m_;
    std::mutex create_m_;  // stream pairs are created one pair at a time
    Heap dev_, host_;
    std::vector<std::pair<hipStream_t, hipStream_t>> streams_;
};

inline hipError_t dev_alloc(void **p, size_t bytes) { return Pool::get().dev_alloc(p, bytes); }
template <typename P> inline hipError_t dev_alloc(P **p, size_t bytes) {
    return Pool::get().dev_alloc(reinterpret_cast<void **>(p), bytes);
}
inline void dev_free(void *p) { Pool::get().dev_free(p); }

}  // namespace thy
