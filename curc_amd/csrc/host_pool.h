// Pinned host memory for block results.
//
// A block's results (pseudogenomes, records, id lists) are copied from the
// device into host arrays that the caller reads until curc_block_free.
// Fresh pageable arrays cost more than the copies: the kernel zero-fills
// and maps every page on first touch, and the runtime stages pageable
// copies through 1 MiB pinned bounce buffers.  Arrays of at least
// POOL_MIN_BYTES come instead from a process-wide pool of pinned chunks
// (hipHostMalloc), recycled when a block is freed: after the first block of
// a given size, results land by direct DMA into pages that are already
// mapped.  Smaller arrays use malloc.  HostAlloc default-initialises
// elements, so resize() does not clear what a copy will overwrite.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <new>
#include <vector>

namespace curc {

class HostPool {
   public:
    static constexpr size_t POOL_MIN_BYTES = 1u << 20;

    static HostPool& get() {
        static HostPool* pool = new HostPool();  // lives until the process ends
        return *pool;
    }

    void* alloc(size_t bytes) {
        std::lock_guard<std::mutex> lk(mu_);
        // best fit among free chunks no more than twice the request
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            used_[p] = it->first;
            free_.erase(it);
            return p;
        }
        const size_t cap = bytes + bytes / 8;
        void* p = nullptr;
        if (hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess || !p) {
            // keep going with pageable memory rather than fail the block
            p = std::malloc(cap);
            if (!p) throw std::bad_alloc();
            pageable_[p] = cap;
            return p;
        }
        used_[p] = cap;
        total_ += cap;
        return p;
    }

    void release(void* p) {
        std::lock_guard<std::mutex> lk(mu_);
        auto pg = pageable_.find(p);
        if (pg != pageable_.end()) {
            pageable_.erase(pg);
            std::free(p);
            return;
        }
        auto it = used_.find(p);
        if (it == used_.end()) return;
        free_.emplace(it->second, p);
        used_.erase(it);
    }

    // Returns the free chunks to the system; chunks in use stay.
    void trim() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& kv : free_) {
            hipHostFree(kv.second);
            total_ -= kv.first;
        }
        free_.clear();
    }

    size_t pinned_bytes() {
        std::lock_guard<std::mutex> lk(mu_);
        return total_;
    }

   private:
    std::mutex mu_;
    std::multimap<size_t, void*> free_;  // capacity -> chunk
    std::map<void*, size_t> used_;       // chunk -> capacity
    std::map<void*, size_t> pageable_;   // fallback allocations
    size_t total_ = 0;
};

template <class T>
struct HostAlloc {
    using value_type = T;
    HostAlloc() noexcept = default;
    template <class U>
    HostAlloc(const HostAlloc<U>&) noexcept {}
    T* allocate(size_t n) {
        const size_t bytes = n * sizeof(T);
        if (bytes < HostPool::POOL_MIN_BYTES) {
            void* p = std::malloc(bytes ? bytes : 1);
            if (!p) throw std::bad_alloc();
            return static_cast<T*>(p);
        }
        return static_cast<T*>(HostPool::get().alloc(bytes));
    }
    void deallocate(T* p, size_t n) noexcept {
        if (n * sizeof(T) < HostPool::POOL_MIN_BYTES)
            std::free(p);
        else
            HostPool::get().release(p);
    }
    // default-initialise (no zero fill for trivial types)
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
    template <class U>
    bool operator==(const HostAlloc<U>&) const noexcept {
        return true;
    }
    template <class U>
    bool operator!=(const HostAlloc<U>&) const noexcept {
        return false;
    }
};

template <class T>
using hvec = std::vector<T, HostAlloc<T>>;

}  // namespace curc
