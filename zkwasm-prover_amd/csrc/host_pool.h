// Bookkeeping of the pinned host pool: slabs of kSlabBytes, each split into
// kChunksPerSlab chunks of kChunkBytes (one Fr column of 2^22 rows).  A
// request of at most a chunk takes a chunk, preferring slabs already partly
// used; a request of at most a slab takes a whole slab; anything larger is
// refused.  Slabs are never returned to the system while the pool lives, so
// a proof's host columns recycle the same pinned pages (the reference's
// pinned buffer cache, src/hugetlb.rs:17-229).  The backing allocation is a
// template parameter so that this logic runs on the CPU alone in tests.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <mutex>
#include <unordered_map>
#include <vector>

namespace zkh {

constexpr size_t kChunkBytes = (size_t)1 << 27;
constexpr int kChunksPerSlab = 2;
constexpr size_t kSlabBytes = kChunkBytes * kChunksPerSlab;

// Backend: static void *alloc(size_t bytes) (null on failure), static void release(void *)
template <class Backend>
class HostPool {
   public:
    ~HostPool() {
        for (auto &s : slabs_) Backend::release(s.base);
    }

    // grow to at least `slabs` slabs; false if the backend refused one
    bool reserve(size_t slabs) {
        std::lock_guard<std::mutex> g(mu_);
        while (slabs_.size() < slabs)
            if (!grow()) return false;
        return true;
    }

    void *alloc(size_t bytes) {
        if (bytes == 0 || bytes > kSlabBytes) return nullptr;
        std::lock_guard<std::mutex> g(mu_);
        if (bytes <= kChunkBytes) {
            int si = -1;
            for (size_t i = 0; i < slabs_.size(); i++)  // a partly used slab first
                if (slabs_[i].used > 0 && slabs_[i].used < kChunksPerSlab && !slabs_[i].whole) {
                    si = (int)i;
                    break;
                }
            if (si < 0) si = free_slab();
            if (si < 0) return nullptr;
            Slab &s = slabs_[si];
            for (int c = 0; c < kChunksPerSlab; c++)
                if (!(s.mask >> c & 1)) {
                    void *p = s.base + (size_t)c * kChunkBytes;
                    owner_[p] = si;  // first: if it throws, nothing is marked
                    s.mask |= 1u << c;
                    s.used++;
                    return p;
                }
            return nullptr;  // unreachable: used < kChunksPerSlab
        }
        int si = free_slab();
        if (si < 0) return nullptr;
        Slab &s = slabs_[si];
        owner_[s.base] = si;  // first: if it throws, nothing is marked
        s.whole = true;
        s.used = kChunksPerSlab;
        s.mask = (1u << kChunksPerSlab) - 1;
        return s.base;
    }

    // false if p is not a live allocation of this pool
    bool release(void *p) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = owner_.find(p);
        if (it == owner_.end()) return false;
        Slab &s = slabs_[it->second];
        if (s.whole) {
            s.whole = false;
            s.used = 0;
            s.mask = 0;
        } else {
            int c = (int)(((uint8_t *)p - s.base) / kChunkBytes);
            s.mask &= ~(1u << c);
            s.used--;
        }
        owner_.erase(it);
        return true;
    }

    void stats(size_t *slabs, size_t *chunks_in_use) {
        std::lock_guard<std::mutex> g(mu_);
        size_t used = 0;
        for (auto &s : slabs_) used += (size_t)s.used;
        if (slabs) *slabs = slabs_.size();
        if (chunks_in_use) *chunks_in_use = used;
    }

   private:
    struct Slab {
        uint8_t *base;
        unsigned mask;  // chunk c in use
        int used;       // chunks in use
        bool whole;     // taken by one slab-sized allocation
    };

    bool grow() {
        void *p = Backend::alloc(kSlabBytes);
        if (!p) return false;
        try {
            slabs_.push_back(Slab{(uint8_t *)p, 0u, 0, false});
        } catch (...) {
            Backend::release(p);
            throw;
        }
        return true;
    }
    int free_slab() {
        for (size_t i = 0; i < slabs_.size(); i++)
            if (slabs_[i].used == 0) return (int)i;
        if (!grow()) return -1;
        return (int)slabs_.size() - 1;
    }

    std::mutex mu_;
    std::vector<Slab> slabs_;
    std::unordered_map<void *, int> owner_;
};

}  // namespace zkh
