// CPU harness for the pinned pool's bookkeeping (csrc/host_pool.h) with a
// backend that hands out distinct, never-dereferenced address ranges and can
// be told to refuse after a number of slabs.
#include <stdint.h>

#include "host_pool.h"

namespace {
uintptr_t g_next = (uintptr_t)1 << 40;
long g_limit = -1, g_allocs = 0, g_releases = 0;
struct Fake {
    static void *alloc(size_t bytes) {
        if (g_limit >= 0 && g_allocs >= g_limit) return nullptr;
        g_allocs++;
        void *p = (void *)g_next;
        g_next += bytes + zkh::kSlabBytes;  // a gap: neighbouring slabs are not contiguous
        return p;
    }
    static void release(void *) { g_releases++; }
};
using Pool = zkh::HostPool<Fake>;
}  // namespace

extern "C" {
void *hp_new(long limit) {
    g_limit = limit;
    g_allocs = g_releases = 0;
    g_next = (uintptr_t)1 << 40;
    return new Pool();
}
void hp_delete(void *h) { delete (Pool *)h; }
int hp_reserve(void *h, size_t slabs) { return ((Pool *)h)->reserve(slabs); }
void *hp_alloc(void *h, size_t bytes) { return ((Pool *)h)->alloc(bytes); }
int hp_free(void *h, void *p) { return ((Pool *)h)->release(p); }
void hp_stats(void *h, size_t *slabs, size_t *used) { ((Pool *)h)->stats(slabs, used); }
long hp_backend_allocs() { return g_allocs; }
long hp_backend_releases() { return g_releases; }
size_t hp_chunk() { return zkh::kChunkBytes; }
size_t hp_slab() { return zkh::kSlabBytes; }
}
