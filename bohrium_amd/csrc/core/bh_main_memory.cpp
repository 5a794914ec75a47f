#include <bohrium/bh_main_memory.hpp>

#include <sys/mman.h>

#include <list>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace {

struct Entry {
    void *mem;
    uint64_t nbytes;
};

// Most-recently-freed first; an exact-size match is reused.
std::list<Entry> g_cache;
// Live allocations (address -> bytes): lets a caller check that a page range is this allocator's alone.
std::unordered_map<void *, uint64_t> g_live;
uint64_t g_cached_bytes = 0;
uint64_t g_limit = 1ull << 30;
std::mutex g_mu;

void *raw_alloc(uint64_t nbytes) {
    void *p = mmap(nullptr, nbytes == 0 ? 1 : nbytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (p == MAP_FAILED) throw std::runtime_error("bh_memory_malloc(): mmap of " + std::to_string(nbytes) + " bytes failed");
    return p;
}

void raw_free(void *mem, uint64_t nbytes) { munmap(mem, nbytes == 0 ? 1 : nbytes); }

void shrink_to(uint64_t limit) {
    while (g_cached_bytes > limit && !g_cache.empty()) {
        Entry e = g_cache.back();
        g_cache.pop_back();
        g_cached_bytes -= e.nbytes;
        raw_free(e.mem, e.nbytes);
    }
}

} // namespace

void *bh_memory_malloc(uint64_t nbytes) {
    {
        std::lock_guard<std::mutex> lock(g_mu);
        for (auto it = g_cache.begin(); it != g_cache.end(); ++it) {
            if (it->nbytes == nbytes) {
                void *m = it->mem;
                g_cached_bytes -= it->nbytes;
                g_cache.erase(it);
                g_live[m] = nbytes;
                return m;
            }
        }
    }
    void *m = raw_alloc(nbytes);
    std::lock_guard<std::mutex> lock(g_mu);
    g_live[m] = nbytes;
    return m;
}

bool bh_memory_owns(const void *mem, uint64_t nbytes) {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_live.find(const_cast<void *>(mem));
    return it != g_live.end() && it->second >= nbytes;
}

void bh_memory_free(void *mem, uint64_t nbytes) {
    if (mem == nullptr) return;
    std::lock_guard<std::mutex> lock(g_mu);
    g_live.erase(mem);
    if (nbytes > g_limit) {
        raw_free(mem, nbytes);
        return;
    }
    g_cache.push_front({mem, nbytes});
    g_cached_bytes += nbytes;
    shrink_to(g_limit);
}

void bh_set_malloc_cache_limit(uint64_t nbytes) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_limit = nbytes;
    shrink_to(g_limit);
}

void bh_data_malloc(bh_base *base) {
    if (base == nullptr || base->getDataPtr() != nullptr) return;
    base->resetDataPtr(bh_memory_malloc(static_cast<uint64_t>(base->nbytes())));
}

void bh_data_free(bh_base *base) {
    if (base == nullptr || base->getDataPtr() == nullptr) return;
    bh_memory_free(base->getDataPtr(), static_cast<uint64_t>(base->nbytes()));
    base->resetDataPtr(nullptr);
}
