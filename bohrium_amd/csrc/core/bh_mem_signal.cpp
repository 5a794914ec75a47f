/* The memory-signal dispatcher (include/bh_mem_signal.h): the reference's interface
 * (core/bh_mem_signal.cpp:121-196) on a POSIX SIGSEGV handler instead of libsigsegv.
 *
 * The attached segments form a sorted, disjoint table that is replaced as a whole on every
 * attach/detach (under a mutex) and read without a lock by the signal handler.  A writer frees
 * the table it replaced only when no handler is reading: handlers count themselves in `readers`
 * around the lookup, and leave before they call the callback, so a callback may itself detach.
 */
#include "../../../include/bh_mem_signal.h"

#include <signal.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace {

struct Segment {
    uintptr_t begin, end;  // [begin, end)
    void *idx;
    bh_mem_signal_callback_t callback;
};
using Table = std::vector<Segment>;

std::mutex g_mu;
std::atomic<const Table *> g_table{nullptr};
std::atomic<int> g_readers{0};
struct sigaction g_prev;
bool g_installed = false;

const Segment *find(const Table &t, uintptr_t a) {
    auto it = std::upper_bound(t.begin(), t.end(), a, [](uintptr_t x, const Segment &s) { return x < s.begin; });
    if (it == t.begin()) return nullptr;
    --it;
    return a < it->end ? &*it : nullptr;
}

// Hand the fault to the handler installed before ours; with none, restore the default action:
// a fault re-executes the access and ends the process, a signal sent by kill() is raised again.
void chain(int sig, siginfo_t *si, void *uc) {
    if ((g_prev.sa_flags & SA_SIGINFO) && g_prev.sa_sigaction != nullptr) {
        g_prev.sa_sigaction(sig, si, uc);
        return;
    }
    if (!(g_prev.sa_flags & SA_SIGINFO) && g_prev.sa_handler != SIG_DFL && g_prev.sa_handler != SIG_IGN) {
        g_prev.sa_handler(sig);
        return;
    }
    struct sigaction dfl;
    sigemptyset(&dfl.sa_mask);
    dfl.sa_flags = 0;
    dfl.sa_handler = SIG_DFL;
    sigaction(sig, &dfl, nullptr);
    if (si == nullptr || si->si_code <= 0) raise(sig);
}

void handler(int sig, siginfo_t *si, void *uc) {
    if (si != nullptr && si->si_code > 0) {  // raised by the kernel for a memory access
        const uintptr_t a = reinterpret_cast<uintptr_t>(si->si_addr);
        Segment hit{0, 0, nullptr, nullptr};
        g_readers.fetch_add(1);
        const Table *t = g_table.load();
        if (t != nullptr) {
            const Segment *s = find(*t, a);
            if (s != nullptr) hit = *s;
        }
        g_readers.fetch_sub(1);
        if (hit.callback != nullptr && hit.callback(si->si_addr, hit.idx) != 0) return;
    }
    chain(sig, si, uc);
}

// Install our handler unless it is the current one (another library may have installed its own
// since, e.g. Python's faulthandler: ours then goes in front of it and chains to it).
void install_locked() {
    struct sigaction cur;
    if (sigaction(SIGSEGV, nullptr, &cur) != 0) return;
    if ((cur.sa_flags & SA_SIGINFO) && cur.sa_sigaction == handler) return;
    struct sigaction sa;
    sigemptyset(&sa.sa_mask);
    sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
    sa.sa_sigaction = handler;
    if (sigaction(SIGSEGV, &sa, &g_prev) == 0) g_installed = true;
}

// Replace the table (caller holds g_mu) and free the old one once no handler reads it.
void publish(Table *next) {
    const Table *old = g_table.exchange(next);
    while (g_readers.load() != 0) {
    }
    delete old;
}

} // namespace

extern "C" {

void bh_mem_signal_init(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    install_locked();
}

void bh_mem_signal_shutdown(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    const Table *t = g_table.load();
    if (t != nullptr && !t->empty() && std::getenv("BH_MEM_WARN") != nullptr) {
        std::printf("MEM_WARN: bh_mem_signal_shutdown() - not all attached memory segments are detached!\n");
        for (const Segment &s : *t)
            std::printf("%p{addr: %p - %p}\n", s.idx, reinterpret_cast<void *>(s.begin), reinterpret_cast<void *>(s.end));
    }
    if (g_installed) {
        struct sigaction cur;
        if (sigaction(SIGSEGV, nullptr, &cur) == 0 && (cur.sa_flags & SA_SIGINFO) && cur.sa_sigaction == handler)
            sigaction(SIGSEGV, &g_prev, nullptr);
        g_installed = false;
    }
}

int bh_mem_signal_attach(void *idx, void *addr, uint64_t size, bh_mem_signal_callback_t callback) {
    if (addr == nullptr || size == 0 || callback == nullptr) return -1;
    std::lock_guard<std::mutex> lock(g_mu);
    install_locked();
    const uintptr_t b = reinterpret_cast<uintptr_t>(addr), e = b + size;
    const Table *t = g_table.load();
    Table *next = t != nullptr ? new Table(*t) : new Table();
    auto it = std::lower_bound(next->begin(), next->end(), b, [](const Segment &s, uintptr_t x) { return s.begin < x; });
    const bool overlaps_next = it != next->end() && it->begin < e;
    const bool overlaps_prev = it != next->begin() && std::prev(it)->end > b;
    if (overlaps_next || overlaps_prev) {
        delete next;
        return -1;
    }
    next->insert(it, Segment{b, e, idx, callback});
    publish(next);
    return 0;
}

void bh_mem_signal_detach(const void *addr) {
    std::lock_guard<std::mutex> lock(g_mu);
    const Table *t = g_table.load();
    if (t == nullptr) return;
    const Segment *s = find(*t, reinterpret_cast<uintptr_t>(addr));
    if (s == nullptr) return;
    Table *next = new Table(*t);
    next->erase(next->begin() + (s - t->data()));
    publish(next);
}

int bh_mem_signal_exist(const void *addr) {
    std::lock_guard<std::mutex> lock(g_mu);
    const Table *t = g_table.load();
    return t != nullptr && find(*t, reinterpret_cast<uintptr_t>(addr)) != nullptr;
}

void bh_mem_signal_pprint_db(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    std::printf("bh_mem_signal contains:\n");
    const Table *t = g_table.load();
    if (t == nullptr) return;
    for (const Segment &s : *t)
        std::printf("%p{addr: %p - %p}\n", s.idx, reinterpret_cast<void *>(s.begin), reinterpret_cast<void *>(s.end));
    std::fflush(stdout);
}

} // extern "C"
