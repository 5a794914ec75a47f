/* Exercises libbh's bh_mem_signal dispatcher (include/bh_mem_signal.h) on the CPU.
 *   mem_signal_check basic     attach/exist/overlap, a fault on PROT_NONE pages and a write to
 *                              PROT_READ pages reach the callback; prints OK
 *   mem_signal_check foreign   a fault outside every segment with no earlier handler ends the
 *                              process with SIGSEGV
 *   mem_signal_check chained   a fault outside every segment goes to the handler installed before
 *                              bh_mem_signal_init (which exits with 42)
 *   mem_signal_check reinstall a handler installed after init (as Python's faulthandler may be) is
 *                              put behind ours at the next attach: segment faults still reach the
 *                              callback, other faults reach that handler (which exits with 43) */
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <unistd.h>

#include "bh_mem_signal.h"

static size_t g_len;
static volatile int g_calls;  /* changed by the fault handler */

static int on_fault(void *addr, void *idx) {
    char *base = (char *)idx;
    if ((char *)addr < base || (char *)addr >= base + g_len) return 0;
    ++g_calls;
    if (mprotect(base, g_len, PROT_READ | PROT_WRITE) != 0) return 0;
    bh_mem_signal_detach(base);
    return 1;
}

static void exit42(int sig) { (void)sig; _exit(42); }
static void exit43(int sig) { (void)sig; _exit(43); }

static char *unmapped_page(size_t page) {
    char *q = mmap(NULL, page, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    munmap(q, page);
    return q;
}

#define CHECK(c) do { if (!(c)) { printf("FAIL line %d: %s\n", __LINE__, #c); return 1; } } while (0)

int main(int argc, char **argv) {
    const char *mode = argc > 1 ? argv[1] : "basic";
    const size_t page = (size_t)sysconf(_SC_PAGESIZE);
    g_len = 2 * page;
    char *p = mmap(NULL, g_len, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    CHECK(p != MAP_FAILED);
    if (strcmp(mode, "chained") == 0) signal(SIGSEGV, exit42);
    bh_mem_signal_init();
    if (strcmp(mode, "foreign") == 0 || strcmp(mode, "chained") == 0) {
        *(volatile int *)unmapped_page(page) = 1;
        printf("NOT REACHED\n");
        return 1;
    }
    if (strcmp(mode, "reinstall") == 0) {
        signal(SIGSEGV, exit43);
        CHECK(bh_mem_signal_attach(p, p, g_len, on_fault) == 0);
        CHECK(mprotect(p, g_len, PROT_NONE) == 0);
        ((volatile int *)p)[1] = 5;
        CHECK(g_calls == 1);
        fflush(stdout);
        *(volatile int *)unmapped_page(page) = 1;
        printf("NOT REACHED\n");
        return 1;
    }
    CHECK(bh_mem_signal_attach(p, p, g_len, on_fault) == 0);
    CHECK(bh_mem_signal_exist(p));
    CHECK(bh_mem_signal_exist(p + g_len - 1));
    CHECK(!bh_mem_signal_exist(p + g_len));
    CHECK(bh_mem_signal_attach(p, p + page, page, on_fault) == -1);  /* overlaps */
    CHECK(bh_mem_signal_attach(p, p - page, page + 1, on_fault) == -1);
    /* A fault anywhere in the segment reaches the callback, which makes the access legal. */
    CHECK(mprotect(p, g_len, PROT_NONE) == 0);
    ((volatile int *)(p + page))[3] = 7;
    CHECK(g_calls == 1);
    CHECK(!bh_mem_signal_exist(p));
    CHECK(((int *)(p + page))[3] == 7);
    /* Read-only pages: reads pass, the first write is caught. */
    CHECK(bh_mem_signal_attach(p, p, g_len, on_fault) == 0);
    CHECK(mprotect(p, g_len, PROT_READ) == 0);
    CHECK(((volatile int *)(p + page))[3] == 7);
    CHECK(g_calls == 1);
    ((volatile int *)p)[0] = 9;
    CHECK(g_calls == 2);
    CHECK(((int *)p)[0] == 9);
    /* Many segments: each is found. */
    char *q = mmap(NULL, 64 * page, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    for (int i = 0; i < 64; i += 2) CHECK(bh_mem_signal_attach(q + i * page, q + i * page, page, on_fault) == 0);
    for (int i = 0; i < 64; ++i) CHECK(bh_mem_signal_exist(q + i * page + 5) == (i % 2 == 0));
    for (int i = 0; i < 64; i += 2) bh_mem_signal_detach(q + i * page + 17);
    for (int i = 0; i < 64; ++i) CHECK(!bh_mem_signal_exist(q + i * page));
    bh_mem_signal_shutdown();
    printf("OK\n");
    return 0;
}
