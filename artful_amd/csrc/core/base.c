/* artful-hip — host memory, the log and the clock (the services of Artful 0.4.0's src/core/base.c:8-62
 * and src/core/threads.c:134-138, with the same contracts: allocation failure aborts, the log says
 * nothing until given a function). */
#define _POSIX_C_SOURCE 200809L
#include "artful/core.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

static ArtfLogFn logSink;
static void* logSinkUser;

void artfSetLog(ArtfLogFn fn, void* user) {
    logSink = fn;
    logSinkUser = user;
}

void artfLog(const char* format, ...) {
    if (!logSink) return;
    char text[512];
    va_list args;
    va_start(args, format);
    vsnprintf(text, sizeof text, format, args);
    va_end(args);
    logSink(logSinkUser, text);
}

static void* orDie(void* p) {
    if (!p) {
        fputs("artful: out of memory\n", stderr);
        abort();
    }
    return p;
}

/* Built with ARTFUL_COUNT_ALLOCATIONS, the blocks alive are counted (as the reference's base.c:32-40
 * does): a test then knows that everything was given back. */
#ifdef ARTFUL_COUNT_ALLOCATIONS
#include <stdatomic.h>
static atomic_long alive;
long artfAllocationsAlive(void) { return atomic_load(&alive); }
#define COUNT(n) atomic_fetch_add(&alive, n)
#else
#define COUNT(n) ((void)0)
#endif

void* artfMalloc(size_t size) {
    COUNT(1);
    return orDie(malloc(size ? size : 1));
}
void* artfCalloc(size_t count, size_t size) {
    COUNT(1);
    return orDie(calloc(count ? count : 1, size ? size : 1));
}
void* artfRealloc(void* p, size_t size) {
    if (!p) COUNT(1);
    return orDie(realloc(p, size ? size : 1));
}
void artfFree(void* p) {
    if (p) COUNT(-1);
    free(p);
}

void* artfAlignedMalloc(size_t size, size_t align) {
    void* p = NULL;
    if (align < sizeof(void*)) align = sizeof(void*);
    if (posix_memalign(&p, align, size ? size : 1)) p = NULL;
    COUNT(1);
    return orDie(p);
}
void artfAlignedFree(void* p) {
    if (p) COUNT(-1);
    free(p);
}

double artfNowMs(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec * 1e3 + (double)t.tv_nsec * 1e-6;
}
