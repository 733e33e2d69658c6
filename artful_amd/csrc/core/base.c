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

void* artfMalloc(size_t size) { return orDie(malloc(size ? size : 1)); }
void* artfCalloc(size_t count, size_t size) { return orDie(calloc(count ? count : 1, size ? size : 1)); }
void* artfRealloc(void* p, size_t size) { return orDie(realloc(p, size ? size : 1)); }
void artfFree(void* p) { free(p); }

void* artfAlignedMalloc(size_t size, size_t align) {
    void* p = NULL;
    if (align < sizeof(void*)) align = sizeof(void*);
    if (posix_memalign(&p, align, size ? size : 1)) p = NULL;
    return orDie(p);
}
void artfAlignedFree(void* p) { free(p); }

double artfNowMs(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec * 1e3 + (double)t.tv_nsec * 1e-6;
}
