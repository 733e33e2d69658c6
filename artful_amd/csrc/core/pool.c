/* artful-hip — the host thread pool (pool.h). Workers sleep on a condition variable between loops;
 * a loop is a generation number, a shared index counter, and a count of workers still inside it.
 * The thread-count rule is Artful 0.4.0's (src/core/threads.c:54, 126-133, 144): the CPUs in the
 * affinity mask, at most 16. */
#define _GNU_SOURCE
#include "pool.h"

#include "artful/core.h"

#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <unistd.h>

#define POOL_MOST 16

struct ArtfPool {
    int size;
    pthread_t threads[POOL_MOST];
    pthread_mutex_t lock;
    pthread_cond_t start, done;
    unsigned generation; /* under lock */
    int inside;          /* workers still running the loop, under lock */
    bool quit;
    ArtfJobFn fn;
    void* user;
    uint32_t count;
    atomic_uint next;
};

static void drain(ArtfPool* p, uint32_t worker) {
    for (;;) {
        const uint32_t i = atomic_fetch_add_explicit(&p->next, 1, memory_order_relaxed);
        if (i >= p->count) return;
        p->fn(p->user, i, worker);
    }
}

typedef struct WorkerArg { ArtfPool* pool; uint32_t index; } WorkerArg;

static void* workerMain(void* arg) {
    WorkerArg* w = arg;
    ArtfPool* p = w->pool;
    const uint32_t index = w->index;
    artfFree(w);
    unsigned seen = 0;
    pthread_mutex_lock(&p->lock);
    for (;;) {
        while (p->generation == seen && !p->quit) pthread_cond_wait(&p->start, &p->lock);
        if (p->quit) break;
        seen = p->generation;
        pthread_mutex_unlock(&p->lock);
        drain(p, index);
        pthread_mutex_lock(&p->lock);
        if (--p->inside == 0) pthread_cond_signal(&p->done);
    }
    pthread_mutex_unlock(&p->lock);
    return NULL;
}

static atomic_int chosenSize;

void artfPoolSetDefaultSize(int threads) { atomic_store(&chosenSize, threads > 0 ? artfClampi(threads, 1, POOL_MOST) : 0); }

int artfPoolDefaultSize(void) {
    const int chosen = atomic_load(&chosenSize);
    if (chosen > 0) return chosen;
    cpu_set_t set;
    int n = 0;
    if (!sched_getaffinity(0, sizeof set, &set)) n = CPU_COUNT(&set);
    if (n <= 0) n = (int)sysconf(_SC_NPROCESSORS_ONLN);
    return artfClampi(n, 1, POOL_MOST);
}

ArtfPool* artfPoolNew(int threads) {
    ArtfPool* p = ARTF_NEW(ArtfPool);
    p->size = threads <= 0 ? artfPoolDefaultSize() : artfClampi(threads, 1, POOL_MOST);
    pthread_mutex_init(&p->lock, NULL);
    pthread_cond_init(&p->start, NULL);
    pthread_cond_init(&p->done, NULL);
    for (int i = 1; i < p->size; ++i) {
        WorkerArg* w = ARTF_NEW(WorkerArg);
        w->pool = p;
        w->index = (uint32_t)i;
        pthread_create(&p->threads[i], NULL, workerMain, w);
    }
    return p;
}

void artfPoolFree(ArtfPool* p) {
    if (!p) return;
    pthread_mutex_lock(&p->lock);
    p->quit = true;
    pthread_cond_broadcast(&p->start);
    pthread_mutex_unlock(&p->lock);
    for (int i = 1; i < p->size; ++i) pthread_join(p->threads[i], NULL);
    pthread_cond_destroy(&p->start);
    pthread_cond_destroy(&p->done);
    pthread_mutex_destroy(&p->lock);
    artfFree(p);
}

int artfPoolSize(const ArtfPool* p) { return p->size; }

void artfPoolRun(ArtfPool* p, ArtfJobFn fn, void* user, uint32_t count) {
    if (p->size == 1 || count <= 1) {
        for (uint32_t i = 0; i < count; ++i) fn(user, i, 0);
        return;
    }
    pthread_mutex_lock(&p->lock);
    p->fn = fn;
    p->user = user;
    p->count = count;
    atomic_store_explicit(&p->next, 0, memory_order_relaxed);
    p->inside = p->size - 1;
    ++p->generation;
    pthread_cond_broadcast(&p->start);
    pthread_mutex_unlock(&p->lock);
    drain(p, 0);
    pthread_mutex_lock(&p->lock);
    while (p->inside > 0) pthread_cond_wait(&p->done, &p->lock);
    pthread_mutex_unlock(&p->lock);
}
