/* artful-hip — a host thread pool for the work that stays on the CPU: the disc's tables and the
 * fluid's rows. Same contract as Artful 0.4.0's src/core/threads.h:7-17 (artfPoolNew/Free/Size/Run):
 * at most 16 threads, the calling thread takes part, indices are handed out from one atomic counter,
 * and every index's work must not depend on which thread runs it. */
#ifndef ARTFUL_POOL_H
#define ARTFUL_POOL_H

#include <stdint.h>

typedef struct ArtfPool ArtfPool;
typedef void (*ArtfJobFn)(void* user, uint32_t index, uint32_t worker);

/* threads <= 0: as many CPUs as this process may use, at most 16. */
ArtfPool* artfPoolNew(int threads);
void artfPoolFree(ArtfPool* p);
int artfPoolSize(const ArtfPool* p);
/* fn over 0..count-1; returns when all are done. */
void artfPoolRun(ArtfPool* p, ArtfJobFn fn, void* user, uint32_t count);
/* What artfPoolNew(0) would pick. */
int artfPoolDefaultSize(void);

#endif
