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
/* threads > 0: what artfPoolNew(0) picks from now on, in this process (at most 16); 0: the rule
 * again. The pools the scene makes for the disc's tables and fluid are new at each build, so this
 * reaches them: the bytes they make are the same whatever the size (tested at 1, 3 and 16). */
void artfPoolSetDefaultSize(int threads);

#endif
