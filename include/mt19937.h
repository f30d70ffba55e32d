/*
 * mt19937.h -- 32-bit Mersenne Twister (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998).
 *
 * The reference's struct Arguments embeds an `mt19937state` (02_capi_graph/include/config/
 * graphConfig.h:45) whose definition lives in the absent OpenGraph submodule; this is an
 * independent implementation of the published generator with its own state layout.
 * Known answer: seeding with 5489 gives 3499211612 as the first output.
 */
#ifndef AG_MT19937_H
#define AG_MT19937_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AG_MT_N 624

typedef struct mt19937state {
    uint32_t mt[AG_MT_N];
    uint32_t mti;
} mt19937state;

void     initializeMersenneState(mt19937state *s, uint32_t seed);
uint32_t generateRandInt(mt19937state *s);
/* uniform double in [0,1) with 32 random bits */
double   generateRandDouble(mt19937state *s);

#ifdef __cplusplus
}
#endif

#endif
