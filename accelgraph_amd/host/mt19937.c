/*
 * mt19937.c -- MT19937 (Matsumoto & Nishimura 1998): 624-word state, period 2^19937-1.
 * Initialisation: mt[0] = seed, mt[i] = 1812433253 * (mt[i-1] ^ (mt[i-1] >> 30)) + i.
 * Tempering: y ^= y>>11; y ^= (y<<7) & 0x9d2c5680; y ^= (y<<15) & 0xefc60000; y ^= y>>18.
 */
#include "mt19937.h"

#define MT_M 397
#define MT_UPPER 0x80000000u
#define MT_LOWER 0x7fffffffu
#define MT_MATRIX 0x9908b0dfu

void initializeMersenneState(mt19937state *s, uint32_t seed)
{
    s->mt[0] = seed;
    for (uint32_t i = 1; i < AG_MT_N; i++)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + i;
    s->mti = AG_MT_N;
}

static void mt_twist(mt19937state *s)
{
    for (uint32_t i = 0; i < AG_MT_N; i++) {
        uint32_t y = (s->mt[i] & MT_UPPER) | (s->mt[(i + 1) % AG_MT_N] & MT_LOWER);
        uint32_t next = s->mt[(i + MT_M) % AG_MT_N] ^ (y >> 1);
        if (y & 1u)
            next ^= MT_MATRIX;
        s->mt[i] = next;
    }
    s->mti = 0;
}

uint32_t generateRandInt(mt19937state *s)
{
    if (s->mti >= AG_MT_N)
        mt_twist(s);
    uint32_t y = s->mt[s->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

double generateRandDouble(mt19937state *s)
{
    return (double)generateRandInt(s) * (1.0 / 4294967296.0);
}
