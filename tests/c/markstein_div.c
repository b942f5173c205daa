/*
 * markstein_div.c -- checks that x*r corrected by one FMA residual step,
 *   q0 = x*r, q = fma(fma(-q0, c, x), r, q0),  r = RN(1/c),
 * equals the IEEE quotient x/c, the identity the fluid sweep uses for its
 * division by the pencil-constant speed (|x| in [2^-60, 2^60], c in
 * [2^-30, 2^30]).  Every divisor significand is tried in three binades, with
 * random dividends and with dividends placed so that the quotient lies within
 * a few ulps of a float (the hard rounding cases).  Exit status 0 iff no
 * mismatch.   usage: markstein_div [per_significand]
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint64_t s = 88172645463325252ull;
static inline uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }

static inline float qdiv(float x, float c, float r)
{
    float q0 = x * r;
    return fmaf(fmaf(-q0, c, x), r, q0);
}

int main(int argc, char **argv)
{
    const int per = argc > 1 ? atoi(argv[1]) : 4;
    long bad = 0, tot = 0;
    for (uint32_t m = 0; m < (1u << 23); m++) {
        for (int ce = -30; ce <= 30; ce += 30) {
            const float c = f_of(((uint32_t)(127 + ce) << 23) | m);
            volatile float one = 1.0f;
            const float r = one / c;
            for (int k = 0; k < per; k++) {
                uint64_t z = xr();
                /* random dividend, |x| in [2^-60, 2^61) */
                float x = f_of(((uint32_t)(127 + (int)(z % 121) - 60) << 23) | ((uint32_t)(z >> 20) & 0x7fffff));
                if (z >> 63) x = -x;
                tot++;
                if (u_of(qdiv(x, c, r)) != u_of(x / c)) bad++;
                /* quotient near a float y: x = RN(c*y) +- a few ulps */
                float y = f_of(((uint32_t)(127 + (int)((z >> 8) % 9) - 4) << 23) | ((uint32_t)(z >> 33) & 0x7fffff));
                float x2 = f_of(u_of(c * y) + (int)((z >> 4) % 5) - 2);
                if (fabsf(x2) >= 0x1p-60f && fabsf(x2) <= 0x1p60f) {
                    tot++;
                    if (u_of(qdiv(x2, c, r)) != u_of(x2 / c)) bad++;
                }
            }
        }
    }
    printf("markstein_div: %ld quotients, %ld mismatches\n", tot, bad);
    return bad != 0;
}
