/* Checks mdiscr_divmagic_mod(x) == x % d for every d in [1,64] and some
 * larger d, per divisor on 12 edge values (0, d-1, d, 2^31, 2^32-1, the last
 * multiple of d, ...), on every x below 2^20, and on 2^22 pseudo-random x. */
#include <stdio.h>
#include <stdint.h>
#include "divmagic.h"

static uint64_t s = 0x1234567;
static uint32_t next32(void) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)((z ^ (z >> 31)) >> 16);
}

static int check(uint32_t d) {
    mdiscr_divmagic g = mdiscr_divmagic_make(d);
    const uint32_t edge[] = {0u, 1u, 2u, d - 1u, d, d + 1u, 0x7FFFFFFFu, 0x80000000u,
                             0xFFFFFFFEu, 0xFFFFFFFFu, 0xFFFFFFFFu - d, 0xFFFFFFFFu / d * d};
    for (unsigned i = 0; i < sizeof(edge) / sizeof(edge[0]); i++) {
        uint32_t x = edge[i];
        if (mdiscr_divmagic_mod(g, x) != x % d || mdiscr_divmagic_div(g, x) != x / d) {
            printf("FAIL d=%u x=%u\n", d, x);
            return 1;
        }
    }
    for (int i = 0; i < (1 << 22); i++) {
        uint32_t x = next32();
        if (mdiscr_divmagic_mod(g, x) != x % d) {
            printf("FAIL d=%u x=%u\n", d, x);
            return 1;
        }
    }
    /* every multiple boundary below 2^20 */
    for (uint32_t x = 0; x < (1u << 20); x++)
        if (mdiscr_divmagic_mod(g, x) != x % d) { printf("FAIL d=%u x=%u\n", d, x); return 1; }
    return 0;
}

int main(void) {
    for (uint32_t d = 1; d <= 64; d++)
        if (check(d)) return 1;
    const uint32_t big[] = {65u, 1000u, 65535u, 65537u, 0x7FFFFFFFu, 0x80000001u, 0xFFFFFFFFu};
    for (unsigned i = 0; i < sizeof(big) / sizeof(big[0]); i++)
        if (check(big[i])) return 1;
    printf("divmagic ok\n");
    return 0;
}
