/*
 * divmagic.h -- unsigned 32-bit division by an invariant divisor.
 *
 * Granlund & Montgomery, "Division by Invariant Integers using
 * Multiplication", PLDI 1994, Figure 4.1 (the round-up method that needs no
 * 33-bit multiplier): for 1 <= d < 2^32 and l = ceil(log2 d),
 *     mprime = floor(2^32 * (2^l - d) / d) + 1
 *     sh1 = min(l, 1), sh2 = max(l - 1, 0)
 *     t1 = mulhi(mprime, x);  q = (t1 + ((x - t1) >> sh1)) >> sh2
 * gives q = floor(x / d) for every 32-bit x.  The remainder x - q*d is the
 * class of the Mod<N> discriminator.
 *
 * Plain C, shared by the HIP kernels (through MDISCR_HD) and the host tests.
 */
#ifndef MDISCR_DIVMAGIC_H
#define MDISCR_DIVMAGIC_H

#include <stdint.h>

#ifndef MDISCR_HD
#define MDISCR_HD
#endif

typedef struct {
    uint32_t mprime;
    uint32_t sh1;
    uint32_t sh2;
    uint32_t d;
} mdiscr_divmagic;

static inline mdiscr_divmagic mdiscr_divmagic_make(uint32_t d) {
    mdiscr_divmagic r;
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < (uint64_t)d) l++;
    uint64_t num = ((1ull << l) - (uint64_t)d) << 32; /* 2^32 (2^l - d) */
    r.mprime = (uint32_t)(num / d + 1u);
    r.sh1 = l < 1 ? l : 1u;
    r.sh2 = l > 0 ? l - 1u : 0u;
    r.d = d;
    return r;
}

static inline MDISCR_HD uint32_t mdiscr_divmagic_div(mdiscr_divmagic g,
                                                     uint32_t x) {
    uint32_t t1 = (uint32_t)(((uint64_t)g.mprime * (uint64_t)x) >> 32);
    return (t1 + ((x - t1) >> g.sh1)) >> g.sh2;
}

static inline MDISCR_HD uint32_t mdiscr_divmagic_mod(mdiscr_divmagic g,
                                                     uint32_t x) {
    return x - mdiscr_divmagic_div(g, x) * g.d;
}

#endif /* MDISCR_DIVMAGIC_H */
