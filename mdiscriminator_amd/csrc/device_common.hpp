// device_common.hpp -- gfx950 device helpers shared by the partition kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MDISCR_HD __host__ __device__
#include "divmagic.h"

namespace mdiscr {

constexpr int kWave = 64;      // CDNA wavefront width
constexpr int kNumXcd = 8;     // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr int kMaxClasses = 64;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---- discriminators (the class of a 32-bit key) ---------------------------

// Mod4 / Mod8 / Mod<2^k> (HelpersCommon.h:593-595, 846-848): x & (m-1).
struct DiscrMask {
    uint32_t mask;
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        return x & mask;
    }
};

// Mod<N> (HelpersCommon.h:1207-1209): x % N by the invariant-divisor magic.
struct DiscrMod {
    mdiscr_divmagic g;
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        return mdiscr_divmagic_mod(g, x);
    }
};

// Two-way filter (parFilter, ParFilter.hs:8-21): class 0 where the predicate
// holds, class 1 elsewhere.  Every predicate is normalised on the host to one
// unsigned window test, holds = ((x & mask) - lo <= hi) != inv, so the device
// spends an AND, a SUB and a compare per element whatever the predicate.
struct DiscrPred {
    uint32_t mask, lo, hi, flip;  // flip = inv ^ 1
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        return (uint32_t)((x & mask) - lo <= hi) ^ flip;
    }
};

// Discriminators whose evaluation costs more than an LDS byte read per
// element (x % m: two quarter-rate 32-bit multiplies; the filter's window).
// Kernels that group a tile by class in LDS keep each slot's class there for
// the write-out instead of evaluating them again.
template <class D>
struct CostlyDiscr {
    static constexpr bool value = false;
};
template <>
struct CostlyDiscr<DiscrMod> {
    static constexpr bool value = true;
};
// The filter's window test (four VALU operations) also costs more than the
// byte read: 2^28 keys, x < 2^30, partition 0.364 -> 0.359 ms with class
// bytes.
template <>
struct CostlyDiscr<DiscrPred> {
    static constexpr bool value = true;
};

// The loaded value already is the class (class sources that apply their own
// discriminator, e.g. a selection between two key columns).
struct DiscrIdent {
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const { return x; }
};

// ---- wave64 primitives ------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// This wave's index in the workgroup, made a scalar (it is wave-uniform), so
// that wave-relative bases and addresses need no vector registers.
__device__ __forceinline__ uint32_t wave_id() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

// Number of set bits of `mask` in lanes strictly below this lane.
__device__ __forceinline__ uint32_t mbcnt64(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot64(bool p) {
    return __builtin_amdgcn_ballot_w64(p);
}

// Exclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t lane) {
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t y = __shfl_up(x, d, kWave);
        if (lane >= (uint32_t)d) x += y;
    }
    return x - v;
}

// Inclusive prefix sum across the 64 lanes of a wave, on DPP moves only (no
// LDS traffic): four row shifts scan each row of 16 lanes, then the last lane
// of row 0 is added into row 1 and of row 2 into row 3 (row_bcast:15), and
// the last lane of row 1 into rows 2 and 3 (row_bcast:31).  Lanes a DPP move
// cannot source from (or whose row is masked off) add 0.
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Rank of this lane's class among the lower lanes holding the same class,
// and the number of lanes holding it, from NBITS ballots of the class bits.
// A lane differs from another iff some class bit differs, so the mismatch
// mask is the OR over bits of (ballot of bit b) xor (own bit b splat).
template <int NBITS>
__device__ __forceinline__ void wave_match_rank(uint32_t c, uint32_t lane,
                                                uint32_t& rank, uint32_t& cnt) {
    uint32_t mlo = 0, mhi = 0;
#pragma unroll
    for (int b = 0; b < NBITS; ++b) {
        // splat = all ones if this lane's bit b is set; mismatching lanes are
        // the ballot bits that differ from it.
        const uint32_t splat = (uint32_t)__builtin_amdgcn_sbfe((int)c, b, 1);
        const uint64_t B = ballot64(splat != 0u);
        mlo |= (uint32_t)B ^ splat;
        mhi |= (uint32_t)(B >> 32) ^ splat;
    }
    rank = lane - __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, 0u));
    cnt = kWave - (uint32_t)(__builtin_popcount(mlo) + __builtin_popcount(mhi));
}

// ---- XCD-aware tile order -------------------------------------------------------
// Workgroup b runs on XCD b % 8.  Give each XCD one contiguous run of tiles so
// that neighbouring tiles (which share partially written output lines at class
// run boundaries) meet in the same L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nblocks) {
    const uint32_t q = nblocks / kNumXcd, r = nblocks % kNumXcd;
    const uint32_t x = b % kNumXcd, i = b / kNumXcd;
    return x * q + (x < r ? x : r) + i;
}

// Same XCD ranges, each walked back to front.
__device__ __forceinline__ uint32_t xcd_tile_reversed(uint32_t b, uint32_t nblocks) {
    const uint32_t q = nblocks / kNumXcd, r = nblocks % kNumXcd;
    const uint32_t x = b % kNumXcd, i = b / kNumXcd;
    const uint32_t cnt = q + (x < r ? 1u : 0u);
    return x * q + (x < r ? x : r) + (cnt - 1u - i);
}

}  // namespace mdiscr
