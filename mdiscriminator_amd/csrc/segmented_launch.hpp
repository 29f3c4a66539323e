// segmented_launch.hpp -- launch sequence and candidate configurations of the
// batched segmented partition, shared by segmented.hip and the per-class-bits
// translation units (seg_cfg_b0.hip .. seg_cfg_b6.hip) that instantiate every
// candidate for one class-bit count, so that they compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "segmented_kernels.hpp"

namespace mdiscr {

constexpr int kPlanThreads = 1024;
constexpr uint32_t kExpandGrid = 1024;  // workgroups writing the T items of long segments
constexpr uint32_t kExpandInPlan = kPlanExpandMax;  // the planner expands them itself

struct SegConfig {
    int nw, r;
    bool staged;
    const char* name;
};

// Autotuner candidates: block size {256, 512, 1024} x elements per thread
// {4, 8, 16, 32} x the scatter variant: staged (a tile is grouped by class in
// LDS, then every class run is written contiguously: the small-m variant) or
// direct (every element is stored from registers to its destination; with
// many classes a tile's class runs are short, and staging buys little
// coalescing for its LDS round trip: the large-m variant).  All 24 are built
// for every class count and both discriminator kinds.  Entry 0 is the
// default; entries 0-7 keep their earlier numbering.
constexpr SegConfig kConfigs[] = {
    {4, 16, true, "b256_e16_staged"},   {4, 16, false, "b256_e16_direct"},
    {8, 8, true, "b512_e8_staged"},     {8, 8, false, "b512_e8_direct"},
    {4, 8, true, "b256_e8_staged"},     {8, 16, true, "b512_e16_staged"},
    {16, 4, true, "b1024_e4_staged"},   {4, 32, true, "b256_e32_staged"},
    {4, 4, true, "b256_e4_staged"},     {8, 4, true, "b512_e4_staged"},
    {8, 32, true, "b512_e32_staged"},   {16, 8, true, "b1024_e8_staged"},
    {16, 16, true, "b1024_e16_staged"}, {16, 32, true, "b1024_e32_staged"},
    {4, 4, false, "b256_e4_direct"},    {4, 8, false, "b256_e8_direct"},
    {4, 32, false, "b256_e32_direct"},  {8, 4, false, "b512_e4_direct"},
    {8, 16, false, "b512_e16_direct"},  {8, 32, false, "b512_e32_direct"},
    {16, 4, false, "b1024_e4_direct"},  {16, 8, false, "b1024_e8_direct"},
    {16, 16, false, "b1024_e16_direct"}, {16, 32, false, "b1024_e32_direct"},
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);
constexpr uint32_t kMinTile = 1024;  // smallest NW*64*R among the candidates

struct SegArgs {
    OffSrc os{};                      // segments by offsets (descriptors built on the device)
    const SegDesc* staged = nullptr;  // batch descriptors in mapped pinned memory
    hipEvent_t staged_read = nullptr;  // recorded once the planner has read them
    hipError_t* staged_err = nullptr;  // where that record's status goes
    const SegDesc* desc;
    SegItem* items;
    SegLong* longs;
    SegPlan* plan;
    uint32_t* counts;
    uint32_t* prefix;   // the scan of counts
    uint32_t ld;
    uint32_t max_rows;  // bound on the plan's row count
    uint32_t nseg;
    uint32_t mrows;
    uint32_t grid_cap;  // no more workgroups than the queue can hold
    uint32_t item_cap;  // items the workspace's queue has room for
    uint32_t num_cu;
    // False when no segment can span more than one tile (every length known
    // on the host to be at most the tile): no count rows, so the count and
    // scan launches are skipped.
    bool rows = true;
};

// Workgroups of `kernel` resident on one CU (the occupancy API), at least 1.
template <typename K>
int resident_per_cu(K kernel, int threads) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess ||
        nb < 1)
        nb = 1;
    return nb;
}

#ifndef MDISCR_SEG_OVERSUB
#define MDISCR_SEG_OVERSUB 16
#endif
// The queue-walking grid: MDISCR_SEG_OVERSUB times what fits at once, so
// each workgroup takes a few items and finished workgroups are replaced by
// the hardware (a grid of exactly the resident capacity leaves the last
// items' imbalance as a tail), and never more workgroups than items.
inline uint32_t persistent_grid(int per_cu, const SegArgs& a) {
    uint64_t g = (uint64_t)per_cu * a.num_cu * MDISCR_SEG_OVERSUB;
    if (g > a.grid_cap) g = a.grid_cap;
    return (uint32_t)(g ? g : 1);
}

template <int NW, int R, int NBITS, bool MASK, bool STAGED>
inline void launch_cfg(const SegArgs& a, hipStream_t s) {
    const bool expand_here = a.nseg <= kExpandInPlan;
    hipLaunchKernelGGL((k_seg_plan<kPlanThreads, NW, R>), dim3(1), dim3(kPlanThreads), 0, s,
                       a.staged, a.os, const_cast<SegDesc*>(a.desc), a.nseg, a.items, a.longs,
                       a.plan, a.counts, a.ld, a.mrows, expand_here ? 1u : 0u, a.item_cap,
                       a.ld - 1u);
    if (a.staged_read) *a.staged_err = hipEventRecord(a.staged_read, s);
    if (!expand_here) {
        const uint32_t eg = a.nseg < kExpandGrid ? (a.nseg ? a.nseg : 1u) : kExpandGrid;
        hipLaunchKernelGGL((k_seg_expand<256>), dim3(eg), dim3(256), 0, s, a.longs, a.plan,
                           a.items, a.os, const_cast<SegDesc*>(a.desc), a.nseg);
    }
    static const int count_per_cu = resident_per_cu(k_seg_count<NW, R, NBITS, MASK>, NW * kWave);
    static const int part_per_cu =
        resident_per_cu(k_seg_partition<NW, R, NBITS, MASK, STAGED>, NW * kWave);
    if (a.rows) {
        hipLaunchKernelGGL((k_seg_count<NW, R, NBITS, MASK>),
                           dim3(persistent_grid(count_per_cu, a)), dim3(NW * kWave), 0, s,
                           a.desc, a.items, a.plan, a.counts, a.ld);
        launch_tile_scan(a.counts, a.prefix, a.max_rows + 1u, a.ld, a.mrows, nullptr, nullptr,
                         &a.plan->n_rows, s);
    }
    hipLaunchKernelGGL((k_seg_partition<NW, R, NBITS, MASK, STAGED>),
                       dim3(persistent_grid(part_per_cu, a)), dim3(NW * kWave), 0, s, a.desc,
                       a.items, a.plan, a.prefix, a.ld);
}

// Launches candidate `cfg` for NBITS class bits; false for an unknown cfg.
template <int NBITS, bool MASK>
bool launch_nbits(int cfg, const SegArgs& a, hipStream_t s) {
    switch (cfg) {
        case 0: launch_cfg<4, 16, NBITS, MASK, true>(a, s); return true;
        case 1: launch_cfg<4, 16, NBITS, MASK, false>(a, s); return true;
        case 2: launch_cfg<8, 8, NBITS, MASK, true>(a, s); return true;
        case 3: launch_cfg<8, 8, NBITS, MASK, false>(a, s); return true;
        case 4: launch_cfg<4, 8, NBITS, MASK, true>(a, s); return true;
        case 5: launch_cfg<8, 16, NBITS, MASK, true>(a, s); return true;
        case 6: launch_cfg<16, 4, NBITS, MASK, true>(a, s); return true;
        case 7: launch_cfg<4, 32, NBITS, MASK, true>(a, s); return true;
        case 8: launch_cfg<4, 4, NBITS, MASK, true>(a, s); return true;
        case 9: launch_cfg<8, 4, NBITS, MASK, true>(a, s); return true;
        case 10: launch_cfg<8, 32, NBITS, MASK, true>(a, s); return true;
        case 11: launch_cfg<16, 8, NBITS, MASK, true>(a, s); return true;
        case 12: launch_cfg<16, 16, NBITS, MASK, true>(a, s); return true;
        case 13: launch_cfg<16, 32, NBITS, MASK, true>(a, s); return true;
        case 14: launch_cfg<4, 4, NBITS, MASK, false>(a, s); return true;
        case 15: launch_cfg<4, 8, NBITS, MASK, false>(a, s); return true;
        case 16: launch_cfg<4, 32, NBITS, MASK, false>(a, s); return true;
        case 17: launch_cfg<8, 4, NBITS, MASK, false>(a, s); return true;
        case 18: launch_cfg<8, 16, NBITS, MASK, false>(a, s); return true;
        case 19: launch_cfg<8, 32, NBITS, MASK, false>(a, s); return true;
        case 20: launch_cfg<16, 4, NBITS, MASK, false>(a, s); return true;
        case 21: launch_cfg<16, 8, NBITS, MASK, false>(a, s); return true;
        case 22: launch_cfg<16, 16, NBITS, MASK, false>(a, s); return true;
        case 23: launch_cfg<16, 32, NBITS, MASK, false>(a, s); return true;
    }
    return false;
}

#define MDISCR_SEG_EXTERN_NBITS(B)                                                   \
    extern template bool launch_nbits<B, true>(int, const SegArgs&, hipStream_t);  \
    extern template bool launch_nbits<B, false>(int, const SegArgs&, hipStream_t);
MDISCR_SEG_EXTERN_NBITS(0)
MDISCR_SEG_EXTERN_NBITS(1)
MDISCR_SEG_EXTERN_NBITS(2)
MDISCR_SEG_EXTERN_NBITS(3)
MDISCR_SEG_EXTERN_NBITS(4)
MDISCR_SEG_EXTERN_NBITS(5)
MDISCR_SEG_EXTERN_NBITS(6)
#undef MDISCR_SEG_EXTERN_NBITS

}  // namespace mdiscr
