// mdiscr.hpp -- C++ host interface in the shape of the reference's templates,
// implemented over the C ABI of libmdiscr.so (include/mdiscr.h).
//
// Reference (asbjornvg/mdiscriminator, cuda_code/):
//   Mod4, Mod8, Mod<N>              HelpersCommon.h:587-596, 840-849, 1200-1210
//   MyInt<N> counts tuple           HelpersCommon.h:855-965 (cardinal, operator[])
//   mdiscr<DISCR>(n, num_hwd_thds, d_in, d_out) -> DISCR::TupleType
//                                   MdiscrHost_optimized_clean.cu.h:9-151
//   ModArray4<LEN>, ModArray<N,LEN> MdiscrHost_arrays.cu.h:14-36
//   mdiscr<DISCR,LEN>(n, num_hwd_thds, d_in, d_out) -> DISCR::TupleType
//                                   MdiscrHost_arrays.cu.h:250-470
//   ModTuple4, ModTuple<N>          MdiscrHost_tuples.cu.h:13-48
//   mdiscr<DISCR>(n, num_hwd_thds, d_in1, d_in2, d_in3, d_out1, d_out2, d_out3)
//                                   MdiscrHost_tuples.cu.h:262-429
//   mdiscrSegm<DISCR>(n, in, segment_sizes, out, new_sizes)
//                                   MdiscrSegmHost.cu.h:36-230
//   tupleToSizeArray(sizes, sizes_array, n)
//                                   HelpersCommon.h:295-311
//
// Same names, argument order and meaning: device pointers in and out, counts
// returned by value after the device work has finished (synchronous, like the
// reference).  Error behaviour also follows the reference's gpuErrchk
// (HelpersHost.cu.h:11-19): print the failure and exit, unless
// MDISCR_HPP_THROW is defined, in which case std::runtime_error is thrown.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime_api.h>

#include "mdiscr.h"

namespace mdiscr_ref {

// m class counts (the reference's MyInt<N>: arr[N], cardinal, operator[]).
template <unsigned int N>
struct MyInt {
    unsigned int arr[N];
    static const unsigned int cardinal = N;
    MyInt() { std::memset(arr, 0, sizeof(arr)); }
    unsigned int& operator[](unsigned int i) { return arr[i]; }
    const unsigned int& operator[](unsigned int i) const { return arr[i]; }
};

// Discriminator policies: InType, TupleType, padelm, apply, plus the runtime
// descriptor the C ABI takes.
struct Mod4 {
    typedef unsigned int InType;
    typedef MyInt<4> TupleType;
    static const InType padelm = 3;
    static unsigned int apply(InType x) { return x & 3; }
    static mdiscr_discr discr() { return mdiscr_discr{MDISCR_DISCR_MASK, 4}; }
};

struct Mod8 {
    typedef unsigned int InType;
    typedef MyInt<8> TupleType;
    static const InType padelm = 7;
    static unsigned int apply(InType x) { return x & 7; }
    static mdiscr_discr discr() { return mdiscr_discr{MDISCR_DISCR_MASK, 8}; }
};

template <unsigned int N>
struct Mod {
    static_assert(N >= 1 && N <= MDISCR_MAX_CLASSES, "1 <= N <= 64");
    typedef unsigned int InType;
    typedef MyInt<N> TupleType;
    static const InType padelm = N - 1;
    static unsigned int apply(InType x) { return x % N; }
    static mdiscr_discr discr() { return mdiscr_discr{MDISCR_DISCR_MOD, N}; }
};

namespace detail {

inline void fail(mdiscr_status s, const char* what, const mdiscr_ctx* ctx) {
    std::string msg = std::string(what) + ": " + mdiscr_status_str(s);
    if (ctx) msg += std::string(" (") + mdiscr_last_error(ctx) + ")";
#ifdef MDISCR_HPP_THROW
    throw std::runtime_error(msg);
#else
    std::fprintf(stderr, "mdiscr: %s\n", msg.c_str());
    std::exit(static_cast<int>(s));
#endif
}

// One context per process for the reference-shaped calls (created on first use).
inline mdiscr_ctx* context() {
    static mdiscr_ctx* ctx = [] {
        mdiscr_ctx* c = nullptr;
        mdiscr_status s = mdiscr_ctx_create(&c, -1);
        if (s != MDISCR_OK) fail(s, "mdiscr_ctx_create", nullptr);
        return c;
    }();
    return ctx;
}

// Device memory for the class counts of the calls whose C entry point writes
// them to the device (64 entries, allocated on first use, kept for the process).
inline uint32_t* device_counts() {
    static uint32_t* dc = [] {
        void* p = nullptr;
        if (hipMalloc(&p, MDISCR_MAX_CLASSES * sizeof(uint32_t)) != hipSuccess)
            fail(MDISCR_ENOMEM, "counts buffer", nullptr);
        return static_cast<uint32_t*>(p);
    }();
    return dc;
}

// Waits for the call on the null stream and copies the m counts to the host.
inline void read_counts(mdiscr_ctx* ctx, unsigned int* h, unsigned int m, const char* what) {
    mdiscr_status s = mdiscr_sync(ctx, nullptr);
    if (s == MDISCR_OK &&
        hipMemcpy(h, device_counts(), m * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
        s = MDISCR_EDEVICE;
    if (s != MDISCR_OK) fail(s, what, ctx);
}

}  // namespace detail

// mdiscr<DISCR>(num_elems, num_hwd_thds, d_in, d_out): stable partition of
// d_in into d_out; returns the class counts.  num_hwd_thds is accepted for
// signature compatibility (the tile geometry is fixed by this library).
template <class DISCR>
typename DISCR::TupleType mdiscr(const unsigned int num_elems, const unsigned int num_hwd_thds,
                                 typename DISCR::InType* d_in, typename DISCR::InType* d_out) {
    typename DISCR::TupleType counts;
    mdiscr_ctx* ctx = detail::context();
    const mdiscr_status s =
        mdiscr_u32(ctx, num_elems, num_hwd_thds, d_in, d_out, DISCR::discr(), counts.arr);
    if (s != MDISCR_OK) detail::fail(s, "mdiscr", ctx);
    return counts;
}

// ---- array-valued elements ----------------------------------------------

// Elements are LEN consecutive unsigned ints; the class comes from word 0.
template <unsigned int LEN>
struct ModArray4 {
    typedef unsigned int InType;
    typedef MyInt<4> TupleType;
    static const InType padelm = 3;
    static unsigned int apply(const InType* x) { return x[0] & 3; }
    static mdiscr_discr discr() { return mdiscr_discr{MDISCR_DISCR_MASK, 4}; }
};

template <unsigned int N, unsigned int LEN>
struct ModArray {
    static_assert(N >= 1 && N <= MDISCR_MAX_CLASSES, "1 <= N <= 64");
    typedef unsigned int InType;
    typedef MyInt<N> TupleType;
    static const InType padelm = N - 1;
    static unsigned int apply(const InType* x) { return x[0] % N; }
    static mdiscr_discr discr() { return mdiscr_discr{MDISCR_DISCR_MOD, N}; }
};

// mdiscr<DISCR, LEN>(num_elems, num_hwd_thds, d_in, d_out): num_elems arrays
// of LEN words each, moved whole to their stable class positions.
template <class DISCR, unsigned int LEN>
typename DISCR::TupleType mdiscr(const unsigned int num_elems, const unsigned int num_hwd_thds,
                                 typename DISCR::InType* d_in, typename DISCR::InType* d_out) {
    (void)num_hwd_thds;
    typename DISCR::TupleType counts;
    mdiscr_ctx* ctx = detail::context();
    const mdiscr_status s = mdiscr_partition_arrays(ctx, d_in, d_out, num_elems, LEN,
                                                    DISCR::discr(), detail::device_counts(),
                                                    nullptr);
    if (s != MDISCR_OK) detail::fail(s, "mdiscr (arrays)", ctx);
    detail::read_counts(ctx, counts.arr, DISCR::TupleType::cardinal, "mdiscr (arrays)");
    return counts;
}

// ---- three-column tuples -------------------------------------------------

// (unsigned int, unsigned char, bool) records classified by the first column.
struct ModTuple4 {
    typedef unsigned int InType1;
    typedef unsigned char InType2;
    typedef bool InType3;
    typedef MyInt<4> TupleType;
    static const bool select = false;
    static unsigned int apply(InType1 x, InType2, InType3) { return x & 3; }
    static mdiscr_discr discr() { return mdiscr_discr{MDISCR_DISCR_MASK, 4}; }
};

template <unsigned int N>
struct ModTuple {
    static_assert(N >= 1 && N <= MDISCR_MAX_CLASSES, "1 <= N <= 64");
    typedef unsigned int InType1;
    typedef unsigned char InType2;
    typedef bool InType3;
    typedef MyInt<N> TupleType;
    static const bool select = false;
    static unsigned int apply(InType1 x, InType2, InType3) { return x % N; }
    static mdiscr_discr discr() { return mdiscr_discr{MDISCR_DISCR_MOD, N}; }
};

// The alternative body the reference keeps commented out in ModTuple4 and
// ModTuple<N> (MdiscrHost_tuples.cu.h:27,46): z ? x % N : y % N.
template <unsigned int N>
struct ModTupleSelect : ModTuple<N> {
    static const bool select = true;
    static unsigned int apply(unsigned int x, unsigned char y, bool z) { return z ? x % N : y % N; }
};

// mdiscr<DISCR>(num_elems, num_hwd_thds, d_in1..3, d_out1..3): the three
// columns move together; the class is DISCR::apply of the record.
template <class DISCR>
typename DISCR::TupleType mdiscr(const unsigned int num_elems, const unsigned int num_hwd_thds,
                                 typename DISCR::InType1* d_in1, typename DISCR::InType2* d_in2,
                                 typename DISCR::InType3* d_in3, typename DISCR::InType1* d_out1,
                                 typename DISCR::InType2* d_out2, typename DISCR::InType3* d_out3) {
    (void)num_hwd_thds;
    static_assert(sizeof(typename DISCR::InType3) == 1, "bool is one byte here");
    typename DISCR::TupleType counts;
    mdiscr_ctx* ctx = detail::context();
    const void* in[3] = {d_in1, d_in2, d_in3};
    void* out[3] = {d_out1, d_out2, d_out3};
    const uint32_t bytes[3] = {sizeof(typename DISCR::InType1), sizeof(typename DISCR::InType2),
                               sizeof(typename DISCR::InType3)};
    const mdiscr_status s =
        DISCR::select
            ? mdiscr_partition_soa_select(ctx, 3, in, bytes, 2, 0, 1, out, num_elems,
                                          DISCR::discr(), detail::device_counts(), nullptr)
            : mdiscr_partition_soa(ctx, 3, in, bytes, 0, out, num_elems, DISCR::discr(),
                                   detail::device_counts(), nullptr);
    if (s != MDISCR_OK) detail::fail(s, "mdiscr (tuples)", ctx);
    detail::read_counts(ctx, counts.arr, DISCR::TupleType::cardinal, "mdiscr (tuples)");
    return counts;
}

// ---- segmented -------------------------------------------------------------

// mdiscrSegm<DISCR>(num_elems, in_array, segment_sizes, out_array, new_sizes):
// every flag-delimited segment partitioned in place of its range; new_sizes
// gets each segment's sizes array.  All pointers are device memory.
template <class DISCR>
void mdiscrSegm(const unsigned int num_elems, typename DISCR::InType* in_array, int* segment_sizes,
                typename DISCR::InType* out_array, int* new_sizes) {
    mdiscr_ctx* ctx = detail::context();
    const mdiscr_status s = mdiscr_segm_flags(ctx, num_elems, in_array, segment_sizes, out_array,
                                              new_sizes, DISCR::discr(), nullptr);
    if (s != MDISCR_OK) detail::fail(s, "mdiscrSegm", ctx);
}

// tupleToSizeArray(sizes, sizes_array, num_elems): host-side conversion of
// the counts tuple to the N-length sizes array (zero except each class's count
// at its start; an empty class is overwritten by the next one).
template <typename TupleType>
void tupleToSizeArray(TupleType sizes, unsigned int* sizes_array, unsigned int num_elems) {
    std::memset(sizes_array, 0, num_elems * sizeof(unsigned int));
    unsigned int i = 0;
    for (unsigned int j = 0; j < TupleType::cardinal && i < num_elems; ++j) {
        sizes_array[i] = sizes[j];
        i += sizes[j];
    }
}

}  // namespace mdiscr_ref
