// mdiscr_ref_driver.cpp -- test executable in the shape of the reference's
// MdiscrTest_optimized_clean.cu / MdiscrSegmTest.cu drivers: argv[1] = N,
// argv[2] = m (4 and 8 use the Mod4/Mod8 policies, anything else Mod<m>),
// argv[3] = "segm" for the flag-array segmented call, "arrays" for 3-word
// array elements (MdiscrTest_arrays.cu), "tuples" / "select" for the
// (uint, uchar, bool) records of MdiscrTest_tuples.cu under ModTuple<m> /
// z ? x % m : y % m.  Seeded input (the
// reference seeds with the wall clock), run through the C++ mirror of the
// reference templates (mdiscriminator_amd/host/mdiscr.hpp), validated bit for
// bit against the oracle; prints VALID RESULT / INVALID RESULT and returns
// EXIT_SUCCESS / EXIT_FAILURE like the reference drivers.
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mdiscr.hpp"
#include "oracle.h"

using namespace mdiscr_ref;

static uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <class DISCR>
int test(unsigned int n, unsigned int m, bool segm) {
    std::vector<unsigned int> h_in(n), h_out(n), ref(n), sizes(n), ref_sizes(n);
    for (unsigned int i = 0; i < n; i++)
        h_in[i] = (unsigned int)(mix(0x5EED0000ull + (i + 1) * 0x9E3779B97F4A7C15ull) >> 33);
    unsigned int *d_in, *d_out;
    hipMalloc((void**)&d_in, (n + 1) * 4ull);
    hipMalloc((void**)&d_out, (n + 1) * 4ull);
    hipMemcpy(d_in, h_in.data(), n * 4ull, hipMemcpyHostToDevice);
    bool ok = true;
    if (!segm) {
        typename DISCR::TupleType counts = mdiscr<DISCR>(n, 1 << 16, d_in, d_out);
        hipMemcpy(h_out.data(), d_out, n * 4ull, hipMemcpyDeviceToHost);
        tupleToSizeArray(counts, sizes.data(), n);
        const uint32_t kind = DISCR::discr().kind;
        oracle_seq_mdiscr_u32(n, h_in.data(), ref.data(), ref_sizes.data(), kind, m, nullptr);
        ok = h_out == ref && sizes == ref_sizes;
    } else {
        // segments of lengths 1, 2, 3, ... (the reference draws 25..74)
        std::vector<int> flags(n, 0), new_sizes(n), ref_new(n);
        for (unsigned int i = 0, len = 1; i < n; i += len, ++len) flags[i] = (int)std::min(len, n - i);
        int *d_f, *d_ns;
        hipMalloc((void**)&d_f, (n + 1) * 4ull);
        hipMalloc((void**)&d_ns, (n + 1) * 4ull);
        hipMemcpy(d_f, flags.data(), n * 4ull, hipMemcpyHostToDevice);
        mdiscrSegm<DISCR>(n, d_in, d_f, d_out, d_ns);
        hipMemcpy(h_out.data(), d_out, n * 4ull, hipMemcpyDeviceToHost);
        hipMemcpy(new_sizes.data(), d_ns, n * 4ull, hipMemcpyDeviceToHost);
        oracle_segm_flags(n, h_in.data(), flags.data(), ref.data(), ref_new.data(),
                          DISCR::discr().kind, m);
        ok = h_out == ref && new_sizes == ref_new;
        hipFree(d_f);
        hipFree(d_ns);
    }
    hipFree(d_in);
    hipFree(d_out);
    std::printf("%s on %u elems (m=%u): %s\n", segm ? "mdiscrSegm" : "mdiscr", n, m,
                ok ? "VALID RESULT!" : "INVALID RESULT!");
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}

static unsigned int seeded(unsigned int i, uint64_t salt) {
    return (unsigned int)(mix(0x5EED0000ull + salt + (i + 1) * 0x9E3779B97F4A7C15ull) >> 33);
}

// mdiscr<ModArray, LEN>: num_elems arrays of LEN words (MdiscrTest_arrays.cu:7-70).
template <class DISCR, unsigned int LEN>
int test_arrays(unsigned int n, unsigned int m) {
    std::vector<unsigned int> h_in((size_t)n * LEN), h_out(h_in.size()), ref(h_in.size());
    std::vector<uint32_t> ref_counts(m);
    for (size_t i = 0; i < h_in.size(); i++) h_in[i] = seeded((unsigned int)i, 1);
    unsigned int *d_in, *d_out;
    hipMalloc((void**)&d_in, h_in.size() * 4ull + 4);
    hipMalloc((void**)&d_out, h_in.size() * 4ull + 4);
    hipMemcpy(d_in, h_in.data(), h_in.size() * 4ull, hipMemcpyHostToDevice);
    typename DISCR::TupleType counts = mdiscr<DISCR, LEN>(n, 1 << 16, d_in, d_out);
    hipMemcpy(h_out.data(), d_out, h_in.size() * 4ull, hipMemcpyDeviceToHost);
    oracle_seq_mdiscr_arrays(n, LEN, h_in.data(), ref.data(), DISCR::discr().kind, m,
                             ref_counts.data());
    bool ok = h_out == ref;
    for (unsigned int k = 0; k < m; k++) ok = ok && counts[k] == ref_counts[k];
    hipFree(d_in);
    hipFree(d_out);
    std::printf("mdiscr (arrays, LEN=%u) on %u elems (m=%u): %s\n", LEN, n, m,
                ok ? "VALID RESULT!" : "INVALID RESULT!");
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}

// mdiscr<ModTuple...>: (unsigned int, unsigned char, bool) columns
// (MdiscrTest_tuples.cu:45-110).
template <class DISCR>
int test_tuples(unsigned int n, unsigned int m) {
    std::vector<unsigned int> x(n), rx(n), ox(n);
    std::vector<unsigned char> y(n), ry(n), oy(n);
    std::vector<unsigned char> z(n), rz(n), oz(n);  // bool columns as bytes
    for (unsigned int i = 0; i < n; i++) {
        x[i] = seeded(i, 2);
        y[i] = (unsigned char)seeded(i, 3);
        z[i] = (unsigned char)(seeded(i, 4) & 1);
    }
    unsigned int *dx, *ox_d;
    unsigned char *dy, *oy_d;
    bool *dz, *oz_d;
    hipMalloc((void**)&dx, n * 4ull + 4);
    hipMalloc((void**)&ox_d, n * 4ull + 4);
    hipMalloc((void**)&dy, n + 1ull);
    hipMalloc((void**)&oy_d, n + 1ull);
    hipMalloc((void**)&dz, n + 1ull);
    hipMalloc((void**)&oz_d, n + 1ull);
    hipMemcpy(dx, x.data(), n * 4ull, hipMemcpyHostToDevice);
    hipMemcpy(dy, y.data(), n, hipMemcpyHostToDevice);
    hipMemcpy(dz, z.data(), n, hipMemcpyHostToDevice);
    typename DISCR::TupleType counts = mdiscr<DISCR>(n, 1 << 16, dx, dy, dz, ox_d, oy_d, oz_d);
    hipMemcpy(ox.data(), ox_d, n * 4ull, hipMemcpyDeviceToHost);
    hipMemcpy(oy.data(), oy_d, n, hipMemcpyDeviceToHost);
    hipMemcpy(oz.data(), oz_d, n, hipMemcpyDeviceToHost);
    const void* cols[3] = {x.data(), y.data(), z.data()};
    void* refs[3] = {rx.data(), ry.data(), rz.data()};
    const uint32_t bytes[3] = {4, 1, 1};
    std::vector<uint32_t> ref_counts(m);
    const uint32_t kind = DISCR::discr().kind;
    if (DISCR::select)
        oracle_seq_mdiscr_soa_select(n, 3, cols, bytes, 2, 0, 1, refs, kind, m, ref_counts.data());
    else
        oracle_seq_mdiscr_soa(n, 3, cols, bytes, 0, refs, kind, m, ref_counts.data());
    bool ok = ox == rx && oy == ry && oz == rz;
    for (unsigned int k = 0; k < m; k++) ok = ok && counts[k] == ref_counts[k];
    // every record's class agrees with DISCR::apply, in non-decreasing order
    unsigned int prev = 0;
    for (unsigned int i = 0; i < n && ok; i++) {
        const unsigned int c = DISCR::apply(ox[i], oy[i], oz[i] != 0);
        ok = c >= prev;
        prev = c;
    }
    hipFree(dx);
    hipFree(ox_d);
    hipFree(dy);
    hipFree(oy_d);
    hipFree(dz);
    hipFree(oz_d);
    std::printf("mdiscr (%s) on %u elems (m=%u): %s\n", DISCR::select ? "select" : "tuples", n, m,
                ok ? "VALID RESULT!" : "INVALID RESULT!");
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}

template <unsigned int M>
int dispatch_mode(const char* mode, unsigned int n) {
    if (std::strcmp(mode, "arrays") == 0) return test_arrays<ModArray<M, 3>, 3>(n, M);
    if (std::strcmp(mode, "tuples") == 0) return test_tuples<ModTuple<M> >(n, M);
    if (std::strcmp(mode, "select") == 0) return test_tuples<ModTupleSelect<M> >(n, M);
    std::printf("Unknown mode %s.\n", mode);
    return EXIT_FAILURE;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("The program takes <num_elems> <m> [segm] as arguments!\n");
        return EXIT_FAILURE;
    }
    const unsigned int n = (unsigned int)std::strtoul(argv[1], nullptr, 10);
    const unsigned int m = (unsigned int)std::strtoul(argv[2], nullptr, 10);
    const bool segm = argc > 3 && std::strcmp(argv[3], "segm") == 0;
    if (argc > 3 && !segm) {
        if (m == 4 && std::strcmp(argv[3], "arrays") == 0) return test_arrays<ModArray4<3>, 3>(n, 4);
        if (m == 4 && std::strcmp(argv[3], "tuples") == 0) return test_tuples<ModTuple4>(n, 4);
        switch (m) {
            case 4: return dispatch_mode<4>(argv[3], n);
            case 5: return dispatch_mode<5>(argv[3], n);
            case 8: return dispatch_mode<8>(argv[3], n);
            case 64: return dispatch_mode<64>(argv[3], n);
            default: std::printf("Unsupported number of equivalence classes.\n"); return EXIT_FAILURE;
        }
    }
    switch (m) {
        case 4: return test<Mod4>(n, m, segm);
        case 8: return test<Mod8>(n, m, segm);
        case 3: return test<Mod<3> >(n, m, segm);
        case 5: return test<Mod<5> >(n, m, segm);
        case 16: return test<Mod<16> >(n, m, segm);
        case 64: return test<Mod<64> >(n, m, segm);
        default: std::printf("Unsupported number of equivalence classes.\n"); return EXIT_FAILURE;
    }
}
