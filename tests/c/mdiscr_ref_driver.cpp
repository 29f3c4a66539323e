// mdiscr_ref_driver.cpp -- test executable in the shape of the reference's
// MdiscrTest_optimized_clean.cu / MdiscrSegmTest.cu drivers: argv[1] = N,
// argv[2] = m (4 and 8 use the Mod4/Mod8 policies, anything else Mod<m>),
// argv[3] = "segm" for the flag-array segmented call.  Seeded input (the
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

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("The program takes <num_elems> <m> [segm] as arguments!\n");
        return EXIT_FAILURE;
    }
    const unsigned int n = (unsigned int)std::strtoul(argv[1], nullptr, 10);
    const unsigned int m = (unsigned int)std::strtoul(argv[2], nullptr, 10);
    const bool segm = argc > 3 && std::strcmp(argv[3], "segm") == 0;
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
