// ref_shim.cpp -- C entry points onto the reference's own SeqMdiscr.
//
// TEST INFRASTRUCTURE ONLY.  This translation unit contains no algorithm: it
// includes the reference headers where they lie (cuda_code/HelpersCommon.h,
// which includes cuda_code/SeqMdiscr.h) and instantiates
// seqMdiscr<Mod<M>,false> (SeqMdiscr.h:10-57), seqMdiscr<Mod4/Mod8,false> and
// tupleToSizeArray (HelpersCommon.h:295-311) behind extern "C" symbols, so the
// oracle's restatement can be checked against the reference itself.  Built by
// oracle/Makefile into oracle/_ref/libseqref.so; never part of the product.
#include <cstring>
#include <cstdint>
#include "HelpersCommon.h"

namespace {

template <unsigned M>
void run_mod(unsigned n, unsigned* in, unsigned* out, unsigned* sizes) {
    seqMdiscr<Mod<M>, false>(n, in, out, sizes);
}

typedef void (*mod_fn)(unsigned, unsigned*, unsigned*, unsigned*);

template <unsigned... Ms>
struct Table {
    static mod_fn get(unsigned m) {
        mod_fn fns[] = {nullptr, &run_mod<Ms>...};
        return m < sizeof(fns) / sizeof(fns[0]) ? fns[m] : nullptr;
    }
};

typedef Table<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19,
              20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36,
              37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53,
              54, 55, 56, 57, 58, 59, 60, 61, 62, 63, 64>
    ModTable;

}  // namespace

extern "C" {

// sizes must have n+1 entries: the reference writes sizes[n] when trailing
// classes are empty (SeqMdiscr.h:47-49).
int ref_seq_mdiscr_mod(unsigned n, unsigned* in, unsigned* out,
                       unsigned* sizes, unsigned m) {
    mod_fn f = ModTable::get(m);
    if (!f) return -1;
    f(n, in, out, sizes);
    return 0;
}

int ref_seq_mdiscr_mod4(unsigned n, unsigned* in, unsigned* out,
                        unsigned* sizes) {
    seqMdiscr<Mod4, false>(n, in, out, sizes);
    return 0;
}

int ref_seq_mdiscr_mod8(unsigned n, unsigned* in, unsigned* out,
                        unsigned* sizes) {
    seqMdiscr<Mod8, false>(n, in, out, sizes);
    return 0;
}

// counts (m <= 64 entries) -> n-length sizes array via the reference's
// tupleToSizeArray on its MyInt<64> tuple type.
int ref_tuple_to_size_array(const unsigned* counts, unsigned m, unsigned n,
                            unsigned* sizes) {
    if (m == 0 || m > 64) return -1;
    MyInt<64> t;
    for (unsigned k = 0; k < m; k++) t[k] = counts[k];
    tupleToSizeArray<MyInt<64> >(t, sizes, n);
    return 0;
}

}  // extern "C"
