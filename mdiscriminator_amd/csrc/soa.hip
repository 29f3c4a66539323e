// soa.hip -- struct-of-arrays partition (placeholder until the SoA kernels land).
#include "api_internal.hpp"

extern "C" mdiscr_status mdiscr_partition_soa(mdiscr_ctx* ctx, uint32_t, const void* const*,
                                              const uint32_t*, uint32_t, void* const*, uint64_t,
                                              mdiscr_discr, uint32_t*, void*) {
    if (!ctx) return MDISCR_EINVAL;
    return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "SoA partition not built yet");
}
