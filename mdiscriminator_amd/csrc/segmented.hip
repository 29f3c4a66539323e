// segmented.hip -- segmented partition (placeholder until the batched kernels land).
#include "api_internal.hpp"

void ctx_release_extra(mdiscr_ctx* ctx) {
    if (ctx->small) (void)hipFree(ctx->small);
    ctx->small = nullptr;
}

extern "C" mdiscr_status mdiscr_partition_segmented(mdiscr_ctx* ctx, const uint32_t*, uint32_t*,
                                                    uint64_t, const uint64_t*, const uint64_t*,
                                                    uint64_t, mdiscr_discr, uint32_t*, void*) {
    if (!ctx) return MDISCR_EINVAL;
    return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "segmented partition not built yet");
}

extern "C" mdiscr_status mdiscr_segm_flags(mdiscr_ctx* ctx, uint64_t, const uint32_t*,
                                           const int32_t*, uint32_t*, int32_t*, mdiscr_discr,
                                           void*) {
    if (!ctx) return MDISCR_EINVAL;
    return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "segmented partition not built yet");
}
