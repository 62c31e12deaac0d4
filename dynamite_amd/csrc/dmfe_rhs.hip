// RHS row launches of a context: rhs_rows_kernel and its tangent form (device code in
// dmfe_device.h).
#include "dmfe_device.h"

namespace dmfe {

void launch_tan_rhs(const Tables& T, const History& H, const Work& W, const Model& M, const RhsArgs& a,
                    const TanArgs& X, hipStream_t s) {
    const dim3 grid((unsigned)(T.N * a.split));
    const int cw = rhs_chunk_cap(T.N, a.split);
    const size_t lds = rhs_lds_bytes(T.N, cw, X.K);
#define DMFE_TR(GG, KK) \
    hipLaunchKernelGGL((rhs_rows_tan_kernel<GG, KK>), grid, dim3(kRowThreads), lds, s, T, H, W, M, a, X)
#define DMFE_TRG(KK)                         \
    switch (a.G) {                           \
        case 4: DMFE_TR(4, KK); break;       \
        case 8: DMFE_TR(8, KK); break;       \
        case 16: DMFE_TR(16, KK); break;     \
        default: DMFE_TR(64, KK); break;     \
    }
    if (X.K == 1) DMFE_TRG(1)
    else DMFE_TRG(2)
#undef DMFE_TRG
#undef DMFE_TR
}

void launch_rhs_rows(const Tables& T, const History& H, const Work& W, const Model& M, const RhsArgs& a,
                     hipStream_t s) {
    const dim3 grid((unsigned)(T.N * a.split));
    const int cw = rhs_chunk_cap(T.N, a.split);
    const size_t lds = rhs_lds_bytes(T.N, cw, 0);
#define DMFE_RHS(LOG, G, NT) \
    hipLaunchKernelGGL((rhs_rows_kernel<LOG, G, NT>), grid, dim3(kRowThreads), lds, s, T, H, W, M, a)
    DMFE_SHAPES(a.log_response, a.G, b2_taps_per_pass(a.G, T.m), DMFE_RHS)
#undef DMFE_RHS
}

}  // namespace dmfe
