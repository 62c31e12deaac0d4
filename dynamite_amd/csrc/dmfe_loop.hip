// RHS row launches of the device-driven loops: the many-instance batch and one context's
// device loop (device code in dmfe_device.h).
#include "dmfe_device.h"

namespace dmfe {

void launch_batch_rhs(InstDesc* D, const int4* map, int nblocks, int n, int log_response, int G, int maxm,
                      size_t lds, hipStream_t s) {
#define DMFE_BRHS(LOG, GG, NT) \
    hipLaunchKernelGGL((batch_rhs_kernel<LOG, GG, NT>), dim3(nblocks), dim3(kRowThreads), lds, s, D, map, n)
    DMFE_SHAPES(log_response, G, b2_taps_per_pass(G, maxm), DMFE_BRHS)
#undef DMFE_BRHS
}

void launch_loop_rhs(const InstDesc& d, int n, int log_response, int G, size_t lds, hipStream_t s) {
    const dim3 grid((unsigned)(d.T.N * d.split));
    if (d.X.K) {
#define DMFE_LRT(GG, KK) hipLaunchKernelGGL((loop_rhs_tan_kernel<GG, KK>), grid, dim3(kRowThreads), lds, s, d, n)
#define DMFE_LRTG(KK)                       \
    switch (G) {                            \
        case 4: DMFE_LRT(4, KK); break;     \
        case 8: DMFE_LRT(8, KK); break;     \
        case 16: DMFE_LRT(16, KK); break;   \
        default: DMFE_LRT(64, KK); break;   \
    }
        if (d.X.K == 1) DMFE_LRTG(1)
        else DMFE_LRTG(2)
#undef DMFE_LRTG
#undef DMFE_LRT
        return;
    }
#define DMFE_LR(LOG, GG, NT) hipLaunchKernelGGL((loop_rhs_kernel<LOG, GG, NT>), grid, dim3(kRowThreads), lds, s, d, n)
    DMFE_SHAPES(log_response, G, b2_taps_per_pass(G, d.T.m), DMFE_LR)
#undef DMFE_LR
}


}  // namespace dmfe
