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


// Batch launches with K tangent directions per instance (batch_stage_tan_kernel,
// batch_rhs_tan_kernel).
void launch_batch_stage_tan(InstDesc* D, const int2* map, int nblocks, int n, int ept, int K, hipStream_t s) {
#define DMFE_BST(E, KK) hipLaunchKernelGGL((batch_stage_tan_kernel<E, KK>), dim3(nblocks), dim3(kStageThreads), 0, s, D, map, n)
#define DMFE_BSTE(KK)                       \
    if (ept <= 1) DMFE_BST(1, KK);          \
    else if (ept <= 2) DMFE_BST(2, KK);     \
    else if (ept <= 4) DMFE_BST(4, KK);     \
    else DMFE_BST(8, KK);
    if (K == 1) { DMFE_BSTE(1) }
    else { DMFE_BSTE(2) }
#undef DMFE_BSTE
#undef DMFE_BST
}

void launch_batch_rhs_tan(InstDesc* D, const int4* map, int nblocks, int n, int G, int K, size_t lds, hipStream_t s) {
#define DMFE_BRT(GG, KK) hipLaunchKernelGGL((batch_rhs_tan_kernel<GG, KK>), dim3(nblocks), dim3(kRowThreads), lds, s, D, map, n)
#define DMFE_BRTG(KK)                       \
    switch (G) {                            \
        case 4: DMFE_BRT(4, KK); break;     \
        case 8: DMFE_BRT(8, KK); break;     \
        case 16: DMFE_BRT(16, KK); break;   \
        default: DMFE_BRT(64, KK); break;   \
    }
    if (K == 1) DMFE_BRTG(1)
    else DMFE_BRTG(2)
#undef DMFE_BRTG
#undef DMFE_BRT
}

}  // namespace dmfe
