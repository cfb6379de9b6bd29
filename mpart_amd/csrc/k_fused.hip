#include "launchers.hpp"
#include "mono_kernels.hpp"

namespace mpd {

hipError_t launch_fused(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                        long prs, long pcs, long N, const double* sens, long srs, long scs, const double* lsens,
                        long lss, int klMode, double* partialGrad, double* partialLoss, double* gtab,
                        double* gradGlobal) {
    PtsArg P{pts, prs, pcs, N};
#define CALL(B, Q, R)                                                                                               \
    hipLaunchKernelGGL((k_fused<B, Q, R>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, sens, srs, scs, lsens, lss, \
                       klMode, partialGrad, partialLoss, gtab, gradGlobal)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

int occupancy_fused(int var, int block, size_t smem) {
    int nb = 0;
#define CALL(B, Q, R) \
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused<B, Q, R>, block, smem)
    switch (var) {
        case 0: CALL(0, 0, 0); break;
        case 1: CALL(0, 0, 1); break;
        case 2: CALL(0, 0, 2); break;
        case 10: CALL(0, 1, 0); break;
        case 11: CALL(0, 1, 1); break;
        case 12: CALL(0, 1, 2); break;
        case 100: CALL(1, 0, 0); break;
        case 101: CALL(1, 0, 1); break;
        case 102: CALL(1, 0, 2); break;
        case 110: CALL(1, 1, 0); break;
        case 111: CALL(1, 1, 1); break;
        case 112: CALL(1, 1, 2); break;
        default: return 1;
    }
#undef CALL
    return nb > 0 ? nb : 1;
}

}  // namespace mpd
