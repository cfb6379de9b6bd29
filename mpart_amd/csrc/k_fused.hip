#include "launchers.hpp"
#include "mono_kernels.hpp"

namespace mpd {

hipError_t launch_fused(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                        long prs, long pcs, long N, const double* sens, long srs, long scs, const double* lsens,
                        long lss, int klMode, double* partialGrad, double* partialLoss, double* gtab,
                        double* gradGlobal) {
    PtsArg P{pts, prs, pcs, N};
#define CALL(B, Q, R, PM)                                                                                               \
    hipLaunchKernelGGL((k_fused<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, sens, srs, scs, lsens, lss, \
                       klMode, partialGrad, partialLoss, gtab, gradGlobal)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

static hipError_t occ_query(int var, int block, size_t smem, int* nb) {
#define CALL(B, Q, R, PM) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, k_fused<B, Q, R, PM>, block, smem)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipSuccess;
}

int occupancy_fused(int var, int block, size_t smem) {
    int nb = 0;
    if (occ_query(var, block, smem, &nb) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace mpd
