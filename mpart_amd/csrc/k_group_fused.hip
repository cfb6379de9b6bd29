// Launchers of the group-per-point fused kernels (G lanes per point).
#include "launchers.hpp"
#include "mono_kernels.hpp"

namespace mpd {

hipError_t launch_fused_grp(int var, int G, const Launch& L, const MapDev& M, const double* w, long ws,
                            const double* pts, long prs, long pcs, long N, const double* sens, long srs, long scs,
                            const double* lsens, long lss, int klMode, double* rows) {
    PtsArg P{pts, prs, pcs, N};
    if (G == 4) {
#define CALL(B, Q, R, PM)                                                                                         \
    hipLaunchKernelGGL((k_fused_grp<B, Q, R, PM, 4>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, sens, srs, \
                       scs, lsens, lss, klMode, rows)
        MPD_DISPATCH(var, CALL)
#undef CALL
    } else if (G == 8) {
#define CALL(B, Q, R, PM)                                                                                         \
    hipLaunchKernelGGL((k_fused_grp<B, Q, R, PM, 8>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, sens, srs, \
                       scs, lsens, lss, klMode, rows)
        MPD_DISPATCH(var, CALL)
#undef CALL
    } else if (G == 64) {
#define CALL(B, Q, R, PM)                                                                                          \
    hipLaunchKernelGGL((k_fused_grp<B, Q, R, PM, 64>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, sens, srs, \
                       scs, lsens, lss, klMode, rows)
        MPD_DISPATCH(var, CALL)
#undef CALL
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int occupancy_fused_grp(int var, int G, int block, size_t smem) {
    int nb = 0;
#define CALL(B, Q, R, PM)                                                                                     \
    if (G == 4)                                                                                               \
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused_grp<B, Q, R, PM, 4>, block, smem);    \
    else if (G == 8)                                                                                          \
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused_grp<B, Q, R, PM, 8>, block, smem);    \
    else                                                                                                      \
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused_grp<B, Q, R, PM, 64>, block, smem);
    switch (var) {
        MPD_CASES_PM(0, 4, CALL)
        MPD_CASES_PM(1, 12, CALL)
        default: return 1;
    }
#undef CALL
    return nb > 0 ? nb : 1;
}

}  // namespace mpd
