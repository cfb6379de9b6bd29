// Launcher of the component-split fused kernel built for four wavefronts per
// SIMD (mono_split.hpp); dispatched from k_split.hip.
#include "launchers.hpp"
#include "mono_split.hpp"

namespace mpd {

hipError_t launch_fused_split2(int var, const Launch& L, const MapDev& M, const double* w, long ws,
                               const double* pts, long prs, long pcs, long N, const double* sens, long srs, long scs,
                               const double* lsens, long lss, int klMode, const int* compOrder, const int* waveStart,
                               double* rows) {
    PtsArg P{pts, prs, pcs, N};
#define CALL(B, Q, R, PM)                                                                                           \
    hipLaunchKernelGGL((k_fused_split<B, Q, R, PM, 4>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, sens, srs, \
                       scs, lsens, lss, klMode, compOrder, waveStart, rows)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

int occupancy_fused_split2(int var, size_t smem) {
    int nb = 0;
#define CALL(B, Q, R, PM) \
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused_split<B, Q, R, PM, 4>, SPLIT_WAVES * 64, smem);
    switch (var) {
        MPD_CASES_PM(0, 4, CALL)
        MPD_CASES_PM(1, 12, CALL)
        default: return 1;
    }
#undef CALL
    return nb > 0 ? nb : 1;
}

}  // namespace mpd
