// Launcher of the component-split fused kernel (mono_split.hpp).
#include "launchers.hpp"
#include "mono_split.hpp"

namespace mpd {

static_assert(FUSED_SPLIT_TILE == SPLIT_TILE && FUSED_SPLIT_WAVES == SPLIT_WAVES, "split schedule shape");

// Two builds: at least two wavefronts per SIMD (minb 1, here) and at least
// four (minb 2, in k_split2.hip, so the two compile in parallel).
hipError_t launch_fused_split2(int var, const Launch& L, const MapDev& M, const double* w, long ws,
                               const double* pts, long prs, long pcs, long N, const double* sens, long srs, long scs,
                               const double* lsens, long lss, int klMode, const int* compOrder, const int* waveStart,
                               double* rows);
int occupancy_fused_split2(int var, size_t smem);

hipError_t launch_fused_split(int var, int minb, const Launch& L, const MapDev& M, const double* w, long ws,
                              const double* pts, long prs, long pcs, long N, const double* sens, long srs, long scs,
                              const double* lsens, long lss, int klMode, const int* compOrder, const int* waveStart,
                              double* rows) {
    if (minb == 2)
        return launch_fused_split2(var, L, M, w, ws, pts, prs, pcs, N, sens, srs, scs, lsens, lss, klMode, compOrder,
                                   waveStart, rows);
    PtsArg P{pts, prs, pcs, N};
#define CALL(B, Q, R, PM)                                                                                           \
    hipLaunchKernelGGL((k_fused_split<B, Q, R, PM, 2>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, sens, srs, \
                       scs, lsens, lss, klMode, compOrder, waveStart, rows)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

int occupancy_fused_split(int var, int minb, size_t smem) {
    if (minb == 2) return occupancy_fused_split2(var, smem);
    int nb = 0;
#define CALL(B, Q, R, PM) \
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused_split<B, Q, R, PM, 2>, SPLIT_WAVES * 64, smem);
    switch (var) {
        MPD_CASES_PM(0, 4, CALL)
        MPD_CASES_PM(1, 12, CALL)
        default: return 1;
    }
#undef CALL
    return nb > 0 ? nb : 1;
}

}  // namespace mpd
