// Launcher of the component-split Evaluate / LogDeterminant kernel (mono_split.hpp).
#include "launchers.hpp"
#include "mono_split.hpp"

namespace mpd {

hipError_t launch_eval_split(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                             long prs, long pcs, long N, double* out, long ors, long ocs, double* ld, long lds,
                             int ldAccumulate, int* nevals, const int* compOrder, const int* waveStart) {
    PtsArg P{pts, prs, pcs, N};
    OutArg O{out, ors, ocs};
#define CALL(B, Q, R, PM)                                                                                       \
    hipLaunchKernelGGL((k_eval_split<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, O, ld, lds, \
                       ldAccumulate, nevals, compOrder, waveStart)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

int occupancy_eval_split(int var, size_t smem) {
    int nb = 0;
#define CALL(B, Q, R, PM) \
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_eval_split<B, Q, R, PM>, SPLIT_WAVES * 64, smem);
    switch (var) {
        MPD_CASES_PM(0, 4, CALL)
        MPD_CASES_PM(1, 12, CALL)
        default: return 1;
    }
#undef CALL
    return nb > 0 ? nb : 1;
}

}  // namespace mpd
