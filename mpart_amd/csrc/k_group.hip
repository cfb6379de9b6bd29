// Launchers of the group-per-point Evaluate / LogDeterminant kernels (G lanes per point).
#include "launchers.hpp"
#include "mono_kernels.hpp"

namespace mpd {

hipError_t launch_eval_grp(int var, int G, const Launch& L, const MapDev& M, const double* w, long ws,
                           const double* pts, long prs, long pcs, long N, double* out, long ors, long ocs, double* ld,
                           long lds, int ldAcc, int* nevals, int cb, int ce) {
    PtsArg P{pts, prs, pcs, N};
    OutArg O{out, ors, ocs};
    if (G == 4) {
#define CALL(B, Q, R, PM)                                                                                        \
    hipLaunchKernelGGL((k_eval_grp<B, Q, R, PM, 4>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, O, ld, lds, \
                       ldAcc, nevals, cb, ce)
        MPD_DISPATCH(var, CALL)
#undef CALL
    } else if (G == 8) {
#define CALL(B, Q, R, PM)                                                                                        \
    hipLaunchKernelGGL((k_eval_grp<B, Q, R, PM, 8>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, O, ld, lds, \
                       ldAcc, nevals, cb, ce)
        MPD_DISPATCH(var, CALL)
#undef CALL
    } else if (G == 64) {
#define CALL(B, Q, R, PM)                                                                                         \
    hipLaunchKernelGGL((k_eval_grp<B, Q, R, PM, 64>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, O, ld, lds, \
                       ldAcc, nevals, cb, ce)
        MPD_DISPATCH(var, CALL)
#undef CALL
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mpd
