#include "launchers.hpp"
#include "mono_kernels.hpp"

namespace mpd {

hipError_t launch_input_grad(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                             long prs, long pcs, long N, const double* sens, long srs, long scs, double* out, long ors,
                             long ocs, int c, int mode, double* gtab) {
    PtsArg P{pts, prs, pcs, N};
    OutArg O{out, ors, ocs};
#define CALL(B, Q, R, PM)                                                                                              \
    hipLaunchKernelGGL((k_input_grad<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, sens, srs, scs, O, \
                       c, mode, gtab)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_input_vjp(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                            long prs, long pcs, long N, int klMode, const double* sens, long srs, long scs,
                            const double* lsens, long lss, double lval, double scale, double* out, long ors, long ocs,
                            double* gtab) {
    PtsArg P{pts, prs, pcs, N};
    OutArg O{out, ors, ocs};
#define CALL(B, Q, R, PM)                                                                                          \
    hipLaunchKernelGGL((k_input_vjp<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, klMode, sens, \
                       srs, scs, lsens, lss, lval, scale, O, gtab)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

}  // namespace mpd
