#include "launchers.hpp"
#include "mono_kernels.hpp"

namespace mpd {

hipError_t launch_inverse(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* xs,
                          long xrows, long xcols, long xrs, long xcs, const double* ys, long yss, long N, double* out,
                          long os, int c, double xtol, double ytol, double* gtab) {
#define CALL(B, Q, R, PM)                                                                                               \
    hipLaunchKernelGGL((k_inverse<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, xs, xrows, xcols, xrs, xcs, \
                       ys, yss, N, out, os, c, xtol, ytol, gtab)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

}  // namespace mpd
