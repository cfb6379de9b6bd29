#include "launchers.hpp"
#include "mono_kernels.hpp"

namespace mpd {

hipError_t launch_coeff_grad(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                             long prs, long pcs, long N, const double* sens, long srs, long scs, double* out, long ors,
                             long ocs, int c, double* gtab) {
    PtsArg P{pts, prs, pcs, N};
    OutArg O{out, ors, ocs};
#define CALL(B, Q, R, PM) \
    hipLaunchKernelGGL((k_coeff_grad<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, sens, srs, scs, O, c, gtab)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_logdet_coeff_grad(int var, const Launch& L, const MapDev& M, const double* w, long ws,
                                    const double* pts, long prs, long pcs, long N, double* out, long ors, long ocs,
                                    int c, double* gtab) {
    PtsArg P{pts, prs, pcs, N};
    OutArg O{out, ors, ocs};
#define CALL(B, Q, R, PM) \
    hipLaunchKernelGGL((k_logdet_coeff_grad<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, O, c, gtab)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_logdet_discrete(int var, const Launch& L, const MapDev& M, const double* w, long ws,
                                  const double* pts, long prs, long pcs, long N, double* ld, long lds, int ldAcc,
                                  int cb, int ce, double* gtab) {
    PtsArg P{pts, prs, pcs, N};
#define CALL(B, Q, R, PM)                                                                                            \
    hipLaunchKernelGGL((k_logdet_discrete<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, ld, lds, ldAcc, \
                       cb, ce, gtab)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_logdet_coeff_grad_discrete(int var, const Launch& L, const MapDev& M, const double* w, long ws,
                                             const double* pts, long prs, long pcs, long N, double* out, long ors,
                                             long ocs, int c, double* gtab) {
    PtsArg P{pts, prs, pcs, N};
    OutArg O{out, ors, ocs};
#define CALL(B, Q, R, PM)                                                                                           \
    hipLaunchKernelGGL((k_logdet_coeff_grad_discrete<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, O, c, \
                       gtab)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

}  // namespace mpd
