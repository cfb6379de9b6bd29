#define MPD_UTILITY_KERNELS
#include "launchers.hpp"
#include "mono_kernels.hpp"

namespace mpd {

hipError_t launch_eval(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts, long prs,
                       long pcs, long N, double* out, long ors, long ocs, double* ld, long lds, int ldAcc, int* nevals,
                       int cb, int ce, double* gtab) {
    PtsArg P{pts, prs, pcs, N};
    OutArg O{out, ors, ocs};
#define CALL(B, Q, R, PM) \
    hipLaunchKernelGGL((k_eval<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, P, O, ld, lds, ldAcc, nevals, cb, ce, gtab)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_reduce_rows(const double* in, long nrows, long K, double scale, double* out, hipStream_t s) {
    return launch_reduce_cols(in, nrows, K, K, scale, out, nullptr, s);
}
hipError_t launch_reduce_cols(const double* in, long nrows, long rowStride, long ncols, double scale, double* out,
                              double* outLast, hipStream_t s) {
    if (ncols <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_cols, dim3((unsigned)((ncols + 7) / 8)), dim3(256), 0, s, in, nrows, rowStride, ncols,
                       scale, out, outLast);
    return hipGetLastError();
}
hipError_t launch_colsum(const double* in, long nrows, long rowStride, long ncols, double scale, double* out,
                         double* outLast, double* work, hipStream_t s) {
    if (ncols <= 0) return hipSuccess;
    const long nchunks = std::max<long>(1, std::min<long>(COLSUM_MAX_CHUNKS, (nrows + 15) / 16));
    const long chunk = std::max<long>(1, (nrows + nchunks - 1) / nchunks);
    hipLaunchKernelGGL(k_colsum_chunks, dim3((unsigned)((ncols + 63) / 64), (unsigned)nchunks), dim3(256), 0, s, in,
                       nrows, rowStride, ncols, chunk, work);
    hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, s, work, nchunks, ncols,
                       scale, out, outLast);
    return hipGetLastError();
}
hipError_t launch_kl_rows(const double* G, const double* LG, long N, long K, double* grad, hipStream_t s) {
    hipLaunchKernelGGL(k_kl_rows, dim3((unsigned)K), dim3(256), 0, s, G, LG, N, K, grad);
    return hipGetLastError();
}
hipError_t launch_kl_loss(const double* T, long M, const double* ld, long N, double* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_kl_loss, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, T, M, ld, N, partial);
    return hipGetLastError();
}
hipError_t launch_fill(double* p, long rows, long cols, long rs, long cs, double v, hipStream_t s) {
    long n = rows * cols;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, rows, cols, rs, cs, v);
    return hipGetLastError();
}
hipError_t launch_copy(const double* src, long rows, long cols, long srs, long scs, double* dst, long drs, long dcs,
                       hipStream_t s) {
    long n = rows * cols;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, rows, cols, srs, scs, dst, drs,
                       dcs);
    return hipGetLastError();
}

}  // namespace mpd
