#include "launchers.hpp"
#include "mono_inverse.hpp"

namespace mpd {

// Zeroes the point counter and the hand-over count of a pass, writes the
// hand-over arguments where the solve pass reads them (Hd), and on the
// call's first pass (component 0) zeroes the hand-over statistics.
__global__ void k_zero_u64(unsigned long long* p, InvHeavy H, InvHeavy* Hd, int first) {
    *p = 0ull;
    if (H.count) *H.count = 0ull;
    if (H.take) *H.take = 0ull;
    if (Hd) *Hd = H;
    if (first && H.stats)
        for (int k = 0; k < HS_COUNT; ++k) H.stats[k] = 0ull;
}

template <int B, int Q, int R, int PM>
static int inverse_occupancy(int block) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_inverse_queue<B, Q, R, PM, false, false>, block, 0) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

int occupancy_inverse(int var, int block) {
#define CALL(B, Q, R, PM) return inverse_occupancy<B, Q, R, PM>(block)
    switch (var) {
        MPD_CASES_PM(0, 4, CALL)
        MPD_CASES_PM(1, 12, CALL)
        default: return 1;
    }
#undef CALL
}

int inverse_record_rows(int var) {
#define CALL(B, Q, R, PM) return inv_rec_rows<PM>()
    switch (var) {
        MPD_CASES_PM(0, 4, CALL)
        MPD_CASES_PM(1, 12, CALL)
        default: return inv_rec_rows<MAXP>();
    }
#undef CALL
}

hipError_t launch_inverse(int var, const Launch& L, dim3 qgrid, dim3 qblock, const MapDev& M, const double* w, long ws,
                          const double* xs, long xrows, long xcols, long xrs, long xcs, const double* ys, long yss,
                          long N, double* out, long os, int c, double xtol, double ytol, double* gtab, double* rec,
                          unsigned long long* counter, const InvHeavy& H, InvHeavy* Hd, int heavyBlocks,
                          size_t grpSmem) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, L.stream, counter, H, Hd, c == 0 ? 1 : 0);
#define CALL(B, Q, R, PM)                                                                                          \
    if (grpSmem > 0)                                                                                             \
        hipLaunchKernelGGL((k_inverse_setup_grp<B, Q, R, PM, 4>), dim3((unsigned)((N + 63) / 64)), dim3(256),     \
                           grpSmem, L.stream, M, w, ws, xs, xrows, xcols, xrs, xcs, ys, yss, N, c, rec);           \
    else                                                                                                         \
        hipLaunchKernelGGL((k_inverse_setup<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, M, w, ws, xs, xrows,  \
                           xcols, xrs, xcs, ys, yss, N, c, gtab, rec);                                             \
    if (M.lin) {                                                                                                 \
        hipLaunchKernelGGL((k_inverse_queue<B, Q, R, PM, true, false>), qgrid, qblock, 0, L.stream, M, rec, N, out, \
                           os, c, xtol, ytol, counter, nullptr, nullptr, Hd, H.solveSteps);                                     \
        launch_inverse_heavy<B, Q, R, PM, true, false>(heavyBlocks, L.stream, M, rec, N, out, os,                    \
            c, xtol, ytol, H, nullptr, nullptr);                                                                     \
    } else {                                                                                                     \
        hipLaunchKernelGGL((k_inverse_queue<B, Q, R, PM, false, false>), qgrid, qblock, 0, L.stream, M, rec, N, out, \
                           os, c, xtol, ytol, counter, nullptr, nullptr, Hd, H.solveSteps);                                     \
        launch_inverse_heavy<B, Q, R, PM, false, false>(heavyBlocks, L.stream, M, rec, N, out, os,                   \
            c, xtol, ytol, H, nullptr, nullptr);                                                                     \
    }
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

}  // namespace mpd
