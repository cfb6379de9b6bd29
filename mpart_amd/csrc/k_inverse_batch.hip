#include "launchers.hpp"
#include "mono_inverse.hpp"

namespace mpd {

// Zeroes the point counter and the hand-over count of a pass, writes the
// hand-over arguments where the solve pass reads them (Hd), and on the
// call's first pass (component 0) zeroes the hand-over statistics.
__global__ void k_zero_u64_batch(unsigned long long* p, InvHeavy H, InvHeavy* Hd, int first) {
    *p = 0ull;
    if (H.count) *H.count = 0ull;
    if (H.take) *H.take = 0ull;
    if (Hd) *Hd = H;
    if (first && H.stats)
        for (int k = 0; k < HS_COUNT; ++k) H.stats[k] = 0ull;
}

template <int B, int Q, int R, int PM>
static int inverse_batch_occupancy(int block) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_inverse_queue<B, Q, R, PM, false, true>, block, 0) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

int occupancy_inverse_batch(int var, int block) {
#define CALL(B, Q, R, PM) return inverse_batch_occupancy<B, Q, R, PM>(block)
    switch (var) {
        MPD_CASES_PM(0, 4, CALL)
        MPD_CASES_PM(1, 12, CALL)
        default: return 1;
    }
#undef CALL
}

hipError_t launch_inverse_batch(int var, const Launch& L, int ntiles, dim3 qgrid, dim3 qblock, const MapDev& M0,
                                const MapDev* maps, const InvSeg* segs, const int* tileSeg, const long* tileStart,
                                const int* chunkSeg, long Ntot, int c, double xtol, double ytol, double* rec,
                                unsigned long long* counter, const InvHeavy& H, InvHeavy* Hd, int heavyBlocks,
                                int ntilesGrp, int tileN, size_t grpSmem) {
    if (ntiles <= 0 || Ntot <= 0) return hipSuccess;
    const int ntiles1 = ntiles - ntilesGrp;
    const int parts = (tileN + 63) / 64;  // 256-lane blocks of 64 points (4 lanes each) per tile
    hipLaunchKernelGGL(k_zero_u64_batch, dim3(1), dim3(1), 0, L.stream, counter, H, Hd, c == 0 ? 1 : 0);
#define CALL(B, Q, R, PM)                                                                                           \
    if (ntiles1 > 0)                                                                                              \
        hipLaunchKernelGGL((k_inverse_setup_batch<B, Q, R, PM>), dim3((unsigned)ntiles1), L.block, L.smem, L.stream, \
                           maps, segs, tileSeg, tileStart, Ntot, c, rec);                                          \
    if (ntilesGrp > 0)                                                                                            \
        hipLaunchKernelGGL((k_inverse_setup_batch_grp<B, Q, R, PM, 4>), dim3((unsigned)(ntilesGrp * parts)),       \
                           dim3(256), grpSmem, L.stream, maps, segs, tileSeg + ntiles1, tileStart + ntiles1,       \
                           tileN, Ntot, c, rec);                                                                   \
    if (M0.lin) {                                                                                                 \
        hipLaunchKernelGGL((k_inverse_queue<B, Q, R, PM, true, true>), qgrid, qblock, 0, L.stream, M0, rec, Ntot,   \
                           nullptr, 0, c, xtol, ytol, counter, segs, chunkSeg, Hd, H.solveSteps);                                \
        launch_inverse_heavy<B, Q, R, PM, true, true>(heavyBlocks, L.stream, M0, rec, Ntot, nullptr, 0,              \
            c, xtol, ytol, H, segs, chunkSeg);                                                                       \
    } else {                                                                                                      \
        hipLaunchKernelGGL((k_inverse_queue<B, Q, R, PM, false, true>), qgrid, qblock, 0, L.stream, M0, rec, Ntot,  \
                           nullptr, 0, c, xtol, ytol, counter, segs, chunkSeg, Hd, H.solveSteps);                                \
        launch_inverse_heavy<B, Q, R, PM, false, true>(heavyBlocks, L.stream, M0, rec, Ntot, nullptr, 0,             \
            c, xtol, ytol, H, segs, chunkSeg);                                                                       \
    }
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

}  // namespace mpd
