#include "launchers.hpp"
#include "mono_batch.hpp"

namespace mpd {

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

template <int B, int Q, int R, int PM>
static int batch_occupancy(int block, size_t smem) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<B, Q, R, PM>, block, smem) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

int occupancy_batch(int var, int block, size_t smem) {
#define CALL(B, Q, R, PM) return batch_occupancy<B, Q, R, PM>(block, smem)
    switch (var) {
        MPD_CASES_PM(0, 4, CALL)
        MPD_CASES_PM(1, 12, CALL)
        default: return 1;
    }
#undef CALL
}

hipError_t launch_batch(int var, const Launch& L, const MapDev* maps, const BatchIO* io, const int* tileMap,
                        const long* tileStart, int ntiles, unsigned* counter, int method, double xtol, double ytol) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, L.stream, counter);
#define CALL(B, Q, R, PM)                                                                                          \
    hipLaunchKernelGGL((k_batch<B, Q, R, PM>), L.grid, L.block, L.smem, L.stream, maps, io, tileMap, tileStart, ntiles, \
                       counter, method, xtol, ytol)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

template <int B, int Q, int R, int PM>
static int batch_grp_occupancy(int block, size_t smem) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_eval_mixed<B, Q, R, PM, 4>, block, smem) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

int occupancy_batch_eval_grp(int var, int block, size_t smem) {
#define CALL(B, Q, R, PM) return batch_grp_occupancy<B, Q, R, PM>(block, smem)
    switch (var) {
        MPD_CASES_PM(0, 4, CALL)
        MPD_CASES_PM(1, 12, CALL)
        default: return 1;
    }
#undef CALL
}

hipError_t launch_batch_eval_mixed(int var, const Launch& L, const MapDev* maps, const BatchIO* io,
                                   const int* tileMap, const long* tileStart, const int* tileLanes, int ntiles,
                                   unsigned* counter, int grpSlots) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, L.stream, counter);
#define CALL(B, Q, R, PM)                                                                                       \
    hipLaunchKernelGGL((k_batch_eval_mixed<B, Q, R, PM, 4>), L.grid, L.block, L.smem, L.stream, maps, io, tileMap, \
                       tileStart, tileLanes, ntiles, counter, grpSlots)
    MPD_DISPATCH(var, CALL)
#undef CALL
    return hipGetLastError();
}

}  // namespace mpd
