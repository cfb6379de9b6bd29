// launchers.hpp -- host-side launchers for the kernel variants (one
// translation unit per kernel family so they compile in parallel).
#pragma once
#include <hip/hip_runtime.h>
#include "batch_io.hpp"
#include "mono_device.hpp"

namespace mpd {

struct PtsArg;
struct OutArg;

struct Launch {
    dim3 grid, block;
    size_t smem;
    hipStream_t stream;
};

// Variant id: bk * 100 + pk * 10 + qk.
hipError_t launch_eval(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts, long prs,
                       long pcs, long N, double* out, long ors, long ocs, double* ld, long lds, int ldAcc, int* nevals,
                       int cb, int ce, double* gtab);
// Group-per-point Evaluate / LogDeterminant (G = 4, 8 or 64 lanes per point;
// L.smem = block / G tables of nslots doubles).
hipError_t launch_eval_grp(int var, int G, const Launch& L, const MapDev& M, const double* w, long ws,
                           const double* pts, long prs, long pcs, long N, double* out, long ors, long ocs, double* ld,
                           long lds, int ldAcc, int* nevals, int cb, int ce);
// Group-per-point fused pass: one partial row of K + 1 per wavefront of the
// grid (rows = grid * block / 64).
hipError_t launch_fused_grp(int var, int G, const Launch& L, const MapDev& M, const double* w, long ws,
                            const double* pts, long prs, long pcs, long N, const double* sens, long srs, long scs,
                            const double* lsens, long lss, int klMode, double* rows);
int occupancy_fused_grp(int var, int G, int block, size_t smem);
constexpr int FUSED_SPLIT_TILE = 64;   // points per tile of the component-split fused pass
constexpr int FUSED_SPLIT_WAVES = 8;   // its wavefronts per block
// Fused pass with the components split over the eight wavefronts of a block
// (mono_split.hpp): block 512, LDS 8 * nslots * 64 bytes, one partial row of
// K + 1 doubles per block.
// minb: 1 for the build at two wavefronts per SIMD (256 registers), 2 for the
// build at four (128 registers, with spills).
hipError_t launch_fused_split(int var, int minb, const Launch& L, const MapDev& M, const double* w, long ws,
                              const double* pts, long prs, long pcs, long N, const double* sens, long srs, long scs,
                              const double* lsens, long lss, int klMode, const int* compOrder, const int* waveStart,
                              double* rows);
int occupancy_fused_split(int var, int minb, size_t smem);
// Evaluate / LogDeterminant of every component with the same split.
hipError_t launch_eval_split(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                             long prs, long pcs, long N, double* out, long ors, long ocs, double* ld, long lds,
                             int ldAccumulate, int* nevals, const int* compOrder, const int* waveStart);
int occupancy_eval_split(int var, size_t smem);
hipError_t launch_coeff_grad(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                             long prs, long pcs, long N, const double* sens, long srs, long scs, double* out, long ors,
                             long ocs, int c, double* gtab);
hipError_t launch_logdet_coeff_grad(int var, const Launch& L, const MapDev& M, const double* w, long ws,
                                    const double* pts, long prs, long pcs, long N, double* out, long ors, long ocs,
                                    int c, double* gtab);
hipError_t launch_logdet_discrete(int var, const Launch& L, const MapDev& M, const double* w, long ws,
                                  const double* pts, long prs, long pcs, long N, double* ld, long lds, int ldAcc,
                                  int cb, int ce, double* gtab);
hipError_t launch_logdet_coeff_grad_discrete(int var, const Launch& L, const MapDev& M, const double* w, long ws,
                                             const double* pts, long prs, long pcs, long N, double* out, long ors,
                                             long ocs, int c, double* gtab);
hipError_t launch_input_grad(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                             long prs, long pcs, long N, const double* sens, long srs, long scs, double* out, long ors,
                             long ocs, int c, int mode, double* gtab);
// Input VJP of all components (klMode: sensitivities T and -1 for the KL
// objective); lsens == nullptr uses lval for every point.
hipError_t launch_input_vjp(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                            long prs, long pcs, long N, int klMode, const double* sens, long srs, long scs,
                            const double* lsens, long lss, double lval, double scale, double* out, long ors, long ocs,
                            double* gtab);
// Inverse of component c: a setup pass (shape L; or, with grpSmem > 0, 4
// lanes per point in 256-lane blocks of grpSmem bytes, for maps with many
// terms) writes per-point records to rec, then the persistent solve pass
// (qgrid x qblock) takes points from the
// counter at rec + records; with H's budget set (adaptive rules), the solves
// it hands over (reading H from Hd, device memory) then run in heavyBlocks
// one-wavefront blocks.
hipError_t launch_inverse(int var, const Launch& L, dim3 qgrid, dim3 qblock, const MapDev& M, const double* w, long ws,
                          const double* xs, long xrows, long xcols, long xrs, long xcs, const double* ys, long yss,
                          long N, double* out, long os, int c, double xtol, double ytol, double* gtab, double* rec,
                          unsigned long long* counter, const InvHeavy& H, InvHeavy* Hd, int heavyBlocks,
                          size_t grpSmem);
int occupancy_inverse(int var, int block);
int inverse_record_rows(int var);
// Batched Inverse of component c over many maps: a counter reset, the setup
// pass (tiles of tileN = L.block points, a block each with L.smem bytes of
// basis tables; the last ntilesGrp tiles, of maps with many terms, with 4
// lanes per point in 256-lane blocks of grpSmem bytes), then the persistent
// solve pass over the Ntot indices of the segments.
struct InvSeg;
hipError_t launch_inverse_batch(int var, const Launch& L, int ntiles, dim3 qgrid, dim3 qblock, const MapDev& M0,
                                const MapDev* maps, const InvSeg* segs, const int* tileSeg, const long* tileStart,
                                const int* chunkSeg, long Ntot, int c, double xtol, double ytol, double* rec,
                                unsigned long long* counter, const InvHeavy& H, InvHeavy* Hd, int heavyBlocks,
                                int ntilesGrp, int tileN, size_t grpSmem);
int occupancy_inverse_batch(int var, int block);
hipError_t launch_fused(int var, const Launch& L, const MapDev& M, const double* w, long ws, const double* pts,
                        long prs, long pcs, long N, const double* sens, long srs, long scs, const double* lsens,
                        long lss, int klMode, double* partialGrad, double* partialLoss, double* gtab,
                        double* gradGlobal);
int occupancy_fused(int var, int block, size_t smem);

struct BatchIO;
hipError_t launch_batch(int var, const Launch& L, const MapDev* maps, const BatchIO* io, const int* tileMap,
                        const long* tileStart, int ntiles, unsigned* counter, int method, double xtol, double ytol);
int occupancy_batch(int var, int block, size_t smem);
// Batched Evaluate / LogDeterminant with a schedule per tile (tileLanes: 1, a
// lane per point on block points, or 4 lanes per point on block / 4 points
// with tables of grpSlots doubles).
hipError_t launch_batch_eval_mixed(int var, const Launch& L, const MapDev* maps, const BatchIO* io,
                                   const int* tileMap, const long* tileStart, const int* tileLanes, int ntiles,
                                   unsigned* counter, int grpSlots);
int occupancy_batch_eval_grp(int var, int block, size_t smem);

hipError_t launch_reduce_rows(const double* in, long nrows, long K, double scale, double* out, hipStream_t s);
hipError_t launch_reduce_cols(const double* in, long nrows, long rowStride, long ncols, double scale, double* out,
                              double* outLast, hipStream_t s);
// Column sums of many partial rows in two fixed-order stages; work holds
// COLSUM_MAX_CHUNKS * ncols doubles.
constexpr long COLSUM_MAX_CHUNKS = 64;
hipError_t launch_colsum(const double* in, long nrows, long rowStride, long ncols, double scale, double* out,
                         double* outLast, double* work, hipStream_t s);
hipError_t launch_kl_rows(const double* G, const double* LG, long N, long K, double* grad, hipStream_t s);
hipError_t launch_kl_loss(const double* T, long M, const double* ld, long N, double* partial, hipStream_t s);
hipError_t launch_fill(double* p, long rows, long cols, long rs, long cs, double v, hipStream_t s);
hipError_t launch_copy(const double* src, long rows, long cols, long srs, long scs, double* dst, long drs, long dcs,
                       hipStream_t s);

}  // namespace mpd

// Expands CALL(BK, PK, QK, PM) for the variant id
//   var = pmClass * 1000 + bk * 100 + pk * 10 + qk,  PM = 4 (pmClass 0) or 12 (pmClass 1),
// bk: ProbabilistHermite 0, HermiteFunctions 1, PhysicistHermite 2; pk: Exp 0,
// SoftPlus 1; qk: adaptive Simpson 0, adaptive Clenshaw-Curtis 3/5 1, adaptive
// Clenshaw-Curtis (run-time rule) 2, fixed Clenshaw-Curtis 3.
#define MPD_CASES_Q(PMC, PMV, B, P, CALL)                     \
    case PMC * 1000 + B * 100 + P * 10 + 0: CALL(B, P, 0, PMV); break; \
    case PMC * 1000 + B * 100 + P * 10 + 1: CALL(B, P, 1, PMV); break; \
    case PMC * 1000 + B * 100 + P * 10 + 2: CALL(B, P, 2, PMV); break; \
    case PMC * 1000 + B * 100 + P * 10 + 3: CALL(B, P, 3, PMV); break;
#define MPD_CASES_PM(PMC, PMV, CALL)      \
    MPD_CASES_Q(PMC, PMV, 0, 0, CALL)    \
    MPD_CASES_Q(PMC, PMV, 0, 1, CALL)    \
    MPD_CASES_Q(PMC, PMV, 1, 0, CALL)    \
    MPD_CASES_Q(PMC, PMV, 1, 1, CALL)    \
    MPD_CASES_Q(PMC, PMV, 2, 0, CALL)    \
    MPD_CASES_Q(PMC, PMV, 2, 1, CALL)

#define MPD_DISPATCH(var, CALL)              \
    switch (var) {                           \
        MPD_CASES_PM(0, 4, CALL)             \
        MPD_CASES_PM(1, 12, CALL)            \
        default: return hipErrorInvalidValue; \
    }
