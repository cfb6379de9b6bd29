// batch_io.hpp -- per-map arguments of a batched call (shared by the host
// code and the batched kernel).
#pragma once

namespace mpd {

constexpr int BATCH_MAX_DIM = 64;  // inverse keeps a point in a per-lane array
constexpr int INV_CHUNK = 64;      // point indices the inverse queue takes from its counter at a time

enum BatchMethod { BM_EVAL = 0, BM_COEFF_GRAD = 1, BM_INVERSE = 2 };

// Per-map arguments of one batched call (device copy; unused fields null).
struct BatchIO {
    const double* w;  // coefficients
    long ws;
    const double* p;  // points (Evaluate, CoeffGrad) or x1 (Inverse)
    long prs, pcs, pcols;
    long xrows;       // Inverse: rows of x1 (a component: all D, the last one the initial guess;
                      // a triangular map: D - outDim, the solved coordinates start at 0)
    long N;           // points of this map
    double* o;        // Evaluate: outDim x N; CoeffGrad: numCoeffs x N; Inverse: outDim x N
    long ors, ocs;
    double* ld;       // LogDeterminant (Evaluate), N
    long lds;
    const double* s;  // CoeffGrad sensitivities (outDim x N) or Inverse targets r (outDim x N)
    long srs, scs;
};

// One map's share of a batched Inverse pass over component c (device copy).
// The pass's points are numbered in one index space in which every map
// starts at a multiple of the queue chunk, so a chunk never spans two maps.
struct InvSeg {
    int map;          // index into the batch's MapDev array
    int p;            // last-coordinate order of component c
    long base;        // first index of this map in the pass's index space
    long N;           // points of this map
    const double* w;  // coefficients
    long ws;
    const double* x;  // inputs of component c (the last row holds the initial guess)
    long xrows, xcols, xrs, xcs;
    const double* y;  // targets r of component c
    long yss;
    double* o;        // root of component c, point n at o[n * os]
    long os;
    double* o2;       // second copy of the root (the caller's output) or null
    long os2;
    int* status;      // the map's MapDev::status
};

// Hand-over of the longest inverse solves to k_inverse_heavy (mono_inverse.hpp):
// the handed-over point indices (count of them, at most cap kept) and the
// step budget (solveSteps 0: off).
struct InvHeavy {
    unsigned long long* count;
    unsigned long long* take;  // next list entry for k_inverse_heavy's blocks to take
    long* list;
    long cap;
    int solveSteps;
    // Totals over a call (zeroed by its first pass; null: not kept): solves
    // handed over, their T evaluations, intervals evaluated level by level,
    // T evaluations integrated depth-first because the tree outgrew the
    // workspace or met a non-finite interval error, the largest tree
    // (intervals), the deepest level reached, the most T evaluations of one
    // solve, and the longest solve in ticks of the 100 MHz device clock.
    unsigned long long* stats;
};
enum { HS_SOLVES = 0, HS_TEVALS, HS_NODES, HS_FALLBACK_SIZE, HS_FALLBACK_NONFINITE, HS_MAX_NODES, HS_MAX_LEVELS,
       HS_MAX_SOLVE_TEVALS, HS_MAX_SOLVE_TICKS, HS_COUNT };
// Intervals one cooperative quadrature keeps in LDS (48 bytes each, plus 64
// node values of a round; a larger tree is integrated depth-first instead).
// 320 intervals (15.5 KB: ten one-wavefront blocks per CU) hold the trees
// of the long solves met so far (at most 289 intervals).
constexpr int COOP_NODES = 320;
constexpr int COOP_LDS_BYTES = COOP_NODES * 48 + 64 * 8;

}  // namespace mpd
