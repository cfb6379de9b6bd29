// batch_io.hpp -- per-map arguments of a batched call (shared by the host
// code and the batched kernel).
#pragma once

namespace mpd {

constexpr int BATCH_MAX_DIM = 64;  // inverse keeps a point in a per-lane array

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

}  // namespace mpd
