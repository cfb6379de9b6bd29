// mono_split.hpp -- fused forward + adjoint for maps with many terms, with the
// components of a map split over the wavefronts of a block.
//
// A block of eight wavefronts takes a tile of 64 points at a time.  The tile's
// basis table (every coordinate's values phi_1..phi_deg, slot-major: entry
// (slot, point) at tab[slot * 64 + point]) sits once in LDS, filled by all
// eight wavefronts, so it costs 8 * nslots bytes per point instead of per lane.
// Each wavefront then runs a lane per point over the components assigned to
// it (a host-side partition of the components balanced by term count): the
// C_q sums, the component's quadrature and the gradient pass, whose term loop
// is uniform over the wavefront (every lane reads the same term record, so
// the records come through the scalar cache).  No quadrature is replicated,
// unlike the lanes-per-point schedules, and no wavefront waits on another
// except at the tile boundary.  Gradient sums go to the block's own partial
// row with return-less atomics: a coefficient belongs to one component and a
// component to one wavefront, so every row entry has a single writer and the
// order of its additions is fixed.
#pragma once
#include "mono_kernels.hpp"

namespace mpd {

constexpr int SPLIT_TILE = 64;   // points per tile
constexpr int SPLIT_WAVES = 8;   // wavefronts per block

// MINW: wavefronts per SIMD the compiler must allow (the second launch
// bound): 2 (up to 256 registers) or 4 (up to 128, with spills).  The tile's
// table allows two blocks of eight wavefronts per CU, so at 4 the LDS and
// the registers both admit four wavefronts per SIMD; the latency of the
// scalar record loads and the LDS reads wants them.
template <int BK, int PK, int QK, int PM, int MINW>
__global__ void __launch_bounds__(SPLIT_WAVES * 64, MINW) k_fused_split(MapDev M, const double* w, long ws, PtsArg P, const double* sens,
                                                     long srs, long scs, const double* lsens, long lss, int klMode,
                                                     const int* compOrder, const int* waveStart, double* rows) {
    extern __shared__ double tab[];
    __shared__ double lossWave[SPLIT_WAVES];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar loads below
    const int K = M.numCoeffs;
    double* row = rows + (long)blockIdx.x * (K + 1);
    for (int j = threadIdx.x; j <= K; j += blockDim.x) row[j] = 0.0;
    __threadfence();  // the zeros are in memory before any wavefront adds to the row
    __syncthreads();
    const int ndims = M.D - 1;
    const int c0 = ro(waveStart)[wave], c1 = ro(waveStart)[wave + 1];
    double lossLane = 0.0;
    const long ntiles = (P.N + SPLIT_TILE - 1) / SPLIT_TILE;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long n0 = tile * SPLIT_TILE + lane;
        const bool active = n0 < P.N;
        const long n = active ? n0 : P.N - 1;
        const double* pt = P.p + n * P.cs;
        __syncthreads();  // the previous tile's table is no longer read
        if (wave == 0) tab[lane] = 1.0;
        for (int d = wave; d < ndims; d += SPLIT_WAVES) {
            const int deg = ro(M.dimDeg)[d];
            if (deg > 0) values_1d<BK>(M, deg, pt[d * P.rs], tab + (long)(ro(M.dimSlot)[d] - 1) * SPLIT_TILE + lane, SPLIT_TILE);
        }
        __syncthreads();
        const double* mytab = tab + lane;
        for (int i = c0; i < c1; ++i) {
            const CompDev cd = ro(M.comps)[ro(compOrder)[i]];
            const double* wc = w + (long)cd.coeffOff * ws;
            double C[PM + 1];
            coeff_groups<PM>(M, cd, wc, ws, mytab, SPLIT_TILE, C);
            const double xk = pt[(long)(cd.dim - 1) * P.rs];
            const double sv = (!klMode && sens) ? sens[(long)cd.outRow * srs + n * scs] : 0.0;
            const double lv = (!klMode && lsens) ? lsens[n * lss] : 0.0;
            double G[PM + 1];
            if (M.discrete) comp_core_discrete<BK, PK, QK, PM>(M, cd, C, xk, active, klMode, sv, lv, G, lossLane);
            else comp_core<BK, PK, QK, PM>(M, cd, C, xk, active, klMode, sv, lv, G, lossLane);
            // Gradient pass: chunks of 16 term records per group, each chunk
            // reduced over the tile's points by the wavefront.
            dispatch_nd(M.nd, [&](auto tag) {
                constexpr int ND = decltype(tag)::value;
                for (int q = 0; q <= cd.p; ++q) {
                    const GroupDev g = ro(M.groups)[cd.groupOff + q];
                    const double Gq = select_q<PM>(G, q);
                    for (int j0 = g.start; j0 < g.start + g.countPad; j0 += 16) {
                        double v[16];
                        chunk16_products<ND>(M, j0, mytab, SPLIT_TILE, Gq, active, v);
                        const double rsum = wave_reduce16(v, lane);
                        if ((lane & 3) == 0) {
                            const uint32_t idx = ro(M.terms)[(long)(j0 + (lane >> 2)) * rec_len<ND>(M)];
                            if (idx != DUMMY) (void)unsafeAtomicAdd(row + cd.coeffOff + idx, rsum);
                        }
                    }
                }
            });
        }
    }
    const double l = wave_sum(lossLane);
    if (lane == 0) lossWave[wave] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < SPLIT_WAVES; ++i) t += lossWave[i];
        row[K] = t;
    }
}

}  // namespace mpd

namespace mpd {

// Evaluate and/or LogDeterminant of every component with the same split: a
// block of eight wavefronts per tile of 64 points, one shared basis table,
// each wavefront a lane per point over its own components.  T goes straight
// to the output row of the component; the log-determinant terms of the four
// wavefronts are added in wavefront order at the end of the tile, so that sum
// runs in another order than the component order of the lane-per-point
// kernels (same value up to rounding).  The sums pass through the table's
// first rows once the tile's components are done, so the kernel needs no LDS
// beyond the table.
template <int BK, int PK, int QK, int PM>
// (four wavefronts per SIMD, as the fused split kernel's second build)
__global__ void __launch_bounds__(SPLIT_WAVES * 64, 4) k_eval_split(MapDev M, const double* w, long ws, PtsArg P, OutArg out,
                                                    double* ld, long lds, int ldAccumulate, int* nevals,
                                                    const int* compOrder, const int* waveStart) {
    extern __shared__ double tab[];  // nslots >= SPLIT_WAVES (the host checks)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar loads below
    const int ndims = M.D - 1;
    const int c0 = ro(waveStart)[wave], c1 = ro(waveStart)[wave + 1];
    const long ntiles = (P.N + SPLIT_TILE - 1) / SPLIT_TILE;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long n0 = tile * SPLIT_TILE + lane;
        const bool active = n0 < P.N;
        const long n = active ? n0 : P.N - 1;
        const double* pt = P.p + n * P.cs;
        __syncthreads();  // the previous tile's table and log-det terms are no longer read
        if (wave == 0) tab[lane] = 1.0;
        for (int d = wave; d < ndims; d += SPLIT_WAVES) {
            const int deg = ro(M.dimDeg)[d];
            if (deg > 0) values_1d<BK>(M, deg, pt[d * P.rs], tab + (long)(ro(M.dimSlot)[d] - 1) * SPLIT_TILE + lane, SPLIT_TILE);
        }
        __syncthreads();
        const double* mytab = tab + lane;
        double ldsum = 0.0;
        for (int i = c0; i < c1; ++i) {
            const int c = ro(compOrder)[i];
            const CompDev cd = ro(M.comps)[c];
            double C[PM + 1];
            coeff_groups<PM>(M, cd, w + (long)cd.coeffOff * ws, ws, mytab, SPLIT_TILE, C);
            const double xk = pt[(long)(cd.dim - 1) * P.rs];
            DiagCoef<PM> dc;
            if (out.p) {
                make_diag<PM>(M, cd, C, xk, 0.0, dc);
                AdjAcc<PM> A;
                int ne = 0;
                const double I = integrate<BK, PK, QK, FL_VAL, PM>(M.quad, dc, A, nevals ? &ne : nullptr);
                if (active) {
                    out.p[(long)cd.outRow * out.rs + n * out.cs] = I + f_at_zero<PM>(M, cd, C);
                    if (nevals) nevals[(long)c * P.N + n] = ne;
                }
            }
            if (ld) {
                make_diag<PM>(M, cd, C, xk, 0.0, dc);
                const double z = diag_eval<BK, false, PM>(dc, xk, nullptr);
                const double g = posfunc<PK>(z).g;
                ldsum += (g <= 0) ? -__builtin_inf() : log_pos(g);
            }
        }
        if (ld) {
            // the table is no longer read: its first rows take the wavefronts' sums
            __syncthreads();
            tab[wave * SPLIT_TILE + lane] = ldsum;
            __syncthreads();
            if (wave == 0 && active) {
                double s = tab[lane];
                for (int v = 1; v < SPLIT_WAVES; ++v) s += tab[v * SPLIT_TILE + lane];
                if (ldAccumulate) ld[n * lds] += s;
                else ld[n * lds] = s;
            }
        }
    }
}

}  // namespace mpd
