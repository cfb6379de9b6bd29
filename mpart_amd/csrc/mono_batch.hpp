// mono_batch.hpp -- one persistent launch over many independent maps.
//
// A batch is a set of maps that share MapOptions.  A call describes, per map,
// its points, outputs and coefficients (BatchIO); the host splits every map's
// points into tiles of blockDim points, orders the tiles by estimated cost
// (largest first) and the kernel's blocks take tiles from a global counter
// until none is left.  Within a tile a lane handles one point and runs the
// same per-point arithmetic as the single-map generic kernels, for every
// component of its map (the components of a triangular map, whatever their
// dimension and term count, go through the same launch).
#pragma once
#include "batch_io.hpp"
#include "mono_kernels.hpp"

namespace mpd {

// Copies a read-only struct (size a multiple of 8 bytes) into LDS with the
// whole block.  A tile's map and arguments are read from there: loads at a
// uniform LDS address give uniform values, so the pointers in them feed
// scalar loads of the term records.
template <class T>
__device__ __forceinline__ void block_copy(T& dst, const T* src) {
    static_assert(sizeof(T) % 8 == 0, "copied as 8-byte words");
    const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
    uint64_t* d = reinterpret_cast<uint64_t*>(&dst);
    for (int i = threadIdx.x; i < (int)(sizeof(T) / 8); i += blockDim.x) d[i] = s[i];
}

template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_batch(const MapDev* maps, const BatchIO* io, const int* tileMap,
                                               const long* tileStart, int ntiles, unsigned* counter, int method,
                                               double xtol, double ytol) {
    extern __shared__ double smem[];
    __shared__ int sTile;
    __shared__ MapDev sMap;
    __shared__ BatchIO sIO;
    const int tid = threadIdx.x;
    double* tab = smem + tid;
    const int stride = blockDim.x;
    while (true) {
        if (tid == 0) sTile = (int)atomicAdd(counter, 1u);
        __syncthreads();  // the previous tile's reads of sMap / sIO are done
        const int t = sTile;
        if (t >= ntiles) break;
        const int mi = ro(tileMap)[t];
        block_copy(sMap, maps + mi);
        block_copy(sIO, io + mi);
        __syncthreads();
        const MapDev& M = sMap;
        const BatchIO& B = sIO;
        const long n = ro(tileStart)[t] + tid;
        if (n >= B.N) continue;
        if (method == BM_EVAL) {
            const double* pt = B.p + n * B.pcs;
            fill_values<BK>(M, M.D - 1, pt, B.prs, tab, stride);
            double ldsum = 0.0;
            for (int c = 0; c < M.ncomp; ++c) {
                const CompDev cd = ro(M.comps)[c];
                double C[PM + 1];
                coeff_groups<PM>(M, cd, B.w + (long)cd.coeffOff * B.ws, B.ws, tab, stride, C);
                const double xk = pt[(long)(cd.dim - 1) * B.prs];
                DiagCoef<PM> dc;
                make_diag<PM>(M, cd, C, xk, 0.0, dc);
                if (B.o) {
                    AdjAcc<PM> A;
                    const double I = integrate<BK, PK, QK, FL_VAL, PM>(M.quad, dc, A, nullptr);
                    B.o[(long)cd.outRow * B.ors + n * B.ocs] = I + f_at_zero<PM>(M, cd, C);
                }
                if (B.ld) {
                    const double z = diag_eval<BK, false, PM>(dc, xk, nullptr);
                    const double g = posfunc<PK>(z).g;
                    ldsum += (g <= 0) ? -__builtin_inf() : log_pos(g);
                }
            }
            if (B.ld) B.ld[n * B.lds] = ldsum;
        } else if (method == BM_COEFF_GRAD) {
            const double* pt = B.p + n * B.pcs;
            fill_values<BK>(M, M.D - 1, pt, B.prs, tab, stride);
            for (int c = 0; c < M.ncomp; ++c) {
                const CompDev cd = ro(M.comps)[c];
                double C[PM + 1];
                coeff_groups<PM>(M, cd, B.w + (long)cd.coeffOff * B.ws, B.ws, tab, stride, C);
                DiagCoef<PM> dc;
                make_diag<PM>(M, cd, C, pt[(long)(cd.dim - 1) * B.prs], M.nugget, dc);
                AdjAcc<PM> A;
                zero_adj<PM>(A);
                integrate<BK, PK, QK, FL_COEF, PM>(M.quad, dc, A, nullptr);
                const double s = B.s[(long)cd.outRow * B.srs + n * B.scs];
                double G[PM + 1];
#pragma unroll
                for (int q = 0; q <= PM; ++q) G[q] = (q <= cd.p) ? (M.phi0[q] + M.rho[q] * A.S[q]) : 0.0;
                double* o = B.o + (long)cd.coeffOff * B.ors + n * B.ocs;
                struct Op {
                    double* o;
                    long rs;
                    const double* G;
                    double s;
                    __device__ void operator()(uint32_t idx, double Pj, int q) { o[(long)idx * rs] = (Pj * G[q]) * s; }
                } op{o, B.ors, G, s};
                visit_terms<PM>(M, cd, tab, stride, op);
            }
        } else {
            // Triangular inverse (TriangularMap.cpp:147-184) for one point:
            // the point starts as [x1, 0, ..., 0]; component c solves for
            // coordinate extra + c, whose initial guess is that 0.
            const int D = M.D;
            const int extra = (int)B.xrows;
            double xf[BATCH_MAX_DIM];
            bool bad = false;
            for (int i = 0; i < BATCH_MAX_DIM; ++i) {
                if (i >= D) break;
                double v = 0.0;
                if (i < extra) {
                    v = B.p[(long)i * B.prs + (B.pcols == 1 ? 0 : n) * B.pcs];
                    bad = bad || isnan(v);
                }
                xf[i] = v;
            }
            for (int c = 0; c < M.ncomp; ++c) {
                const CompDev cd = ro(M.comps)[c];
                double x = NAN;
                if (!bad) {
                    fill_values<BK>(M, cd.dim - 1, xf, 1, tab, stride);
                    double C[PM + 1];
                    coeff_groups<PM>(M, cd, B.w + (long)cd.coeffOff * B.ws, B.ws, tab, stride, C);
                    const double f0 = f_at_zero<PM>(M, cd, C);
                    const double x0 = xf[cd.dim - 1];
                    DiagCoef<PM> dc;
                    make_diag<PM>(M, cd, C, x0, M.nugget, dc);
                    x = inverse_solve<BK, PK, QK, PM>(M, dc, f0, x0, B.s[(long)cd.outRow * B.srs + n * B.scs], xtol,
                                                      ytol);
                }
                xf[cd.dim - 1] = x;
                bad = bad || isnan(x);
                B.o[(long)cd.outRow * B.ors + n * B.ocs] = x;
            }
        }
    }
}

// Batched Evaluate / LogDeterminant with a schedule per tile: tileLanes[t] = 1
// (blockDim points, a lane per point, the per-lane table of at most
// laneSlots doubles) or G (blockDim / G points of a map with many terms, G
// lanes per point as k_eval_grp, one table of at most grpSlots doubles per
// group).  The LDS holds the larger of the two layouts.
template <int BK, int PK, int QK, int PM, int G>
__global__ void __launch_bounds__(256) k_batch_eval_mixed(const MapDev* maps, const BatchIO* io, const int* tileMap,
                                                          const long* tileStart, const int* tileLanes, int ntiles,
                                                          unsigned* counter, int grpSlots) {
    extern __shared__ double smem[];
    __shared__ int sTile;
    __shared__ MapDev sMap;
    __shared__ BatchIO sIO;
    const int tid = threadIdx.x;
    const int gl = tid & (G - 1);
    const int grp = tid / G;
    while (true) {
        if (tid == 0) sTile = (int)atomicAdd(counter, 1u);
        __syncthreads();  // the previous tile's reads of sMap / sIO and its tables are done
        const int t = sTile;
        if (t >= ntiles) break;
        const int mi = ro(tileMap)[t];
        block_copy(sMap, maps + mi);
        block_copy(sIO, io + mi);
        __syncthreads();
        const MapDev& M = sMap;
        const BatchIO& B = sIO;
        const bool group = ro(tileLanes)[t] > 1;  // uniform over the block
        const long n = ro(tileStart)[t] + (group ? grp : tid);
        if (n >= B.N) continue;
        const double* pt = B.p + n * B.pcs;
        double* tab;
        int stride;
        if (group) {
            tab = smem + (long)grp * grpSlots;
            stride = 1;
            group_fill_values<BK, G>(M, M.D - 1, pt, B.prs, tab, gl);
        } else {
            tab = smem + tid;
            stride = blockDim.x;
            fill_values<BK>(M, M.D - 1, pt, B.prs, tab, stride);
        }
        double ldsum = 0.0;
        for (int c = 0; c < M.ncomp; ++c) {
            const CompDev cd = ro(M.comps)[c];
            double C[PM + 1];
            if (group) group_coeff_groups<PM, G>(M, cd, B.w + (long)cd.coeffOff * B.ws, B.ws, tab, gl, C);
            else coeff_groups<PM>(M, cd, B.w + (long)cd.coeffOff * B.ws, B.ws, tab, stride, C);
            const double xk = pt[(long)(cd.dim - 1) * B.prs];
            DiagCoef<PM> dc;
            make_diag<PM>(M, cd, C, xk, 0.0, dc);
            if (B.o) {
                AdjAcc<PM> A;
                const double I = integrate<BK, PK, QK, FL_VAL, PM>(M.quad, dc, A, nullptr);
                if (!group || gl == 0) B.o[(long)cd.outRow * B.ors + n * B.ocs] = I + f_at_zero<PM>(M, cd, C);
            }
            if (B.ld) {
                const double z = diag_eval<BK, false, PM>(dc, xk, nullptr);
                const double g = posfunc<PK>(z).g;
                ldsum += (g <= 0) ? -__builtin_inf() : log_pos(g);
            }
        }
        if (B.ld && (!group || gl == 0)) B.ld[n * B.lds] = ldsum;
        // tables are rewritten for the next tile only after the block-wide
        // barrier at the top of the loop
    }
}

}  // namespace mpd
