// mono_kernels.hpp -- HIP kernels of the monotone-component path (lane = point).
//
// Template parameters: BK basis (ProbabilistHermite / Hermite functions), PK
// positive bijector (Exp / SoftPlus), QK quadrature (adaptive Simpson, adaptive
// Clenshaw-Curtis with a 5-point fine rule, or with a runtime fine rule <= 17).
//
// The per-lane off-diagonal basis table lives in LDS (tab[slot * blockDim +
// tid]) unless the launcher passes a global table (large maps), in which case
// it is gtab[slot * stride + global thread index].
#pragma once
#include "mono_device.hpp"

namespace mpd {

struct PtsArg {
    const double* p;
    long rs, cs, N;
};
struct OutArg {
    double* p;
    long rs, cs;
};

__device__ __forceinline__ double* lane_table(double* smem_tab, double* gtab, long& stride) {
    if (gtab) {
        stride = (long)gridDim.x * blockDim.x;
        return gtab + (long)blockIdx.x * blockDim.x + threadIdx.x;
    }
    stride = blockDim.x;
    return smem_tab + threadIdx.x;
}

// Largest coordinate whose table entries the components [cb, ce) read.
__device__ __forceinline__ int table_dims(const MapDev& M, int cb, int ce) {
    int nd = 0;
    for (int c = cb; c < ce; ++c) nd = max(nd, ro(M.comps)[c].dim - 1);
    return nd;
}

// ---------------------------------------------------------------------------
// Evaluate (MonotoneComponent::EvaluateImpl, MonotoneComponent.h:255-299, with
// the nugget dropped as at line 287) and/or LogDeterminant
// (ContinuousDerivative + log, 105-125, 488-531) for components [cb, ce).
// logdet is summed over the components (TriangularMap.cpp:100-125); with
// ldAccumulate the launch adds to the existing values.
template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_eval(MapDev M, const double* w, long ws, PtsArg P, OutArg out, double* ld,
                                              long lds, int ldAccumulate, int* nevals, int cb, int ce, double* gtab) {
    extern __shared__ double smem[];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.N) return;
    long stride;
    double* tab = lane_table(smem, gtab, stride);
    const double* pt = P.p + n * P.cs;
    fill_values<BK>(M, table_dims(M, cb, ce), pt, P.rs, tab, (int)stride);
    double ldsum = 0.0;
    for (int c = cb; c < ce; ++c) {
        const CompDev cd = ro(M.comps)[c];
        double C[PM + 1];
        coeff_groups<PM>(M, cd, w + (long)cd.coeffOff * ws, ws, tab, (int)stride, C);
        const double xk = pt[(long)(cd.dim - 1) * P.rs];
        DiagCoef<PM> dc;
        if (out.p) {
            make_diag<PM>(M, cd, C, xk, 0.0, dc);
            AdjAcc<PM> A;
            int ne = 0;
            double I = integrate<BK, PK, QK, FL_VAL, PM>(M.quad, dc, A, nevals ? &ne : nullptr);
            out.p[(long)cd.outRow * out.rs + n * out.cs] = I + f_at_zero<PM>(M, cd, C);
            if (nevals) nevals[(long)c * P.N + n] = ne;
        }
        if (ld) {
            make_diag<PM>(M, cd, C, xk, 0.0, dc);
            double z = diag_eval<BK, false, PM>(dc, xk, nullptr);
            double g = posfunc<PK>(z).g;
            ldsum += (g <= 0) ? -__builtin_inf() : log_pos(g);
        }
    }
    if (ld) {
        if (ldAccumulate) ld[n * lds] += ldsum;
        else ld[n * lds] = ldsum;
    }
}

// ---------------------------------------------------------------------------
// Group-per-point schedule for maps with many terms per component: G lanes
// (G = 4, 8 or 64) share one point.  The point's basis table sits once in LDS
// (filled by the group, one coordinate per lane), the lanes split each
// group of terms (C_q by a butterfly sum, bitwise the same in every lane of
// the group), and every lane of the group runs the same quadrature, so the
// group stays converged.  Same outputs as k_eval up to the summation order
// of C_q.

// Orders all LDS traffic of the wavefront before what follows.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int BK, int G>
__device__ __forceinline__ void group_fill_values(const MapDev& M, int ndims, const double* pt, long prs, double* tab,
                                                  int gl) {
    if (gl == 0) tab[0] = 1.0;
    for (int d = gl; d < ndims; d += G) {
        const int deg = ro(M.dimDeg)[d];
        if (deg > 0) values_1d<BK>(M, deg, pt[d * prs], tab + ro(M.dimSlot)[d] - 1, 1);
    }
    wave_lds_sync();
}

template <int ND, int PM, int G>
__device__ __forceinline__ void group_coeff_groups_nd(const MapDev& M, const CompDev& c, const double* w, long ws,
                                                      const double* tab, int gl, double* C) {
#pragma unroll
    for (int q = 0; q <= PM; ++q) {
        if (q > c.p) break;
        const GroupDev g = ro(M.groups)[c.groupOff + q];
        double acc = 0.0;
        for (int j = g.start + gl; j < g.start + g.count; j += G) {
            const auto rec = ro(M.terms) + (long)j * rec_len<ND>(M);
            acc += ro(w)[(long)rec[0] * ws] * rec_product<ND>(M, rec, tab, 1);
        }
        C[q] = group_sum<G>(acc);
    }
}

template <int PM, int G>
__device__ __forceinline__ void group_coeff_groups(const MapDev& M, const CompDev& c, const double* w, long ws,
                                                   const double* tab, int gl, double* C) {
    switch (M.nd) {
        case 1: group_coeff_groups_nd<1, PM, G>(M, c, w, ws, tab, gl, C); break;
        case 2: group_coeff_groups_nd<2, PM, G>(M, c, w, ws, tab, gl, C); break;
        case 3: group_coeff_groups_nd<3, PM, G>(M, c, w, ws, tab, gl, C); break;
        case 4: group_coeff_groups_nd<4, PM, G>(M, c, w, ws, tab, gl, C); break;
        case 5: group_coeff_groups_nd<5, PM, G>(M, c, w, ws, tab, gl, C); break;
        case 6: group_coeff_groups_nd<6, PM, G>(M, c, w, ws, tab, gl, C); break;
        case 7: group_coeff_groups_nd<7, PM, G>(M, c, w, ws, tab, gl, C); break;
        case 8: group_coeff_groups_nd<8, PM, G>(M, c, w, ws, tab, gl, C); break;
        default: group_coeff_groups_nd<0, PM, G>(M, c, w, ws, tab, gl, C); break;
    }
}

template <int BK, int PK, int QK, int PM, int G>
__global__ void __launch_bounds__(256) k_eval_grp(MapDev M, const double* w, long ws, PtsArg P, OutArg out, double* ld,
                                                  long lds, int ldAccumulate, int* nevals, int cb, int ce) {
    extern __shared__ double smem[];
    const int gl = threadIdx.x & (G - 1);
    const int grp = threadIdx.x / G;
    const long n = (long)blockIdx.x * (blockDim.x / G) + grp;
    if (n >= P.N) return;  // the whole group leaves together
    double* tab = smem + (long)grp * M.nslots;
    const double* pt = P.p + n * P.cs;
    group_fill_values<BK, G>(M, table_dims(M, cb, ce), pt, P.rs, tab, gl);
    double ldsum = 0.0;
    for (int c = cb; c < ce; ++c) {
        const CompDev cd = ro(M.comps)[c];
        double C[PM + 1];
        group_coeff_groups<PM, G>(M, cd, w + (long)cd.coeffOff * ws, ws, tab, gl, C);
        const double xk = pt[(long)(cd.dim - 1) * P.rs];
        DiagCoef<PM> dc;
        if (out.p) {
            make_diag<PM>(M, cd, C, xk, 0.0, dc);
            AdjAcc<PM> A;
            int ne = 0;
            double I = integrate<BK, PK, QK, FL_VAL, PM>(M.quad, dc, A, nevals ? &ne : nullptr);
            if (gl == 0) {
                out.p[(long)cd.outRow * out.rs + n * out.cs] = I + f_at_zero<PM>(M, cd, C);
                if (nevals) nevals[(long)c * P.N + n] = ne;
            }
        }
        if (ld) {
            make_diag<PM>(M, cd, C, xk, 0.0, dc);
            double z = diag_eval<BK, false, PM>(dc, xk, nullptr);
            double g = posfunc<PK>(z).g;
            ldsum += (g <= 0) ? -__builtin_inf() : log_pos(g);
        }
    }
    if (ld && gl == 0) {
        if (ldAccumulate) ld[n * lds] += ldsum;
        else ld[n * lds] = ldsum;
    }
}

// Fused forward + adjoint (as k_fused) with G lanes per point.  A wavefront
// takes 64 / G points per step from a static tile sequence; after each
// component's quadrature the lanes split the terms, sum P_j G_q over the
// wave's points by a butterfly across the groups, and the lanes of the first
// group add the sums to the wave's own partial row (K gradient sums, then the
// loss), so every row has one writer and the order of additions is fixed.
// The additions are hardware FP64 atomics on ordinary (coarse-grained) device
// memory, which the wave does not wait for.
template <int BK, int PK, int QK, int PM, int G>
__global__ void __launch_bounds__(256) k_fused_grp(MapDev M, const double* w, long ws, PtsArg P, const double* sens,
                                                   long srs, long scs, const double* lsens, long lss, int klMode,
                                                   double* rows) {
    extern __shared__ double smem[];
    constexpr int GPW = 64 / G;  // points per wavefront step
    const int lane = threadIdx.x & 63;
    const int gl = threadIdx.x & (G - 1);
    const int gw = lane / G;
    const long waveId = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long nwavesTotal = (long)gridDim.x * (blockDim.x >> 6);
    const int K = M.numCoeffs;
    double* tab = smem + (long)(threadIdx.x / G) * M.nslots;
    double* row = rows + waveId * (long)(K + 1);
    for (int j = lane; j <= K; j += 64) row[j] = 0.0;
    double lossLane = 0.0;
    const int ndims = table_dims(M, 0, M.ncomp);
    const long ntiles = (P.N + GPW - 1) / GPW;
    for (long tile = waveId; tile < ntiles; tile += nwavesTotal) {
        const long n0 = tile * GPW + gw;
        const bool active = n0 < P.N;
        const long n = active ? n0 : P.N - 1;
        const double* pt = P.p + n * P.cs;
        wave_lds_sync();
        group_fill_values<BK, G>(M, ndims, pt, P.rs, tab, gl);
        for (int c = 0; c < M.ncomp; ++c) {
            const CompDev cd = ro(M.comps)[c];
            const double* wc = w + (long)cd.coeffOff * ws;
            double C[PM + 1];
            group_coeff_groups<PM, G>(M, cd, wc, ws, tab, gl, C);
            const double xk = pt[(long)(cd.dim - 1) * P.rs];
            const double sv = (!klMode && sens) ? sens[(long)cd.outRow * srs + n * scs] : 0.0;
            const double lv = (!klMode && lsens) ? lsens[n * lss] : 0.0;
            double Gq[PM + 1];
            double lp = 0.0;
            comp_core<BK, PK, QK, PM>(M, cd, C, xk, active, klMode, sv, lv, Gq, lp);
            if (gl == 0) lossLane += lp;
#pragma unroll
            for (int q = 0; q <= PM; ++q) {
                if (q > cd.p) break;
                const GroupDev g = ro(M.groups)[cd.groupOff + q];
                const int end = g.start + g.count;
                // four chunks of G terms per step, so their loads overlap
                constexpr int U = 4;
                for (int j0 = g.start; j0 < end; j0 += U * G) {
                    double v[U];
                    uint32_t idx[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int j = j0 + u * G + gl;
                        v[u] = 0.0;
                        idx[u] = 0;
                        if (j < end) {
                            const auto rec = ro(M.terms) + (long)j * M.trec;
                            double Pj = tab[rec[1]];
                            for (int e = 2; e <= M.nd; ++e) Pj *= tab[rec[e]];
                            v[u] = Pj * Gq[q];
                            idx[u] = rec[0];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
#pragma unroll
                        for (int off = G; off < 64; off <<= 1) v[u] += __shfl_xor(v[u], off, 64);
                    }
                    // hardware add at L2, no return value: the wave does not
                    // wait for the row; one writer per row keeps the order fixed
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (gw == 0 && j0 + u * G + gl < end)
                            (void)unsafeAtomicAdd(row + cd.coeffOff + idx[u], v[u]);
                }
            }
        }
    }
    const double l = wave_sum(lossLane);
    if (lane == 0) row[K] = l;
}

// ---------------------------------------------------------------------------
// CoeffGrad of one component (CoeffJacobian + sensitivity scaling,
// MonotoneComponent.h:166-182, 685-741): rows coeffOff.. of out receive
//   sens_n * P_j (phi~_q(0) + rho_q S'_q).
template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_coeff_grad(MapDev M, const double* w, long ws, PtsArg P, const double* sens,
                                                    long srs, long scs, OutArg out, int c, double* gtab) {
    extern __shared__ double smem[];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.N) return;
    long stride;
    double* tab = lane_table(smem, gtab, stride);
    const double* pt = P.p + n * P.cs;
    const CompDev cd = ro(M.comps)[c];
    fill_values<BK>(M, cd.dim - 1, pt, P.rs, tab, (int)stride);
    double C[PM + 1];
    coeff_groups<PM>(M, cd, w + (long)cd.coeffOff * ws, ws, tab, (int)stride, C);
    DiagCoef<PM> dc;
    make_diag<PM>(M, cd, C, pt[(long)(cd.dim - 1) * P.rs], M.nugget, dc);
    AdjAcc<PM> A;
    zero_adj<PM>(A);
    integrate<BK, PK, QK, FL_COEF, PM>(M.quad, dc, A, nullptr);
    const double s = sens[(long)cd.outRow * srs + n * scs];
    double G[PM + 1];
#pragma unroll
    for (int q = 0; q <= PM; ++q) G[q] = (q <= cd.p) ? (M.phi0[q] + M.rho[q] * A.S[q]) : 0.0;
    double* o = out.p + (long)cd.coeffOff * out.rs + n * out.cs;
    struct Op {
        double* o;
        long rs;
        const double* G;
        double s;
        __device__ void operator()(uint32_t idx, double Pj, int q) { o[(long)idx * rs] = (Pj * G[q]) * s; }
    } op{o, out.rs, G, s};
    visit_terms<PM>(M, cd, tab, (int)stride, op);
}

// LogDeterminantCoeffGrad of one component with contDeriv (ContinuousMixedJacobian
// + ContinuousDerivative, MonotoneComponent.h:185-206, 825-874):
//   P_j phi~'_q(x_k) g'(d_k f) * (1 / g(d_k f)).
template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_logdet_coeff_grad(MapDev M, const double* w, long ws, PtsArg P, OutArg out,
                                                           int c, double* gtab) {
    extern __shared__ double smem[];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.N) return;
    long stride;
    double* tab = lane_table(smem, gtab, stride);
    const double* pt = P.p + n * P.cs;
    const CompDev cd = ro(M.comps)[c];
    fill_values<BK>(M, cd.dim - 1, pt, P.rs, tab, (int)stride);
    double C[PM + 1];
    coeff_groups<PM>(M, cd, w + (long)cd.coeffOff * ws, ws, tab, (int)stride, C);
    DiagCoef<PM> dc;
    const double xk = pt[(long)(cd.dim - 1) * P.rs];
    make_diag<PM>(M, cd, C, xk, 0.0, dc);
    const double z = diag_eval<BK, false, PM>(dc, xk, nullptr);
    const PosF<PK> gv = posfunc<PK>(z);
    const double inv = 1.0 / gv.g;
    double bq[PM + 1];
    diag_basis<BK, PM>(dc, xk, bq);
    double G[PM + 1];
#pragma unroll
    for (int q = 0; q <= PM; ++q) G[q] = (q <= cd.p) ? M.rho[q] * bq[q] : 0.0;
    double* o = out.p + (long)cd.coeffOff * out.rs + n * out.cs;
    struct Op {
        double* o;
        long rs;
        const double* G;
        double gp, inv;
        __device__ void operator()(uint32_t idx, double Pj, int q) { o[(long)idx * rs] = ((Pj * G[q]) * gp) * inv; }
    } op{o, out.rs, G, gv.gp, inv};
    visit_terms<PM>(M, cd, tab, (int)stride, op);
}

// ---------------------------------------------------------------------------
// Discrete-derivative mode (contDeriv = false).  LogDeterminant uses
// DiscreteDerivative (MonotoneComponent.h:586-638): dT/dx_k as the integral
// U = int_0^1 (g + nugget) + x t g' d2f dt on the tree of x (g + nugget);
// logdet is the sum of log U over components [cb, ce).
template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_logdet_discrete(MapDev M, const double* w, long ws, PtsArg P, double* ld,
                                                         long lds, int ldAccumulate, int cb, int ce, double* gtab) {
    extern __shared__ double smem[];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.N) return;
    long stride;
    double* tab = lane_table(smem, gtab, stride);
    const double* pt = P.p + n * P.cs;
    fill_values<BK>(M, table_dims(M, cb, ce), pt, P.rs, tab, (int)stride);
    double ldsum = 0.0;
    for (int c = cb; c < ce; ++c) {
        const CompDev cd = ro(M.comps)[c];
        double C[PM + 1];
        coeff_groups<PM>(M, cd, w + (long)cd.coeffOff * ws, ws, tab, (int)stride, C);
        DiagCoef<PM> dc;
        make_diag<PM>(M, cd, C, pt[(long)(cd.dim - 1) * P.rs], M.nugget, dc);
        AdjAcc<PM> A;
        zero_adj<PM>(A);
        integrate<BK, PK, QK, FL_INPUT, PM>(M.quad, dc, A, nullptr);
        ldsum += (A.U <= 0) ? -__builtin_inf() : log_pos(A.U);
    }
    if (ldAccumulate) ld[n * lds] += ldsum;
    else ld[n * lds] = ldsum;
}

// LogDeterminantCoeffGrad of one component without contDeriv
// (DiscreteMixedJacobian + DiscreteDerivative, MonotoneComponent.h:185-206,
// 928-980): P_j rho_q S'_q * (1 / U) with S'_q of the MixedCoeff integrand.
template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_logdet_coeff_grad_discrete(MapDev M, const double* w, long ws, PtsArg P,
                                                                    OutArg out, int c, double* gtab) {
    extern __shared__ double smem[];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.N) return;
    long stride;
    double* tab = lane_table(smem, gtab, stride);
    const double* pt = P.p + n * P.cs;
    const CompDev cd = ro(M.comps)[c];
    fill_values<BK>(M, cd.dim - 1, pt, P.rs, tab, (int)stride);
    double C[PM + 1];
    coeff_groups<PM>(M, cd, w + (long)cd.coeffOff * ws, ws, tab, (int)stride, C);
    DiagCoef<PM> dc;
    make_diag<PM>(M, cd, C, pt[(long)(cd.dim - 1) * P.rs], M.nugget, dc);
    AdjAcc<PM> A;
    zero_adj<PM>(A);
    integrate<BK, PK, QK, FL_DCOEF, PM>(M.quad, dc, A, nullptr);
    const double inv = 1.0 / A.U;
    double G[PM + 1];
#pragma unroll
    for (int q = 0; q <= PM; ++q) G[q] = (q <= cd.p) ? M.rho[q] * A.S[q] : 0.0;
    double* o = out.p + (long)cd.coeffOff * out.rs + n * out.cs;
    struct Op {
        double* o;
        long rs;
        const double* G;
        double inv;
        __device__ void operator()(uint32_t idx, double Pj, int q) { o[(long)idx * rs] = (Pj * G[q]) * inv; }
    } op{o, out.rs, G, inv};
    visit_terms<PM>(M, cd, tab, (int)stride, op);
}

// ---------------------------------------------------------------------------
// Input gradients of one component, accumulated into out (inDim x N):
//  mode 0: Gradient      sens_n * dT/dx  (InputJacobian, MonotoneComponent.h:148-164, 758-818)
//  mode 1: LogDeterminantInputGrad       (ContinuousMixedInputJacobian, 208-232, 877-926)
// dT/dx_i = sum_q D^(i)_q (phi~_q(0) + rho_q S'_q) for i < k, dT/dx_k = U;
// with D^(i)_q = sum_{j in group q} w_j dP_j/dx_i.
// One factor of dP_j/dx_wrt: the derivative table for the factor in x_wrt
// (has = true), the value table otherwise (slot 0 is 1.0).
__device__ __forceinline__ void input_factor(const MapDev& M, uint32_t s, const double* tab, const double* dtab,
                                             int stride, int wrt, bool& has, double& Pj) {
    if (s != 0 && ro(M.slotDim)[s] == wrt) {
        has = true;
        Pj *= dtab[(long)s * stride];
    } else {
        Pj *= tab[(long)s * stride];
    }
}

template <int ND, int PM>
__device__ __forceinline__ void input_groups_nd(const MapDev& M, const CompDev& c, const double* w, long ws,
                                                const double* tab, const double* dtab, int stride, int wrt,
                                                double* Dq) {
#pragma unroll
    for (int q = 0; q <= PM; ++q) {
        if (q > c.p) break;
        const GroupDev g = ro(M.groups)[c.groupOff + q];
        double acc = 0.0;
        for (int j = g.start; j < g.start + g.count; ++j) {
            const auto rec = ro(M.terms) + (long)j * rec_len<ND>(M);
            bool has = false;
            double Pj = 1.0;
            if constexpr (ND > 0) {
#pragma unroll
                for (int e = 1; e <= ND; ++e) input_factor(M, rec[e], tab, dtab, stride, wrt, has, Pj);
            } else {
                for (int e = 1; e <= M.nd; ++e) input_factor(M, rec[e], tab, dtab, stride, wrt, has, Pj);
            }
            if (has) acc += ro(w)[(long)rec[0] * ws] * Pj;
        }
        Dq[q] = acc;
    }
}

template <int PM>
__device__ __forceinline__ void input_groups(const MapDev& M, const CompDev& c, const double* w, long ws,
                                             const double* tab, const double* dtab, int stride, int wrt, double* Dq) {
    switch (M.nd) {
        case 1: input_groups_nd<1, PM>(M, c, w, ws, tab, dtab, stride, wrt, Dq); break;
        case 2: input_groups_nd<2, PM>(M, c, w, ws, tab, dtab, stride, wrt, Dq); break;
        case 3: input_groups_nd<3, PM>(M, c, w, ws, tab, dtab, stride, wrt, Dq); break;
        case 4: input_groups_nd<4, PM>(M, c, w, ws, tab, dtab, stride, wrt, Dq); break;
        case 5: input_groups_nd<5, PM>(M, c, w, ws, tab, dtab, stride, wrt, Dq); break;
        case 6: input_groups_nd<6, PM>(M, c, w, ws, tab, dtab, stride, wrt, Dq); break;
        case 7: input_groups_nd<7, PM>(M, c, w, ws, tab, dtab, stride, wrt, Dq); break;
        case 8: input_groups_nd<8, PM>(M, c, w, ws, tab, dtab, stride, wrt, Dq); break;
        default: input_groups_nd<0, PM>(M, c, w, ws, tab, dtab, stride, wrt, Dq); break;
    }
}

template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_input_grad(MapDev M, const double* w, long ws, PtsArg P, const double* sens,
                                                    long srs, long scs, OutArg out, int c, int mode, double* gtab) {
    extern __shared__ double smem[];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.N) return;
    long stride;
    double* tab;
    if (gtab) {
        stride = (long)gridDim.x * blockDim.x;
        tab = gtab + n;
    } else {
        stride = blockDim.x;
        tab = smem + threadIdx.x;
    }
    double* dtab = tab + (long)M.nslots * stride;
    const double* pt = P.p + n * P.cs;
    const CompDev cd = ro(M.comps)[c];
    const int k = cd.dim - 1;
    fill_values_derivs<BK>(M, k, pt, P.rs, tab, dtab, (int)stride);
    const double* wc = w + (long)cd.coeffOff * ws;
    double C[PM + 1];
    coeff_groups<PM>(M, cd, wc, ws, tab, (int)stride, C);
    const double xk = pt[(long)k * P.rs];
    DiagCoef<PM> dc;
    double* o = out.p + n * out.cs;
    if (mode == 0) {
        make_diag<PM>(M, cd, C, xk, M.nugget, dc);
        AdjAcc<PM> A;
        zero_adj<PM>(A);
        integrate<BK, PK, QK, FL_INPUT, PM>(M.quad, dc, A, nullptr);
        const double s = sens[(long)cd.outRow * srs + n * scs];
        double G[PM + 1];
#pragma unroll
        for (int q = 0; q <= PM; ++q) G[q] = (q <= cd.p) ? (M.phi0[q] + M.rho[q] * A.S[q]) : 0.0;
        for (int i = 0; i < k; ++i) {
            double Dq[PM + 1];
            input_groups<PM>(M, cd, wc, ws, tab, dtab, (int)stride, i, Dq);
            double v = 0.0;
#pragma unroll
            for (int q = 0; q <= PM; ++q) {
                if (q > cd.p) break;
                v += Dq[q] * G[q];
            }
            o[(long)i * out.rs] += v * s;
        }
        o[(long)k * out.rs] += A.U * s;
    } else {
        make_diag<PM>(M, cd, C, xk, 0.0, dc);
        double z2;
        const double z = diag_eval<BK, true, PM>(dc, xk, &z2);
        const PosF<PK> gv = posfunc<PK>(z);
        const double inv = 1.0 / gv.g;
        double bq[PM + 1];
        diag_basis<BK, PM>(dc, xk, bq);
        for (int i = 0; i < k; ++i) {
            double Dq[PM + 1];
            input_groups<PM>(M, cd, wc, ws, tab, dtab, (int)stride, i, Dq);
            double v = 0.0;
#pragma unroll
            for (int q = 0; q <= PM; ++q) {
                if (q > cd.p) break;
                v += Dq[q] * (M.rho[q] * bq[q]);
            }
            o[(long)i * out.rs] += (v * gv.gp) * inv;
        }
        o[(long)k * out.rs] += (z2 * gv.gp) * inv;
    }
}

// ---------------------------------------------------------------------------
// Input vector-Jacobian product, all components in one pass:
//   out(:, n) = scale * grad_x [ sum_c s_c T_c(x_n) + l log det grad T(x_n) ],
// s_c = sens(c, n) (absent: 0), l = lsens[n] (absent: lval), i.e.
// GradientImpl(sens) + l LogDeterminantInputGradImpl (InputJacobian,
// MonotoneComponent.h:148-164, 758-818; ContinuousMixedInputJacobian,
// 208-232, 877-926) summed over the components.  In KL mode s_c = T_c and
// l = -1: the point gradient of |T|^2 / 2 - log det, without storing T.
template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_input_vjp(MapDev M, const double* w, long ws, PtsArg P, int klMode,
                                                   const double* sens, long srs, long scs, const double* lsens,
                                                   long lss, double lval, double scale, OutArg out, double* gtab) {
    extern __shared__ double smem[];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.N) return;
    long stride;
    double* tab;
    if (gtab) {
        stride = (long)gridDim.x * blockDim.x;
        tab = gtab + n;
    } else {
        stride = blockDim.x;
        tab = smem + threadIdx.x;
    }
    double* dtab = tab + (long)M.nslots * stride;
    const double* pt = P.p + n * P.cs;
    fill_values_derivs<BK>(M, table_dims(M, 0, M.ncomp), pt, P.rs, tab, dtab, (int)stride);
    double* o = out.p + n * out.cs;
    for (int d = 0; d < M.D; ++d) o[(long)d * out.rs] = 0.0;
    const bool hasL = lsens != nullptr || lval != 0.0;
    const double lv = lsens ? lsens[n * lss] : lval;
    for (int c = 0; c < M.ncomp; ++c) {
        const CompDev cd = ro(M.comps)[c];
        const int k = cd.dim - 1;
        const double* wc = w + (long)cd.coeffOff * ws;
        double C[PM + 1];
        coeff_groups<PM>(M, cd, wc, ws, tab, (int)stride, C);
        const double sv = sens ? sens[(long)cd.outRow * srs + n * scs] : 0.0;
        double G[PM + 1], gk;
        if (klMode)
            ivjp_core<BK, PK, QK, PM, true>(M, cd, C, pt[(long)k * P.rs], false, 0.0, true, -1.0, G, gk);
        else
            ivjp_core<BK, PK, QK, PM, false>(M, cd, C, pt[(long)k * P.rs], sens != nullptr, sv, hasL, lv, G, gk);
        for (int i = 0; i < k; ++i) {
            double Dq[PM + 1];
            input_groups<PM>(M, cd, wc, ws, tab, dtab, (int)stride, i, Dq);
            double v = 0.0;
#pragma unroll
            for (int q = 0; q <= PM; ++q) {
                if (q > cd.p) break;
                v += Dq[q] * G[q];
            }
            o[(long)i * out.rs] += scale * v;
        }
        o[(long)k * out.rs] += scale * gk;
    }
}

// ---------------------------------------------------------------------------
// Inverse of one component (MonotoneComponent::InverseImpl, MonotoneComponent.h:331-443;
// RootFinding.h:20-178).  xs has xrows rows and 1 or N columns; its last row
// holds the initial guess.  out may alias that row.
template <int BK, int PK, int QK, int PM>
__device__ double eval_T(const MapDev& M, const DiagCoef<PM>& dc0, double f0, double x) {
    DiagCoef<PM> dc = dc0;
    dc.x = x;
    AdjAcc<PM> A;
    return integrate<BK, PK, QK, FL_VAL, PM>(M.quad, dc, A, nullptr) + f0;
}

// One ITP iterate (Find_x_ITP, RootFinding.h:92-106; k1 = 0.1, k2 = 2, n0 = 1):
// interpolation xf, truncation towards the bisection point xb by
// delta = min(k1 (xub - xlb)^2, |xb - xf|), projection onto the ball of
// radius xtol 2^(nhalf + n0 - it) - (xub - xlb)/2 around xb.  The powers are
// a product and a power of two (pow(w, 2) and pow(2, k) are exactly these in
// a correctly rounded libm).
constexpr unsigned ITP_MAX_ITS = 10000;

// An ITP iterate equal (bit for bit) to an end of the bracket needs no
// quadrature: T there is the value computed when that end was set, and the
// same evaluation gives the same bits.  Where the bracket has shrunk to two
// neighbouring doubles wider than xtol (roots of large magnitude: at 1e13 an
// ulp is 2e-3), every iterate is such an end and the reference's loop
// (RootFinding.h:158-174) runs to its 10 000 iterations without moving;
// inverse_solve_with takes those iterations too, without integrating T again
// (the per-lane queue hands such a solve over, mono_inverse.hpp).
__device__ __forceinline__ bool same_bits(double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b);
}
__device__ __forceinline__ double itp_nhalf(double xlb, double xub, double xtol) {
    return ceil(log2(0.5 * (xub - xlb) / xtol));
}
__device__ __forceinline__ double itp_point(double xlb, double xub, double ylb, double yub, double nhalf, int it,
                                            double xtol) {
#pragma clang fp contract(off)  // the iterate as RootFinding.h:96-104 writes it, unfused
    const double k1 = 0.1, n0 = 1.0;
    const double xb = 0.5 * (xub + xlb);
    double xf = (xub * ylb - xlb * yub) / (ylb - yub);
    const double sigma = ((xb - xf) > 0) ? 1.0 : -1.0;
    const double w = xub - xlb;
    const double delta = fmin(k1 * (w * w), fabs(xb - xf));
    xf += delta * sigma;
    const double e = nhalf + n0 - (double)it;
    const double p2 = e < -2000.0 ? 0.0 : (e > 2000.0 ? __builtin_inf() : ldexp(1.0, (int)e));
    const double rho = fmin(xtol * p2 - 0.5 * (xub - xlb), fabs(xf - xb));
    return xb - sigma * rho;
}

// Bracket (FindBracket, RootFinding.h:20-80) and ITP (Find_x_ITP, 82-178) for
// T(x) = yd, T(x) = evT(x), starting at x0.  Returns NaN when no bracket is
// found, and when T is not finite at an iterate (with finite inputs, an
// infinite g, where the reference fails).  The reference evaluates T(x0)
// three times (InverseSingleBracket, then FindBracket twice); the values are
// identical, so it is evaluated once.
template <class EvT>
__device__ double inverse_solve_with(EvT&& evT, double x0, double yd, double xtol, double ytol) {
    double xlb = x0, xub = x0;
    double ylb = evT(xlb);
    if (nonfinite_err(fabs(ylb))) return NAN;
    double yub = ylb;
    int info = 0;
    {
        const unsigned maxIts = 128;
        double stepSize = 1.0;
        if (ylb > yd) {
            double t = ylb; ylb = yub; yub = t;
            t = xlb; xlb = xub; xub = t;
            unsigned i;
            for (i = 0; i < maxIts; ++i) {
                xlb = xub - stepSize;
                ylb = evT(xlb);
                if (nonfinite_err(fabs(ylb))) return NAN;
                if ((fabs((yub - ylb) / (xub - xlb)) < 1e-12) && ((xub - xlb) > 10)) { info = -1; break; }
                if (ylb > yd) { yub = ylb; xub = xlb; stepSize *= 2.0; } else break;
            }
            if (info == 0 && i >= maxIts) info = -2;
        } else {
            unsigned i;
            for (i = 0; i < maxIts; ++i) {
                xub = xlb + stepSize;
                yub = evT(xub);
                if (nonfinite_err(fabs(yub))) return NAN;
                if ((fabs((yub - ylb) / (xub - xlb)) < 1e-12) && ((xub - xlb) > 10)) { info = -1; break; }
                if (yub < yd) { ylb = yub; xlb = xub; stepSize *= 2.0; } else break;
            }
            if (info == 0 && i >= maxIts) info = -2;
        }
    }
    if ((info < 0) || ((ylb > yd) || (yub < yd))) return NAN;
    const double nhalf = itp_nhalf(xlb, xub, xtol);
    for (unsigned it = 0; it < ITP_MAX_ITS; ++it) {
        const double xc = itp_point(xlb, xub, ylb, yub, nhalf, (int)it, xtol);
        // T at an end of the bracket is the value held for that end (itp_known)
        const double yc = same_bits(xc, xlb) ? ylb : (same_bits(xc, xub) ? yub : evT(xc));
        if (nonfinite_err(fabs(yc))) return NAN;
        if (fabs(yc - yd) < ytol) {
            return xc;
        } else if (yc > yd) {
            xub = xc;
            yub = yc;
        } else {
            xlb = xc;
            ylb = yc;
        }
        if (((xub - xlb) < xtol) || ((yub - ylb) < ytol)) break;
    }
    return 0.5 * (xub + xlb);
}

template <int BK, int PK, int QK, int PM>
__device__ double inverse_solve(const MapDev& M, const DiagCoef<PM>& dc, double f0, double x0, double yd, double xtol,
                                double ytol) {
    return inverse_solve_with([&](double x) { return eval_T<BK, PK, QK, PM>(M, dc, f0, x); }, x0, yd, xtol, ytol);
}

// ---------------------------------------------------------------------------
// Fused forward + adjoint over every component of a map (one launch).
//
// For each point n and component c it computes T_c, log g(d_c f), the
// adjoint integrals S'_q on the forward tree, and
//   G_q = s_c (phi~_q(0) + rho_q S'_q) + l_n rho_q B'_q(x_c) g'/g,
// then reduces sum_n P_j(x_n) G_q(x_n) over the wavefront (wave_reduce16)
// into per-wave LDS accumulators.  klMode uses
// s_c = T_c and l_n = -1 (gradient of mean(0.5 |T|^2 - logdet) up to 1/N) and
// also sums 0.5 |T|^2 - logdet.  Blocks loop over point tiles (persistent
// grid) and write one partial gradient row per block.
//
// LDS layout: [acc: (blockDim/64) * numCoeffs] [table: nslots * blockDim].
// Compile-time factor-slot count for the term-record loops (ND as in
// rec_product; 0: more than eight factors, a run-time loop).
template <int N>
struct NdTag {
    static constexpr int value = N;
};
template <class F>
__device__ __forceinline__ void dispatch_nd(int nd, F&& f) {
    switch (nd) {
        case 1: f(NdTag<1>{}); break;
        case 2: f(NdTag<2>{}); break;
        case 3: f(NdTag<3>{}); break;
        case 4: f(NdTag<4>{}); break;
        case 5: f(NdTag<5>{}); break;
        case 6: f(NdTag<6>{}); break;
        case 7: f(NdTag<7>{}); break;
        case 8: f(NdTag<8>{}); break;
        default: f(NdTag<0>{}); break;
    }
}

// G[q] for a run-time q (a select chain over the register array).
template <int PM>
__device__ __forceinline__ double select_q(const double* G, int q) {
    double g = G[0];
#pragma unroll
    for (int i = 1; i <= PM; ++i) g = (i == q) ? G[i] : g;
    return g;
}

// Products G_q P_j of the 16 term records j0 .. j0+15, zero on an inactive
// lane.  The records are the same for every lane of the wavefront, so with
// ND known they come through scalar loads, several in flight at once.
template <int ND>
__device__ __forceinline__ void chunk16_products(const MapDev& M, int j0, const double* tab, long stride, double Gq,
                                                 bool active, double* v) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const auto rec = ro(M.terms) + (long)(j0 + u) * rec_len<ND>(M);
        const double Pj = rec_product<ND>(M, rec, tab, stride);
        v[u] = active ? Pj * Gq : 0.0;
    }
}

template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_fused(MapDev M, const double* w, long ws, PtsArg P, const double* sens,
                                               long srs, long scs, const double* lsens, long lss, int klMode,
                                               double* partialGrad, double* partialLoss, double* gtab,
                                               double* gradGlobal) {
    extern __shared__ double smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nwaves = blockDim.x >> 6;
    const int K = M.numCoeffs;
    const bool ldsAcc = (gradGlobal == nullptr);
    double* acc = smem;
    double* tabBase = smem + (ldsAcc ? (long)nwaves * K : 0);
    if (ldsAcc) {
        for (int i = tid; i < nwaves * K; i += blockDim.x) acc[i] = 0.0;
    }
    __syncthreads();
    double* myAcc = acc + (long)wave * K;
    long stride;
    double* tab;
    if (gtab) {
        stride = (long)gridDim.x * blockDim.x;
        tab = gtab + (long)blockIdx.x * blockDim.x + tid;
    } else {
        stride = blockDim.x;
        tab = tabBase + tid;
    }
    const int ndims = M.D - 1;
    double lossLane = 0.0;
    const long ntiles = (P.N + blockDim.x - 1) / blockDim.x;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long n0 = tile * blockDim.x + tid;
        const bool active = n0 < P.N;
        const long n = active ? n0 : P.N - 1;
        const double* pt = P.p + n * P.cs;
        fill_values<BK>(M, ndims, pt, P.rs, tab, (int)stride);
        for (int c = 0; c < M.ncomp; ++c) {
            const CompDev cd = ro(M.comps)[c];
            const double* wc = w + (long)cd.coeffOff * ws;
            double C[PM + 1];
            coeff_groups<PM>(M, cd, wc, ws, tab, (int)stride, C);
            const double xk = pt[(long)(cd.dim - 1) * P.rs];
            const double sv = (!klMode && sens) ? sens[(long)cd.outRow * srs + n * scs] : 0.0;
            const double lv = (!klMode && lsens) ? lsens[n * lss] : 0.0;
            double G[PM + 1];
            if (M.discrete) comp_core_discrete<BK, PK, QK, PM>(M, cd, C, xk, active, klMode, sv, lv, G, lossLane);
            else comp_core<BK, PK, QK, PM>(M, cd, C, xk, active, klMode, sv, lv, G, lossLane);
            // Gradient pass: chunks of 16 term records per group.
            dispatch_nd(M.nd, [&](auto tag) {
                constexpr int ND = decltype(tag)::value;
                for (int q = 0; q <= cd.p; ++q) {
                    const GroupDev g = ro(M.groups)[cd.groupOff + q];
                    const double Gq = select_q<PM>(G, q);
                    for (int j0 = g.start; j0 < g.start + g.countPad; j0 += 16) {
                        double v[16];
                        chunk16_products<ND>(M, j0, tab, stride, Gq, active, v);
                        const double rsum = wave_reduce16(v, lane);
                        if ((lane & 3) == 0) {
                            const uint32_t idx = ro(M.terms)[(long)(j0 + (lane >> 2)) * rec_len<ND>(M)];
                            if (idx != DUMMY) {
                                if (ldsAcc) myAcc[cd.coeffOff + idx] += rsum;
                                else atomicAdd(gradGlobal + cd.coeffOff + idx, rsum);
                            }
                        }
                    }
                }
            });
        }
    }
    // Loss: wave sum, then block sum.
    double lw = wave_sum(lossLane);
    __shared__ double lossWave[4];
    if (lane == 0) lossWave[wave] = lw;
    __syncthreads();
    if (tid == 0) {
        double l = 0.0;
        for (int i = 0; i < nwaves; ++i) l += lossWave[i];
        partialLoss[blockIdx.x] = l;
    }
    if (ldsAcc) {
        for (int k = tid; k < K; k += blockDim.x) {
            double s = 0.0;
            for (int i = 0; i < nwaves; ++i) s += acc[(long)i * K + k];
            partialGrad[(long)blockIdx.x * K + k] = s;
        }
    }
}

#ifdef MPD_UTILITY_KERNELS
// Column sums of a block of per-block partial rows: for c < ncols,
// out[c] = scale * sum_r in[r * rowStride + c], except that the last column
// goes to *outLast when outLast is given.  Eight columns per block, 32 lanes
// over the rows, then a fixed-order sum of the 32 lane partials (the result
// does not depend on the launch).
__global__ void __launch_bounds__(256) k_reduce_cols(const double* in, long nrows, long rowStride, long ncols,
                                                     double scale, double* out, double* outLast) {
    __shared__ double red[32][9];
    const int cc = threadIdx.x & 7, rl = threadIdx.x >> 3;
    const long c = (long)blockIdx.x * 8 + cc;
    double s = 0.0;
    if (c < ncols)
        for (long r = rl; r < nrows; r += 32) s += in[r * rowStride + c];
    red[rl][cc] = s;
    __syncthreads();
    if (rl == 0 && c < ncols) {
        double t = 0.0;
        for (int i = 0; i < 32; ++i) t += red[i][cc];
        if (outLast && c == ncols - 1) *outLast = scale * t;
        else out[c] = scale * t;
    }
}

// Two-stage column sums of a row-major partial matrix, for many rows.  Stage
// one: block (x = group of 64 columns, y = chunk of rows) sums its chunk with
// one lane per column (coalesced 512-byte rows per wave); its four waves take
// every fourth row and are combined in wave order.  Stage two sums the chunk
// rows of each column in order.  Both orders are fixed, so results repeat
// bitwise for a given launch shape.
__global__ void __launch_bounds__(256) k_colsum_chunks(const double* in, long nrows, long rowStride, long ncols,
                                                       long chunk, double* work) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long c = (long)blockIdx.x * 64 + lane;
    const long r0 = (long)blockIdx.y * chunk;
    const long r1 = r0 + chunk < nrows ? r0 + chunk : nrows;
    double s = 0.0;
    if (c < ncols) {
#pragma unroll 4
        for (long r = r0 + wv; r < r1; r += 4) s += in[r * rowStride + c];
    }
    red[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && c < ncols)
        work[(long)blockIdx.y * ncols + c] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

__global__ void __launch_bounds__(256) k_colsum_final(const double* work, long nchunks, long ncols, double scale,
                                                      double* out, double* outLast) {
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncols) return;
    double t = 0.0;
#pragma unroll 8
    for (long i = 0; i < nchunks; ++i) t += work[i * ncols + c];
    if (outLast && c == ncols - 1) *outLast = scale * t;
    else out[c] = scale * t;
}

// Plain KL reduction (FastGaussianReverseKLObjective, MapObjective.cpp:146-215):
// grad[k] = (1/N) sum_n (G[k][n] - LG[k][n]), one block per coefficient row.
__global__ void __launch_bounds__(256) k_kl_rows(const double* G, const double* LG, long N, long K, double* grad) {
    __shared__ double red[256];
    const long k = blockIdx.x;
    double s = 0.0;
    for (long n = threadIdx.x; n < N; n += blockDim.x) s += G[k * N + n] - LG[k * N + n];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[k] = red[0] / (double)N;
}

// Per-block partial sums of sum_i 0.5 T_i^2 - ld (T is M x N row-major).
__global__ void __launch_bounds__(256) k_kl_loss(const double* T, long M, const double* ld, long N, double* partial) {
    __shared__ double red[256];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (n < N) {
        for (long i = 0; i < M; ++i) v += 0.5 * T[i * N + n] * T[i * N + n];
        v -= ld[n];
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Fill a strided matrix with a value (used instead of memset so that captured
// graphs hold no memset nodes).
__global__ void k_fill(double* p, long rows, long cols, long rs, long cs, double v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows * cols) p[(i / cols) * rs + (i % cols) * cs] = v;
}

// Strided copy (SetCoeffs deep copy, Inverse buffers).
__global__ void k_copy(const double* src, long rows, long cols, long srs, long scs, double* dst, long drs, long dcs) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows * cols) {
        const long r = i / cols, c = i % cols;
        dst[r * drs + c * dcs] = src[r * srs + c * scs];
    }
}

#endif  // MPD_UTILITY_KERNELS

}  // namespace mpd
