// mono_inverse.hpp -- inverse of one component with per-lane solve scheduling.
//
// MonotoneComponent::InverseImpl (MonotoneComponent.h:331-443) solves
// T(x_<k, x) = r for every point with a bracket search (FindBracket,
// RootFinding.h:20-90) and ITP iterations (InverseSingleBracket, 126-176),
// each step an adaptive quadrature.  Per point the work varies by two orders
// of magnitude (a few hundred to tens of thousands of integrand evaluations
// on random targets), so a lane-per-point kernel that runs each wavefront
// until its slowest lane is done spends most of its time masked.
//
// Here a lane advances its own solve one quadrature interval at a time: each
// loop iteration every busy lane evaluates the integrand at the nodes of its
// current interval (independent evaluations, so they overlap), applies the
// interval's acceptance test and, when its quadrature is complete, advances
// its bracket / ITP state.  A lane whose solve ends writes the root and takes
// the next point from a wavefront-local range of point indices, refilled 64
// at a time from a global counter.  The nodes, the accumulations and every
// comparison are those of integrate_as / integrate_acc and inverse_solve, so
// the results are those of the sequential solver bit for bit; only the
// interleaving differs.
//
// The per-point setup (basis values, C_q sums, f(x_<k, 0)) runs in a separate
// lane-per-point pass that stores a record per point, so that refilling a
// lane costs a few loads.
#pragma once
#include "batch_io.hpp"
#include "mono_kernels.hpp"

namespace mpd {

// Per-point record, structure of arrays with N doubles per row:
// rows 0..PM: C_q rho_q (DiagCoef::cd), PM+1: f(x_<k, 0), PM+2: the initial
// guess (NaN when any input coordinate is NaN), PM+3: the target r.
template <int PM>
constexpr int inv_rec_rows() {
    return PM + 4;
}

// Record of one point: pt is the point's input column (xrows entries, stride
// xrs; the last one the initial guess), r its record column (rows R apart).
template <int BK, int PK, int QK, int PM>
__device__ __forceinline__ void inv_setup_point(const MapDev& M, const double* w, long ws, const double* pt,
                                                long xrows, long xrs, double yv, int c, double* tab, long stride,
                                                double* r, long R) {
    bool bad = false;
    for (long i = 0; i < xrows; ++i) bad |= isnan(pt[i * xrs]);
    r[(PM + 3) * R] = yv;
    if (bad) {
        r[(PM + 2) * R] = NAN;
        return;
    }
    const double x0 = pt[(xrows - 1) * xrs];
    const CompDev cd = ro(M.comps)[c];
    fill_values<BK>(M, cd.dim - 1, pt, xrs, tab, (int)stride);
    double C[PM + 1];
    coeff_groups<PM>(M, cd, w + (long)cd.coeffOff * ws, ws, tab, (int)stride, C);
    DiagCoef<PM> dc;
    make_diag<PM>(M, cd, C, x0, M.nugget, dc);
    for (int q = 0; q <= PM; ++q) r[q * R] = dc.cd[q];
    r[(PM + 1) * R] = f_at_zero<PM>(M, cd, C);
    r[(PM + 2) * R] = x0;
}

template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_inverse_setup(MapDev M, const double* w, long ws, const double* xs,
                                                       long xrows, long xcols, long xrs, long xcs, const double* ys,
                                                       long yss, long N, int c, double* gtab, double* rec) {
    extern __shared__ double smem[];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const long xi = (xcols == 1) ? 0 : n;
    long stride;
    double* tab = lane_table(smem, gtab, stride);
    inv_setup_point<BK, PK, QK, PM>(M, w, ws, xs + xi * xcs, xrows, xrs, ys[n * yss], c, tab, stride, rec + n, N);
}

// Batched setup: block b covers points tileStart[b] .. + blockDim.x - 1 of
// segment tileSeg[b]; records go to the pass's index space (rows R apart).
template <int BK, int PK, int QK, int PM>
__global__ void __launch_bounds__(256) k_inverse_setup_batch(const MapDev* __restrict__ maps,
                                                             const InvSeg* __restrict__ segs,
                                                             const int* __restrict__ tileSeg,
                                                             const long* __restrict__ tileStart, long R, int c,
                                                             double* rec) {
    extern __shared__ double smem[];
    const InvSeg& S = segs[tileSeg[blockIdx.x]];
    const long n = tileStart[blockIdx.x] + threadIdx.x;
    if (n >= S.N) return;
    const long xi = (S.xcols == 1) ? 0 : n;
    long stride;
    double* tab = lane_table(smem, nullptr, stride);
    inv_setup_point<BK, PK, QK, PM>(maps[S.map], S.w, S.ws, S.x + xi * S.xcs, S.xrows, S.xrs, S.y[n * S.yss], c, tab,
                                    stride, rec + S.base + n, R);
}

// Batched setup with G lanes per point, for maps with many terms (the
// tiles after the first ntiles1 of a pass, each of tileN points of one map):
// block b covers blockDim / G points of tile b / parts, the G lanes of a
// point share one basis table in LDS (M.nslots doubles, at most the batch's
// grpSlots) and split the term sums (group_coeff_groups, a butterfly sum of
// G partial sums: the same C_q up to the order of summation); lane 0 of the
// group writes the record.  Lane-per-point tables of such maps fill the LDS
// with a few wavefronts per CU; a table per G lanes gives G times as many.
template <int BK, int PK, int QK, int PM, int G>
__global__ void __launch_bounds__(256) k_inverse_setup_batch_grp(const MapDev* __restrict__ maps,
                                                                 const InvSeg* __restrict__ segs,
                                                                 const int* __restrict__ tileSeg,
                                                                 const long* __restrict__ tileStart, int tileN, long R,
                                                                 int c, double* rec) {
    extern __shared__ double smem[];
    const int per = blockDim.x / G;
    const int parts = (tileN + per - 1) / per;
    const int tile = blockIdx.x / parts, part = blockIdx.x - tile * parts;
    const InvSeg& S = segs[tileSeg[tile]];
    const int gl = threadIdx.x & (G - 1), grp = threadIdx.x / G;
    const int k = part * per + grp;
    const long n = tileStart[tile] + k;
    if (k >= tileN || n >= S.N) return;  // the whole group leaves together
    const MapDev& M = maps[S.map];
    double* tab = smem + (long)grp * M.nslots;
    const long xi = (S.xcols == 1) ? 0 : n;
    const double* pt = S.x + xi * S.xcs;
    double* r = rec + S.base + n;
    bool bad = false;
    for (long i = 0; i < S.xrows; ++i) bad |= isnan(pt[i * S.xrs]);
    if (gl == 0) r[(PM + 3) * R] = S.y[n * S.yss];
    if (bad) {
        if (gl == 0) r[(PM + 2) * R] = NAN;
        return;
    }
    const double x0 = pt[(S.xrows - 1) * S.xrs];
    const CompDev cd = ro(M.comps)[c];
    group_fill_values<BK, G>(M, cd.dim - 1, pt, S.xrs, tab, gl);
    double C[PM + 1];
    group_coeff_groups<PM, G>(M, cd, S.w + (long)cd.coeffOff * S.ws, S.ws, tab, gl, C);
    if (gl == 0) {
        DiagCoef<PM> dc;
        make_diag<PM>(M, cd, C, x0, M.nugget, dc);
        for (int q = 0; q <= PM; ++q) r[q * R] = dc.cd[q];
        r[(PM + 1) * R] = f_at_zero<PM>(M, cd, C);
        r[(PM + 2) * R] = x0;
    }
}

// The same for one map (the single-map Inverse of a map with many terms):
// block b covers points b * blockDim / G onwards.  The arithmetic is that
// of k_inverse_setup_batch_grp, so the records, and the roots, are those of
// the batched Inverse.
template <int BK, int PK, int QK, int PM, int G>
__global__ void __launch_bounds__(256) k_inverse_setup_grp(MapDev M, const double* w, long ws, const double* xs,
                                                           long xrows, long xcols, long xrs, long xcs,
                                                           const double* ys, long yss, long N, int c, double* rec) {
    extern __shared__ double smem[];
    const int per = blockDim.x / G;
    const int gl = threadIdx.x & (G - 1), grp = threadIdx.x / G;
    const long n = (long)blockIdx.x * per + grp;
    if (n >= N) return;  // the whole group leaves together
    double* tab = smem + (long)grp * M.nslots;
    const long xi = (xcols == 1) ? 0 : n;
    const double* pt = xs + xi * xcs;
    double* r = rec + n;
    bool bad = false;
    for (long i = 0; i < xrows; ++i) bad |= isnan(pt[i * xrs]);
    if (gl == 0) r[(PM + 3) * N] = ys[n * yss];
    if (bad) {
        if (gl == 0) r[(PM + 2) * N] = NAN;
        return;
    }
    const double x0 = pt[(xrows - 1) * xrs];
    const CompDev cd = ro(M.comps)[c];
    group_fill_values<BK, G>(M, cd.dim - 1, pt, xrs, tab, gl);
    double C[PM + 1];
    group_coeff_groups<PM, G>(M, cd, w + (long)cd.coeffOff * ws, ws, tab, gl, C);
    if (gl == 0) {
        DiagCoef<PM> dc;
        make_diag<PM>(M, cd, C, x0, M.nugget, dc);
        for (int q = 0; q <= PM; ++q) r[q * N] = dc.cd[q];
        r[(PM + 1) * N] = f_at_zero<PM>(M, cd, C);
        r[(PM + 2) * N] = x0;
    }
}

// Pending right half of the depth-first quadrature (the entries of
// integrate_as / integrate_acc: the interval's left end and its value are
// those of the leaf accepted before it is popped, its width 2^-level; Simpson
// also keeps the midpoint value).
template <int QK>
struct InvStk {
    double fb;
    int level;
};
template <>
struct InvStk<QK_AS> {
    double fm, fb;
    int level;
};

// Solver phases.
enum { IP_START = 0, IP_DOWN = 1, IP_UP = 2, IP_ITP = 3 };

// ---------------------------------------------------------------------------
// Hand-over of the longest solves.  A few solves cost thousands of times the
// median (T evaluated where it spans hundreds of orders of magnitude, each
// quadrature subdividing into thousands of intervals); one lane running such
// a solve is the whole tail of a pass.  With an adaptive rule, once the
// wavefront finds no point left to take, a queue lane whose solve exceeds
// solveSteps interval steps hands the point over: it appends the point index to list (count is the
// list length; appends past cap are not made, and that lane keeps its
// solve) and goes idle.  (Long solves met earlier stay in their lanes: the
// other lanes keep taking points meanwhile, and a wavefront spends more
// time on a solve of a small tree than its lanes' share of it.)  k_inverse_heavy then runs
// every handed-over solve again from its record, one wavefront per solve,
// each T evaluation spread over the 64 lanes (coop_integrate).  Budgets of
// 0 turn the hand-over off.

// Workspace of one wavefront (LDS, COOP_LDS_BYTES): per processed interval
// (node) of one quadrature, in level order, its left end index, end and
// midpoint values, fine sum, split rank, leaf count and slot in the round's
// speculative node values; and the node values of one round (64).  The
// leaves in left-to-right order and their offsets are written only once the
// tree is complete, when the left-end values and the slots are no longer
// read: they share those arrays.
struct CoopWs {
    double *fa, *fm, *fb, *fine, *leaf, *spec;
    unsigned* idx;
    int *srank, *cnt, *off, *sp;
    int cap;
};
__device__ __forceinline__ CoopWs coop_ws(double* base, int cap) {
    static_assert(COOP_LDS_BYTES == COOP_NODES * (4 * sizeof(double) + 4 * sizeof(int)) + 64 * sizeof(double),
                  "workspace layout");
    CoopWs W;
    W.fa = base;
    W.fm = base + cap;
    W.fb = base + 2L * cap;
    W.fine = base + 3L * cap;
    W.leaf = W.fa;  // after the rounds
    W.spec = base + 4L * cap;
    int* ib = reinterpret_cast<int*>(base + 4L * cap + 64);
    W.idx = reinterpret_cast<unsigned*>(ib);
    W.srank = ib + cap;
    W.cnt = ib + 2L * cap;
    W.sp = ib + 3L * cap;
    W.off = W.sp;  // after the rounds
    W.cap = cap;
    return W;
}

__device__ __forceinline__ double lane_bcast(double v, int src) {
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Hand-over statistics of one wavefront, kept in registers and added to the
// call's totals (InvHeavy::stats) once, when the wavefront is done: atomics
// on those few words at every T evaluation serialize in the L2.
struct HeavyStats {
    unsigned long long v[HS_COUNT];
    __device__ void add(int k, unsigned long long x) { v[k] += x; }
    __device__ void max(int k, unsigned long long x) { v[k] = x > v[k] ? x : v[k]; }
    __device__ void flush(unsigned long long* st) const {
        if (!st || threadIdx.x != 0) return;
#pragma unroll
        for (int k = 0; k < HS_COUNT; ++k) {
            if (k == HS_MAX_NODES || k == HS_MAX_LEVELS || k == HS_MAX_SOLVE_TEVALS || k == HS_MAX_SOLVE_TICKS) {
                if (v[k]) atomicMax(st + k, v[k]);
            } else if (v[k]) {
                atomicAdd(st + k, v[k]);
            }
        }
    }
};

// The nested Clenshaw-Curtis rule on one interval, as integrate_acc and
// integrate_acc_big form it.  Its E = acc_nodes(Q) new nodes are numbered in
// their order of evaluation there: the midpoint, the coarse rule's interior
// nodes left of it and right of it, then the fine-only nodes.
struct AccRule {
    const double *cw, *cp, *fw, *fp;
    int nf, cRight, cMid, fRight, fMid;
};
__device__ __forceinline__ AccRule acc_rule(const QuadDev& Q) {
    AccRule R;
    const int nc = Q.nc;
    R.nf = Q.nf;
    R.cw = Q.big ? Q.big : Q.cw;
    R.cp = Q.big ? Q.big + nc : Q.cp;
    R.fw = Q.big ? Q.big + 2 * nc : Q.fw;
    R.fp = Q.big ? Q.big + 2 * nc + R.nf : Q.fp;
    R.cRight = nc - 1;
    R.cMid = R.cRight / 2;
    R.fRight = R.nf - 1;
    R.fMid = R.fRight / 2;
    return R;
}
// coarse interior nodes left of the midpoint and in all (the midpoint aside)
__device__ __forceinline__ int acc_left(const AccRule& R) { return R.cMid > 1 ? R.cMid - 1 : 0; }
__device__ __forceinline__ int acc_interior(const AccRule& R) {
    return acc_left(R) + (R.cRight - R.cMid - 1 > 0 ? R.cRight - R.cMid - 1 : 0);
}
__device__ __forceinline__ int acc_nodes(const AccRule& R) {
    return 1 + acc_interior(R) + (R.nf - 1) / 2;  // midpoint, coarse interior, fine-only
}
// Position of node e of [a, b].
__device__ __forceinline__ double acc_node_pos(const AccRule& R, int e, double a, double b) {
#pragma clang fp contract(off)  // as in integrate_acc
    const double mid = 0.5 * (a + b);
    const double scale = 0.5 * (b - a);
    if (e == 0) return mid;
    const int nci = acc_interior(R);
    if (e <= nci) {
        const int k = e - 1;  // k-th coarse interior node: i = 1 .. cMid-1, then cMid+1 .. cRight-1
        const int c1 = acc_left(R);
        const int ci = (k < c1) ? k + 1 : R.cMid + 1 + (k - c1);
        return mid + scale * R.cp[ci];
    }
    const int k = e - 1 - nci;  // k-th fine-only node: i = 1, 3, 5, ...
    return mid + scale * R.fp[2 * k + 1];
}
// Coarse and fine sums of [a, b] from its end values and its node values v[e].
__device__ __forceinline__ void acc_sums(const AccRule& R, double a, double b, double fa, double fb, const double* v,
                                         double& coarse, double& fine) {
#pragma clang fp contract(off)  // as in integrate_acc
    const double scale = 0.5 * (b - a);
    const double fm = v[0];
    coarse = scale * R.cw[0] * fa;
    coarse += scale * R.cw[R.cMid] * fm;
    coarse += scale * R.cw[R.cRight] * fb;
    fine = scale * R.fw[0] * fa;
    fine += scale * R.fw[R.fMid] * fm;
    fine += scale * R.fw[R.fRight] * fb;
    int e = 1;
    for (int i = 1; i < R.cMid; ++i, ++e) {
        coarse += scale * R.cw[i] * v[e];
        fine += scale * R.fw[2 * i] * v[e];
    }
    for (int ii = R.cMid + 1; ii < R.cRight; ++ii, ++e) {
        coarse += scale * R.cw[ii] * v[e];
        fine += scale * R.fw[2 * ii] * v[e];
    }
    for (int i = 1; i < R.nf; i += 2, ++e) fine += scale * R.fw[i] * v[e];
}

// The adaptive rules (integrate_as, integrate_acc; FL_VAL) computed by the 64
// lanes of a one-wavefront block.  The tree is built level by level, in
// rounds (below): the new nodes of the pending intervals (Simpson's two
// quarter points, or the nested rule's midpoint and interior nodes) are
// evaluated a lane per node, the intervals are tested a lane per interval,
// and the children of the intervals that fail are written to the next level
// in left-to-right order.  The intervals are the dyadic ones the
// depth-first loop visits (positions idx 2^-level are exact), with the same
// node values, sums and tests, so the tree is the same.  The accepted
// intervals are then ranked left to right (leaf counts bottom-up, offsets
// top-down) and summed in that order from 0, as the depth-first loop adds
// them: the result is the same bit for bit.  A non-finite interval error
// (where the depth-first loop stops early) or a tree beyond the workspace
// falls back to the depth-first loop in every lane; the level pass records
// nothing in the status word, the fallback records what the loop records.
// Every lane returns the same value.  lvl: LDS, MAXSUB + 2 level starts.
template <int BK, int PK, int QK, int PM>
__device__ double coop_integrate(const QuadDev& Q, const DiagCoef<PM>& dcIn, const CoopWs& W, int* lvl,
                                 HeavyStats& st) {
#pragma clang fp contract(off)  // as in integrate_as
    constexpr bool AS = (QK == QK_AS);
    const int lane = threadIdx.x;
    DiagCoef<PM> dc = dcIn;
    dc.status = nullptr;
    AccRule R{};
    int E = 2;  // new nodes per interval
    if constexpr (!AS) {
        R = acc_rule(Q);
        E = acc_nodes(R);
    }
    bool fallback = E > 64, nonfinite = false;
    int start = 0, n = 1, L = 0;
    if (!fallback) {
        // the depth-first loops evaluate F(0), F(1) and (Simpson) F(1/2) first
        const double t = lane == 0 ? 0.0 : (lane == 1 ? 1.0 : 0.5);
        const double v = eval_node<BK, PK, FL_VAL, PM>(dc, t).F;
        const double fa = lane_bcast(v, 0), fb = lane_bcast(v, 1), fm = lane_bcast(v, 2);
        if (lane == 0) {
            W.idx[0] = 0u;
            W.fa[0] = fa;
            W.fm[0] = fm;
            W.fb[0] = fb;
            lvl[0] = 0;
        }
        __syncthreads();
    }
    // Rounds.  A round takes the pending intervals of level L (n of them, from
    // lvl[L] = start) and, when n E <= 32, the speculative subtrees of depth
    // d below them (the largest d <= 6 with n E (2^d - 1) <= 64 that stays
    // above quadMaxSub): every new node of every interval of those subtrees
    // gets a lane and is evaluated at once (into W.spec).  Levels L .. L+d-1
    // are then tested one after the other from those values, a lane per
    // interval, and only the children of intervals that split are kept, so
    // the tree is the depth-first loop's.  With n E > 32 a round covers the
    // one level L, 64 / E of its intervals at a time.  W.sp holds the slot of
    // an interval's node values in W.spec (its pending ancestor's index in
    // the chunk times 2^d - 1, plus its place in the subtree in heap order).
    while (n > 0 && !fallback) {
        int d = 1;
        if (n * E <= 32) {
            while (d < 6 && n * E * ((2 << d) - 1) <= 64 && L + d < Q.maxSub) ++d;
        }
        const int S = (1 << d) - 1;
        const int chunk = (d == 1) ? 64 / E : n;
        int nsplit = 0;  // splits so far in level L (the chunks of a one-level round)
        for (int c0 = 0; c0 < n && !fallback; c0 += chunk) {
            const int cn = min(chunk, n - c0);
            {  // evaluate: task t -> pending interval i of the chunk, subtree interval h (heap order), node e
                const int t = lane;
                if (t < cn * S * E) {
                    const int ii = t / (S * E);
                    const int r = t - ii * (S * E);
                    const int h = r / E, e = r - h * E;
                    const int l = 31 - __builtin_clz((unsigned)(h + 1));
                    const unsigned id = (W.idx[start + c0 + ii] << l) + (unsigned)(h + 1 - (1 << l));
                    const double a = ldexp((double)id, -(L + l));
                    const double b = ldexp((double)(id + 1u), -(L + l));
                    double x;
                    if constexpr (AS) {
                        const double m = 0.5 * (a + b);
                        x = (e == 0) ? 0.5 * (a + m) : 0.5 * (m + b);
                    } else {
                        x = acc_node_pos(R, e, a, b);
                    }
                    W.spec[t] = eval_node<BK, PK, FL_VAL, PM>(dc, x).F;
                }
                if (lane < cn) W.sp[start + c0 + lane] = lane * S;
            }
            __syncthreads();
            // test the round's levels: intervals ls .. ls + ln - 1 of the
            // level that starts at lvlS and holds lvlN intervals
            int lvlS = start, lvlN = n, ls = start + c0, ln = cn;
            for (int l = 0; l < d; ++l) {
                const int Lc = L;  // the level under test (L moves down with l)
                const bool last = (l == d - 1);
                const int childS = lvlS + lvlN;
                const int k = lane;
                const bool act = k < ln;
                const int node = ls + k;
                unsigned id = 0u;
                int sp = 0;
                double a = 0.0, b = 1.0, fa = 0.0, fm = 0.0, fb = 0.0, coarse = 0.0, fine = 0.0, v0 = 0.0, v1 = 0.0;
                if (act) {
                    id = W.idx[node];
                    sp = W.sp[node];
                    fa = W.fa[node];
                    fb = W.fb[node];
                    a = ldexp((double)id, -Lc);
                    b = ldexp((double)(id + 1u), -Lc);
                    v0 = W.spec[sp * E];
                    if constexpr (AS) {
                        fm = W.fm[node];
                        v1 = W.spec[sp * E + 1];
                        const double m = 0.5 * (a + b);
                        coarse = div6(b - a) * (fa + 4.0 * fm + fb);
                        const double wl = div6(m - a), wr = div6(b - m);
                        fine = wl * (fa + 4.0 * v0 + fm);
                        fine += wr * (fm + 4.0 * v1 + fb);
                    } else {
                        acc_sums(R, a, b, fa, fb, W.spec + sp * E, coarse, fine);
                    }
                }
                const double err = fabs(fine - coarse);
                if (__ballot(act && nonfinite_err(err))) {
                    fallback = nonfinite = true;
                    break;
                }
                const double tol = fmax(Q.relTol * fabs(coarse), Q.absTol);
                const bool accept = act && (((err < tol) || (Lc == Q.maxSub - 1)) && (Lc >= Q.minSub));
                const bool split = act && !accept;
                const unsigned long long mask = __ballot(split);
                const int ns = __popcll(mask);
                const int base = nsplit;  // splits before this chunk in level Lc
                if (childS + 2 * (base + ns) > W.cap || Lc + 1 > MAXSUB) {
                    fallback = true;
                    break;
                }
                const int rank =
                    __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
                if (act) {
                    W.fine[node] = fine;
                    W.srank[node] = accept ? -1 : base + rank;
                }
                if (split) {
                    // Simpson: [a, m] holds (fa, f(lm), fm), [m, b] (fm, f(rm), fb);
                    // the nested rule: [a, mid] holds (fa, f(mid)), [mid, b] (f(mid), fb)
                    const int ch = childS + 2 * (base + rank);
                    const int h = sp % S;
                    W.idx[ch] = 2u * id;
                    W.fa[ch] = fa;
                    W.fb[ch] = AS ? fm : v0;
                    W.idx[ch + 1] = 2u * id + 1u;
                    W.fa[ch + 1] = AS ? fm : v0;
                    W.fb[ch + 1] = fb;
                    if constexpr (AS) {
                        W.fm[ch] = v0;
                        W.fm[ch + 1] = v1;
                    }
                    if (!last) {
                        W.sp[ch] = sp + h + 1;
                        W.sp[ch + 1] = sp + h + 2;
                    }
                }
                nsplit += ns;
                __syncthreads();
                if (last) break;
                // level L + l is complete (a subtree round has one chunk): move down
                ++L;
                if (lane == 0) lvl[L] = childS;
                lvlS = childS;
                lvlN = 2 * ns;
                ls = childS;
                ln = lvlN;
                start = childS;
                n = lvlN;
                nsplit = 0;
                if (ln == 0) break;
            }
            if (d > 1) break;  // the subtree round is done
        }
        if (fallback) break;
        if (d == 1) {
            // the last tested level (L, in chunks) is complete
            ++L;
            const int next = start + n;
            if (lane == 0) lvl[L] = next;
            __syncthreads();
            start = next;
            n = 2 * nsplit;
        } else if (n > 0) {
            // the round's last tested level is L; its children were written
            // after it, nsplit of its intervals split
            const int next = start + n;
            ++L;
            if (lane == 0) lvl[L] = next;
            __syncthreads();
            start = next;
            n = 2 * nsplit;
        }
    }
    st.add(HS_TEVALS, 1);
    st.max(HS_MAX_LEVELS, (unsigned long long)L);
    if (fallback) {
        st.add(nonfinite ? HS_FALLBACK_NONFINITE : HS_FALLBACK_SIZE, 1);
        __syncthreads();
        AdjAcc<PM> A;
        return integrate<BK, PK, QK, FL_VAL, PM>(Q, dcIn, A, nullptr);
    }
    st.add(HS_NODES, (unsigned long long)start);
    st.max(HS_MAX_NODES, (unsigned long long)start);
    __syncthreads();  // lvl[L]
    // Levels 0 .. L-1 hold the tree (level k at lvl[k] .. lvl[k+1]-1; the
    // children of level k start at lvl[k+1]).  Leaf counts, bottom-up.
    for (int k = L - 1; k >= 0; --k) {
        const int s = lvl[k], e = lvl[k + 1];
        for (int j = s + lane; j < e; j += 64) {
            const int r = W.srank[j];
            W.cnt[j] = (r < 0) ? 1 : W.cnt[e + 2 * r] + W.cnt[e + 2 * r + 1];
        }
        __syncthreads();
    }
    // Offsets of the leaves in left-to-right order, top-down.
    if (lane == 0) W.off[0] = 0;
    __syncthreads();
    for (int k = 0; k < L; ++k) {
        const int s = lvl[k], e = lvl[k + 1];
        for (int j = s + lane; j < e; j += 64) {
            const int r = W.srank[j];
            const int o = W.off[j];
            if (r < 0) {
                W.leaf[o] = W.fine[j];
            } else {
                const int c = e + 2 * r;
                W.off[c] = o;
                W.off[c + 1] = o + W.cnt[c];
            }
        }
        __syncthreads();
    }
    const int nLeaves = W.cnt[0];
    double res = 0.0;
#pragma unroll 8
    for (int k = 0; k < nLeaves; ++k) res += W.leaf[k];
    __syncthreads();  // the workspace is rewritten by the next call
    return res;
}

// The handed-over solves (see InvHeavy): one wavefront (block of 64) per
// solve, blocks taking list entries from a counter until none is left.  Each solve runs from the point's
// record with inverse_solve_with and the cooperative T; all lanes hold the
// same values, and lane 0 writes the root.
template <int BK, int PK, int QK, int PM, bool LIN, bool BATCH>
__global__ void __launch_bounds__(64) k_inverse_heavy(MapDev M, const double* __restrict__ rec, long N, double* out,
                                                      long os, int c, double xtol, double ytol, InvHeavy H,
                                                      const InvSeg* __restrict__ segs,
                                                      const int* __restrict__ chunkSeg) {
    __shared__ int lvl[MAXSUB + 2];
    extern __shared__ double coopLds[];  // COOP_LDS_BYTES
    const unsigned long long total = *H.count;
    const long cnt = (long)(total < (unsigned long long)H.cap ? total : (unsigned long long)H.cap);
    __shared__ long eTake;
    const CoopWs W = coop_ws(coopLds, COOP_NODES);
    HeavyStats hs{};
    while (true) {
        // solves are taken one at a time, so that long ones spread over the blocks
        if (threadIdx.x == 0) eTake = (long)atomicAdd(H.take, 1ull);
        __syncthreads();
        const long e = eTake;
        __syncthreads();
        if (e >= cnt) break;
        const long n = H.list[e];
        hs.add(HS_SOLVES, 1);
        DiagCoef<PM> dc;
        dc.nug = M.nugget;
        dc.x = 0.0;
        dc.lin = LIN ? M.lin : 0;
        dc.lb = M.lb;
        dc.ub = M.ub;
        double* op = nullptr;
        double* op2 = nullptr;
        if constexpr (BATCH) {
            const InvSeg& S = segs[chunkSeg[n / INV_CHUNK]];
            const long local = n - S.base;
            op = S.o + local * S.os;
            op2 = S.o2 ? S.o2 + local * S.os2 : nullptr;
            dc.p = S.p;
            dc.status = S.status;
        } else {
            dc.p = ro(M.comps)[c].p;
            dc.status = M.status;
        }
#pragma unroll
        for (int q = 0; q <= PM; ++q) dc.cd[q] = rec[q * N + n];
        const double f0 = rec[(PM + 1) * N + n];
        const double x0 = rec[(PM + 2) * N + n];
        const double yd = rec[(PM + 3) * N + n];
        const unsigned long long t0 = wall_clock64();
        unsigned long long nT = 0;
        const double root = inverse_solve_with(
            [&](double x) {
                DiagCoef<PM> d = dc;
                d.x = x;
                ++nT;
                return coop_integrate<BK, PK, QK, PM>(M.quad, d, W, lvl, hs) + f0;
            },
            x0, yd, xtol, ytol);
        hs.max(HS_MAX_SOLVE_TEVALS, nT);
        hs.max(HS_MAX_SOLVE_TICKS, wall_clock64() - t0);
        if (threadIdx.x == 0) {
            if constexpr (BATCH) {
                *op = root;
                if (op2) *op2 = root;
            } else {
                out[n * os] = root;
            }
        }
    }
    hs.flush(H.stats);
}


// Launches k_inverse_heavy after a queue pass (the adaptive rules with the
// hand-over on; no kernel is built for the fixed rule, whose solves take one
// step per T evaluation).
template <int BK, int PK, int QK, int PM, bool LIN, bool BATCH>
inline void launch_inverse_heavy(int blocks, hipStream_t s, const MapDev& M, const double* rec, long N, double* out,
                                 long os, int c, double xtol, double ytol, const InvHeavy& H, const InvSeg* segs,
                                 const int* chunkSeg) {
    if constexpr (QK != QK_CC) {
        if (H.solveSteps > 0 && blocks > 0)
            hipLaunchKernelGGL((k_inverse_heavy<BK, PK, QK, PM, LIN, BATCH>), dim3((unsigned)blocks), dim3(64),
                               COOP_LDS_BYTES, s, M, rec, N, out, os, c, xtol, ytol, H, segs, chunkSeg);
    }
}

// LIN: the map has LinearizedBasis bounds.  Without them the kernel is built
// with the bound checks folded away (they cost ~15 % of the integrand work).
// BATCH: the points of many maps (segs, one per map; chunkSeg: segment of
// each chunk of INV_CHUNK indices), each map starting at a chunk boundary;
// M then only supplies the options every map of a batch shares (quadrature,
// nugget, basis bounds), and each lane takes p and its output rows from the
// segment of its point.  Indices past a map's points are skipped.
template <int BK, int PK, int QK, int PM, bool LIN, bool BATCH>
__global__ void __launch_bounds__(256) k_inverse_queue(MapDev M, const double* __restrict__ rec, long N, double* out,
                                                       long os, int c, double xtol, double ytol,
                                                       unsigned long long* counter, const InvSeg* __restrict__ segs,
                                                       const int* __restrict__ chunkSeg,
                                                       const InvHeavy* __restrict__ Hd, int solveSteps) {
    const QuadDev& Q = M.quad;
    const int p = BATCH ? 0 : ro(M.comps)[c].p;
    // Clenshaw-Curtis node table of one interval, in the order integrate_acc
    // visits the nodes after the midpoint: position (cp or fp), coarse weight
    // index (-1: fine-only node) and fine weight index.
    __shared__ double nPos[2 * MAXNF];
    __shared__ int nCw[2 * MAXNF], nFw[2 * MAXNF];
    __shared__ double sCw[MAXNF], sFw[MAXNF];
    const int nf = (QK == QK_ACC5) ? 5 : Q.nf;
    const int nc = (nf - 1) / 2 + 1;
    const int cRight = nc - 1, cMid = cRight / 2;
    const int fRight = nf - 1, fMid = fRight / 2;
    const int nNodes = cRight - 1 + (nf - 1) / 2;  // per interval, midpoint included
    // A rule of more than MAXNF points (Q.big) is read from device memory by
    // the steps below; the tables here hold rules up to MAXNF points.
    if ((QK == QK_ACC5 || QK == QK_ACC) && !Q.big) {
        if (threadIdx.x == 0) {
            int j = 1;
            for (int i = 1; i < cMid; ++i, ++j) {
                nPos[j] = Q.cp[i];
                nCw[j] = i;
                nFw[j] = 2 * i;
            }
            for (int ii = cMid + 1; ii < cRight; ++ii, ++j) {
                nPos[j] = Q.cp[ii];
                nCw[j] = ii;
                nFw[j] = 2 * ii;
            }
            for (int i = 1; i < nf; i += 2, ++j) {
                nPos[j] = Q.fp[i];
                nCw[j] = -1;
                nFw[j] = i;
            }
        }
        for (int i = threadIdx.x; i < MAXNF; i += blockDim.x) {
            sCw[i] = Q.cw[i];
            sFw[i] = Q.fw[i];
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const double nug = M.nugget;

    // Point and solver state.
    long n = -1;
    double* op = nullptr;   // BATCH: the lane's output rows
    double* op2 = nullptr;
    bool busy = false, finished = false;
    DiagCoef<PM> dc;
    dc.p = p;
    dc.nug = nug;
    dc.x = 0.0;
#pragma unroll
    for (int q = 0; q <= PM; ++q) dc.cd[q] = 0.0;
    dc.lin = LIN ? M.lin : 0;
    dc.lb = M.lb;
    dc.ub = M.ub;
    dc.status = BATCH ? nullptr : M.status;
    double f0 = 0.0, yd = 0.0, x0 = 0.0;
    int phase = IP_START, bi = 0, it = 0;
    double step = 1.0, xlb = 0.0, xub = 0.0, ylb = 0.0, yub = 0.0, nhalf = 0.0, xc = 0.0;
    // Quadrature state.
    double qa = 0.0, qb = 1.0, res = 0.0, fa = 0.0, fb = 0.0, fm = 0.0;
    int level = 0, sp = 0, k = 0;
    // Depth-first stack: the top entry in registers, the others in scratch
    // (stk[0 .. sp-2]), so that popping does not wait for memory; the load
    // that refills the top completes while the next interval is evaluated.
    // topInMem: the top was refilled from stk[sp-1] and is unchanged, so a
    // push need not store it again.
    InvStk<QK> stk[MAXSUB];
    InvStk<QK> top{};
    bool topInMem = false;
    // Wavefront-uniform range of point indices not yet handed to a lane.
    long qNext = 0, qEnd = 0;
    bool exhausted = false;
    // Interval steps the current solve may still take before it is handed
    // over once no point is left (InvHeavy); __INT_MAX__: never.
    const bool handOver = (QK != QK_CC) && solveSteps > 0;
    int budget = __INT_MAX__;

    auto start_quad = [&](double x) {
        dc.x = x;
        qa = 0.0;
        qb = 1.0;
        res = 0.0;
        level = 0;
        sp = 0;
        k = (QK == QK_AS) ? -2 : -1;
    };

    while (true) {
        // Hand-over (InvHeavy), only once no point is left to take: before
        // that a lane's long solve costs the wavefront nothing it could spend
        // elsewhere.  `exhausted` is the same in every lane.
        if (handOver && exhausted) {
            if (busy && budget < 0) {
                const InvHeavy H = *Hd;  // read here only, not held through the loop
                const unsigned long long slot = atomicAdd(H.count, 1ull);
                if (slot < (unsigned long long)H.cap) {
                    H.list[slot] = n;  // k_inverse_heavy writes this root
                    busy = false;
                } else {
                    budget = __INT_MAX__;  // the list is full: the solve stays here
                }
            }
        }
        // Refill lanes without a point.
        while (true) {
            const bool need = !busy && !finished;
            const unsigned long long mask = __ballot(need);
            if (mask == 0) break;
            const int cnt = __popcll(mask);
            const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
            const long avail = qEnd - qNext;
            long base = 0;
            if (!exhausted && cnt > avail) {
                const int first = __builtin_ctzll(mask);
                unsigned long long b = 0;
                if (lane == first) b = atomicAdd(counter, (unsigned long long)INV_CHUNK);
                const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, first);
                const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), first);
                base = (long)(((unsigned long long)hi << 32) | lo);
            }
            if (need) {
                long idx = N;
                if (rank < avail) idx = qNext + rank;
                else if (!exhausted) idx = base + (rank - avail);
                bool skip = false;
                if (idx >= N) {
                    finished = true;
                } else if constexpr (BATCH) {
                    const InvSeg& S = segs[chunkSeg[idx / INV_CHUNK]];
                    const long local = idx - S.base;
                    if (local >= S.N) {
                        skip = true;  // padding after the map's points: stays without a point
                    } else {
                        op = S.o + local * S.os;
                        op2 = S.o2 ? S.o2 + local * S.os2 : nullptr;
                        dc.p = S.p;
                        dc.status = S.status;
                    }
                }
                if (idx < N && !skip) {
                    n = idx;
                    x0 = rec[(PM + 2) * N + n];
                    if (isnan(x0)) {
                        // stays without a point: refilled on the next pass
                        if constexpr (BATCH) {
                            *op = NAN;
                            if (op2) *op2 = NAN;
                        } else {
                            out[n * os] = NAN;
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q <= PM; ++q) dc.cd[q] = rec[q * N + n];
                        f0 = rec[(PM + 1) * N + n];
                        yd = rec[(PM + 3) * N + n];
                        busy = true;
                        phase = IP_START;
                        budget = handOver ? solveSteps : __INT_MAX__;
                        start_quad(x0);
                    }
                }
            }
            if (!exhausted && cnt > avail) {
                qNext = base + (cnt - avail);
                qEnd = base + INV_CHUNK;
                if (base >= N) exhausted = true;
            } else {
                qNext += cnt;
            }
        }
        if (__ballot(busy) == 0) break;
        if (!busy) continue;
        if (handOver) --budget;

        // One step of this lane's quadrature: the integrand at every node of
        // its current interval (E evaluations, independent of each other),
        // then the interval's acceptance test.  Starting a quadrature takes
        // one step for F(0) and F(1) (and, for Simpson, one for F(1/2)).
        bool qdone = false;
        if constexpr (QK == QK_CC) {
            // The fixed rule: one step evaluates every node (integrate_cc).
#pragma clang fp contract(off)  // sums and node positions as integrate_cc forms them
            double r = 0.0;
            if (Q.big) {
                const auto bw = ro(Q.big) + 2 * Q.nc;
                for (int i = 0; i < Q.nf; ++i) {
                    const double t = 0.5 * (1.0 + 0.0 + (1.0 - 0.0) * bw[Q.nf + i]);
                    r += (0.5 * (1.0 - 0.0) * bw[i]) * eval_node<BK, PK, FL_VAL, PM>(dc, t).F;
                }
            } else {
#pragma unroll
                for (int i = 0; i < MAXNF; ++i) {
                    if (i >= Q.nf) break;
                    const double t = 0.5 * (1.0 + 0.0 + (1.0 - 0.0) * Q.fp[i]);
                    r += (0.5 * (1.0 - 0.0) * Q.fw[i]) * eval_node<BK, PK, FL_VAL, PM>(dc, t).F;
                }
            }
            res = r;
            qdone = true;
        } else if constexpr (QK == QK_AS) {
#pragma clang fp contract(off)  // as in integrate_as
            const double m = 0.5 * (qa + qb);
            const double t0 = (k == -2) ? 0.0 : (k == -1) ? 0.5 : 0.5 * (qa + m);
            const double t1 = (k == -2) ? 1.0 : (k == -1) ? 0.5 : 0.5 * (m + qb);
            const double F0 = eval_node<BK, PK, FL_VAL, PM>(dc, t0).F;
            const double F1 = eval_node<BK, PK, FL_VAL, PM>(dc, t1).F;
            if (k == -2) {
                fa = F0;
                fb = F1;
                k = -1;
            } else if (k == -1) {
                fm = F0;
                k = 0;
            } else {
                const double a = qa, b = qb;
                const double flm = F0, frm = F1;
                const double coarse_ = div6(b - a) * (fa + 4.0 * fm + fb);
                const double wl = div6(m - a), wr = div6(b - m);
                double fine_ = wl * (fa + 4.0 * flm + fm);
                fine_ += wr * (fm + 4.0 * frm + fb);
                const double err = fabs(fine_ - coarse_);
                const double tol = fmax(Q.relTol * fabs(coarse_), Q.absTol);
                if (nonfinite_err(err)) {
                    res += fine_;  // non-finite already: ends the quadrature (integrate_as)
                    qdone = true;
                } else if (((err < tol) || (level == Q.maxSub - 1)) && (level >= Q.minSub)) {
                    res += fine_;
                    if (sp == 0) {
                        qdone = true;
                    } else {
                        // the next interval starts where this one ends
                        qa = b;
                        fa = fb;
                        level = top.level;
                        qb = qa + ldexp(1.0, -level);
                        fm = top.fm;
                        fb = top.fb;
                        if (--sp > 0) {
                            top = stk[sp - 1];
                            topInMem = true;
                        }
                    }
                } else {
                    if (sp > 0 && !topInMem) stk[sp - 1] = top;
                    top.level = level + 1;
                    top.fm = frm;
                    top.fb = fb;
                    topInMem = false;
                    ++sp;
                    qb = m;
                    fb = fm;
                    fm = flm;
                    ++level;
                }
            }
        } else {
#pragma clang fp contract(off)  // as in integrate_acc
            constexpr int NF = (QK == QK_ACC5) ? 5 : MAXNF;
            const double mid = 0.5 * (qa + qb);
            const double scale = 0.5 * (qb - qa);
            const bool start = k < 0;
            double coarse_ = 0.0, fine_ = 0.0;
            if (QK == QK_ACC && Q.big) {
                // A rule in device memory: the nodes in the order of the
                // tables above, each added to the sums as it is evaluated.
                if (start) {
                    fa = eval_node<BK, PK, FL_VAL, PM>(dc, 0.0).F;
                    fb = eval_node<BK, PK, FL_VAL, PM>(dc, 1.0).F;
                } else {
                    const int nc_ = Q.nc;
                    const auto bcw = ro(Q.big);
                    const auto bcp = bcw + nc_;
                    const auto bfw = bcp + nc_;
                    const auto bfp = bfw + nf;
                    fm = eval_node<BK, PK, FL_VAL, PM>(dc, mid).F;
                    coarse_ = scale * bcw[0] * fa;
                    coarse_ += scale * bcw[cMid] * fm;
                    coarse_ += scale * bcw[cRight] * fb;
                    fine_ = scale * bfw[0] * fa;
                    fine_ += scale * bfw[fMid] * fm;
                    fine_ += scale * bfw[fRight] * fb;
                    for (int i = 1; i < cMid; ++i) {
                        const double F = eval_node<BK, PK, FL_VAL, PM>(dc, mid + scale * bcp[i]).F;
                        coarse_ += scale * bcw[i] * F;
                        fine_ += scale * bfw[2 * i] * F;
                    }
                    for (int ii = cMid + 1; ii < cRight; ++ii) {
                        const double F = eval_node<BK, PK, FL_VAL, PM>(dc, mid + scale * bcp[ii]).F;
                        coarse_ += scale * bcw[ii] * F;
                        fine_ += scale * bfw[2 * ii] * F;
                    }
                    for (int i = 1; i < nf; i += 2)
                        fine_ += scale * bfw[i] * eval_node<BK, PK, FL_VAL, PM>(dc, mid + scale * bfp[i]).F;
                }
            } else {
                double Fv[NF];
#pragma unroll
                for (int j = 0; j < NF; ++j) {
                    if (j >= nNodes) break;
                    const double t = start ? (j == 0 ? 0.0 : 1.0) : (j == 0 ? mid : mid + scale * nPos[j]);
                    Fv[j] = (start && j > 1) ? 0.0 : eval_node<BK, PK, FL_VAL, PM>(dc, t).F;
                }
                if (start) {
                    fa = Fv[0];
                    fb = Fv[1];
                } else {
                    fm = Fv[0];
                    coarse_ = scale * sCw[0] * fa;
                    coarse_ += scale * sCw[cMid] * fm;
                    coarse_ += scale * sCw[cRight] * fb;
                    fine_ = scale * sFw[0] * fa;
                    fine_ += scale * sFw[fMid] * fm;
                    fine_ += scale * sFw[fRight] * fb;
#pragma unroll
                    for (int j = 1; j < NF; ++j) {
                        if (j >= nNodes) break;
                        const int ci = nCw[j];
                        if (ci >= 0) coarse_ += scale * sCw[ci] * Fv[j];
                        fine_ += scale * sFw[nFw[j]] * Fv[j];
                    }
                }
            }
            if (start) {
                k = 0;
            } else {
                const double err = fabs(fine_ - coarse_);
                const double tol = fmax(Q.relTol * fabs(coarse_), Q.absTol);
                if (nonfinite_err(err)) {
                    res += fine_;  // as above
                    qdone = true;
                } else if (((err < tol) || (level == Q.maxSub - 1)) && (level >= Q.minSub)) {
                    res += fine_;
                    if (sp == 0) {
                        qdone = true;
                    } else {
                        qa = qb;
                        fa = fb;
                        level = top.level;
                        qb = qa + ldexp(1.0, -level);
                        fb = top.fb;
                        if (--sp > 0) {
                            top = stk[sp - 1];
                            topInMem = true;
                        }
                    }
                } else {
                    if (sp > 0 && !topInMem) stk[sp - 1] = top;
                    top.level = level + 1;
                    top.fb = fb;
                    topInMem = false;
                    ++sp;
                    qb = mid;
                    fb = fm;
                    ++level;
                }
            }
        }
        if (!qdone) continue;

        // T at the current iterate; advance the bracket search or ITP.  A
        // non-finite T ends the solve with NaN: with finite inputs it comes
        // from an infinite g, where the reference fails.
        const double y = res + f0;
        bool fail = nonfinite_err(fabs(y)), done = false;
        double root = 0.0;
        if (fail) {
        } else if (phase == IP_START) {
            // InverseSingleBracket and FindBracket evaluate T(x0) three times;
            // the values are identical.
            xlb = xub = x0;
            ylb = yub = y;
            step = 1.0;
            bi = 0;
            if (ylb > yd) {
                phase = IP_DOWN;
                xlb = xub - step;
                start_quad(xlb);
            } else {
                phase = IP_UP;
                xub = xlb + step;
                start_quad(xub);
            }
        } else if (phase == IP_DOWN || phase == IP_UP) {
            bool bracketed = false;
            if (phase == IP_DOWN) ylb = y;
            else yub = y;
            if ((fabs((yub - ylb) / (xub - xlb)) < 1e-12) && ((xub - xlb) > 10)) {
                fail = true;
            } else if (phase == IP_DOWN && ylb > yd) {
                yub = ylb;
                xub = xlb;
                step *= 2.0;
                if (++bi >= 128) fail = true;
                else {
                    xlb = xub - step;
                    start_quad(xlb);
                }
            } else if (phase == IP_UP && yub < yd) {
                ylb = yub;
                xlb = xub;
                step *= 2.0;
                if (++bi >= 128) fail = true;
                else {
                    xub = xlb + step;
                    start_quad(xub);
                }
            } else {
                bracketed = true;
            }
            if (bracketed) {
                if ((ylb > yd) || (yub < yd)) {
                    fail = true;
                } else {
                    phase = IP_ITP;
                    it = 0;
                    nhalf = itp_nhalf(xlb, xub, xtol);
                    xc = itp_point(xlb, xub, ylb, yub, nhalf, it, xtol);
                    start_quad(xc);
                }
            }
        } else {
            const double yc = y;
            if (fabs(yc - yd) < ytol) {
                done = true;
                root = xc;
            } else {
                if (yc > yd) {
                    xub = xc;
                    yub = yc;
                } else {
                    xlb = xc;
                    ylb = yc;
                }
                if (((xub - xlb) < xtol) || ((yub - ylb) < ytol) || (unsigned)(it + 1) >= ITP_MAX_ITS) {
                    done = true;
                    root = 0.5 * (xub + xlb);
                } else {
                    ++it;
                    xc = itp_point(xlb, xub, ylb, yub, nhalf, it, xtol);
                    start_quad(xc);
                }
            }
        }
        if (fail || done) {
            const double v = fail ? NAN : root;
            if constexpr (BATCH) {
                *op = v;
                if (op2) *op2 = v;
            } else {
                out[n * os] = v;
            }
            busy = false;
        }
    }
}

}  // namespace mpd
