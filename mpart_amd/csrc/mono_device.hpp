// mono_device.hpp -- device building blocks of the monotone-component path.
//
// Lane-per-point design.  For one point x and one component of input dimension
// k+1 the expansion f(x_0..x_k) = sum_j w_j prod_e phi_{d_e,o_e}(x_{d_e}) * phi_q(x_k)
// is regrouped by the order q of the last coordinate:
//     f(x_<k, s) = sum_q C_q phi~_q(s),   C_q = sum_{j: q_j = q} w_j P_j(x_<k),
// where P_j is the product of the term's off-diagonal factors and phi~_0 = 1
// (a term without a last-coordinate factor).  Every integrand evaluation of
//     T(x) = f(x_<k, 0) + int_0^1 x_k g(d_k f(x_<k, t x_k)) dt
// then costs O(p) instead of O(#terms), and the coefficient adjoint factorizes:
//     dT/dw_j = P_j (phi~_q(0) + S_q),   S_q = int_0^1 x_k g'(d_k f) phi~'_q(t x_k) dt.
// The S_q are accumulated at accepted leaves of the same adaptive tree the
// forward integral uses (the reference's error estimate looks only at the
// first integrand component, MParT/Quadrature.h:448-451, so its vector
// integrals share the scalar tree).
//
// The basis, bijector, quadrature traversal and acceptance rule follow
// MParT/OrthogonalPolynomial.h, MParT/HermiteFunction.h, MParT/PositiveBijectors.h
// and MParT/Quadrature.h:553-676 (adaptive Simpson), 819-968 (adaptive
// Clenshaw-Curtis).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#else
typedef __hip_internal::uint32_t uint32_t;
#endif

// Helpers usable from host code as well (a host harness runs the same
// per-point arithmetic on the CPU); the wavefront primitives stay device-only.
#define MPD_HD __host__ __device__ __forceinline__

// Read-only data of a launch (term records, groups, components, basis-slot
// maps, coefficients) is read through the constant address space on the
// device: a wavefront-uniform address then becomes a scalar load through the
// scalar cache instead of one vector load per lane.  None of it changes while
// a kernel runs.
#if defined(__HIP_DEVICE_COMPILE__)
#define MPD_RO_AS __attribute__((address_space(4)))
#else
#define MPD_RO_AS
#endif

namespace mpd {

template <class T>
MPD_HD const MPD_RO_AS T* ro(const T* p) {
    return (const MPD_RO_AS T*)p;
}

constexpr int MAXP = 12;        // largest last-coordinate degree held in registers
constexpr int MAXND = 8;        // largest factor count with an unrolled term-record loop (more: run-time loop)
constexpr int MAXSUB = 32;      // largest adaptive depth (quadMaxSub)
constexpr int MAXNF = 17;       // largest nested Clenshaw-Curtis fine rule
constexpr int MAXORD = 32;      // largest off-diagonal degree in the basis table
constexpr uint32_t DUMMY = 0xFFFFFFFFu;

enum BasisKind { BK_PH = 0, BK_HF = 1, BK_PHYS = 2 };
enum PosKind { PK_EXP = 0, PK_SP = 1 };
enum QuadKind { QK_AS = 0, QK_ACC5 = 1, QK_ACC = 2, QK_CC = 3 };

// Orthogonal-polynomial bases p_k = (a x + 0) p_{k-1} - c_k p_{k-2} with
// p_0 = 1, p_1 = a x (OrthogonalPolynomial.h:25-156): ProbabilistHermite
// a = 1, c_k = k - 1 (301-306); PhysicistHermite a = 2, c_k = 2 (k - 1)
// (319-324).  p_k' = a k p_{k-1} and p_k'' = a^2 k (k-1) p_{k-2} for both.
template <int BK>
constexpr double poly_a() {
    return BK == BK_PHYS ? 2.0 : 1.0;
}

// LinearizedBasis (MParT/LinearizedBasis.h): outside [lb, ub] every 1-D basis
// function continues linearly with its value and slope at the nearer bound.
// The bounds are run-time data (MapDev::lin/lb/ub, DiagCoef::lb/ub); kernels
// generated for a map without bounds define MPD_LINEARIZED 0 and drop the
// checks.
#ifndef MPD_LINEARIZED
#define MPD_LINEARIZED 1
#endif
// Integrand flavours: value only; coefficient adjoint; input adjoint (and the
// x_k-derivative U); discrete-derivative mode (U and its coefficient adjoint).
enum Flavor { FL_VAL = 0, FL_COEF = 1, FL_INPUT = 2, FL_DCOEF = 3 };

// Nested Clenshaw-Curtis pair and adaptive parameters; the fixed
// Clenshaw-Curtis rule (QK_CC) uses fw / fp with nf points.
struct QuadDev {
    double absTol, relTol;
    int maxSub, minSub;
    int nc, nf;  // coarse / fine rule sizes (2^L+1, 2^(L+1)+1)
    double cw[MAXNF], cp[MAXNF], fw[MAXNF], fp[MAXNF];
    // Rules of more than MAXNF points live in device memory instead: cw, cp
    // (nc each; nc = 0 for the fixed rule), then fw, fp (nf each).  NULL when
    // the arrays above hold the rule.
    const double* big;
};

// One monotone component of a map.
struct CompDev {
    int dim;       // input dimension of the component (k+1)
    int nterms;
    int coeffOff;  // offset of its coefficients in the map's coefficient vector
    int outRow;    // output row of the component in the map
    int p;         // largest order of the last coordinate
    int groupOff;  // index of its q = 0 group in the group table
    int pad0, pad1;
};

// Terms of one component with a given last-coordinate order q.  The term
// records [start, start+countPad) are padded with dummies to a multiple of 16.
struct GroupDev {
    int start, count, countPad, pad;
};

// Everything a kernel needs to know about a map (passed by value).
struct MapDev {
    const CompDev* comps;
    const GroupDev* groups;
    const uint32_t* terms;   // records of trec uint32: [coeff index, slot_0 .. slot_{nd-1}]
    const int* slotDim;      // basis-table slot -> coordinate (slot 0 holds 1.0)
    const int* dimSlot;      // coordinate -> first slot (order 1)
    const int* dimDeg;       // coordinate -> largest off-diagonal order in the table
    int trec, nd;            // record length and number of factor slots (trec = nd + 1)
    int nslots;              // table slots including slot 0
    int D;                   // input dimension of the map
    int ncomp, numCoeffs;
    int normalize;           // ProbabilistHermite basisNorm
    int pmax;                // largest p over components
    double nugget;
    double norm[MAXORD + 1];   // ProbabilistHermite normalization sqrt(sqrt(2 pi) k!)
    double rho[MAXP + 1];      // phi~'_q = rho_q * B'_q (folded normalization)
    double phi0[MAXP + 1];     // phi~_q(0), with phi~_0 = 1
    double hfA[MAXORD + 1];    // Hermite-function recurrence: sqrt(0.5 (i-1))
    double hfB[MAXORD + 1];    //                              sqrt(0.5 i)
    double hfS[MAXORD + 1];    // (2^i i!)^(-1/2)
    double dfac[MAXORD + 1];   // orthogonal polynomials: d phi_o / dx = dfac[o] phi_{o-1} (phi_0 = 1)
    double hfPi;               // pi^(-1/4)
    double hfSqrt2;            // sqrt(2)
    int discrete;              // contDeriv = false: log-determinants use the discrete derivative
    int lin;                   // LinearizedBasis bounds in use
    double lb, ub;             // its bounds (-inf / +inf when unused)
    QuadDev quad;
    int* status;               // set to 1 when g is infinite at a quadrature node (NULL: not recorded)
};

// ---------------------------------------------------------------------------
// Positive bijector g, g', g'' (MParT/PositiveBijectors.h:12-57).
template <int PK>
struct PosF {
    double g, gp;
};

// 1/d for d in [1, 2.5]: hardware reciprocal estimate refined by two Newton
// steps (within an ulp of the correctly rounded quotient; no scaling needed
// on this range).
MPD_HD double recip_1_2(double d) {
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
#else
    return 1.0 / d;
#endif
}

// x / 6, correctly rounded, in three operations.  q = RN(x RN(1/6)) is within
// 1.5 ulp of x/6, so the remainder r = x - 6q is exact (a multiple of 2 ulp(q)
// below 20 ulp(q)) and q + r RN(1/6) differs from x/6 by under 2^-50 ulp.  For
// normal x, x/6 is an integer multiple of ulp(x/6)/3, hence at least ulp/6 from
// every rounding midpoint, so the fused correction rounds to RN(x/6) exactly.
MPD_HD double div6(double x) {
    const double c = 1.0 / 6.0;
    const double q = x * c;
    return fma(fma(-q, 6.0, x), c, q);
}

// Logarithms by the method of fdlibm's __ieee754_log: x = 2^k m with m in
// [sqrt(1/2), sqrt(2)), f = m - 1 (exact), s = f / (2 + f) and
//   log(1 + f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R)),  R = sum_{j>=1} 2 s^(2j) / (2j + 1),
// with the series cut after ten terms (|s| <= 0.1716: the remainder is below
// 1e-18 of the result) and ln 2 split into a head with 12 trailing zero bits
// (k * head is exact) and a tail.  Within an ulp of the correctly rounded
// logarithm, at about a quarter of the instructions of the general
// double-double device log.
__device__ __forceinline__ double log_reduced(double f, double k) {
    const double hfsq = 0.5 * f * f;
    const double s = f * recip_1_2(2.0 + f);
    const double z = s * s;
    double R = 0.09523809523809523;  // 2/21
    R = fma(R, z, 0.10526315789473684);  // 2/19
    R = fma(R, z, 0.11764705882352941);  // 2/17
    R = fma(R, z, 0.13333333333333333);  // 2/15
    R = fma(R, z, 0.15384615384615385);  // 2/13
    R = fma(R, z, 0.18181818181818182);  // 2/11
    R = fma(R, z, 0.2222222222222222);   // 2/9
    R = fma(R, z, 0.2857142857142857);   // 2/7
    R = fma(R, z, 0.4);                  // 2/5
    R = fma(R, z, 0.6666666666666666);   // 2/3
    R *= z;
    const double ln2Hi = 0.693147180559663, ln2Lo = 2.8235290563031577e-13;
    return k * ln2Hi - ((hfsq - (s * (hfsq + R) + k * ln2Lo)) - f);
}

// log(x) for x in [1, 2] (the SoftPlus argument 1 + exp(-|z|)).
MPD_HD double log_1_2(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    const bool up = x > 1.4142135623730951;
    return log_reduced((up ? 0.5 * x : x) - 1.0, up ? 1.0 : 0.0);
#else
    return log(x);
#endif
}

// log(x) for any x: the reduction above for positive normal numbers, the
// general device log for zero, subnormals, infinities, NaN and negatives.
MPD_HD double log_pos(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308)) return log(x);
    int e;
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < 0.7071067811865476) {
        m *= 2.0;
        e -= 1;
    }
    return log_reduced(m - 1.0, (double)e);
#else
    return log(x);
#endif
}

// g'' as well (SecondDerivative, MParT/PositiveBijectors.h:25-28, 48-50):
// SoftPlus g'' = g' (1 - g'), Exp g'' = exp.
template <int PK>
struct PosF2 {
    double g, gp, gpp;
};

template <int PK>
MPD_HD PosF<PK> posfunc(double z);

template <int PK>
MPD_HD PosF2<PK> posfunc2(double z) {
    const PosF<PK> a = posfunc<PK>(z);
    PosF2<PK> r;
    r.g = a.g;
    r.gp = a.gp;
    if constexpr (PK == PK_EXP) r.gpp = a.g;
    else r.gpp = a.gp * (1.0 - a.gp);
    return r;
}

template <int PK>
MPD_HD PosF<PK> posfunc(double z) {
    PosF<PK> r;
    if constexpr (PK == PK_EXP) {
        r.g = exp(z);
        r.gp = r.g;
    } else {
        // log(1 + exp(-|z|)) + max(z, 0); derivative shares e = exp(-|z|):
        // g' = e / (1 + e) for z < 0, 1 / (1 + e) otherwise.  g' only enters
        // derivatives, so its quotient uses the refined reciprocal.
        double e = exp(-fabs(z));
        r.g = log_1_2(1.0 + e) + fmax(z, 0.0);
        const double q = recip_1_2(1.0 + e);
        r.gp = (z < 0) ? e * q : q;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Diagonal (last-coordinate) basis with folded normalization.
//
// ProbabilistHermite: phi'_q = He'_q / n_q = q He_{q-1} / n_q, so with
// Cd_q = C_q / n_q:   d_k f(y) = sum_{r=0}^{p-1} (r+1) Cd_{r+1} He_r(y).
// Hermite functions (q >= 2, k = q-2): phi'_q = pi^-1/4 s_k e^{-y^2/2} (2k H_{k-1} - y H_k)
// with physicist H; phi'_1 = 1 and phi'_0 = 0.
template <int PM>
struct DiagCoef {
    double cd[PM + 1];  // C_q * rho_q
    int p;
    double x;   // x_k
    double nug; // nugget added to g inside the integrand
    int lin;        // LinearizedBasis in use (the same for every lane)
    double lb, ub;  // its bounds
    int* status;    // MapDev::status
};

// The argument of the diagonal basis derivatives: a linearized basis has the
// slope of its nearer bound outside [lb, ub] (NaN stays NaN).
template <int PM>
MPD_HD double lin_arg(const DiagCoef<PM>& dc, double y) {
#if MPD_LINEARIZED
    if (!dc.lin) return y;
    return (y < dc.lb) ? dc.lb : ((y > dc.ub) ? dc.ub : y);
#else
    return y;
#endif
}
template <int PM>
MPD_HD bool lin_outside(const DiagCoef<PM>& dc, double y) {
#if MPD_LINEARIZED
    return dc.lin && ((y < dc.lb) || (y > dc.ub));
#else
    return false;
#endif
}

// d_k f at y (and optionally the second derivative).
template <int BK, bool SECOND, int PM>
MPD_HD double diag_eval(const DiagCoef<PM>& dc, double y0, double* z2out) {
    const int p = dc.p;
    const double y = lin_arg(dc, y0);
    if constexpr (BK != BK_HF) {
        // p_0 = 1, p_1 = a y, p_r = a y p_{r-1} - a (r-1) p_{r-2};  p'_{r+1} = a (r+1) p_r
        constexpr double A = poly_a<BK>();
        double hm2 = 0.0, hm1 = 1.0;  // p_{r-2}, p_{r-1}
        double z = 0.0, z2 = 0.0;
#pragma unroll
        for (int r = 0; r < PM; ++r) {
            if (r >= p) break;
            double hr = (r == 0) ? 1.0 : ((r == 1) ? A * y : (A * y) * hm1 - (A * (r - 1.0)) * hm2);
            z += (A * (r + 1.0)) * dc.cd[r + 1] * hr;
            if constexpr (SECOND) {
                if (r + 2 <= p) z2 += (A * A * (r + 2.0) * (r + 1.0)) * dc.cd[r + 2] * hr;
            }
            if (r >= 1) {
                hm2 = hm1;
                hm1 = hr;
            } else {
                hm1 = hr;
            }
        }
        if constexpr (SECOND) *z2out = lin_outside(dc, y0) ? 0.0 : z2;
        return z;
    } else {
        double z = (p >= 1) ? dc.cd[1] : 0.0;
        if (p < 2) {
            if constexpr (SECOND) *z2out = 0.0;
            return z;
        }
        double ex = exp(-0.5 * y * y);
        // H_0 = 1, H_1 = 2y, H_k = 2y H_{k-1} - 2(k-1) H_{k-2}
        double hkm1 = 0.0, hkm2 = 0.0, s = 0.0, s2 = 0.0;
#pragma unroll
        for (int k = 0; k < PM - 1; ++k) {
            if (k + 2 > p) break;
            double hk = (k == 0) ? 1.0 : ((k == 1) ? 2.0 * y : 2.0 * y * hkm1 - 2.0 * (k - 1.0) * hkm2);
            double dk = (k == 0) ? 0.0 : 2.0 * k * hkm1;  // H'_k = 2k H_{k-1}
            s += dc.cd[k + 2] * (dk - y * hk);
            if constexpr (SECOND) s2 += dc.cd[k + 2] * (y * y - (2.0 * k + 1.0)) * hk;
            hkm2 = hkm1;
            hkm1 = hk;
        }
        if constexpr (SECOND) *z2out = lin_outside(dc, y0) ? 0.0 : ex * s2;
        return z + ex * s;
    }
}

// B'_q(y) for q = 0..p (raw derivative basis; phi~'_q = rho_q B'_q).
template <int BK, int PM>
MPD_HD void diag_basis(const DiagCoef<PM>& dc, double y0, double* bq) {
    const int p = dc.p;
    const double y = lin_arg(dc, y0);
    bq[0] = 0.0;
    if constexpr (BK != BK_HF) {
        constexpr double A = poly_a<BK>();
        double hm2 = 0.0, hm1 = 1.0;
#pragma unroll
        for (int r = 0; r < PM; ++r) {
            if (r >= p) break;
            double hr = (r == 0) ? 1.0 : ((r == 1) ? A * y : (A * y) * hm1 - (A * (r - 1.0)) * hm2);
            bq[r + 1] = (A * (r + 1.0)) * hr;
            if (r >= 1) {
                hm2 = hm1;
                hm1 = hr;
            } else {
                hm1 = hr;
            }
        }
    } else {
        if (p >= 1) bq[1] = 1.0;
        if (p < 2) return;
        double ex = exp(-0.5 * y * y);
        double hkm1 = 0.0, hkm2 = 0.0;
#pragma unroll
        for (int k = 0; k < PM - 1; ++k) {
            if (k + 2 > p) break;
            double hk = (k == 0) ? 1.0 : ((k == 1) ? 2.0 * y : 2.0 * y * hkm1 - 2.0 * (k - 1.0) * hkm2);
            double dk = (k == 0) ? 0.0 : 2.0 * k * hkm1;
            bq[k + 2] = ex * (dk - y * hk);
            hkm2 = hkm1;
            hkm1 = hk;
        }
    }
}

// B''_q(y) (raw second-derivative basis, same folding as B'_q); zero outside
// the bounds of a linearized basis.
template <int BK, int PM>
MPD_HD void diag_basis2(const DiagCoef<PM>& dc, double y, double* bq) {
    const int p = dc.p;
    bq[0] = 0.0;
    if (p >= 1) bq[1] = 0.0;
    if (lin_outside(dc, y)) {
#pragma unroll
        for (int q = 2; q <= PM; ++q) bq[q] = 0.0;
        return;
    }
    if constexpr (BK != BK_HF) {
        // p''_q = a^2 q (q-1) p_{q-2}
        constexpr double A = poly_a<BK>();
        double hm2 = 0.0, hm1 = 1.0;
#pragma unroll
        for (int r = 0; r < PM - 1; ++r) {
            if (r + 2 > p) break;
            double hr = (r == 0) ? 1.0 : ((r == 1) ? A * y : (A * y) * hm1 - (A * (r - 1.0)) * hm2);
            bq[r + 2] = (A * A * (r + 2.0) * (r + 1.0)) * hr;
            if (r >= 1) {
                hm2 = hm1;
                hm1 = hr;
            } else {
                hm1 = hr;
            }
        }
    } else {
        if (p < 2) return;
        double ex = exp(-0.5 * y * y);
        double hkm1 = 0.0, hkm2 = 0.0;
#pragma unroll
        for (int k = 0; k < PM - 1; ++k) {
            if (k + 2 > p) break;
            double hk = (k == 0) ? 1.0 : ((k == 1) ? 2.0 * y : 2.0 * y * hkm1 - 2.0 * (k - 1.0) * hkm2);
            bq[k + 2] = ex * (y * y - (2.0 * k + 1.0)) * hk;
            hkm2 = hkm1;
            hkm1 = hk;
        }
    }
}

// Normalized value and first derivative of the order-o basis function
// (o >= 1) at x, with the arithmetic of fill_values_derivs (used at the bound
// of a linearized basis).
template <int BK>
MPD_HD void basis_vd(const MapDev& M, int o, double x, double& v, double& d) {
    if constexpr (BK != BK_HF) {
        constexpr double A = poly_a<BK>();
        double vm2 = 1.0, vm1 = A * x, dm2 = 0.0, dm1 = A;
        for (int k = 2; k <= o; ++k) {
            const double vv = (A * x + 0.0) * vm1 - (A * (k - 1.0)) * vm2;
            const double dd = A * vm1 + (A * x + 0.0) * dm1 - (A * (k - 1.0)) * dm2;
            vm2 = vm1;
            vm1 = vv;
            dm2 = dm1;
            dm1 = dd;
        }
        if (M.normalize) {
            v = vm1 / M.norm[o];
            d = dm1 / M.norm[o];
        } else {
            v = vm1;
            d = dm1;
        }
    } else {
        if (o == 1) {
            v = x;
            d = 1.0;
            return;
        }
        const int k = o - 2;
        const double base = M.hfPi * exp(-0.5 * x * x);
        double hm1 = 0.0, hm2 = 0.0, dm1 = 0.0, dm2 = 0.0, h = 1.0, dh = 0.0;
        for (int j = 0; j <= k; ++j) {
            if (j == 0) {
                h = 1.0;
                dh = 0.0;
            } else if (j == 1) {
                h = 2.0 * x;
                dh = 2.0;
            } else {
                h = (2.0 * x + 0.0) * hm1 - 2.0 * (j - 1.0) * hm2;
                dh = 2.0 * hm1 + (2.0 * x + 0.0) * dm1 - 2.0 * (j - 1.0) * dm2;
            }
            hm2 = hm1;
            hm1 = h;
            dm2 = dm1;
            dm1 = dh;
        }
        const double scale = (k == 0) ? base : base * M.hfS[k];
        v = h * scale;
        d = (dh - x * h) * scale;
    }
}

// Whether x lies outside the bounds of a linearized basis (and which bound).
MPD_HD bool lin_bound(const MapDev& M, double x, double& c) {
#if MPD_LINEARIZED
    if (M.lin && (x < M.lb || x > M.ub)) {
        c = (x < M.lb) ? M.lb : M.ub;
        return true;
    }
#endif
    (void)M;
    (void)x;
    (void)c;
    return false;
}

// ---------------------------------------------------------------------------
// Off-diagonal basis table.  Slot 0 holds 1.0; coordinate d occupies slots
// dimSlot[d] .. dimSlot[d]+dimDeg[d]-1 for orders 1..dimDeg[d].  The table of a
// lane lives at tab[slot * stride].
//
// Values: EvaluateAll (OrthogonalPolynomial.h:25-42, HermiteFunction.h:22-43).
// Values of orders 1..deg of one coordinate into out[o * stride].
template <int BK>
MPD_HD void values_1d(const MapDev& M, int deg, double x, double* out, long stride) {
    double c;
    if (lin_bound(M, x, c)) {
        // LinearizedBasis::EvaluateAll (LinearizedBasis.h:41-61)
        for (int o = 1; o <= deg; ++o) {
            double v, dv;
            basis_vd<BK>(M, o, c, v, dv);
            out[o * stride] = v + dv * (x - c);
        }
        return;
    }
    if constexpr (BK != BK_HF) {
        constexpr double A = poly_a<BK>();
        double vm2 = 1.0, vm1 = A * x;
        for (int o = 1; o <= deg; ++o) {
            double v = (o == 1) ? A * x : (A * x + 0.0) * vm1 - (A * (o - 1.0)) * vm2;
            if (o >= 2) {
                vm2 = vm1;
                vm1 = v;
            }
            out[o * stride] = M.normalize ? v / M.norm[o] : v;
        }
    } else {
        // phi_1 = x, phi_2 = pi^-1/4 exp(-x^2/2), phi_3 = sqrt(2) x phi_2,
        // phi_{i+2} = (x phi_{i+1} - sqrt((i-1)/2) phi_i) / sqrt(i/2)
        double pm1 = 0.0, pm2 = 0.0;
        for (int o = 1; o <= deg; ++o) {
            double v;
            if (o == 1) v = x;
            else if (o == 2) v = M.hfPi * exp(-0.5 * x * x);
            else if (o == 3) v = M.hfSqrt2 * x * pm1;
            else {
                int i = o - 2;
                v = (x * pm1 - M.hfA[i] * pm2) / M.hfB[i];
            }
            pm2 = pm1;
            pm1 = v;
            out[o * stride] = v;
        }
    }
}

template <int BK>
__host__ __device__ void fill_values(const MapDev& M, int ndims, const double* pt, long prs, double* tab, int stride) {
    tab[0] = 1.0;
    for (int d = 0; d < ndims; ++d) {
        const int deg = ro(M.dimDeg)[d];
        if (deg == 0) continue;
        values_1d<BK>(M, deg, pt[d * prs], tab + (long)ro(M.dimSlot)[d] * stride - stride, stride);
    }
}


// Off-diagonal basis values phi_1..phi_DEG at x into b[1..DEG] (b[0] = 1),
// the same arithmetic as fill_values; used by generated kernels that keep the
// values in registers.
template <int BK, int DEG>
MPD_HD void basis_vals(const MapDev& M, double x, double* b) {
    b[0] = 1.0;
    double c;
    if (lin_bound(M, x, c)) {
#pragma unroll
        for (int o = 1; o <= DEG; ++o) {
            double v, dv;
            basis_vd<BK>(M, o, c, v, dv);
            b[o] = v + dv * (x - c);
        }
        return;
    }
    if constexpr (BK != BK_HF) {
        constexpr double A = poly_a<BK>();
        double vm2 = 1.0, vm1 = A * x;
#pragma unroll
        for (int o = 1; o <= DEG; ++o) {
            double v = (o == 1) ? A * x : (A * x + 0.0) * vm1 - (A * (o - 1.0)) * vm2;
            if (o >= 2) {
                vm2 = vm1;
                vm1 = v;
            }
            b[o] = M.normalize ? v / M.norm[o] : v;
        }
    } else {
        double pm1 = 0.0, pm2 = 0.0;
#pragma unroll
        for (int o = 1; o <= DEG; ++o) {
            double v;
            if (o == 1) v = x;
            else if (o == 2) v = M.hfPi * exp(-0.5 * x * x);
            else if (o == 3) v = M.hfSqrt2 * x * pm1;
            else v = (x * pm1 - M.hfA[o - 2] * pm2) / M.hfB[o - 2];
            pm2 = pm1;
            pm1 = v;
            b[o] = v;
        }
    }
}

// Values and first derivatives as EvaluateDerivatives computes them
// (OrthogonalPolynomial.h:87-118, HermiteFunction.h:45-80) of orders 1..deg
// of one coordinate, into out[o * stride] and dout[o * stride].
template <int BK>
MPD_HD void values_derivs_1d(const MapDev& M, int deg, double x, double* out, double* dout, long stride) {
    double c;
    if (lin_bound(M, x, c)) {
        // LinearizedBasis::EvaluateDerivatives (LinearizedBasis.h:63-87)
        for (int o = 1; o <= deg; ++o) {
            double v, dv;
            basis_vd<BK>(M, o, c, v, dv);
            out[o * stride] = v + dv * (x - c);
            dout[o * stride] = dv;
        }
        return;
    }
    if constexpr (BK != BK_HF) {
        constexpr double A = poly_a<BK>();
        double vm2 = 1.0, vm1 = A * x, dm2 = 0.0, dm1 = A;
        for (int o = 1; o <= deg; ++o) {
            double v, dv;
            if (o == 1) {
                v = A * x;
                dv = A;
            } else {
                v = (A * x + 0.0) * vm1 - (A * (o - 1.0)) * vm2;
                dv = A * vm1 + (A * x + 0.0) * dm1 - (A * (o - 1.0)) * dm2;
                vm2 = vm1;
                vm1 = v;
                dm2 = dm1;
                dm1 = dv;
            }
            if (M.normalize) {
                double n = M.norm[o];
                v /= n;
                dv /= n;
            }
            out[o * stride] = v;
            dout[o * stride] = dv;
        }
    } else {
        // phi_1 = x (derivative 1); phi_{k+2}: physicist H_k and H'_k scaled
        // by pi^-1/4 e^{-x^2/2} (2^k k!)^-1/2, derivative (H'_k - x H_k) * scale.
        const double base = M.hfPi * exp(-0.5 * x * x);
        double hm1 = 0.0, hm2 = 0.0, dm1 = 0.0, dm2 = 0.0;
        for (int o = 1; o <= deg; ++o) {
            if (o == 1) {
                out[o * stride] = x;
                dout[o * stride] = 1.0;
                continue;
            }
            int k = o - 2;
            double h, dh;
            if (k == 0) {
                h = 1.0;
                dh = 0.0;
            } else if (k == 1) {
                h = 2.0 * x;
                dh = 2.0;
            } else {
                h = (2.0 * x + 0.0) * hm1 - 2.0 * (k - 1.0) * hm2;
                dh = 2.0 * hm1 + (2.0 * x + 0.0) * dm1 - 2.0 * (k - 1.0) * dm2;
            }
            hm2 = hm1;
            hm1 = h;
            dm2 = dm1;
            dm1 = dh;
            double scale = (k == 0) ? base : base * M.hfS[k];
            double dd = dh - x * h;
            out[o * stride] = h * scale;
            dout[o * stride] = dd * scale;
        }
    }
}

// Value and derivative tables of coordinates 0..ndims-1 (FillCache1 with
// DerivativeFlags::Input); used by the input-gradient paths.
template <int BK>
__host__ __device__ void fill_values_derivs(const MapDev& M, int ndims, const double* pt, long prs, double* tab, double* dtab,
                                   int stride) {
    tab[0] = 1.0;
    dtab[0] = 0.0;
    for (int d = 0; d < ndims; ++d) {
        const int deg = ro(M.dimDeg)[d];
        if (deg == 0) continue;
        values_derivs_1d<BK>(M, deg, pt[d * prs], tab + (long)ro(M.dimSlot)[d] * stride - stride,
                             dtab + (long)ro(M.dimSlot)[d] * stride - stride, stride);
    }
}

// The same values and derivatives into b[1..DEG], db[1..DEG] (b[0] = 1,
// db[0] = 0), for generated kernels that keep them in registers.
template <int BK, int DEG>
MPD_HD void basis_vals_derivs(const MapDev& M, double x, double* b, double* db) {
    b[0] = 1.0;
    db[0] = 0.0;
    values_derivs_1d<BK>(M, DEG, x, b, db, 1);
}

// ---------------------------------------------------------------------------
// Group passes over the term records of one component.

// Record length and off-diagonal product of a term record.  ND > 0 is the
// map's factor-slot count known at compile time (the loop unrolls); ND = 0
// reads it from the map, for terms with more factors than the unrolled cases.
template <int ND>
MPD_HD long rec_len(const MapDev& M) {
    if constexpr (ND > 0) return ND + 1;
    else return M.trec;
}
template <int ND, class RP>
MPD_HD double rec_product(const MapDev& M, RP rec, const double* tab, long stride) {
    double P = tab[(long)rec[1] * stride];
    if constexpr (ND > 0) {
#pragma unroll
        for (int e = 2; e <= ND; ++e) P *= tab[(long)rec[e] * stride];
    } else {
        for (int e = 2; e <= M.nd; ++e) P *= tab[(long)rec[e] * stride];
    }
    return P;
}

// C_q = sum_{j in group q} w_j P_j for q = 0..p.
template <int ND, int PM>
MPD_HD void coeff_groups_nd(const MapDev& M, const CompDev& c, const double* w, long ws,
                                                const double* tab, int stride, double* C) {
#pragma unroll
    for (int q = 0; q <= PM; ++q) {
        if (q > c.p) break;
        const GroupDev g = ro(M.groups)[c.groupOff + q];
        double acc = 0.0;
        for (int j = g.start; j < g.start + g.count; ++j) {
            const auto rec = ro(M.terms) + (long)j * rec_len<ND>(M);
            acc += ro(w)[(long)rec[0] * ws] * rec_product<ND>(M, rec, tab, stride);
        }
        C[q] = acc;
    }
}

template <int PM>
MPD_HD void coeff_groups(const MapDev& M, const CompDev& c, const double* w, long ws,
                                             const double* tab, int stride, double* C) {
    switch (M.nd) {
        case 1: coeff_groups_nd<1, PM>(M, c, w, ws, tab, stride, C); break;
        case 2: coeff_groups_nd<2, PM>(M, c, w, ws, tab, stride, C); break;
        case 3: coeff_groups_nd<3, PM>(M, c, w, ws, tab, stride, C); break;
        case 4: coeff_groups_nd<4, PM>(M, c, w, ws, tab, stride, C); break;
        case 5: coeff_groups_nd<5, PM>(M, c, w, ws, tab, stride, C); break;
        case 6: coeff_groups_nd<6, PM>(M, c, w, ws, tab, stride, C); break;
        case 7: coeff_groups_nd<7, PM>(M, c, w, ws, tab, stride, C); break;
        case 8: coeff_groups_nd<8, PM>(M, c, w, ws, tab, stride, C); break;
        default: coeff_groups_nd<0, PM>(M, c, w, ws, tab, stride, C); break;
    }
}

// Visit every (real) term: op(coeff index, P_j, q).
template <int ND, int PM, class Op>
MPD_HD void visit_terms_nd(const MapDev& M, const CompDev& c, const double* tab, int stride,
                                               Op& op) {
#pragma unroll
    for (int q = 0; q <= PM; ++q) {
        if (q > c.p) break;
        const GroupDev g = ro(M.groups)[c.groupOff + q];
        for (int j = g.start; j < g.start + g.count; ++j) {
            const auto rec = ro(M.terms) + (long)j * rec_len<ND>(M);
            op(rec[0], rec_product<ND>(M, rec, tab, stride), q);
        }
    }
}

template <int PM, class Op>
MPD_HD void visit_terms(const MapDev& M, const CompDev& c, const double* tab, int stride,
                                            Op& op) {
    switch (M.nd) {
        case 1: visit_terms_nd<1, PM>(M, c, tab, stride, op); break;
        case 2: visit_terms_nd<2, PM>(M, c, tab, stride, op); break;
        case 3: visit_terms_nd<3, PM>(M, c, tab, stride, op); break;
        case 4: visit_terms_nd<4, PM>(M, c, tab, stride, op); break;
        case 5: visit_terms_nd<5, PM>(M, c, tab, stride, op); break;
        case 6: visit_terms_nd<6, PM>(M, c, tab, stride, op); break;
        case 7: visit_terms_nd<7, PM>(M, c, tab, stride, op); break;
        case 8: visit_terms_nd<8, PM>(M, c, tab, stride, op); break;
        default: visit_terms_nd<0, PM>(M, c, tab, stride, op); break;
    }
}

// Diagonal coefficients of component c from its group sums.
template <int PM>
MPD_HD void make_diag(const MapDev& M, const CompDev& c, const double* C, double xk, double nug,
                                          DiagCoef<PM>& dc) {
    dc.p = c.p;
    dc.x = xk;
    dc.nug = nug;
    dc.lin = M.lin;
    dc.lb = M.lb;
    dc.ub = M.ub;
    dc.status = M.status;
#pragma unroll
    for (int q = 0; q <= PM; ++q) dc.cd[q] = (q <= c.p) ? C[q] * M.rho[q] : 0.0;
}

// f(x_<k, 0) = sum_q C_q phi~_q(0).
template <int PM>
MPD_HD double f_at_zero(const MapDev& M, const CompDev& c, const double* C) {
    double f = 0.0;
#pragma unroll
    for (int q = 0; q <= PM; ++q) {
        if (q > c.p) break;
        f += C[q] * M.phi0[q];
    }
    return f;
}

// ---------------------------------------------------------------------------
// Adaptive quadrature of F(t) = x (g(d_k f(x_<k, t x)) + nugget) on [0, 1]
// with optional adjoint accumulation at accepted leaves.
//
// FL_COEF accumulates S'_q = int h(t) B'_q(t x) dt with h = x g'(d_k f).
// FL_INPUT additionally accumulates U = int (g + nugget) + x t g'(d_k f) d^2_k f dt,
// the x_k-derivative of the integral (MonotoneIntegrand.h:193-202).
template <int FL>
struct NodeV;
template <>
struct NodeV<FL_VAL> {
    double F;
};
template <>
struct NodeV<FL_COEF> {
    double F, h;
};
template <>
struct NodeV<FL_INPUT> {
    double F, h, g1;
};
template <>
struct NodeV<FL_DCOEF> {
    double F, gp, gpp, g1;
};

// The reference checks g at every integrand evaluation and fails when it is
// infinite (MonotoneIntegrand.h:183-190: a runtime_error on the host, an
// abort on a device); here the node records it in the map's status word,
// which the host reads where a call is synchronous.
template <int PM>
MPD_HD void note_infinite_g(const DiagCoef<PM>& dc, double g) {
    if (__builtin_isinf(g) && dc.status) *dc.status = 1;
}

template <int BK, int PK, int FL, int PM>
MPD_HD NodeV<FL> eval_node(const DiagCoef<PM>& dc, double t) {
    double z = diag_eval<BK, false, PM>(dc, t * dc.x, nullptr);
    NodeV<FL> v;
    if constexpr (FL == FL_DCOEF) {
        const PosF2<PK> gv = posfunc2<PK>(z);
        note_infinite_g<PM>(dc, gv.g);
        v.F = dc.x * (gv.g + dc.nug);
        v.gp = gv.gp;
        v.gpp = gv.gpp;
        v.g1 = gv.g + dc.nug;
    } else {
        PosF<PK> gv = posfunc<PK>(z);
        note_infinite_g<PM>(dc, gv.g);
        v.F = dc.x * (gv.g + dc.nug);
        if constexpr (FL >= FL_COEF) v.h = dc.x * gv.gp;
        if constexpr (FL == FL_INPUT) v.g1 = gv.g + dc.nug;
    }
    return v;
}

template <int PM>
struct AdjAcc {
    double S[PM + 1];
    double U;
};

// Adds w * v(t) at one node to the adjoint accumulators.
template <int BK, int FL, int PM>
MPD_HD void adj_node(const DiagCoef<PM>& dc, double t, double w, const NodeV<FL>& v, AdjAcc<PM>& A) {
    if constexpr (FL == FL_DCOEF) {
        // Diagonal and MixedCoeff integrand modes (MonotoneIntegrand.h:148-165,
        // 193-202): U += (g + nugget) + x t g' d2f, and
        // S'_q += (g' + x t g'' d2f) B'_q + x t g' B''_q at y = t x.
        const double y = t * dc.x;
        double z2;
        diag_eval<BK, true, PM>(dc, y, &z2);
        double bq[PM + 1], bq2[PM + 1];
        diag_basis<BK, PM>(dc, y, bq);
        diag_basis2<BK, PM>(dc, y, bq2);
        const double xt = dc.x * t;
        A.U += w * (v.g1 + xt * v.gp * z2);
        const double a = v.gp + xt * z2 * v.gpp;
        const double b = xt * v.gp;
#pragma unroll
        for (int q = 1; q <= PM; ++q) {
            if (q > dc.p) break;
            A.S[q] += w * (a * bq[q] + b * bq2[q]);
        }
    } else if constexpr (FL >= FL_COEF) {
        double bq[PM + 1];
        const double y = t * dc.x;
        diag_basis<BK, PM>(dc, y, bq);
        const double wh = w * v.h;
#pragma unroll
        for (int q = 1; q <= PM; ++q) {
            if (q > dc.p) break;
            A.S[q] += wh * bq[q];
        }
        if constexpr (FL == FL_INPUT) {
            double z2;
            diag_eval<BK, true, PM>(dc, y, &z2);
            A.U += w * (v.g1 + t * v.h * z2);
        }
    }
}

// Depth-first stack of pending right halves for the adaptive rules.  All
// intervals are dyadic subintervals of [0, 1]: a pending interval is fully
// described by its level (width 2^-level) once its left end is known, and in
// depth-first order its left end is the right end of the leaf accepted just
// before it is popped, whose right-end integrand value is the pending
// interval's left-end value (the same computed value, propagated by copy).
// So an entry keeps only the level and the values the left end does not
// give: the midpoint and right end (Simpson) or the right end
// (Clenshaw-Curtis).  The top entry lives in registers; deeper entries go to
// the private array, written only when a push covers a top entry that exists
// in registers alone and read only when a pop needs an entry held in memory.
template <int FL>
struct AsEnt {
    NodeV<FL> fm, fb;
    int level;
};
template <int FL>
struct AccEnt {
    NodeV<FL> fb;
    int level;
};

// Push and pop of that stack: stk[0 .. sp-2] in the private array, entry
// sp-1 in `top` (inReg) and/or stk[sp-1] (inMem).  `top` is a variable of its
// own, apart from the dynamically indexed array, so that it stays in registers.
template <class E>
__host__ __device__ __forceinline__ void qpush(E* stk, E& top, int& sp, bool& inReg, bool& inMem, const E& e) {
    if (sp > 0 && !inMem) stk[sp - 1] = top;
    top = e;
    inReg = true;
    inMem = false;
    ++sp;
}
template <class E>
__host__ __device__ __forceinline__ E qpop(const E* stk, E& top, int& sp, bool& inReg, bool& inMem) {
    if (!inReg) top = stk[sp - 1];
    --sp;
    inReg = false;
    inMem = true;
    return top;
}

// A non-finite interval error means a non-finite node value: an infinite g
// (where the reference fails, MonotoneIntegrand.h:183-190) or a NaN input.
// The integral is then non-finite whatever follows, so the adaptive rules end
// at once with it instead of subdividing every remaining interval down to
// quadMaxSub (2^quadMaxSub intervals, as the reference's loop would for NaN
// inputs), and the adjoint sums become NaN.
template <int PM>
MPD_HD void nan_adj(AdjAcc<PM>& A) {
#pragma unroll
    for (int q = 0; q <= PM; ++q) A.S[q] = __builtin_nan("");
    A.U = __builtin_nan("");
}
MPD_HD inline bool nonfinite_err(double err) { return !(err <= __DBL_MAX__); }

// AdaptiveSimpson::Integrate (Quadrature.h:553-676): depth-first, left first,
// accepted intervals add intFine in left-to-right order.
template <int BK, int PK, int FL, int PM>
__host__ __device__ double integrate_as(const QuadDev& Q, const DiagCoef<PM>& dc, AdjAcc<PM>& A, int* nevals,
                                        const NodeV<FL>* at1) {
    // The interval sums, node positions and the acceptance test are computed
    // as written, without FMA contraction, like the reference's x86-64 build:
    // a fused sum rounds differently and can flip a subdivision decision.
#pragma clang fp contract(off)
    AsEnt<FL> stk[MAXSUB];
    AsEnt<FL> top;
    int sp = 0;
    bool inReg = false, inMem = false;
    NodeV<FL> fa = eval_node<BK, PK, FL, PM>(dc, 0.0);
    NodeV<FL> fb = at1 ? *at1 : eval_node<BK, PK, FL, PM>(dc, 1.0);
    NodeV<FL> fm = eval_node<BK, PK, FL, PM>(dc, 0.5);
    int ne = 3;
    double a = 0.0, b = 1.0, res = 0.0;
    int level = 0;
    while (true) {
        const double m = 0.5 * (a + b);
        const double lm = 0.5 * (a + m);
        const double rm = 0.5 * (m + b);
        NodeV<FL> flm = eval_node<BK, PK, FL, PM>(dc, lm);
        NodeV<FL> frm = eval_node<BK, PK, FL, PM>(dc, rm);
        ne += 2;
        const double coarse = div6(b - a) * (fa.F + 4.0 * fm.F + fb.F);
        const double wl = div6(m - a), wr = div6(b - m);
        double fine = wl * (fa.F + 4.0 * flm.F + fm.F);
        fine += wr * (fm.F + 4.0 * frm.F + fb.F);
        const double err = fabs(fine - coarse);
        if (nonfinite_err(err)) {
            res += fine;
            if constexpr (FL >= FL_COEF) nan_adj<PM>(A);
            break;
        }
        const double tol = fmax(Q.relTol * fabs(coarse), Q.absTol);
        if (((err < tol) || (level == Q.maxSub - 1)) && (level >= Q.minSub)) {
            res += fine;
            if constexpr (FL >= FL_COEF) {
                adj_node<BK, FL, PM>(dc, a, wl, fa, A);
                adj_node<BK, FL, PM>(dc, lm, 4.0 * wl, flm, A);
                adj_node<BK, FL, PM>(dc, m, wl + wr, fm, A);
                adj_node<BK, FL, PM>(dc, rm, 4.0 * wr, frm, A);
                adj_node<BK, FL, PM>(dc, b, wr, fb, A);
            }
            if (sp == 0) break;
            const AsEnt<FL> e = qpop(stk, top, sp, inReg, inMem);
            // the next interval starts where this one ends
            a = b;
            fa = fb;
            level = e.level;
            b = a + ldexp(1.0, -level);
            fm = e.fm;
            fb = e.fb;
        } else {
            qpush(stk, top, sp, inReg, inMem, AsEnt<FL>{frm, fb, level + 1});
            b = m;
            fb = fm;
            fm = flm;
            ++level;
        }
    }
    if (nevals) *nevals += ne;
    return res;
}

// AdaptiveClenshawCurtis::Integrate (Quadrature.h:819-968).  NF is the fine
// rule size when known at compile time (5 for quadPts = 5), or MAXNF with the
// runtime size in Q.nf.
// The same for a rule of more than MAXNF points (Q.big): run-time loops over
// the rule in device memory, the same nodes, sums and order.  The interior
// node values are not kept for the adjoint; an accepted interval evaluates
// them again (the same positions, so the same values; not counted in nevals).
template <int BK, int PK, int FL, int PM>
__host__ __device__ double integrate_acc_big(const QuadDev& Q, const DiagCoef<PM>& dc, AdjAcc<PM>& A, int* nevals,
                                             const NodeV<FL>* at1) {
#pragma clang fp contract(off)  // as in integrate_as
    AccEnt<FL> stk[MAXSUB];
    AccEnt<FL> top;
    int sp = 0;
    bool inReg = false, inMem = false;
    const int nf = Q.nf, nc = Q.nc;
    const auto cw = ro(Q.big);
    const auto cp = cw + nc;
    const auto fw = cp + nc;
    const auto fp = fw + nf;
    const int cRight = nc - 1, cMid = cRight / 2;
    const int fRight = nf - 1, fMid = fRight / 2;
    NodeV<FL> fa = eval_node<BK, PK, FL, PM>(dc, 0.0);
    NodeV<FL> fb = at1 ? *at1 : eval_node<BK, PK, FL, PM>(dc, 1.0);
    int ne = 2;
    double a = 0.0, b = 1.0, res = 0.0;
    int level = 0;
    while (true) {
        const double mid = 0.5 * (a + b);
        const double scale = 0.5 * (b - a);
        NodeV<FL> fm = eval_node<BK, PK, FL, PM>(dc, mid);
        ne += 1;
        double coarse = scale * cw[0] * fa.F;
        coarse += scale * cw[cMid] * fm.F;
        coarse += scale * cw[cRight] * fb.F;
        double fine = scale * fw[0] * fa.F;
        fine += scale * fw[fMid] * fm.F;
        fine += scale * fw[fRight] * fb.F;
        for (int i = 1; i < cMid; ++i) {
            const NodeV<FL> v = eval_node<BK, PK, FL, PM>(dc, mid + scale * cp[i]);
            ne += 1;
            coarse += scale * cw[i] * v.F;
            fine += scale * fw[2 * i] * v.F;
        }
        for (int ii = cMid + 1; ii < cRight; ++ii) {
            const NodeV<FL> v = eval_node<BK, PK, FL, PM>(dc, mid + scale * cp[ii]);
            ne += 1;
            coarse += scale * cw[ii] * v.F;
            fine += scale * fw[2 * ii] * v.F;
        }
        for (int i = 1; i < nf; i += 2) {
            const NodeV<FL> v = eval_node<BK, PK, FL, PM>(dc, mid + scale * fp[i]);
            ne += 1;
            fine += scale * fw[i] * v.F;
        }
        const double err = fabs(fine - coarse);
        if (nonfinite_err(err)) {  // as in integrate_as
            res += fine;
            if constexpr (FL >= FL_COEF) nan_adj<PM>(A);
            break;
        }
        const double tol = fmax(Q.relTol * fabs(coarse), Q.absTol);
        if (((err < tol) || (level == Q.maxSub - 1)) && (level >= Q.minSub)) {
            res += fine;
            if constexpr (FL >= FL_COEF) {
                adj_node<BK, FL, PM>(dc, a, scale * fw[0], fa, A);
                adj_node<BK, FL, PM>(dc, mid, scale * fw[fMid], fm, A);
                adj_node<BK, FL, PM>(dc, b, scale * fw[fRight], fb, A);
                for (int i = 1; i < nf - 1; ++i) {
                    if (i == fMid) continue;
                    const double t = mid + scale * fp[i];
                    adj_node<BK, FL, PM>(dc, t, scale * fw[i], eval_node<BK, PK, FL, PM>(dc, t), A);
                }
            }
            if (sp == 0) break;
            const AccEnt<FL> e = qpop(stk, top, sp, inReg, inMem);
            a = b;
            fa = fb;
            level = e.level;
            b = a + ldexp(1.0, -level);
            fb = e.fb;
        } else {
            qpush(stk, top, sp, inReg, inMem, AccEnt<FL>{fb, level + 1});
            b = mid;
            fb = fm;
            ++level;
        }
    }
    if (nevals) *nevals += ne;
    return res;
}

template <int BK, int PK, int FL, int NF, int PM>
__host__ __device__ double integrate_acc(const QuadDev& Q, const DiagCoef<PM>& dc, AdjAcc<PM>& A, int* nevals,
                                         const NodeV<FL>* at1) {
#pragma clang fp contract(off)  // as in integrate_as
    if constexpr (NF == MAXNF) {
        if (Q.big) return integrate_acc_big<BK, PK, FL, PM>(Q, dc, A, nevals, at1);
    }
    AccEnt<FL> stk[MAXSUB];
    AccEnt<FL> top;
    int sp = 0;
    bool inReg = false, inMem = false;
    const int nf = (NF == MAXNF) ? Q.nf : NF;
    const int nc = (nf - 1) / 2 + 1;
    const int cRight = nc - 1, cMid = cRight / 2;
    const int fRight = nf - 1, fMid = fRight / 2;
    NodeV<FL> fa = eval_node<BK, PK, FL, PM>(dc, 0.0);
    NodeV<FL> fb = at1 ? *at1 : eval_node<BK, PK, FL, PM>(dc, 1.0);
    int ne = 2;
    double a = 0.0, b = 1.0, res = 0.0;
    int level = 0;
    NodeV<FL> inner[NF];
    while (true) {
        const double mid = 0.5 * (a + b);
        const double scale = 0.5 * (b - a);
        NodeV<FL> fm = eval_node<BK, PK, FL, PM>(dc, mid);
        ne += 1;
        double coarse = scale * Q.cw[0] * fa.F;
        coarse += scale * Q.cw[cMid] * fm.F;
        coarse += scale * Q.cw[cRight] * fb.F;
        double fine = scale * Q.fw[0] * fa.F;
        fine += scale * Q.fw[fMid] * fm.F;
        fine += scale * Q.fw[fRight] * fb.F;
#pragma unroll
        for (int i = 1; i < NF; ++i) {
            if (i >= cMid) break;
            NodeV<FL> v = eval_node<BK, PK, FL, PM>(dc, mid + scale * Q.cp[i]);
            ne += 1;
            coarse += scale * Q.cw[i] * v.F;
            fine += scale * Q.fw[2 * i] * v.F;
            inner[2 * i] = v;
        }
#pragma unroll
        for (int i = 1; i < NF; ++i) {
            if (i + cMid + 1 >= cRight + 1) break;
            const int ii = cMid + i;
            if (ii >= cRight) break;
            NodeV<FL> v = eval_node<BK, PK, FL, PM>(dc, mid + scale * Q.cp[ii]);
            ne += 1;
            coarse += scale * Q.cw[ii] * v.F;
            fine += scale * Q.fw[2 * ii] * v.F;
            inner[2 * ii] = v;
        }
#pragma unroll
        for (int i = 1; i < NF; i += 2) {
            if (i >= nf) break;
            NodeV<FL> v = eval_node<BK, PK, FL, PM>(dc, mid + scale * Q.fp[i]);
            ne += 1;
            fine += scale * Q.fw[i] * v.F;
            inner[i] = v;
        }
        const double err = fabs(fine - coarse);
        if (nonfinite_err(err)) {  // as in integrate_as
            res += fine;
            if constexpr (FL >= FL_COEF) nan_adj<PM>(A);
            break;
        }
        const double tol = fmax(Q.relTol * fabs(coarse), Q.absTol);
        if (((err < tol) || (level == Q.maxSub - 1)) && (level >= Q.minSub)) {
            res += fine;
            if constexpr (FL >= FL_COEF) {
                adj_node<BK, FL, PM>(dc, a, scale * Q.fw[0], fa, A);
                adj_node<BK, FL, PM>(dc, mid, scale * Q.fw[fMid], fm, A);
                adj_node<BK, FL, PM>(dc, b, scale * Q.fw[fRight], fb, A);
#pragma unroll
                for (int i = 1; i < NF - 1; ++i) {
                    if (i >= fRight) break;
                    if (i == fMid) continue;
                    adj_node<BK, FL, PM>(dc, mid + scale * Q.fp[i], scale * Q.fw[i], inner[i], A);
                }
            }
            if (sp == 0) break;
            const AccEnt<FL> e = qpop(stk, top, sp, inReg, inMem);
            a = b;
            fa = fb;
            level = e.level;
            b = a + ldexp(1.0, -level);
            fb = e.fb;
        } else {
            qpush(stk, top, sp, inReg, inMem, AccEnt<FL>{fb, level + 1});
            b = mid;
            fb = fm;
            ++level;
        }
    }
    if (nevals) *nevals += ne;
    return res;
}

// ClenshawCurtisQuadrature::Integrate (Quadrature.h:279-307) on [0, 1] with
// the fixed rule of Q.nf points (nodes 0.5 (1 + p_i), weights 0.5 w_i, summed
// in rule order); the last node is t = 1.
template <int BK, int PK, int FL, int PM>
MPD_HD double integrate_cc(const QuadDev& Q, const DiagCoef<PM>& dc, AdjAcc<PM>& A, int* nevals,
                           const NodeV<FL>* at1) {
#pragma clang fp contract(off)  // as in integrate_as
    const int n = Q.nf;
    double res = 0.0;
    auto node = [&](double fpi, double fwi) {
#pragma clang fp contract(off)
        const double t = 0.5 * (1.0 + 0.0 + (1.0 - 0.0) * fpi);
        const NodeV<FL> v = (at1 && fpi == 1.0) ? *at1 : eval_node<BK, PK, FL, PM>(dc, t);
        const double w = 0.5 * (1.0 - 0.0) * fwi;
        res += w * v.F;
        if constexpr (FL >= FL_COEF) adj_node<BK, FL, PM>(dc, t, w, v, A);
    };
    if (Q.big) {  // more than MAXNF points: the rule in device memory
        const auto fw = ro(Q.big) + 2 * Q.nc;
        for (int i = 0; i < n; ++i) node(fw[n + i], fw[i]);
    } else {
        for (int i = 0; i < MAXNF; ++i) {
            if (i >= n) break;
            node(Q.fp[i], Q.fw[i]);
        }
    }
    if (nevals) *nevals += n;
    return res;
}

// at1, when given, is the integrand at t = 1 (computed by the caller, which
// also needs g and g' at x_k); it is the value eval_node would produce there.
template <int BK, int PK, int QK, int FL, int PM>
MPD_HD double integrate(const QuadDev& Q, const DiagCoef<PM>& dc, AdjAcc<PM>& A, int* nevals,
                        const NodeV<FL>* at1 = nullptr) {
    if constexpr (QK == QK_AS) return integrate_as<BK, PK, FL, PM>(Q, dc, A, nevals, at1);
    else if constexpr (QK == QK_ACC5) return integrate_acc<BK, PK, FL, 5, PM>(Q, dc, A, nevals, at1);
    else if constexpr (QK == QK_CC) return integrate_cc<BK, PK, FL, PM>(Q, dc, A, nevals, at1);
    else return integrate_acc<BK, PK, FL, MAXNF, PM>(Q, dc, A, nevals, at1);
}

template <int PM>
MPD_HD void zero_adj(AdjAcc<PM>& A) {
#pragma unroll
    for (int q = 0; q <= PM; ++q) A.S[q] = 0.0;
    A.U = 0.0;
}


// ---------------------------------------------------------------------------
// Forward + adjoint of one component at one point for the fused pass: T_c,
// log g(d_c f), the adjoint integrals and
//   G_q = s (phi~_q(0) + rho_q S'_q) + l rho_q B'_q(x_c) g'/g,
// with s = T_c and l = -1 in KL mode (then loss += 0.5 T_c^2 - logdet), or the
// given sensitivities otherwise.  G[0..p] is written; G is zero for an
// inactive lane.
// With IN, the integrand also carries the input adjoint U and gk receives
// the x_c entry of the input gradient, s U + l (g'/g) d^2_c f (the other
// entries are G_q dC_q/dx_i, as for the input VJP below).
template <int BK, int PK, int QK, int PM, bool IN = false>
MPD_HD void comp_core(const MapDev& M, const CompDev& cd, const double* C, double xk, bool active, int klMode,
                      double s_in, double l_in, double* G, double& loss, double* gk = nullptr,
                      double* Tout = nullptr) {
    constexpr int FLV = IN ? FL_INPUT : FL_COEF;
#pragma unroll
    for (int q = 0; q <= PM; ++q) G[q] = 0.0;
    if constexpr (IN) *gk = 0.0;
    if (!active) return;
    const double f0 = f_at_zero<PM>(M, cd, C);
    DiagCoef<PM> dc;
    make_diag<PM>(M, cd, C, xk, M.nugget, dc);
    // g and g' at x_k (log-determinant) are the integrand's values at t = 1.
    double z2 = 0.0;
    const double z = diag_eval<BK, IN, PM>(dc, 1.0 * xk, IN ? &z2 : nullptr);
    const PosF<PK> gv = posfunc<PK>(z);
    if (QK != QK_CC || M.quad.nf > 1) note_infinite_g<PM>(dc, gv.g);  // t = 1 is a node of the rule
    NodeV<FLV> n1;
    n1.F = dc.x * (gv.g + dc.nug);
    n1.h = dc.x * gv.gp;
    if constexpr (IN) n1.g1 = gv.g + dc.nug;
    AdjAcc<PM> A;
    zero_adj<PM>(A);
    double T = integrate<BK, PK, QK, FLV, PM>(M.quad, dc, A, nullptr, &n1) + f0;
    if (M.nugget != 0.0) {
        // Evaluate drops the nugget (MonotoneComponent.h:287): its tree may differ.
        DiagCoef<PM> dv = dc;
        dv.nug = 0.0;
        NodeV<FL_VAL> v1;
        v1.F = dv.x * (gv.g + dv.nug);
        AdjAcc<PM> Av;
        T = integrate<BK, PK, QK, FL_VAL, PM>(M.quad, dv, Av, nullptr, &v1) + f0;
    }
    if (Tout) *Tout = T;
    const double ldv = (gv.g <= 0) ? -__builtin_inf() : log_pos(gv.g);
    const double r = gv.gp * (1.0 / gv.g);
    double s = s_in, l = l_in;
    if (klMode) {
        s = T;
        l = -1.0;
        loss += 0.5 * T * T - ldv;
    }
    double bq[PM + 1];
    diag_basis<BK, PM>(dc, xk, bq);
    G[0] = s;
#pragma unroll
    for (int q = 1; q <= PM; ++q) {
        if (q > cd.p) break;
        G[q] = s * (M.phi0[q] + M.rho[q] * A.S[q]) + (l * r) * (M.rho[q] * bq[q]);
    }
    if constexpr (IN) *gk = s * A.U + (l * r) * z2;
}
// The same with the discrete derivative (contDeriv = false): the
// log-determinant term is log D with D the derivative of the quadrature
// approximation of T in x_c (DiscreteDerivative, MonotoneComponent.h:586-638:
// the Diagonal integrand, nugget included), and its coefficient gradient
// (DiscreteMixedJacobian, 928-980: the MixedCoeff integrand) is
// rho_q S''_q / D, both from one quadrature of the FL_DCOEF integrand; T and
// its adjoint come from comp_core's own quadrature.  G_q = s (phi~_q(0) +
// rho_q S'_q) + l rho_q S''_q / D; in KL mode s = T_c, l = -1.
template <int BK, int PK, int QK, int PM>
MPD_HD void comp_core_discrete(const MapDev& M, const CompDev& cd, const double* C, double xk, bool active,
                               int klMode, double s_in, double l_in, double* G, double& loss) {
    double lossT = 0.0, T = 0.0;
    // T and phi~_q(0) + rho_q S'_q: comp_core with s = 1 and no log-det term
    comp_core<BK, PK, QK, PM>(M, cd, C, xk, active, 0, 1.0, 0.0, G, lossT, nullptr, &T);
    if (!active) return;
    DiagCoef<PM> dc;
    make_diag<PM>(M, cd, C, xk, M.nugget, dc);
    AdjAcc<PM> A;
    zero_adj<PM>(A);
    integrate<BK, PK, QK, FL_DCOEF, PM>(M.quad, dc, A, nullptr);
    const double D = A.U;
    double s = s_in, l = l_in;
    if (klMode) {
        s = T;
        l = -1.0;
        loss += 0.5 * T * T - ((D <= 0) ? -__builtin_inf() : log_pos(D));
    }
#pragma unroll
    for (int q = 0; q <= PM; ++q) G[q] *= s;
    const double li = l / D;
#pragma unroll
    for (int q = 1; q <= PM; ++q) {
        if (q > cd.p) break;
        G[q] += li * (M.rho[q] * A.S[q]);
    }
}

// Per-component factors of an input vector-Jacobian product at one point,
//   d/dx [s T_c + l log g(d_c f)] = sum_q G_q dC_q/dx_i  (i < c)  and  gk  (i = c),
// with G_q = s (phi~_q(0) + rho_q S_q) + l (g'/g) rho_q B'_q(x_c) and
// gk = s U + l (g'/g) d^2_c f (InputJacobian, MonotoneComponent.h:148-164,
// 758-818, and ContinuousMixedInputJacobian, 208-232, 877-926).  S_q and U
// carry the nugget as the reference's input Jacobian does.  In KL mode
// s = T_c (the Evaluate value, without the nugget, 287) and l = -1, the
// gradient of T_c^2 / 2 - log g.  Without s (hasS false) the quadrature is
// skipped; without l (hasL false) the log-determinant term is left out.
// G[0..PM] is written.
template <int BK, int PK, int QK, int PM, bool KL>
MPD_HD void ivjp_core(const MapDev& M, const CompDev& cd, const double* C, double xk, bool hasS, double s_in,
                      bool hasL, double l_in, double* G, double& gk) {
    constexpr bool klMode = KL;
    DiagCoef<PM> dc;
    make_diag<PM>(M, cd, C, xk, M.nugget, dc);
    // g and g' at x_k are the integrand's values at t = 1.
    double z2;
    const double z = diag_eval<BK, true, PM>(dc, 1.0 * xk, &z2);
    const PosF<PK> gv = posfunc<PK>(z);
    AdjAcc<PM> A;
    zero_adj<PM>(A);
    double s = 0.0, l = hasL ? l_in : 0.0;
    if (klMode || hasS) {
        const double f0 = f_at_zero<PM>(M, cd, C);
        if (QK != QK_CC || M.quad.nf > 1) note_infinite_g<PM>(dc, gv.g);  // t = 1 is a node of the rule
        NodeV<FL_INPUT> n1;
        n1.F = dc.x * (gv.g + dc.nug);
        n1.h = dc.x * gv.gp;
        n1.g1 = gv.g + dc.nug;
        double T = integrate<BK, PK, QK, FL_INPUT, PM>(M.quad, dc, A, nullptr, &n1) + f0;
        if constexpr (KL) {
            if (M.nugget != 0.0) {
                DiagCoef<PM> dv = dc;
                dv.nug = 0.0;
                NodeV<FL_VAL> v1;
                v1.F = dv.x * gv.g;
                AdjAcc<PM> Av;
                T = integrate<BK, PK, QK, FL_VAL, PM>(M.quad, dv, Av, nullptr, &v1) + f0;
            }
            s = T;
            l = -1.0;
        } else {
            s = s_in;
        }
    }
    const double lr = (klMode || hasL) ? l * (gv.gp * (1.0 / gv.g)) : 0.0;
    double bq[PM + 1];
    diag_basis<BK, PM>(dc, xk, bq);
    G[0] = s;
#pragma unroll
    for (int q = 1; q <= PM; ++q)
        G[q] = (q <= cd.p) ? s * (M.phi0[q] + M.rho[q] * A.S[q]) + lr * (M.rho[q] * bq[q]) : 0.0;
    gk = s * A.U + lr * z2;
}
// ---------------------------------------------------------------------------
// Wavefront reduction of 16 per-lane values v[0..15] (one per term of a
// chunk) into the 16 sums over the 64 lanes.
//
// Stages, each halving the number of values a lane holds and doubling the
// lanes each value covers:
//   A  v_permlane32_swap  (lanes l, l^32):   16 -> 8 values
//   B  v_permlane16_swap  (lanes l, l^16):    8 -> 4
//   C  DPP row_ror:8      (lanes l, l^8):     4 -> 2
//   D  DPP row_half_mirror (lanes l, l^7):    2 -> 1
//   E  DPP quad_perm      (lanes l^1, l^2):   sum of the remaining value
// The swaps exchange whole register halves between two values, so A and B
// need one add per value pair and no lane selects.  Afterwards every lane
// holds the total of term (lane >> 2) & 15.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    int lo = __double2loint(v);
    int hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// a + b after exchanging lanes 32..63 of a with lanes 0..31 of b.
__device__ __forceinline__ double swap32_add(double a, double b) {
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

// a + b after exchanging the odd 16-lane rows of a with the even rows of b.
__device__ __forceinline__ double swap16_add(double a, double b) {
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

// The intermediate values are named scalars: selects between elements of a
// local array can be turned into address selects, which forces the array
// into scratch memory.
__device__ __forceinline__ double wave_reduce16(const double* v, int lane) {
    const double w0 = swap32_add(v[0], v[8]), w1 = swap32_add(v[1], v[9]);  // term i + 8*b5
    const double w2 = swap32_add(v[2], v[10]), w3 = swap32_add(v[3], v[11]);
    const double w4 = swap32_add(v[4], v[12]), w5 = swap32_add(v[5], v[13]);
    const double w6 = swap32_add(v[6], v[14]), w7 = swap32_add(v[7], v[15]);
    const double x0 = swap16_add(w0, w4), x1 = swap16_add(w1, w5);  // + 4*b4
    const double x2 = swap16_add(w2, w6), x3 = swap16_add(w3, w7);
    const bool b3 = (lane & 8) != 0;  // + 2*b3
    const double y0 = (b3 ? x2 : x0) + dpp_mov<0x128>(b3 ? x0 : x2);
    const double y1 = (b3 ? x3 : x1) + dpp_mov<0x128>(b3 ? x1 : x3);
    const bool b2 = (lane & 4) != 0;  // + b2
    double r = (b2 ? y1 : y0) + dpp_mov<0x141>(b2 ? y0 : y1);
    r += dpp_mov<0xB1>(r);
    r += dpp_mov<0x4E>(r);
    return r;
}

// Reduces 16 per-lane values over the wavefront and adds sum i to dst[i]
// for i < nvalid.  dst is in LDS and owned by this wave; the update is an LDS
// atomic add without return, so the wave does not wait for it (the 16 writer
// lanes hit distinct addresses, and one wave's updates apply in program order).
__device__ __forceinline__ void reduce16_add(double* v, int lane, double* dst, int nvalid) {
    const double r = wave_reduce16(v, lane);
    const int u = lane >> 2;
    if ((lane & 3) == 0 && u < nvalid) __hip_atomic_fetch_add(dst + u, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace mpd
