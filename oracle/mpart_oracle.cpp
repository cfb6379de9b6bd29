// mpart_oracle.cpp -- CPU restatement of MParT's monotone-component path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check
// and the CPU-baseline leg of bench.py call it through ctypes.  The product
// (mpart_amd/) never links, loads or calls it.
//
// Every routine below restates, operation for operation, the reference routine
// cited in its comment (paths are relative to the MParT source tree).  The
// arithmetic order is kept so that this oracle reproduces the reference's
// adaptive subdivision trees and ITP iterates, up to libm last-ulp effects.
// It is compiled with -ffp-contract=off so that no FMA is formed.
//
// Data layout of the C entry points: point matrices are D x N with explicit
// (rowStride, colStride) in units of doubles, exactly like the reference's
// StridedMatrix views (MParT/Utilities/ArrayConversions.h:102-130).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include <algorithm>
#include <omp.h>
#include <atomic>
#include <exception>

namespace orc {

// An exception may not leave an OpenMP region (the process would end): each
// point's work catches it, the first one is kept, the remaining points are
// skipped, and it is thrown again after the region, as the serial loop of
// the reference would throw it.
class OmpError {
   public:
    bool failed() const { return failed_.load(std::memory_order_relaxed); }
    void keep() {
#pragma omp critical(orc_omp_error)
        {
            if (!e_) e_ = std::current_exception();
        }
        failed_.store(true, std::memory_order_relaxed);
    }
    void rethrow() {
        if (e_) std::rethrow_exception(e_);
    }

   private:
    std::atomic<bool> failed_{false};
    std::exception_ptr e_;
};

enum BasisType { PROB_HERMITE = 0, PHYS_HERMITE = 1, HERMITE_FN = 2 };
enum PosFuncType { POS_EXP = 0, POS_SOFTPLUS = 1 };
enum QuadType { QUAD_CC = 0, QUAD_AS = 1, QUAD_ACC = 2 };

// Mirrors MParT/MapOptions.h:58-123 (only the fields used on this path).
struct Options {
    int32_t basisType;
    int32_t posFuncType;
    int32_t quadType;
    int32_t contDeriv;
    int32_t basisNorm;
    uint32_t quadMaxSub;
    uint32_t quadMinSub;
    uint32_t quadPts;
    double quadAbsTol;
    double quadRelTol;
    double nugget;
    double basisLB;  // LinearizedBasis bounds (MapOptions.h:76-77); +-inf: not linearized
    double basisUB;
};

// Derivative flags, MParT/DerivativeFlags.h:7-15.
enum Deriv { D_NONE, D_PARAMETERS, D_DIAGONAL, D_DIAGONAL2, D_INPUT, D_MIXEDINPUT, D_MIXEDCOEFF };

// ---------------------------------------------------------------------------
// Multi-index set in compressed (CSR) form.
// src/MultiIndices/FixedMultiIndexSet.cpp:208-233 (total-order ctor),
// 487-549 (TotalOrderSize / FillTotalOrder), 160-187 (CalculateMaxDegrees).
struct MSet {
    unsigned dim = 0;
    std::vector<uint32_t> starts, dims, orders, maxDeg;
    unsigned size() const { return unsigned(starts.size()) - 1; }

    void calc_max_degrees() {
        maxDeg.assign(dim, 0u);
        for (size_t i = 0; i < orders.size(); ++i)
            maxDeg[dims[i]] = std::max(maxDeg[dims[i]], orders[i]);
    }
};

// Recursion of FillTotalOrder (FixedMultiIndexSet.cpp:507-549): dimension 0 is
// the outermost loop, the last dimension the innermost one; only nonzero
// orders are stored, in increasing dimension.
static void fill_total_order(MSet& s, unsigned maxOrder, unsigned minOrder,
                             std::vector<unsigned>& work, unsigned currDim) {
    if (currDim < s.dim - 1) {
        for (unsigned pw = 0; pw <= maxOrder; ++pw) {
            work[currDim] = pw;
            fill_total_order(s, maxOrder - pw, minOrder, work, currDim + 1);
        }
    } else {
        for (unsigned pw = 0; pw <= maxOrder; ++pw) {
            work[currDim] = pw;
            unsigned currOrder = 0;
            for (unsigned i = 0; i < currDim + 1; ++i) currOrder += work[i];
            if (currOrder >= minOrder) {
                s.starts.push_back(uint32_t(s.dims.size()));
                for (unsigned i = 0; i < s.dim; ++i) {
                    if (work[i] > 0) {
                        s.dims.push_back(i);
                        s.orders.push_back(work[i]);
                    }
                }
            }
        }
    }
}

static MSet total_order_set(unsigned dim, unsigned maxOrder, unsigned minOrder = 0) {
    MSet s;
    s.dim = dim;
    std::vector<unsigned> work(dim, 0u);
    fill_total_order(s, maxOrder, minOrder, work, 0);
    s.starts.push_back(uint32_t(s.dims.size()));
    s.calc_max_degrees();
    return s;
}

// CSR constructor (FixedMultiIndexSet.cpp:189-206): entries of each term are
// sorted so that dims increase.
static MSet csr_set(unsigned dim, unsigned nterms, const uint32_t* starts,
                    const uint32_t* dims, const uint32_t* orders) {
    MSet s;
    s.dim = dim;
    s.starts.assign(starts, starts + nterms + 1);
    unsigned nnz = starts[nterms];
    s.dims.assign(dims, dims + nnz);
    s.orders.assign(orders, orders + nnz);
    for (unsigned t = 0; t < nterms; ++t) {
        std::vector<std::pair<uint32_t, uint32_t>> e;
        for (unsigned i = s.starts[t]; i < s.starts[t + 1]; ++i) e.push_back({s.dims[i], s.orders[i]});
        std::stable_sort(e.begin(), e.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (unsigned i = s.starts[t], k = 0; i < s.starts[t + 1]; ++i, ++k) {
            s.dims[i] = e[k].first;
            s.orders[i] = e[k].second;
        }
    }
    s.calc_max_degrees();
    return s;
}

// ---------------------------------------------------------------------------
// 1-D bases.
// Factorial in 32-bit unsigned arithmetic, MParT/Utilities/MathFunctions.h:16-22.
static unsigned factorial_u32(unsigned d) {
    unsigned out = 1;
    for (unsigned i = 2; i <= d; ++i) out *= i;
    return out;
}

// Three-term recurrence p_k = (a_k x + b_k) p_{k-1} - c_k p_{k-2}
// (MParT/OrthogonalPolynomial.h:25-156) with the ProbabilistHermite
// (301-306) or PhysicistHermite (319-324) coefficients.
struct OrthoPoly {
    bool phys = false;
    bool normalize = false;
    double ak(unsigned) const { return phys ? 2.0 : 1.0; }
    double bk(unsigned) const { return 0.0; }
    double ck(unsigned k) const { return phys ? 2.0 * (k - 1.0) : k - 1.0; }
    double phi0(double) const { return 1.0; }
    double phi1(double x) const { return phys ? 2.0 * x : x; }
    double phi1_deriv(double) const { return phys ? 2.0 : 1.0; }
    double norm(unsigned k) const {
        if (phys) return std::sqrt(std::sqrt(M_PI) * std::pow(2.0, double(k)) * factorial_u32(k));
        return std::sqrt(std::sqrt(2.0 * M_PI) * factorial_u32(k));
    }

    void eval_all(double* out, unsigned maxOrder, double x) const {
        out[0] = phi0(x);
        if (maxOrder > 0) out[1] = phi1(x);
        for (unsigned k = 2; k <= maxOrder; ++k)
            out[k] = (ak(k) * x + bk(k)) * out[k - 1] - ck(k) * out[k - 2];
        if (normalize)
            for (unsigned k = 0; k <= maxOrder; ++k) out[k] /= norm(k);
    }
    void eval_derivs(double* v, double* d, unsigned maxOrder, double x) const {
        v[0] = phi0(x);
        d[0] = 0.0;
        if (maxOrder > 0) {
            v[1] = phi1(x);
            d[1] = phi1_deriv(x);
        }
        for (unsigned k = 2; k <= maxOrder; ++k) {
            double a = ak(k), b = bk(k), c = ck(k);
            v[k] = (a * x + b) * v[k - 1] - c * v[k - 2];
            d[k] = a * v[k - 1] + (a * x + b) * d[k - 1] - c * d[k - 2];
        }
        if (normalize)
            for (unsigned k = 0; k <= maxOrder; ++k) {
                double n = norm(k);
                v[k] /= n;
                d[k] /= n;
            }
    }
    void eval_second(double* v, double* d, double* d2, unsigned maxOrder, double x) const {
        v[0] = phi0(x);
        d[0] = 0.0;
        d2[0] = 0.0;
        if (maxOrder > 0) {
            v[1] = phi1(x);
            d[1] = phi1_deriv(x);
            d2[1] = 0.0;
        }
        for (unsigned k = 2; k <= maxOrder; ++k) {
            double a = ak(k), b = bk(k), c = ck(k);
            v[k] = (a * x + b) * v[k - 1] - c * v[k - 2];
            d[k] = a * v[k - 1] + (a * x + b) * d[k - 1] - c * d[k - 2];
            d2[k] = a * d[k - 1] + a * d[k - 1] + (a * x + b) * d2[k - 1] - c * d2[k - 2];
        }
        if (normalize)
            for (unsigned k = 0; k <= maxOrder; ++k) {
                double n = norm(k);
                v[k] /= n;
                d[k] /= n;
                d2[k] /= n;
            }
    }
    // Derivative(order, x) (OrthogonalPolynomial.h:192-226): upward recurrence.
    double derivative(unsigned order, double x) const {
        if (order == 0) return 0.0;
        if (order == 1) return normalize ? phi1_deriv(x) / norm(1) : phi1_deriv(x);
        double lag2_val, lag1_val = phi0(x), next_val = phi1(x);
        double lag2_deriv, lag1_deriv = 0.0, next_deriv = phi1_deriv(x);
        for (unsigned i = 2; i <= order; ++i) {
            double a = ak(i), b = bk(i), c = ck(i);
            lag2_val = lag1_val;
            lag1_val = next_val;
            next_val = (a * x + b) * lag1_val - c * lag2_val;
            lag2_deriv = lag1_deriv;
            lag1_deriv = next_deriv;
            next_deriv = a * lag1_val + (a * x + b) * lag1_deriv - c * lag2_deriv;
        }
        return normalize ? next_deriv / norm(order) : next_deriv;
    }
    // Evaluate(order, x) (OrthogonalPolynomial.h:160-190): downward Clenshaw sum.
    double evaluate(unsigned order, double x) const {
        double nrm = normalize ? norm(order) : 1.0;
        if (order == 0) return phi0(x) / nrm;
        if (order == 1) return phi1(x) / nrm;
        double yk2 = 0.0, yk1 = 0.0, yk = 1.0;
        for (int k = (int)order - 1; k >= 0; k--) {
            yk2 = yk1;
            yk1 = yk;
            double alpha = ak(k + 1) * x + bk(k + 1);
            double beta = -ck(k + 2);
            yk = alpha * yk1 + beta * yk2;
        }
        double beta = -ck(2);
        return (yk1 * phi1(x) + beta * phi0(x) * yk2) / nrm;
    }
};

// Hermite functions, MParT/HermiteFunction.h:22-104: phi_0 = 1, phi_1 = x and
// phi_{k+2} = psi_k.  Values use the orthonormal recurrence (EvaluateAll),
// derivatives use physicist polynomials plus exp/pow scaling.
struct HermiteFn {
    OrthoPoly base{true, false};
    void eval_all(double* out, unsigned maxOrder, double x) const {
        out[0] = 1.0;
        if (maxOrder > 0) {
            out[1] = x;
            if (maxOrder > 1) {
                out[2] = std::pow(M_PI, -0.25) * std::exp(-0.5 * x * x);
                if (maxOrder > 2) {
                    out[3] = std::sqrt(2.0) * x * out[2];
                    for (unsigned i = 2; i <= maxOrder - 2; ++i)
                        out[i + 2] = (x * out[i + 1] - std::sqrt(0.5 * (i - 1)) * out[i]) / std::sqrt(0.5 * i);
                }
            }
        }
    }
    void eval_derivs(double* v, double* d, unsigned maxOrder, double x) const {
        v[0] = 1.0;
        d[0] = 0.0;
        if (maxOrder > 0) {
            v[1] = x;
            d[1] = 1.0;
            if (maxOrder > 1) {
                base.eval_derivs(&v[2], &d[2], maxOrder - 2, x);
                const double baseScaling = std::pow(M_PI, -0.25) * std::exp(-0.5 * x * x);
                double scale;
                double currFactorial = 1;
                scale = baseScaling;
                d[2] -= x * v[2];
                d[2] *= scale;
                v[2] *= scale;
                for (unsigned i = 1; i <= maxOrder - 2; ++i) {
                    currFactorial *= i;
                    scale = baseScaling * std::pow(std::pow(2, i) * currFactorial, -0.5);
                    d[i + 2] -= x * v[i + 2];
                    d[i + 2] *= scale;
                    v[i + 2] *= scale;
                }
            }
        }
    }
    // Derivative(order, x) (HermiteFunction.h:117-129).
    double derivative(unsigned order, double x) const {
        if (order == 0) return 0.0;
        if (order == 1) return 1.0;
        const double scaling = std::pow(std::pow(2, order - 2) * factorial_u32(order - 2) * std::sqrt(M_PI), -0.5);
        const double expPart = std::exp(-0.5 * x * x);
        return scaling * (-x * expPart * base.evaluate(order - 2, x) + expPart * base.derivative(order - 2, x));
    }
    void eval_second(double* v, double* d, double* d2, unsigned maxOrder, double x) const {
        eval_derivs(v, d, maxOrder, x);
        d2[0] = 0.0;
        if (maxOrder > 0) {
            d2[1] = 0.0;
            if (maxOrder > 1)
                for (unsigned i = 0; i <= maxOrder - 2; ++i) d2[i + 2] = -(2.0 * i + 1.0 - x * x) * v[i + 2];
        }
    }
};

struct Basis1D {
    int type = PROB_HERMITE;
    OrthoPoly poly;
    HermiteFn hf;
    // LinearizedBasis (MParT/LinearizedBasis.h:17-160): linear outside [lb, ub]
    // with the value and slope of the basis at the nearer bound.
    bool lin = false;
    double lb = -INFINITY, ub = INFINITY;

    void raw_all(double* o, unsigned m, double x) const {
        if (type == HERMITE_FN) hf.eval_all(o, m, x); else poly.eval_all(o, m, x);
    }
    void raw_derivs(double* v, double* d, unsigned m, double x) const {
        if (type == HERMITE_FN) hf.eval_derivs(v, d, m, x); else poly.eval_derivs(v, d, m, x);
    }
    void raw_second(double* v, double* d, double* d2, unsigned m, double x) const {
        if (type == HERMITE_FN) hf.eval_second(v, d, d2, m, x); else poly.eval_second(v, d, d2, m, x);
    }
    double raw_derivative(unsigned k, double x) const {
        return type == HERMITE_FN ? hf.derivative(k, x) : poly.derivative(k, x);
    }
    // LinearizedBasis::EvaluateAll (LinearizedBasis.h:41-61).
    void eval_all(double* o, unsigned m, double x) const {
        if (lin && x < lb) {
            raw_all(o, m, lb);
            for (unsigned i = 0; i <= m; ++i) o[i] += raw_derivative(i, lb) * (x - lb);
        } else if (lin && x > ub) {
            raw_all(o, m, ub);
            for (unsigned i = 0; i <= m; ++i) o[i] += raw_derivative(i, ub) * (x - ub);
        } else {
            raw_all(o, m, x);
        }
    }
    // LinearizedBasis::EvaluateDerivatives (63-87).
    void eval_derivs(double* v, double* d, unsigned m, double x) const {
        if (lin && x < lb) {
            raw_derivs(v, d, m, lb);
            for (unsigned i = 0; i <= m; ++i) v[i] += d[i] * (x - lb);
        } else if (lin && x > ub) {
            raw_derivs(v, d, m, ub);
            for (unsigned i = 0; i <= m; ++i) v[i] += d[i] * (x - ub);
        } else {
            raw_derivs(v, d, m, x);
        }
    }
    // LinearizedBasis::EvaluateSecondDerivatives (90-109).
    void eval_second(double* v, double* d, double* d2, unsigned m, double x) const {
        if (lin && (x < lb || x > ub)) {
            eval_derivs(v, d, m, x);
            for (unsigned i = 0; i <= m; ++i) d2[i] = 0.0;
        } else {
            raw_second(v, d, d2, m, x);
        }
    }
};

// ---------------------------------------------------------------------------
// Positive bijectors, MParT/PositiveBijectors.h:12-57.
struct PosFunc {
    int type = POS_SOFTPLUS;
    double eval(double x) const {
        if (type == POS_EXP) return std::exp(x);
        return std::log(1.0 + std::exp(-1.0 * std::abs(x))) + std::fmax(x, 0.0);
    }
    double deriv(double x) const {
        if (type == POS_EXP) return std::exp(x);
        return x < 0 ? std::exp(x) / (std::exp(x) + 1.0) : 1.0 / (1.0 + std::exp(-1.0 * x));
    }
    double second(double x) const {
        if (type == POS_EXP) return std::exp(x);
        double fx = deriv(x);
        return fx * (1.0 - fx);
    }
};

// ---------------------------------------------------------------------------
// Strided views.
struct CMat {
    const double* p;
    long rows, cols, rs, cs;
    double operator()(long i, long j) const { return p[i * rs + j * cs]; }
};
struct Mat {
    double* p;
    long rows, cols, rs, cs;
    double& operator()(long i, long j) const { return p[i * rs + j * cs]; }
};

// A point column x_{0..dim-1} of a point matrix.
struct PtView {
    const double* p;
    long stride;
    unsigned n;
    double operator()(unsigned i) const { return p[i * stride]; }
};

// ---------------------------------------------------------------------------
// Expansion worker, MParT/MultivariateExpansionWorker.h:94-554 (homogeneous
// basis, identity rectifier).
struct Worker {
    unsigned dim = 0;
    MSet ms;
    Basis1D basis;
    std::vector<unsigned> startPos;
    unsigned cacheSize = 0;

    void init() {
        dim = ms.dim;
        startPos.assign(2 * dim + 2, 0u);
        unsigned upd = 0;
        for (unsigned i = 0; i < 2 * dim + 2; ++i) {
            startPos[i] = upd;
            upd += (i < 2 * dim) ? ms.maxDeg[i % dim] + 1 : ms.maxDeg[dim - 1] + 1;
        }
        cacheSize = startPos[2 * dim + 1];
    }
    unsigned nterms() const { return ms.size(); }

    void fill_cache1(double* c, const PtView& pt, int dt) const {
        if (dt == D_INPUT || dt == D_MIXEDINPUT) {
            for (unsigned d = 0; d + 1 < dim; ++d)
                basis.eval_derivs(&c[startPos[d]], &c[startPos[d + dim]], ms.maxDeg[d], pt(d));
        } else {
            for (unsigned d = 0; d + 1 < dim; ++d) basis.eval_all(&c[startPos[d]], ms.maxDeg[d], pt(d));
        }
    }
    void fill_cache2(double* c, double xd, int dt) const {
        if (dt == D_NONE || dt == D_PARAMETERS) {
            basis.eval_all(&c[startPos[dim - 1]], ms.maxDeg[dim - 1], xd);
        } else if (dt == D_DIAGONAL || dt == D_INPUT || dt == D_MIXEDCOEFF) {
            basis.eval_derivs(&c[startPos[dim - 1]], &c[startPos[2 * dim - 1]], ms.maxDeg[dim - 1], xd);
        } else if (dt == D_DIAGONAL2 || dt == D_MIXEDINPUT) {
            basis.eval_second(&c[startPos[dim - 1]], &c[startPos[2 * dim - 1]], &c[startPos[2 * dim]],
                              ms.maxDeg[dim - 1], xd);
        }
    }

    bool is_const(unsigned t) const { return ms.starts[t] == ms.starts[t + 1]; }

    double term_val(unsigned t, const double* c) const {
        double v = 1.0;
        unsigned e = ms.starts[t + 1] - 1;
        for (unsigned i = ms.starts[t]; i < e; ++i) v *= c[startPos[ms.dims[i]] + ms.orders[i]];
        double last = c[startPos[ms.dims[e]] + ms.orders[e]];
        v *= last;
        return v;
    }
    double term_val_diag(unsigned t, const double* c, unsigned posIndex) const {
        double v = 1.0;
        unsigned e = ms.starts[t + 1] - 1;
        if (ms.dims[e] != dim - 1) return 0.;
        for (unsigned i = ms.starts[t]; i < e; ++i) v *= c[startPos[ms.dims[i]] + ms.orders[i]];
        double dv = c[startPos[posIndex] + ms.orders[e]];
        v *= dv;
        return v;
    }
    double term_val_input(unsigned t, int wrt, const double* c) const {
        double v = 1.0, wrtDeriv = (wrt == -1);
        unsigned e = ms.starts[t + 1] - 1;
        for (unsigned i = ms.starts[t]; i < e; ++i) {
            if (int(ms.dims[i]) == wrt) {
                wrtDeriv = c[startPos[dim + wrt] + ms.orders[i]];
            } else {
                v *= c[startPos[ms.dims[i]] + ms.orders[i]];
            }
        }
        int d = ms.dims[e];
        double last;
        if (wrt == d) {
            last = 1;
            wrtDeriv = c[startPos[dim + d] + ms.orders[e]];
        } else {
            last = c[startPos[d] + ms.orders[e]];
        }
        v *= last * wrtDeriv;
        return v;
    }
    double term_val_mixed(unsigned t, int wrt, const double* c) const {
        double v = 1.0;
        double wrtDeriv = (wrt == int(dim) - 1 || wrt == -1);
        unsigned e = ms.starts[t + 1] - 1;
        if (ms.dims[e] != dim - 1) return 0.;
        for (unsigned i = ms.starts[t]; i < e; ++i) {
            if (int(ms.dims[i]) == wrt) {
                wrtDeriv = c[startPos[dim + wrt] + ms.orders[i]];
            } else {
                v *= c[startPos[ms.dims[i]] + ms.orders[i]];
            }
        }
        int posInd = (wrt == int(dim) - 1) ? (2 * dim) : (2 * dim - 1);
        double dv = c[startPos[posInd] + ms.orders[e]];
        v *= (dv * wrtDeriv);
        return v;
    }
    double term_val_mixed_coeff(unsigned t, const double* c, unsigned posIndex) const {
        double v = 1.0;
        if (is_const(t)) return 0.;
        unsigned e = ms.starts[t + 1] - 1;
        if (ms.dims[e] != dim - 1) return 0.;
        for (unsigned i = ms.starts[t]; i < e; ++i) v *= c[startPos[ms.dims[i]] + ms.orders[i]];
        double dv = c[startPos[posIndex] + ms.orders[e]];
        v *= dv;
        return v;
    }

    // Evaluate (204-223)
    double evaluate(const double* c, const double* w) const {
        double out = 0.0;
        for (unsigned t = 0; t < nterms(); ++t) {
            if (is_const(t)) { out += w[t]; continue; }
            out += term_val(t, c) * w[t];
        }
        return out;
    }
    // DiagonalDerivative (233-257)
    double diag_deriv(const double* c, const double* w, unsigned order) const {
        double out = 0.0;
        unsigned pos = 2 * dim + order - 2;
        for (unsigned t = 0; t < nterms(); ++t) {
            if (is_const(t)) continue;
            out += term_val_diag(t, c, pos) * w[t];
        }
        return out;
    }
    // CoeffDerivative (272-292)
    double coeff_deriv(const double* c, const double* w, double* grad, long gs) const {
        double f = 0;
        for (unsigned t = 0; t < nterms(); ++t) {
            if (is_const(t)) { f += w[t]; grad[t * gs] = 1.0; continue; }
            double v = term_val(t, c);
            f += v * w[t];
            grad[t * gs] = v;
        }
        return f;
    }
    // InputDerivative (300-329)
    double input_deriv(const double* c, const double* w, double* grad, long gs) const {
        double f = 0.0;
        for (int wrt = -1; wrt < int(dim); ++wrt) {
            if (wrt >= 0) grad[wrt * gs] = 0;
            for (unsigned t = 0; t < nterms(); ++t) {
                if (is_const(t)) { if (wrt < 0) f += w[t]; continue; }
                double v = term_val_input(t, wrt, c);
                if (wrt < 0) f += v * w[t]; else grad[wrt * gs] += v * w[t];
            }
        }
        return f;
    }
    // MixedInputDerivative (337-366)
    double mixed_input_deriv(const double* c, const double* w, double* grad, long gs) const {
        double df = 0.0;
        for (int wrt = -1; wrt < int(dim); ++wrt) {
            if (wrt >= 0) grad[wrt * gs] = 0;
            for (unsigned t = 0; t < nterms(); ++t) {
                if (is_const(t)) continue;
                double v = term_val_mixed(t, wrt, c);
                if (wrt < 0) df += v * w[t]; else grad[wrt * gs] += v * w[t];
            }
        }
        return df;
    }
    // MixedCoeffDerivative (374-397)
    double mixed_coeff_deriv(const double* c, const double* w, unsigned order, double* grad, long gs) const {
        double df = 0;
        unsigned pos = 2 * dim + order - 2;
        for (unsigned t = 0; t < nterms(); ++t) {
            double v = term_val_mixed_coeff(t, c, pos);
            df += v * w[t];
            grad[t * gs] = v;
        }
        return df;
    }
};

// ---------------------------------------------------------------------------
// Integrand g(d_d f(x_{<d}, t x_d)) x_d and its mode-dependent derivatives,
// MParT/MonotoneIntegrand.h:118-203.
struct Integrand {
    double* cache;
    const Worker* ex;
    PtView pt;
    double xd;
    const double* w;
    int mode;
    double nugget;
    PosFunc g;
    double* work;  // MixedCoeff scratch
    long* nevals;

    unsigned fdim() const {
        unsigned n = 1;
        if (mode == D_DIAGONAL) n++;
        if (mode == D_PARAMETERS || mode == D_MIXEDCOEFF) n += ex->nterms();
        if (mode == D_INPUT || mode == D_MIXEDINPUT) n += ex->dim;
        return n;
    }

    void operator()(double t, double* out) const {
        if (nevals) (*nevals)++;
        const unsigned nt = ex->nterms();
        const unsigned dim = pt.n;
        if (mode == D_DIAGONAL || mode == D_MIXEDCOEFF || mode == D_INPUT)
            ex->fill_cache2(cache, t * xd, D_DIAGONAL2);
        else
            ex->fill_cache2(cache, t * xd, D_DIAGONAL);

        double df = 0;
        if (mode == D_PARAMETERS) {
            df = ex->mixed_coeff_deriv(cache, w, 1, &out[1], 1);
            double scale = xd * g.deriv(df);
            for (unsigned i = 0; i < nt; ++i) out[1 + i] *= scale;
        } else if (mode == D_MIXEDCOEFF) {
            df = ex->diag_deriv(cache, w, 1);
            double dgdf = g.deriv(df);
            double df2 = ex->mixed_coeff_deriv(cache, w, 2, work, 1);
            double scale = xd * t * dgdf;
            for (unsigned i = 0; i < nt; ++i) work[i] *= scale;
            df = ex->mixed_coeff_deriv(cache, w, 1, &out[1], 1);
            scale = xd * t * df2 * g.second(df) + dgdf;
            for (unsigned i = 0; i < nt; ++i) out[1 + i] = scale * out[1 + i] + work[i];
        } else if (mode == D_INPUT) {
            df = ex->mixed_input_deriv(cache, w, &out[1], 1);
            double scale = xd * g.deriv(df);
            for (unsigned i = 0; i + 1 < dim; ++i) out[1 + i] *= scale;
        } else {
            df = ex->diag_deriv(cache, w, 1);
        }
        double gf = g.eval(df);
        out[0] = xd * (gf + nugget);
        if (std::isinf(gf)) throw std::runtime_error("MonotoneIntegrand: value of g(df(...)) is inf");
        if (mode == D_DIAGONAL || mode == D_INPUT) {
            unsigned ind = (mode == D_DIAGONAL) ? 1 : dim;
            out[ind] = ex->diag_deriv(cache, w, 2);
            out[ind] *= xd * t * g.deriv(df);
            out[ind] += (gf + nugget);
        }
    }
};

// ---------------------------------------------------------------------------
// Quadrature.  Clenshaw-Curtis rule (MParT/Quadrature.h:212-258).
static void cc_rule(unsigned numPts, double* wts, double* pts) {
    if (numPts == 0) return;
    if (numPts == 1) {
        pts[0] = 0.0;
        wts[0] = 2.0;
        return;
    }
    for (unsigned i = 0; i < numPts; ++i) pts[i] = std::cos((numPts - (i + 1)) * M_PI / (numPts - 1));
    pts[0] = -1.0;
    if (numPts % 2 == 1) pts[(numPts - 1) / 2] = 0.0;
    pts[numPts - 1] = 1.0;
    double theta;
    for (unsigned i = 0; i < numPts; ++i) {
        wts[i] = 1.0;
        theta = i * M_PI / (numPts - 1);
        for (unsigned j = 0; j < (numPts - 1) / 2; ++j) {
            double b = (2 * (j + 1) == (numPts - 1)) ? 1.0 : 2.0;
            wts[i] = wts[i] - b * std::cos(2.0 * (j + 1) * theta) / (4 * (j + 1) * (j + 1) - 1);
        }
    }
    wts[0] = wts[0] / (numPts - 1.0);
    for (unsigned i = 1; i < (numPts - 1); ++i) wts[i] = 2.0 * wts[i] / (numPts - 1.0);
    wts[numPts - 1] = wts[numPts - 1] / (numPts - 1.0);
}

struct Quad {
    int type = QUAD_AS;
    unsigned maxSub = 30, minSub = 0;
    double absTol = 1e-6, relTol = 1e-6;
    std::vector<double> cPts, cWts, fPts, fWts;  // ACC nested rules

    void init_acc(unsigned quadPts) {
        // level = floor(log2(quadPts-2)) (src/MapFactoryImpl4.cpp:18).
        unsigned level = (unsigned)std::log2(quadPts - 2);
        unsigned nc = (unsigned)std::pow(2, level) + 1, nf = (unsigned)std::pow(2, level + 1) + 1;
        cPts.resize(nc); cWts.resize(nc); fPts.resize(nf); fWts.resize(nf);
        cc_rule(nc, cWts.data(), cPts.data());
        cc_rule(nf, fWts.data(), fPts.data());
    }
    size_t workspace(unsigned fdim) const {
        if (type == QUAD_CC) return fdim;
        if (type == QUAD_AS) return (2 * maxSub + 5) * fdim + 2 * maxSub;
        return (maxSub + 5) * fdim + 2 * maxSub;
    }

    // RecursiveQuadratureBase::EstimateError with QuadError::First
    // (Quadrature.h:441-481); every factory registration uses First.
    void est_error(const double* coarse, const double* fine, double& err, double& tol) const {
        err = std::fabs(fine[0] - coarse[0]);
        double ref = std::fabs(coarse[0]);
        tol = std::fmax(relTol * ref, absTol);
    }

    // AdaptiveSimpson::Integrate (Quadrature.h:553-676) + UpdateValues (698-714).
    template <class F>
    void integrate_as(double* ws, const F& f, unsigned fdim, double lb, double ub, double* res) const {
        for (unsigned i = 0; i < fdim; ++i) res[i] = 0;
        unsigned currLevel = 0, currSegment = 0;
        double midPt = 0.5 * (lb + ub);
        double* leftFunc = &ws[0];
        f(lb, leftFunc);
        double* rightFunc = &ws[fdim];
        f(ub, rightFunc);
        double* midFunc = &ws[2 * fdim];
        f(midPt, midFunc);
        double* intCoarse = &ws[3 * fdim];
        double* intFine = &ws[4 * fdim];
        double* leftPt = &ws[5 * fdim];
        *leftPt = lb;
        double* rightPt = &ws[5 * fdim + 1];
        *rightPt = ub;
        double leftMidPt, rightMidPt;
        double *leftMidFunc, *rightMidFunc;
        unsigned workStart = 5 * fdim, prevStart = 0;
        double error, errorTol;
        while (true) {
            leftPt = &ws[workStart];
            rightPt = &ws[workStart + 1];
            midPt = 0.5 * ((*leftPt) + (*rightPt));
            leftMidPt = 0.5 * ((*leftPt) + midPt);
            rightMidPt = 0.5 * (midPt + (*rightPt));
            leftMidFunc = &ws[workStart + 2];
            f(leftMidPt, leftMidFunc);
            rightMidFunc = &ws[workStart + 2 + fdim];
            f(rightMidPt, rightMidFunc);
            for (unsigned i = 0; i < fdim; ++i) {
                intCoarse[i] = (((*rightPt) - (*leftPt)) / 6.0) * (leftFunc[i] + 4.0 * midFunc[i] + rightFunc[i]);
                intFine[i] = ((midPt - (*leftPt)) / 6.0) * (leftFunc[i] + 4.0 * leftMidFunc[i] + midFunc[i]);
                intFine[i] += (((*rightPt) - midPt) / 6.0) * (midFunc[i] + 4.0 * rightMidFunc[i] + rightFunc[i]);
            }
            est_error(intCoarse, intFine, error, errorTol);
            if ((((error < errorTol) || (currLevel == maxSub - 1)) && (currLevel >= minSub)) ||
                (std::abs(ub - lb) < 1e-14)) {
                for (unsigned i = 0; i < fdim; ++i) res[i] += intFine[i];
                while (((currSegment >> currLevel) & 1U) && (currLevel > 0)) {
                    currSegment &= ~(1UL << currLevel);
                    currLevel--;
                }
                if (currLevel == 0) break;
                currSegment |= 1UL << currLevel;
                workStart = (2 * currLevel + 5) * fdim + 2 * currLevel;
                prevStart = (2 * (currLevel - 1) + 5) * fdim + 2 * (currLevel - 1);
                ws[workStart] = 0.5 * (ws[prevStart] + ws[prevStart + 1]);
                ws[workStart + 1] = ws[prevStart + 1];
                // UpdateValues
                leftFunc = &ws[0];
                midFunc = &ws[2 * fdim];
                rightFunc = &ws[fdim];
                for (unsigned level = 0; level < currLevel; ++level) {
                    if ((currSegment >> (level + 1)) & 1U) {
                        leftFunc = midFunc;
                        midFunc = &ws[(2 * level + 5 + 1) * fdim + 2 * level + 2];
                    } else {
                        rightFunc = midFunc;
                        midFunc = &ws[(2 * level + 5) * fdim + 2 * level + 2];
                    }
                }
            } else {
                currLevel++;
                rightFunc = midFunc;
                midFunc = leftMidFunc;
                workStart = (2 * currLevel + 5) * fdim + 2 * currLevel;
                ws[workStart] = *leftPt;
                ws[workStart + 1] = midPt;
            }
        }
    }

    // AdaptiveClenshawCurtis::Integrate (Quadrature.h:819-968) + UpdateValues (990-1004).
    template <class F>
    void integrate_acc(double* ws, const F& f, unsigned fdim, double lb, double ub, double* res) const {
        for (unsigned i = 0; i < fdim; ++i) res[i] = 0;
        unsigned currLevel = 0, currSegment = 0;
        double midPt = 0.5 * (lb + ub);
        double scale = (ub - lb);
        double* leftFunc = &ws[0];
        f(lb, leftFunc);
        double* rightFunc = &ws[fdim];
        f(ub, rightFunc);
        double* fval = &ws[2 * fdim];
        double* intCoarse = &ws[3 * fdim];
        double* intFine = &ws[4 * fdim];
        double* leftPt = &ws[5 * fdim];
        *leftPt = lb;
        double* rightPt = &ws[5 * fdim + 1];
        *rightPt = ub;
        double* midFunc;
        unsigned workStart = 5 * fdim, prevStart = 0;
        double error, errorTol;
        const unsigned cRight = unsigned(cWts.size()) - 1, cMid = cRight / 2;
        const unsigned fRight = unsigned(fWts.size()) - 1, fMid = fRight / 2;
        while (true) {
            leftPt = &ws[workStart];
            rightPt = &ws[workStart + 1];
            midPt = 0.5 * ((*leftPt) + (*rightPt));
            scale = 0.5 * ((*rightPt) - (*leftPt));
            midFunc = &ws[workStart + 2];
            f(midPt, midFunc);
            for (unsigned i = 0; i < fdim; ++i) {
                intCoarse[i] = scale * cWts[0] * leftFunc[i];
                intCoarse[i] += scale * cWts[cMid] * midFunc[i];
                intCoarse[i] += scale * cWts[cRight] * rightFunc[i];
                intFine[i] = scale * fWts[0] * leftFunc[i];
                intFine[i] += scale * fWts[fMid] * midFunc[i];
                intFine[i] += scale * fWts[fRight] * rightFunc[i];
            }
            for (unsigned i = 1; i < cMid; ++i) {
                f(midPt + scale * cPts[i], fval);
                for (unsigned j = 0; j < fdim; ++j) {
                    intCoarse[j] += scale * cWts[i] * fval[j];
                    intFine[j] += scale * fWts[2 * i] * fval[j];
                }
            }
            for (unsigned i = cMid + 1; i < cRight; ++i) {
                f(midPt + scale * cPts[i], fval);
                for (unsigned j = 0; j < fdim; ++j) {
                    intCoarse[j] += scale * cWts[i] * fval[j];
                    intFine[j] += scale * fWts[2 * i] * fval[j];
                }
            }
            for (unsigned i = 1; i < fPts.size(); i += 2) {
                f(midPt + scale * fPts[i], fval);
                for (unsigned j = 0; j < fdim; ++j) intFine[j] += scale * fWts[i] * fval[j];
            }
            est_error(intCoarse, intFine, error, errorTol);
            if ((((error < errorTol) || (currLevel == maxSub - 1)) && (currLevel >= minSub)) ||
                (std::abs(ub - lb) < 1e-14)) {
                for (unsigned i = 0; i < fdim; ++i) res[i] += intFine[i];
                while (((currSegment >> currLevel) & 1U) && (currLevel > 0)) {
                    currSegment &= ~(1UL << currLevel);
                    currLevel--;
                }
                if (currLevel == 0) break;
                currSegment |= 1UL << currLevel;
                workStart = (currLevel + 5) * fdim + 2 * currLevel;
                prevStart = ((currLevel - 1) + 5) * fdim + 2 * (currLevel - 1);
                ws[workStart] = 0.5 * (ws[prevStart] + ws[prevStart + 1]);
                ws[workStart + 1] = ws[prevStart + 1];
                leftFunc = &ws[0];
                rightFunc = &ws[fdim];
                for (unsigned level = 0; level < currLevel; ++level) {
                    if ((currSegment >> (level + 1)) & 1U)
                        leftFunc = &ws[(level + 5) * fdim + 2 * (level + 1)];
                    else
                        rightFunc = &ws[(level + 5) * fdim + 2 * (level + 1)];
                }
            } else {
                currLevel++;
                rightFunc = midFunc;
                workStart = (currLevel + 5) * fdim + 2 * currLevel;
                ws[workStart] = *leftPt;
                ws[workStart + 1] = midPt;
            }
        }
    }

    // ClenshawCurtisQuadrature::Integrate (Quadrature.h:279-307) with the rule
    // of GetRule(quadPts) (212-258).
    std::vector<double> ccPts, ccWts;
    void init_cc(unsigned quadPts) {
        ccPts.resize(quadPts);
        ccWts.resize(quadPts);
        cc_rule(quadPts, ccWts.data(), ccPts.data());
    }
    template <class F>
    void integrate_cc(double* ws, const F& f, unsigned fdim, double lb, double ub, double* res) const {
        double midpoint_tol = 15.0 * std::numeric_limits<double>::epsilon();
        if ((ub - lb) < midpoint_tol) {
            f(0.5 * (ub + lb), res);
            for (unsigned i = 0; i < fdim; ++i) res[i] *= (ub - lb);
            return;
        }
        for (unsigned j = 0; j < fdim; ++j) res[j] = 0.0;
        double* fval = ws;
        for (unsigned i = 0; i < ccPts.size(); ++i) {
            f(0.5 * (ub + lb + (ub - lb) * ccPts[i]), fval);
            for (unsigned j = 0; j < fdim; ++j) res[j] += 0.5 * (ub - lb) * ccWts[i] * fval[j];
        }
    }

    template <class F>
    void integrate(double* ws, const F& f, unsigned fdim, double lb, double ub, double* res) const {
        if (type == QUAD_AS) integrate_as(ws, f, fdim, lb, ub, res);
        else if (type == QUAD_CC) integrate_cc(ws, f, fdim, lb, ub, res);
        else integrate_acc(ws, f, fdim, lb, ub, res);
    }
};

// ---------------------------------------------------------------------------
// ITP bracketed root finding, MParT/Utilities/RootFinding.h:20-178.
template <class F>
static void find_bracket(F& f, double& xlb, double& ylb, double& xub, double& yub, double yd, int& info) {
    const unsigned maxIts = 128;
    double stepSize = 1.0;
    info = 0;
    ylb = f(xlb);
    yub = f(xub);
    if (ylb > yd) {
        std::swap(ylb, yub);
        std::swap(xlb, xub);
        unsigned i;
        for (i = 0; i < maxIts; ++i) {
            xlb = xub - stepSize;
            ylb = f(xlb);
            if ((std::fabs((yub - ylb) / (xub - xlb)) < 1e-12) && ((xub - xlb) > 10)) { info = -1; return; }
            if (ylb > yd) { yub = ylb; xub = xlb; stepSize *= 2.0; } else break;
        }
        if (i >= maxIts) info = -2;
    } else {
        unsigned i;
        for (i = 0; i < maxIts; ++i) {
            xub = xlb + stepSize;
            yub = f(xub);
            if ((std::fabs((yub - ylb) / (xub - xlb)) < 1e-12) && ((xub - xlb) > 10)) { info = -1; return; }
            if (yub < yd) { ylb = yub; xlb = xub; stepSize *= 2.0; } else break;
        }
        if (i >= maxIts) info = -2;
    }
}

static double find_x_itp(double xlb, double xub, double yd, double ylb, double yub, double k1, double k2,
                         double nhalf, double n0, int it, double xtol) {
    double xb = 0.5 * (xub + xlb);
    double xf = (xub * ylb - xlb * yub) / (ylb - yub);
    double sigma = ((xb - xf) > 0) ? 1.0 : -1.0;
    double delta = std::fmin(k1 * std::pow((xub - xlb), k2), std::fabs(xb - xf));
    xf += delta * sigma;
    double rho = std::fmin(xtol * std::pow(2.0, nhalf + n0 - it) - 0.5 * (xub - xlb), std::fabs(xf - xb));
    return xb - sigma * rho;
}

template <class F>
static double inverse_single_bracket(double yd, F& f, double x0, double xtol, double ftol, int& info) {
    const unsigned maxIts = 10000;
    double xlb, xub, ylb, yub, xc, yc;
    info = 0;
    xlb = xub = x0;
    ylb = yub = f(xlb);
    int binfo = 0;
    find_bracket(f, xlb, ylb, xub, yub, yd, binfo);
    if ((binfo < 0) || ((ylb > yd) || (yub < yd))) {
        info = -2;
        return std::numeric_limits<double>::quiet_NaN();
    }
    const double k1 = 0.1, k2 = 2.0, n0 = 1.0;
    const double nhalf = std::ceil(std::log2(0.5 * (xub - xlb) / xtol));
    unsigned it;
    for (it = 0; it < maxIts; ++it) {
        xc = find_x_itp(xlb, xub, yd, ylb, yub, k1, k2, nhalf, n0, it, xtol);
        yc = f(xc);
        if (std::fabs(yc - yd) < ftol) return xc;
        else if (yc > yd) { std::swap(xc, xub); std::swap(yc, yub); }
        else { std::swap(xc, xlb); std::swap(yc, ylb); }
        if (((xub - xlb) < xtol) || ((yub - ylb) < ftol)) break;
    }
    return 0.5 * (xub + xlb);
}

// ---------------------------------------------------------------------------
// Maps.
struct Map {
    unsigned inDim = 0, outDim = 0, numCoeffs = 0;
    std::vector<double> coeffs;  // owned copy (SetCoeffs semantics)
    virtual ~Map() {}
};

// MonotoneComponent, MParT/MonotoneComponent.h:47-1091.
struct Component : Map {
    Worker ex;
    Quad quad;
    PosFunc g;
    bool contDeriv = true;
    double nugget = 0.0;

    unsigned dim() const { return ex.dim; }

    struct Scratch {
        std::vector<double> cache, ws, integral, work;
    };
    Scratch make_scratch(unsigned fdim) const {
        Scratch s;
        s.cache.assign(ex.cacheSize, 0.0);
        s.ws.assign(quad.workspace(fdim), 0.0);
        s.integral.assign(fdim, 0.0);
        s.work.assign(ex.nterms() + 1, 0.0);
        return s;
    }

    Integrand integrand(Scratch& s, const PtView& pt, double xd, const double* w, int mode, double nug,
                        long* ne) const {
        Integrand I;
        I.cache = s.cache.data(); I.ex = &ex; I.pt = pt; I.xd = xd; I.w = w; I.mode = mode;
        I.nugget = nug; I.g = g; I.work = s.work.data(); I.nevals = ne;
        return I;
    }

    // EvaluateSingle (993-1017)
    double evaluate_single(Scratch& s, const PtView& pt, double xd, const double* w, double nug, long* ne) const {
        double out = 0.0;
        Integrand I = integrand(s, pt, xd, w, D_NONE, nug, ne);
        quad.integrate(s.ws.data(), I, 1, 0.0, 1.0, &out);
        ex.fill_cache2(s.cache.data(), 0.0, D_NONE);
        out += ex.evaluate(s.cache.data(), w);
        return out;
    }

    // EvaluateImpl (255-299); note: the nugget is not passed (line 287).
    void evaluate(const CMat& pts, const double* w, double* out, long os, long* nevals) const {
        long N = pts.cols;
        OmpError err;
#pragma omp parallel
        {
            Scratch s = make_scratch(1);
#pragma omp for schedule(dynamic, 64)
            for (long n = 0; n < N; ++n) {
                if (err.failed()) continue;
                try {
                    PtView pt{&pts.p[n * pts.cs], pts.rs, dim()};
                    ex.fill_cache1(s.cache.data(), pt, D_NONE);
                    long ne = 0;
                    out[n * os] = evaluate_single(s, pt, pt(dim() - 1), w, 0.0, nevals ? &ne : nullptr);
                    if (nevals) nevals[n] = ne;
                } catch (...) {
                    err.keep();
                }
            }
        }
        err.rethrow();
    }

    // ContinuousDerivative (488-531)
    void continuous_derivative(const CMat& pts, const double* w, double* derivs) const {
        long N = pts.cols;
        OmpError err;
#pragma omp parallel
        {
            Scratch s = make_scratch(1);
#pragma omp for schedule(dynamic, 64)
            for (long n = 0; n < N; ++n) {
                if (err.failed()) continue;
                try {
                    PtView pt{&pts.p[n * pts.cs], pts.rs, dim()};
                    ex.fill_cache1(s.cache.data(), pt, D_NONE);
                    ex.fill_cache2(s.cache.data(), pt(dim() - 1), D_DIAGONAL);
                    derivs[n] = ex.diag_deriv(s.cache.data(), w, 1);
                    derivs[n] = g.eval(derivs[n]);
                } catch (...) {
                    err.keep();
                }
            }
        }
        err.rethrow();
    }

    // DiscreteDerivative (586-638)
    void discrete_derivative(const CMat& pts, const double* w, double* evals, double* derivs) const {
        long N = pts.cols;
        OmpError err;
#pragma omp parallel
        {
            Scratch s = make_scratch(2);
#pragma omp for schedule(dynamic, 64)
            for (long n = 0; n < N; ++n) {
                if (err.failed()) continue;
                try {
                    PtView pt{&pts.p[n * pts.cs], pts.rs, dim()};
                    ex.fill_cache1(s.cache.data(), pt, D_NONE);
                    Integrand I = integrand(s, pt, pt(dim() - 1), w, D_DIAGONAL, nugget, nullptr);
                    double both[2];
                    quad.integrate(s.ws.data(), I, 2, 0.0, 1.0, both);
                    evals[n] = both[0];
                    derivs[n] = both[1];
                    ex.fill_cache2(s.cache.data(), 0.0, D_NONE);
                    evals[n] += ex.evaluate(s.cache.data(), w);
                } catch (...) {
                    err.keep();
                }
            }
        }
        err.rethrow();
    }

    // LogDeterminantImpl (105-125)
    void log_determinant(const CMat& pts, const double* w, double* out, long os) const {
        long N = pts.cols;
        std::vector<double> d(N);
        if (contDeriv) continuous_derivative(pts, w, d.data());
        else {
            std::vector<double> ev(N);
            discrete_derivative(pts, w, ev.data(), d.data());
        }
        for (long n = 0; n < N; ++n)
            out[n * os] = (d[n] <= 0) ? -std::numeric_limits<double>::infinity() : std::log(d[n]);
    }

    // CoeffJacobian (685-741)
    void coeff_jacobian(const CMat& pts, const double* w, double* evals, const Mat& jac) const {
        long N = pts.cols;
        unsigned nt = ex.nterms();
        OmpError err;
#pragma omp parallel
        {
            Scratch s = make_scratch(nt + 1);
#pragma omp for schedule(dynamic, 64)
            for (long n = 0; n < N; ++n) {
                if (err.failed()) continue;
                try {
                    PtView pt{&pts.p[n * pts.cs], pts.rs, dim()};
                    double* jv = &jac.p[n * jac.cs];
                    ex.fill_cache1(s.cache.data(), pt, D_NONE);
                    Integrand I = integrand(s, pt, pt(dim() - 1), w, D_PARAMETERS, nugget, nullptr);
                    quad.integrate(s.ws.data(), I, nt + 1, 0.0, 1.0, s.integral.data());
                    evals[n] = s.integral[0];
                    ex.fill_cache2(s.cache.data(), 0.0, D_NONE);
                    evals[n] += ex.coeff_deriv(s.cache.data(), w, jv, jac.rs);
                    for (unsigned t = 0; t < nt; ++t) jv[t * jac.rs] += s.integral[t + 1];
                } catch (...) {
                    err.keep();
                }
            }
        }
        err.rethrow();
    }

    // InputJacobian (758-818)
    void input_jacobian(const CMat& pts, const double* w, double* evals, const Mat& jac) const {
        long N = pts.cols;
        unsigned D = dim();
        OmpError err;
#pragma omp parallel
        {
            Scratch s = make_scratch(D + 1);
#pragma omp for schedule(dynamic, 64)
            for (long n = 0; n < N; ++n) {
                if (err.failed()) continue;
                try {
                    PtView pt{&pts.p[n * pts.cs], pts.rs, D};
                    double* jv = &jac.p[n * jac.cs];
                    ex.fill_cache1(s.cache.data(), pt, D_INPUT);
                    Integrand I = integrand(s, pt, pt(D - 1), w, D_INPUT, nugget, nullptr);
                    quad.integrate(s.ws.data(), I, D + 1, 0.0, 1.0, s.integral.data());
                    evals[n] = s.integral[0];
                    ex.fill_cache2(s.cache.data(), 0.0, D_INPUT);
                    evals[n] += ex.input_deriv(s.cache.data(), w, jv, jac.rs);
                    for (unsigned d = 0; d + 1 < D; ++d) jv[d * jac.rs] += s.integral[d + 1];
                    jv[(D - 1) * jac.rs] = s.integral[D];
                } catch (...) {
                    err.keep();
                }
            }
        }
        err.rethrow();
    }

    // ContinuousMixedJacobian (825-874)
    void continuous_mixed_jacobian(const CMat& pts, const double* w, const Mat& jac) const {
        long N = pts.cols;
        OmpError err;
#pragma omp parallel
        {
            Scratch s = make_scratch(1);
#pragma omp for schedule(dynamic, 64)
            for (long n = 0; n < N; ++n) {
                if (err.failed()) continue;
                try {
                    PtView pt{&pts.p[n * pts.cs], pts.rs, dim()};
                    double* jv = &jac.p[n * jac.cs];
                    ex.fill_cache1(s.cache.data(), pt, D_NONE);
                    ex.fill_cache2(s.cache.data(), pt(dim() - 1), D_DIAGONAL);
                    double df = ex.mixed_coeff_deriv(s.cache.data(), w, 1, jv, jac.rs);
                    double dgdf = g.deriv(df);
                    for (unsigned i = 0; i < ex.nterms(); ++i) jv[i * jac.rs] *= dgdf;
                } catch (...) {
                    err.keep();
                }
            }
        }
        err.rethrow();
    }

    // ContinuousMixedInputJacobian (877-926)
    void continuous_mixed_input_jacobian(const CMat& pts, const double* w, const Mat& jac) const {
        long N = pts.cols;
        unsigned D = dim();
        OmpError err;
#pragma omp parallel
        {
            Scratch s = make_scratch(1);
#pragma omp for schedule(dynamic, 64)
            for (long n = 0; n < N; ++n) {
                if (err.failed()) continue;
                try {
                    PtView pt{&pts.p[n * pts.cs], pts.rs, D};
                    double* jv = &jac.p[n * jac.cs];
                    ex.fill_cache1(s.cache.data(), pt, D_MIXEDINPUT);
                    ex.fill_cache2(s.cache.data(), pt(D - 1), D_MIXEDINPUT);
                    double df = ex.mixed_input_deriv(s.cache.data(), w, jv, jac.rs);
                    double dgdf = g.deriv(df);
                    for (unsigned i = 0; i < D; ++i) jv[i * jac.rs] *= dgdf;
                } catch (...) {
                    err.keep();
                }
            }
        }
        err.rethrow();
    }

    // DiscreteMixedJacobian (929-980)
    void discrete_mixed_jacobian(const CMat& pts, const double* w, const Mat& jac) const {
        long N = pts.cols;
        unsigned nt = ex.nterms();
        OmpError err;
#pragma omp parallel
        {
            Scratch s = make_scratch(nt + 1);
#pragma omp for schedule(dynamic, 64)
            for (long n = 0; n < N; ++n) {
                if (err.failed()) continue;
                try {
                    PtView pt{&pts.p[n * pts.cs], pts.rs, dim()};
                    double* jv = &jac.p[n * jac.cs];
                    ex.fill_cache1(s.cache.data(), pt, D_NONE);
                    Integrand I = integrand(s, pt, pt(dim() - 1), w, D_MIXEDCOEFF, nugget, nullptr);
                    quad.integrate(s.ws.data(), I, nt + 1, 0.0, 1.0, s.integral.data());
                    for (unsigned t = 0; t < nt; ++t) jv[t * jac.rs] += s.integral[t + 1];
                } catch (...) {
                    err.keep();
                }
            }
        }
        err.rethrow();
    }

    // GradientImpl (148-164)
    void gradient(const CMat& pts, const double* w, const double* sens, long ss, const Mat& out) const {
        long N = pts.cols;
        std::vector<double> ev(N);
        input_jacobian(pts, w, ev.data(), out);
        for (long n = 0; n < N; ++n)
            for (unsigned i = 0; i < dim(); ++i) out(i, n) *= sens[n * ss];
    }
    // CoeffGradImpl (166-182)
    void coeff_grad(const CMat& pts, const double* w, const double* sens, long ss, const Mat& out) const {
        long N = pts.cols;
        std::vector<double> ev(N);
        coeff_jacobian(pts, w, ev.data(), out);
        for (long n = 0; n < N; ++n)
            for (unsigned i = 0; i < ex.nterms(); ++i) out(i, n) *= sens[n * ss];
    }
    // LogDeterminantCoeffGradImpl (185-206)
    void log_determinant_coeff_grad(const CMat& pts, const double* w, const Mat& out) const {
        long N = pts.cols;
        std::vector<double> d(N);
        if (contDeriv) {
            continuous_mixed_jacobian(pts, w, out);
            continuous_derivative(pts, w, d.data());
        } else {
            std::vector<double> ev(N);
            for (long n = 0; n < N; ++n)
                for (unsigned i = 0; i < ex.nterms(); ++i) out(i, n) = 0.0;
            discrete_mixed_jacobian(pts, w, out);
            discrete_derivative(pts, w, ev.data(), d.data());
        }
        for (long n = 0; n < N; ++n)
            for (unsigned i = 0; i < ex.nterms(); ++i) out(i, n) *= 1.0 / d[n];
    }
    // LogDeterminantInputGradImpl (208-232)
    void log_determinant_input_grad(const CMat& pts, const double* w, const Mat& out) const {
        if (!contDeriv) throw std::invalid_argument("Discrete derivative version is not implemented yet (To Do)");
        long N = pts.cols;
        std::vector<double> d(N);
        continuous_mixed_input_jacobian(pts, w, out);
        continuous_derivative(pts, w, d.data());
        for (long n = 0; n < N; ++n)
            for (unsigned i = 0; i < dim(); ++i) out(i, n) *= 1.0 / d[n];
    }

    // InverseImpl (331-443); xs has 1 or N columns, its last row is x0.
    void inverse_n(const CMat& xs, const double* ys, long ys_s, long numPts, const double* w, double* out, long os,
                   double xtol, double ytol) const {
        if ((ytol <= std::numeric_limits<double>::epsilon()) && (xtol <= std::numeric_limits<double>::epsilon()))
            throw std::invalid_argument("Either xtol or ytol must be nonzero");
        if (xtol < 0 || ytol < 0) throw std::invalid_argument("Invalid tolerance");
        long numXs = xs.cols;
        if ((numXs != 1) && (numXs != numPts)) throw std::invalid_argument("Invalid argument sizes given to MonotoneComponent::Inverse");
        OmpError err;
#pragma omp parallel
        {
            Scratch s = make_scratch(1);
#pragma omp for schedule(dynamic, 16)
            for (long n = 0; n < numPts; ++n) {
                if (err.failed()) continue;
                try {
                    long xi = (numXs == 1) ? 0 : n;
                    PtView pt{&xs.p[xi * xs.cs], xs.rs, unsigned(xs.rows)};
                    bool hasNan = false;
                    for (unsigned i = 0; i < pt.n; ++i)
                        if (std::isnan(pt(i))) { hasNan = true; break; }
                    if (hasNan) { out[n * os] = std::numeric_limits<double>::quiet_NaN(); continue; }
                    ex.fill_cache1(s.cache.data(), pt, D_NONE);
                    PtView ptd{pt.p, pt.stride, dim()};
                    auto f = [&](double x) { return evaluate_single(s, ptd, x, w, nugget, nullptr); };
                    int info;
                    out[n * os] = inverse_single_bracket(ys[n * ys_s], f, pt(pt.n - 1), xtol, ytol, info);
                } catch (...) {
                    err.keep();
                }
            }
        }
        err.rethrow();
    }
};

// TriangularMap, src/TriangularMap.cpp:10-310.
struct Triangular : Map {
    std::vector<std::unique_ptr<Component>> comps;
    std::vector<unsigned> coeffOff;

    void finalize() {
        inDim = comps.back()->inDim;
        outDim = 0;
        numCoeffs = 0;
        coeffOff.clear();
        for (auto& c : comps) {
            coeffOff.push_back(numCoeffs);
            outDim += c->outDim;
            numCoeffs += c->numCoeffs;
        }
        for (size_t i = 1; i < comps.size(); ++i)
            if (comps[i]->inDim != comps[i - 1]->inDim + comps[i]->outDim)
                throw std::invalid_argument("TriangularMap: inconsistent component input dimensions");
        coeffs.assign(numCoeffs, 0.0);
    }
};

static CMat sub_rows(const CMat& m, long r) { return CMat{m.p, r, m.cols, m.rs, m.cs}; }

}  // namespace orc

// ---------------------------------------------------------------------------
// C entry points (ctypes).
using namespace orc;

static thread_local std::string g_err;

static Component* make_component(const Options* o, const MSet& ms) {
    auto* c = new Component();
    c->ex.ms = ms;
    c->ex.basis.type = o->basisType;
    c->ex.basis.poly.phys = (o->basisType == PHYS_HERMITE);
    c->ex.basis.poly.normalize = (o->basisType != HERMITE_FN) && o->basisNorm;
    // isLinearized = !isinf(basisLB) || !isinf(basisUB) (MParT/MapFactory.h:226)
    c->ex.basis.lin = !std::isinf(o->basisLB) || !std::isinf(o->basisUB);
    c->ex.basis.lb = o->basisLB;
    c->ex.basis.ub = o->basisUB;
    c->ex.init();
    c->g.type = o->posFuncType;
    c->quad.type = o->quadType;
    c->quad.maxSub = o->quadMaxSub;
    c->quad.minSub = o->quadMinSub;
    c->quad.absTol = o->quadAbsTol;
    c->quad.relTol = o->quadRelTol;
    if (o->quadType == QUAD_ACC) c->quad.init_acc(o->quadPts);
    if (o->quadType == QUAD_CC) c->quad.init_cc(o->quadPts);
    c->contDeriv = o->contDeriv != 0;
    c->nugget = o->nugget;
    c->inDim = ms.dim;
    c->outDim = 1;
    c->numCoeffs = ms.size();
    c->coeffs.assign(c->numCoeffs, 0.0);
    return c;
}

#define ORC_TRY try {
#define ORC_CATCH                                   \
    }                                               \
    catch (std::invalid_argument & e) {             \
        g_err = e.what();                           \
        return -1;                                  \
    }                                               \
    catch (std::exception & e) {                    \
        g_err = e.what();                           \
        return -2;                                  \
    }                                               \
    return 0;

static Triangular* as_tri(void* m) { return dynamic_cast<Triangular*>(static_cast<Map*>(m)); }
static Component* as_comp(void* m) { return dynamic_cast<Component*>(static_cast<Map*>(m)); }

// Per-component dispatch helper: calls fn(component, coefficient pointer,
// component index, output-row offset).
template <class F>
static void for_each_comp(void* m, F fn) {
    if (Component* c = as_comp(m)) { fn(*c, c->coeffs.data(), 0u, 0u); return; }
    Triangular* t = as_tri(m);
    unsigned row = 0;
    for (size_t i = 0; i < t->comps.size(); ++i) {
        fn(*t->comps[i], t->coeffs.data() + t->coeffOff[i], unsigned(i), row);
        row += t->comps[i]->outDim;
    }
}

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_max_threads() { return omp_get_max_threads(); }

// Multi-index enumeration (bit-exact check target).
int orc_total_order_size(unsigned dim, unsigned order, unsigned* nterms, unsigned* nnz) {
    MSet s = total_order_set(dim, order);
    *nterms = s.size();
    *nnz = unsigned(s.dims.size());
    return 0;
}
int orc_total_order_fill(unsigned dim, unsigned order, uint32_t* starts, uint32_t* dims, uint32_t* orders,
                         uint32_t* maxDeg) {
    MSet s = total_order_set(dim, order);
    std::copy(s.starts.begin(), s.starts.end(), starts);
    std::copy(s.dims.begin(), s.dims.end(), dims);
    std::copy(s.orders.begin(), s.orders.end(), orders);
    std::copy(s.maxDeg.begin(), s.maxDeg.end(), maxDeg);
    return 0;
}

// 1-D basis evaluation: mode 0 = EvaluateAll, 1 = EvaluateDerivatives, 2 = EvaluateSecondDerivatives.
int orc_basis_eval(int basisType, int normalize, int mode, unsigned maxOrder, double x, double* v, double* d,
                   double* d2) {
    Basis1D b;
    b.type = basisType;
    b.poly.phys = (basisType == PHYS_HERMITE);
    b.poly.normalize = (basisType != HERMITE_FN) && normalize;
    if (mode == 0) b.eval_all(v, maxOrder, x);
    else if (mode == 1) b.eval_derivs(v, d, maxOrder, x);
    else b.eval_second(v, d, d2, maxOrder, x);
    return 0;
}

// The same through LinearizedBasis with bounds [lb, ub].
int orc_basis_eval_lin(int basisType, int normalize, int mode, unsigned maxOrder, double x, double lb, double ub,
                       double* v, double* d, double* d2) {
    Basis1D b;
    b.type = basisType;
    b.poly.phys = (basisType == PHYS_HERMITE);
    b.poly.normalize = (basisType != HERMITE_FN) && normalize;
    b.lin = true;
    b.lb = lb;
    b.ub = ub;
    if (mode == 0) b.eval_all(v, maxOrder, x);
    else if (mode == 1) b.eval_derivs(v, d, maxOrder, x);
    else b.eval_second(v, d, d2, maxOrder, x);
    return 0;
}

int orc_posfunc_eval(int type, double x, double* g, double* dg, double* d2g) {
    PosFunc p;
    p.type = type;
    *g = p.eval(x);
    *dg = p.deriv(x);
    *d2g = p.second(x);
    return 0;
}

int orc_cc_rule(unsigned n, double* wts, double* pts) {
    cc_rule(n, wts, pts);
    return 0;
}

// Integrate a test integrand (mode 0: exp(x); mode 1: discontinuous at 0.5 as in
// tests/Test_Quadrature.cpp:165-186) with AS (quadType 1) or ACC (2).
int orc_quad_test(int quadType, unsigned level, unsigned maxSub, double absTol, double relTol, int mode, double lb,
                  double ub, double* result, long* nevals) {
    Quad q;
    q.type = quadType;
    q.maxSub = maxSub;
    q.absTol = absTol;
    q.relTol = relTol;
    if (quadType == QUAD_ACC) {
        unsigned nc = (unsigned)std::pow(2, level) + 1, nf = (unsigned)std::pow(2, level + 1) + 1;
        q.cPts.resize(nc); q.cWts.resize(nc); q.fPts.resize(nf); q.fWts.resize(nf);
        cc_rule(nc, q.cWts.data(), q.cPts.data());
        cc_rule(nf, q.fWts.data(), q.fPts.data());
    }
    if (quadType == QUAD_CC) q.init_cc(level);  // level = number of points for the fixed rule
    long ne = 0;
    auto f = [&](double x, double* out) {
        ne++;
        if (mode == 0) out[0] = std::exp(x);
        else out[0] = (x < 0.5) ? std::exp(x) : 1.0 + std::exp(x);
    };
    std::vector<double> ws(q.workspace(1));
    q.integrate(ws.data(), f, 1, lb, ub, result);
    *nevals = ne;
    return 0;
}

// Root finding on the reference's unit functions (tests/Test_RootFinding.cpp:23-25):
// fn 0 identity, 1 sigmoid, 2 sigmoid combination, 3 constant -1.
static double itp_fn(int fn, double x) {
    auto sig = [](double v) { return 1. / (std::exp(-v) + 1); };
    if (fn == 0) return x;
    if (fn == 1) return sig(x);
    if (fn == 2) return sig(2 * (x - 1)) + 3 * sig(x * 0.5) + 0.5 * sig(1.5 * (x + 1));
    return -1.0;
}
int orc_itp_test(int fn, double yd, double x0, double xtol, double ftol, int* info, double* x) {
    auto f = [&](double v) { return itp_fn(fn, v); };
    *x = inverse_single_bracket(yd, f, x0, xtol, ftol, *info);
    return 0;
}
int orc_find_bracket_test(int fn, double yd, double* xlb, double* ylb, double* xub, double* yub, int* info) {
    auto f = [&](double v) { return itp_fn(fn, v); };
    find_bracket(f, *xlb, *ylb, *xub, *yub, yd, *info);
    return 0;
}

int orc_create_component(const Options* o, unsigned dim, unsigned nterms, const uint32_t* starts,
                         const uint32_t* dims, const uint32_t* orders, void** out) {
    ORC_TRY
    *out = static_cast<Map*>(make_component(o, csr_set(dim, nterms, starts, dims, orders)));
    ORC_CATCH
}

int orc_create_total_order_component(const Options* o, unsigned dim, unsigned order, void** out) {
    ORC_TRY
    *out = static_cast<Map*>(make_component(o, total_order_set(dim, order)));
    ORC_CATCH
}

// MapFactory::CreateTriangular (src/MapFactory.cpp:105-125).
int orc_create_triangular(const Options* o, unsigned inDim, unsigned outDim, unsigned order, void** out) {
    ORC_TRY
    auto* t = new Triangular();
    unsigned extra = inDim - outDim;
    for (unsigned i = 0; i < outDim; ++i)
        t->comps.emplace_back(make_component(o, total_order_set(i + extra + 1, order)));
    t->finalize();
    *out = static_cast<Map*>(t);
    ORC_CATCH
}

// Triangular map from per-component (dim, order) total-order sets.
int orc_create_triangular_orders(const Options* o, unsigned ncomp, const unsigned* dims, const unsigned* orders,
                                 void** out) {
    ORC_TRY
    auto* t = new Triangular();
    for (unsigned i = 0; i < ncomp; ++i) t->comps.emplace_back(make_component(o, total_order_set(dims[i], orders[i])));
    t->finalize();
    *out = static_cast<Map*>(t);
    ORC_CATCH
}

void orc_destroy(void* m) { delete static_cast<Map*>(m); }

int orc_info(void* m, unsigned* inDim, unsigned* outDim, unsigned* numCoeffs) {
    Map* mp = static_cast<Map*>(m);
    *inDim = mp->inDim;
    *outDim = mp->outDim;
    *numCoeffs = mp->numCoeffs;
    return 0;
}

int orc_set_coeffs(void* m, const double* c) {
    Map* mp = static_cast<Map*>(m);
    std::copy(c, c + mp->numCoeffs, mp->coeffs.begin());
    return 0;
}

int orc_evaluate(void* m, const double* pts, long D, long N, long rs, long cs, double* out, long ors, long ocs,
                 long* nevals) {
    ORC_TRY
    CMat P{pts, D, N, rs, cs};
    for_each_comp(m, [&](Component& c, const double* w, unsigned ci, unsigned row) {
        c.evaluate(sub_rows(P, c.inDim), w, out + row * ors, ocs, nevals ? nevals + long(ci) * N : nullptr);
    });
    ORC_CATCH
}

// TriangularMap::LogDeterminantImpl (100-125): component log-dets are summed.
int orc_log_determinant(void* m, const double* pts, long D, long N, long rs, long cs, double* out) {
    ORC_TRY
    CMat P{pts, D, N, rs, cs};
    std::vector<double> tmp(N);
    bool first = true;
    for_each_comp(m, [&](Component& c, const double* w, unsigned, unsigned) {
        if (first) { c.log_determinant(sub_rows(P, c.inDim), w, out, 1); first = false; }
        else {
            c.log_determinant(sub_rows(P, c.inDim), w, tmp.data(), 1);
            for (long j = 0; j < N; ++j) out[j] += tmp[j];
        }
    });
    ORC_CATCH
}

// TriangularMap::CoeffGradImpl (227-254). out is numCoeffs x N.
int orc_coeff_grad(void* m, const double* pts, long D, long N, long rs, long cs, const double* sens, long srs,
                   long scs, double* out, long ors, long ocs) {
    ORC_TRY
    CMat P{pts, D, N, rs, cs};
    unsigned par = 0;
    for_each_comp(m, [&](Component& c, const double* w, unsigned, unsigned row) {
        Mat O{out + par * ors, c.numCoeffs, N, ors, ocs};
        c.coeff_grad(sub_rows(P, c.inDim), w, sens + row * srs, scs, O);
        par += c.numCoeffs;
    });
    ORC_CATCH
}

// TriangularMap::LogDeterminantCoeffGradImpl (257-276).
int orc_log_determinant_coeff_grad(void* m, const double* pts, long D, long N, long rs, long cs, double* out,
                                   long ors, long ocs) {
    ORC_TRY
    CMat P{pts, D, N, rs, cs};
    unsigned par = 0;
    for_each_comp(m, [&](Component& c, const double* w, unsigned, unsigned) {
        Mat O{out + par * ors, c.numCoeffs, N, ors, ocs};
        c.log_determinant_coeff_grad(sub_rows(P, c.inDim), w, O);
        par += c.numCoeffs;
    });
    ORC_CATCH
}

// TriangularMap::GradientImpl (188-224): out (inDim x N) = sum_i J_i^T sens_i.
int orc_gradient(void* m, const double* pts, long D, long N, long rs, long cs, const double* sens, long srs,
                 long scs, double* out, long ors, long ocs) {
    ORC_TRY
    CMat P{pts, D, N, rs, cs};
    for (long n = 0; n < N; ++n)
        for (long d = 0; d < D; ++d) out[d * ors + n * ocs] = 0.0;
    for_each_comp(m, [&](Component& c, const double* w, unsigned, unsigned row) {
        std::vector<double> sub(size_t(c.inDim) * N);
        Mat S{sub.data(), c.inDim, N, N, 1};
        c.gradient(sub_rows(P, c.inDim), w, sens + row * srs, scs, S);
        for (long n = 0; n < N; ++n)
            for (unsigned d = 0; d < c.inDim; ++d) out[d * ors + n * ocs] += S(d, n);
    });
    ORC_CATCH
}

// TriangularMap::LogDeterminantInputGradImpl (279-310).
int orc_log_determinant_input_grad(void* m, const double* pts, long D, long N, long rs, long cs, double* out,
                                   long ors, long ocs) {
    ORC_TRY
    CMat P{pts, D, N, rs, cs};
    for (long n = 0; n < N; ++n)
        for (long d = 0; d < D; ++d) out[d * ors + n * ocs] = 0.0;
    for_each_comp(m, [&](Component& c, const double* w, unsigned, unsigned) {
        std::vector<double> sub(size_t(c.inDim) * N);
        Mat S{sub.data(), c.inDim, N, N, 1};
        c.log_determinant_input_grad(sub_rows(P, c.inDim), w, S);
        for (long n = 0; n < N; ++n)
            for (unsigned d = 0; d < c.inDim; ++d) out[d * ors + n * ocs] += S(d, n);
    });
    ORC_CATCH
}

// Inverse.  For a component: MonotoneComponent::InverseImpl with xs (rows x numXs)
// and ys[N].  For a triangular map: TriangularMap::InverseImpl (147-184) with
// x1 holding the (inDim - outDim) conditioning rows and r (outDim x N).
int orc_inverse(void* m, const double* x1, long xrows, long xcols, long xrs, long xcs, const double* r, long rrows,
                long N, long rrs, long rcs, double* out, long ors, long ocs, double xtol, double ytol) {
    ORC_TRY
    if (Component* c = as_comp(m)) {
        CMat X{x1, xrows, xcols, xrs, xcs};
        c->inverse_n(X, r, rcs, N, c->coeffs.data(), out, ocs, xtol, ytol);
        return 0;
    }
    Triangular* t = as_tri(m);
    long ip = t->inDim, op = t->outDim, extra = ip - op;
    (void)rrows;
    std::vector<double> full(size_t(ip) * N, 0.0);  // row-major ip x N
    for (long i = 0; i < xrows; ++i)
        for (long n = 0; n < N; ++n) full[i * N + n] = x1[i * xrs + n * xcs];
    unsigned startOut = 0;
    for (size_t i = 0; i < t->comps.size(); ++i) {
        Component& c = *t->comps[i];
        CMat X{full.data(), c.inDim, N, N, 1};
        std::vector<double> res(N);
        c.inverse_n(X, r + startOut * rrs, rcs, N, t->coeffs.data() + t->coeffOff[i], res.data(), 1, xtol, ytol);
        for (long n = 0; n < N; ++n) full[(extra + startOut) * N + n] = res[n];
        startOut += c.outDim;
    }
    for (long i = 0; i < op; ++i)
        for (long n = 0; n < N; ++n) out[i * ors + n * ocs] = full[(extra + i) * N + n];
    ORC_CATCH
}

// KL objective against N(0, I), in the structure of FastGaussianReverseKLObjective
// (src/MapObjective.cpp:107-215) generalized to every output row:
//   loss = mean_n( 0.5 |T(x_n)|^2 - logdet(x_n) ) [+ 0.5 d log(2 pi) if withConst]
//   grad = mean_n( CoeffGrad(sens = T)[:, n] - LogDeterminantCoeffGrad[:, n] ).
// withConst = 1 gives KLObjective::ObjectivePlusCoeffGradImpl (55-91) with a
// GaussianSamplerDensity(dim) reference (src/Distributions/GaussianSamplerDensity.cpp:22-51).
int orc_kl(void* m, const double* pts, long D, long N, long rs, long cs, int withConst, double* loss, double* grad) {
    ORC_TRY
    Map* mp = static_cast<Map*>(m);
    long M = mp->outDim, K = mp->numCoeffs;
    std::vector<double> T(size_t(M) * N), ld(N), G(size_t(K) * N), LG(size_t(K) * N);
    int rc = orc_evaluate(m, pts, D, N, rs, cs, T.data(), N, 1, nullptr);
    if (rc) return rc;
    rc = orc_log_determinant(m, pts, D, N, rs, cs, ld.data());
    if (rc) return rc;
    rc = orc_coeff_grad(m, pts, D, N, rs, cs, T.data(), N, 1, G.data(), N, 1);
    if (rc) return rc;
    rc = orc_log_determinant_coeff_grad(m, pts, D, N, rs, cs, LG.data(), N, 1);
    if (rc) return rc;
    double l = 0.0;
    for (long n = 0; n < N; ++n) {
        double v = 0.0;
        for (long i = 0; i < M; ++i) v += 0.5 * T[i * N + n] * T[i * N + n];
        l += v - ld[n];
    }
    l /= N;
    if (withConst) l += 0.5 * M * std::log(2 * M_PI);
    *loss = l;
    for (long k = 0; k < K; ++k) {
        double s = 0.0;
        for (long n = 0; n < N; ++n) s += G[k * N + n] - LG[k * N + n];
        grad[k] = s / N;
    }
    ORC_CATCH
}

}  // extern "C"
