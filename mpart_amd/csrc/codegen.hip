// codegen.hip -- source generation and hipRTC compilation (host code).
#include "codegen.hpp"

#include <hip/hiprtc.h>

#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <sstream>
#include <unordered_map>

namespace mpg {

// Embedded copy of mono_device.hpp (written by the build into _mono_device_src.inc).
static const char* kDeviceHeader =
#include "_mono_device_src.inc"
    ;

namespace {

// Product of the off-diagonal basis factors of term t ("1.0" if none).
std::string term_product(const GenComp& c, unsigned t, unsigned* q) {
    std::string expr;
    *q = 0;
    for (unsigned i = c.starts[t]; i < c.starts[t + 1]; ++i) {
        if (c.dims[i] == c.dim - 1) {
            *q = c.orders[i];
            continue;
        }
        std::string f = "b" + std::to_string(c.dims[i]) + "_" + std::to_string(c.orders[i]);
        expr = expr.empty() ? f : "(" + expr + " * " + f + ")";
    }
    return expr.empty() ? std::string("1.0") : expr;
}

// Off-diagonal basis values of the point as scalars b<d>_<k> (zero for an
// inactive lane).
void emit_basis(std::ostringstream& o, const GenSpec& s, unsigned ndims) {
    for (unsigned d = 0; d < ndims && d < s.dimDeg.size(); ++d) {
        const int deg = s.dimDeg[d];
        if (deg == 0) continue;
        o << "    double b" << d << "_1";
        for (int k = 2; k <= deg; ++k) o << ", b" << d << "_" << k;
        o << ";\n    {\n        double t[" << deg + 1 << "];\n";
        o << "        basis_vals<GBK, " << deg << ">(M, active ? pt[" << d << " * prs] : 0.0, t);\n";
        for (int k = 1; k <= deg; ++k) o << "        b" << d << "_" << k << " = active ? t[" << k << "] : 0.0;\n";
        o << "    }\n";
    }
}

// Empty asm with the basis values as in/out operands.  It costs no
// instruction but makes the values opaque at this point, so the products of
// the term expressions are formed where they are used: without it LICM hoists
// every (loop-invariant) product out of the component loop and GVN keeps the
// forward products alive across the quadrature for the gradient pass, which
// multiplies the live registers.
void emit_pin(std::ostringstream& o, const GenSpec& s, unsigned ndims, const char* indent) {
    for (unsigned d = 0; d < ndims && d < s.dimDeg.size(); ++d)
        for (int k = 1; k <= s.dimDeg[d]; ++k)
            o << indent << "asm volatile(\"\" : \"+v\"(b" << d << "_" << k << "));\n";
}

// The same with first derivatives db<d>_<k> (FillCache1 with Input flags).
// With s.derivFromValues the derivatives are not held: term_deriv forms them
// from the values (d phi_o / dx = M.dfac[o] phi_{o-1}), which keeps one
// register pair per basis value instead of two.
void emit_basis_derivs(std::ostringstream& o, const GenSpec& s, unsigned ndims) {
    if (s.derivFromValues) {
        emit_basis(o, s, ndims);
        return;
    }
    for (unsigned d = 0; d < ndims && d < s.dimDeg.size(); ++d) {
        const int deg = s.dimDeg[d];
        if (deg == 0) continue;
        o << "    double b" << d << "_1, db" << d << "_1";
        for (int k = 2; k <= deg; ++k) o << ", b" << d << "_" << k << ", db" << d << "_" << k;
        o << ";\n    {\n        double t[" << deg + 1 << "], dt[" << deg + 1 << "];\n";
        o << "        basis_vals_derivs<GBK, " << deg << ">(M, pt[" << d << " * prs], t, dt);\n";
        for (int k = 1; k <= deg; ++k)
            o << "        b" << d << "_" << k << " = t[" << k << "]; db" << d << "_" << k << " = dt[" << k << "];\n";
        o << "    }\n";
    }
}

void emit_pin_derivs(std::ostringstream& o, const GenSpec& s, unsigned ndims, const char* indent) {
    emit_pin(o, s, ndims, indent);
    if (s.derivFromValues) return;
    for (unsigned d = 0; d < ndims && d < s.dimDeg.size(); ++d)
        for (int k = 1; k <= s.dimDeg[d]; ++k)
            o << indent << "asm volatile(\"\" : \"+v\"(db" << d << "_" << k << "));\n";
}

// d/dx_{dims[j]} of the off-diagonal product of term t (entry j of the term).
std::string term_deriv(const GenSpec& s, const GenComp& c, unsigned t, unsigned j) {
    const std::string d = std::to_string(c.dims[j]);
    const unsigned k = c.orders[j];
    std::string expr;
    if (!s.derivFromValues) expr = "db" + d + "_" + std::to_string(k);
    else if (k == 1) expr = "M.dfac[1]";
    else expr = "(M.dfac[" + std::to_string(k) + "] * b" + d + "_" + std::to_string(k - 1) + ")";
    for (unsigned i = c.starts[t]; i < c.starts[t + 1]; ++i) {
        if (i == j || c.dims[i] == c.dim - 1) continue;
        expr = "(" + expr + " * b" + std::to_string(c.dims[i]) + "_" + std::to_string(c.orders[i]) + ")";
    }
    return expr;
}

void emit_coeff_sums(std::ostringstream& o, const GenComp& c) {
    for (unsigned q = 0; q <= c.p; ++q) o << "        C[" << q << "] = 0.0;\n";
    const unsigned nt = unsigned(c.starts.size()) - 1;
    for (unsigned t = 0; t < nt; ++t) {
        unsigned q;
        std::string P = term_product(c, t, &q);
        o << "        C[" << q << "] += w[" << c.coeffOff + t << "] * " << P << ";\n";
    }
}

// Input VJP / KL input gradient kernels (all components).
void emit_input_vjp(std::ostringstream& o, const GenSpec& s) {
    const unsigned ncomp = unsigned(s.comps.size());
    // ---------------- input VJP / KL input gradient (all components) ----------------
    // The body is a template on KL mode (sensitivities T and -1 known at
    // compile time), instantiated by two kernels.
    o << R"(
template <bool KL>
__device__ __forceinline__ void ivjp_body(const MapDev& M, const double* __restrict__ w,
        const double* __restrict__ pts, long prs, long pcs, long N, const double* sens, long srs,
        long scs, const double* lsens, long lss, double lval, double scale, double* out, long ors, long ocs) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const bool active = true;
    const double* pt = pts + n * pcs;
    const bool hasL = lsens != nullptr || lval != 0.0;
    const double lv = lsens ? lsens[n * lss] : lval;
)";
    emit_basis_derivs(o, s, s.D - 1);
    for (unsigned d = 0; d < s.D; ++d) o << "    double gx" << d << " = 0.0;\n";
    o << "#pragma unroll 1\n    for (int c = 0; c < " << ncomp << "; ++c) {\n";
    emit_pin_derivs(o, s, s.D - 1, "        ");
    o << "        const CompDev cd = ro(M.comps)[c];\n        double C[GPM + 1];\n        switch (c) {\n";
    for (unsigned c = 0; c < ncomp; ++c) {
        o << "        case " << c << ": {\n";
        emit_coeff_sums(o, s.comps[c]);
        o << "        break; }\n";
    }
    o << R"(        }
        const double sv = sens ? sens[(long)cd.outRow * srs + n * scs] : 0.0;
        double G[GPM + 1], gk;
        ivjp_core<GBK, GPK, GQK, GPM, KL>(M, cd, C, pt[(long)(cd.dim - 1) * prs], sens != nullptr, sv, hasL, lv,
                                          G, gk);
)";
    emit_pin_derivs(o, s, s.D - 1, "        ");
    o << "        switch (c) {\n";
    for (unsigned c = 0; c < ncomp; ++c) {
        const GenComp& gc = s.comps[c];
        const unsigned nt = unsigned(gc.starts.size()) - 1;
        o << "        case " << c << ": {\n";
        for (unsigned t = 0; t < nt; ++t) {
            unsigned q;
            term_product(gc, t, &q);
            bool any = false;
            for (unsigned j = gc.starts[t]; j < gc.starts[t + 1]; ++j) {
                if (gc.dims[j] == gc.dim - 1) continue;
                if (!any) o << "            { const double a = w[" << gc.coeffOff + t << "] * G[" << q << "];\n";
                any = true;
                o << "              gx" << gc.dims[j] << " += a * " << term_deriv(s, gc, t, j) << ";\n";
            }
            if (any) o << "            }\n";
        }
        o << "            gx" << gc.dim - 1 << " += gk;\n        break; }\n";
    }
    o << "        }\n    }\n    double* o = out + n * ocs;\n";
    for (unsigned d = 0; d < s.D; ++d) o << "    o[" << d << "L * ors] = scale * gx" << d << ";\n";
    o << R"(}
#define MPD_IVJP_ARGS MapDev M, const double* __restrict__ w, const double* __restrict__ pts, long prs, long pcs, \
        long N, const double* sens, long srs, long scs, const double* lsens, long lss, double lval, double scale, \
        double* out, long ors, long ocs
extern "C" __global__ void __launch_bounds__(256) mpart_gen_ivjp(MPD_IVJP_ARGS) {
    ivjp_body<false>(M, w, pts, prs, pcs, N, sens, srs, scs, lsens, lss, lval, scale, out, ors, ocs);
}
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) mpart_gen_klx(MPD_IVJP_ARGS) {
    ivjp_body<true>(M, w, pts, prs, pcs, N, nullptr, 0, 0, nullptr, 0, -1.0, scale, out, ors, ocs);
}
)";
}

}  // namespace

std::string generate_source(const GenSpec& s, int part) {
    std::ostringstream o;
    o << "#define MPD_LINEARIZED " << (s.lin ? 1 : 0) << "\n#include \"mono_device.hpp\"\nusing namespace mpd;\n";
    o << "#define GBK " << s.bk << "\n#define GPK " << s.pk << "\n#define GQK " << s.qk << "\n#define GPM " << s.pm
      << "\n";
    const unsigned ncomp = unsigned(s.comps.size());

    // ---------------- fused forward + adjoint (KL / VJP) ----------------
    // The body is one device function; two kernels instantiate it with
    // different occupancy targets (the compiler's choice, and three waves per
    // SIMD), between which the host picks by timing the first calls.  A
    // second body (fused_body_in) also forms the per-point input gradient of
    // the same sensitivities from the same quadrature (the input-adjoint
    // integrand adds U to the coefficient adjoint S_q).
    for (int in = 0; in < 2; ++in) {
        if (in != part) continue;
        o << "\n__device__ __forceinline__ void " << (in ? "fused_body_in" : "fused_body") << R"((const MapDev& M, const double* __restrict__ w,
        const double* __restrict__ pts, long prs, long pcs, long N, const double* sens, long srs, long scs,
        const double* lsens, long lss, int klMode, double* partial)";
        if (in) o << ", double gscale, double* gin, long gors, long gocs";
        o << R"() {
    extern __shared__ double acc[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nwaves = blockDim.x >> 6;
)";
        o << "    const int K = " << s.numCoeffs << ";\n";
        o << R"(    for (int i = tid; i < nwaves * K; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
    double* myAcc = acc + (long)wave * K;
    double lossLane = 0.0;
    const long ntiles = (N + blockDim.x - 1) / blockDim.x;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long n0 = tile * blockDim.x + tid;
        const bool active = n0 < N;
        const long n = active ? n0 : N - 1;
        const double* pt = pts + n * pcs;
)";
        if (in) {
            emit_basis_derivs(o, s, s.D - 1);
            for (unsigned d = 0; d < s.D; ++d) o << "    double gx" << d << " = 0.0;\n";
        } else {
            emit_basis(o, s, s.D - 1);
        }
        o << "#pragma unroll 1\n        for (int c = 0; c < " << ncomp << "; ++c) {\n";
        if (in) emit_pin_derivs(o, s, s.D - 1, "        ");
        else emit_pin(o, s, s.D - 1, "        ");
        o << "        const CompDev cd = ro(M.comps)[c];\n        double C[GPM + 1];\n        switch (c) {\n";
        for (unsigned c = 0; c < ncomp; ++c) {
            o << "        case " << c << ": {\n";
            emit_coeff_sums(o, s.comps[c]);
            o << "        break; }\n";
        }
        o << R"(        }
        const double xk = pt[(long)(cd.dim - 1) * prs];
        const double sv = (!klMode && sens) ? sens[(long)cd.outRow * srs + n * scs] : 0.0;
        const double lv = (!klMode && lsens) ? lsens[n * lss] : 0.0;
        double G[GPM + 1];
)";
        if (in)
            o << "        double gk;\n        comp_core<GBK, GPK, GQK, GPM, true>(M, cd, C, xk, active, klMode, sv, lv, G, "
                 "lossLane, &gk);\n";
        else
            o << "        comp_core<GBK, GPK, GQK, GPM>(M, cd, C, xk, active, klMode, sv, lv, G, lossLane);\n";
        if (in) emit_pin_derivs(o, s, s.D - 1, "        ");
        else emit_pin(o, s, s.D - 1, "        ");
        o << "        switch (c) {\n";
        for (unsigned c = 0; c < ncomp; ++c) {
            const GenComp& gc = s.comps[c];
            const unsigned nt = unsigned(gc.starts.size()) - 1;
            o << "        case " << c << ": {\n";
            for (unsigned t0 = 0; t0 < nt; t0 += 16) {
                const unsigned nv = std::min(16u, nt - t0);
                o << "            { double v[16];\n";
                for (unsigned u = 0; u < 16; ++u) {
                    if (u < nv) {
                        unsigned q;
                        std::string P = term_product(gc, t0 + u, &q);
                        o << "              v[" << u << "] = G[" << q << "] * " << P << ";\n";
                    } else {
                        o << "              v[" << u << "] = 0.0;\n";
                    }
                }
                o << "              reduce16_add(v, lane, myAcc + " << gc.coeffOff + t0 << ", " << nv << "); }\n";
            }
            if (in) {
                for (unsigned t = 0; t < nt; ++t) {
                    unsigned q;
                    term_product(gc, t, &q);
                    bool any = false;
                    for (unsigned j = gc.starts[t]; j < gc.starts[t + 1]; ++j) {
                        if (gc.dims[j] == gc.dim - 1) continue;
                        if (!any) o << "            { const double a = w[" << gc.coeffOff + t << "] * G[" << q << "];\n";
                        any = true;
                        o << "              gx" << gc.dims[j] << " += a * " << term_deriv(s, gc, t, j) << ";\n";
                    }
                    if (any) o << "            }\n";
                }
                o << "            gx" << gc.dim - 1 << " += gk;\n";
            }
            o << "        break; }\n";
        }
        o << "        }\n        }\n";
        if (in) {
            o << "        if (active) {\n            double* go = gin + n * gocs;\n";
            for (unsigned d = 0; d < s.D; ++d) o << "            go[" << d << "L * gors] = gscale * gx" << d << ";\n";
            o << "        }\n";
        }
        o << R"(    }
    const double lw = wave_sum(lossLane);
    __shared__ double lossWave[4];
    if (lane == 0) lossWave[wave] = lw;
    __syncthreads();
    // Partial row of the block: K gradient sums, then the loss sum.
    double* row = partial + (long)blockIdx.x * (K + 1);
    if (tid == 0) {
        double l = 0.0;
        for (int i = 0; i < nwaves; ++i) l += lossWave[i];
        row[K] = l;
    }
    for (int k = tid; k < K; k += blockDim.x) {
        double t = 0.0;
        for (int i = 0; i < nwaves; ++i) t += acc[(long)i * K + k];
        row[k] = t;
    }
}
)";
    }
    o << R"(#define MPD_FUSED_ARGS MapDev M, const double* __restrict__ w, const double* __restrict__ pts, long prs, long pcs, \
        long N, const double* sens, long srs, long scs, const double* lsens, long lss, int klMode, double* partial
)";
    if (part == 1) {
        o << R"(extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
mpart_gen_fused_in(MPD_FUSED_ARGS, double gscale, double* gin, long gors, long gocs) {
    fused_body_in(M, w, pts, prs, pcs, N, sens, srs, scs, lsens, lss, klMode, partial, gscale, gin, gors, gocs);
}
)";
        emit_input_vjp(o, s);
        return o.str();
    }
    o << R"(extern "C" __global__ void __launch_bounds__(256) mpart_gen_fused(MPD_FUSED_ARGS) {
    fused_body(M, w, pts, prs, pcs, N, sens, srs, scs, lsens, lss, klMode, partial);
}
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
mpart_gen_fused_w3(MPD_FUSED_ARGS) {
    fused_body(M, w, pts, prs, pcs, N, sens, srs, scs, lsens, lss, klMode, partial);
}
)";

    // ---------------- Evaluate / LogDeterminant ----------------
    o << R"(
extern "C" __global__ void __launch_bounds__(256) mpart_gen_eval(MapDev M, const double* __restrict__ w,
        const double* __restrict__ pts, long prs, long pcs, long N, double* out, long ors, long ocs, double* ld,
        long lds, int ldAcc, int* nevals, int cb, int ce) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const bool active = true;
    const double* pt = pts + n * pcs;
)";
    emit_basis(o, s, s.D - 1);
    o << "    double ldsum = 0.0;\n#pragma unroll 1\n    for (int c = cb; c < ce; ++c) {\n";
    emit_pin(o, s, s.D - 1, "        ");
    o << "        const CompDev cd = ro(M.comps)[c];\n        double C[GPM + 1];\n        switch (c) {\n";
    for (unsigned c = 0; c < ncomp; ++c) {
        o << "        case " << c << ": {\n";
        emit_coeff_sums(o, s.comps[c]);
        o << "        break; }\n";
    }
    o << R"(        }
        const double xk = pt[(long)(cd.dim - 1) * prs];
        DiagCoef<GPM> dc;
        make_diag<GPM>(M, cd, C, xk, 0.0, dc);
        if (out) {
            AdjAcc<GPM> A;
            int ne = 0;
            const double I = integrate<GBK, GPK, GQK, FL_VAL, GPM>(M.quad, dc, A, nevals ? &ne : nullptr);
            out[(long)cd.outRow * ors + n * ocs] = I + f_at_zero<GPM>(M, cd, C);
            if (nevals) nevals[(long)c * N + n] = ne;
        }
        if (ld) {
            const double z = diag_eval<GBK, false, GPM>(dc, xk, nullptr);
            const double g = posfunc<GPK>(z).g;
            ldsum += (g <= 0) ? -__builtin_inf() : log_pos(g);
        }
    }
    if (ld) {
        if (ldAcc) ld[n * lds] += ldsum;
        else ld[n * lds] = ldsum;
    }
}
)";

    // ---------------- CoeffGrad / LogDeterminantCoeffGrad (one component) ----------------
    for (int kind = 0; kind < 2; ++kind) {
        if (kind == 0)
            o << R"(
extern "C" __global__ void __launch_bounds__(256) mpart_gen_coef(MapDev M, const double* __restrict__ w,
        const double* __restrict__ pts, long prs, long pcs, long N, const double* sens, long srs, long scs,
        double* out, long ors, long ocs, int c) {
)";
        else
            o << R"(
extern "C" __global__ void __launch_bounds__(256) mpart_gen_ldcoef(MapDev M, const double* __restrict__ w,
        const double* __restrict__ pts, long prs, long pcs, long N, double* out, long ors, long ocs, int c) {
)";
        o << R"(    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const bool active = true;
    const double* pt = pts + n * pcs;
    const CompDev cd = ro(M.comps)[c];
)";
        // Only the coordinates of component c are needed; the basis of every
        // coordinate below the largest component is computed (cheap).
        emit_basis(o, s, s.D - 1);
        o << "    double C[GPM + 1];\n    switch (c) {\n";
        for (unsigned c = 0; c < ncomp; ++c) {
            o << "    case " << c << ": {\n";
            emit_coeff_sums(o, s.comps[c]);
            o << "    break; }\n";
        }
        o << "    }\n    const double xk = pt[(long)(cd.dim - 1) * prs];\n    DiagCoef<GPM> dc;\n    double G[GPM + 1];\n";
        if (kind == 0) {
            o << R"(    make_diag<GPM>(M, cd, C, xk, M.nugget, dc);
    AdjAcc<GPM> A;
    zero_adj<GPM>(A);
    integrate<GBK, GPK, GQK, FL_COEF, GPM>(M.quad, dc, A, nullptr);
    const double s = sens[(long)cd.outRow * srs + n * scs];
#pragma unroll
    for (int q = 0; q <= GPM; ++q) G[q] = (q <= cd.p) ? (M.phi0[q] + M.rho[q] * A.S[q]) : 0.0;
)";
        } else {
            o << R"(    make_diag<GPM>(M, cd, C, xk, 0.0, dc);
    const double z = diag_eval<GBK, false, GPM>(dc, xk, nullptr);
    const PosF<GPK> gv = posfunc<GPK>(z);
    const double inv = 1.0 / gv.g;
    double bq[GPM + 1];
    diag_basis<GBK, GPM>(dc, xk, bq);
#pragma unroll
    for (int q = 0; q <= GPM; ++q) G[q] = (q <= cd.p) ? M.rho[q] * bq[q] : 0.0;
)";
        }
        emit_pin(o, s, s.D - 1, "    ");
        o << "    double* o = out + n * ocs;\n    switch (c) {\n";
        for (unsigned c = 0; c < ncomp; ++c) {
            const GenComp& gc = s.comps[c];
            const unsigned nt = unsigned(gc.starts.size()) - 1;
            o << "    case " << c << ": {\n";
            for (unsigned t = 0; t < nt; ++t) {
                unsigned q;
                std::string P = term_product(gc, t, &q);
                if (kind == 0)
                    o << "        o[" << gc.coeffOff + t << "L * ors] = (" << P << " * G[" << q << "]) * s;\n";
                else
                    o << "        o[" << gc.coeffOff + t << "L * ors] = ((" << P << " * G[" << q
                      << "]) * gv.gp) * inv;\n";
            }
            o << "    break; }\n";
        }
        o << "    }\n}\n";
    }

    return o.str();
}

namespace {
std::mutex g_mu;
std::unordered_map<std::string, std::unique_ptr<GenModule>> g_cache;

const char* const kCompileOpts[] = {"-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};

// 64-bit FNV-1a over several byte strings (separated by a zero byte), from
// the offset basis `h`.
uint64_t fnv1a(std::initializer_list<const std::string*> parts, uint64_t h) {
    for (const std::string* s : parts) {
        for (unsigned char c : *s) h = (h ^ c) * 1099511628211ull;
        h = (h ^ 0xffu) * 1099511628211ull;
    }
    return h;
}

// Path of a module's code object in the directory MPART_CODEGEN_CACHE names
// ("" when unset): the name carries two independent 64-bit hashes of the
// generated source, the embedded device header, the target and the options.
std::string cache_path(const std::string& src, const std::string& arch) {
    const char* dir = std::getenv("MPART_CODEGEN_CACHE");
    if (!dir || !*dir) return "";
    const std::string hdr(kDeviceHeader);
    std::string opts;
    for (const char* o : kCompileOpts) opts += std::string(o) + " ";
    const uint64_t h1 = fnv1a({&src, &hdr, &arch, &opts}, 14695981039346656037ull);
    const uint64_t h2 = fnv1a({&opts, &arch, &hdr, &src}, 0x9e3779b97f4a7c15ull);
    char name[96];
    std::snprintf(name, sizeof(name), "/mpart_gen_%016llx%016llx_", (unsigned long long)h1, (unsigned long long)h2);
    return std::string(dir) + name + arch + ".co";
}

bool read_cached(const std::string& path, std::vector<char>& code) {
    if (path.empty()) return false;
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    // an ELF code object
    if (buf.size() < 64 || buf[0] != 0x7f || buf[1] != 'E' || buf[2] != 'L' || buf[3] != 'F') return false;
    code.swap(buf);
    return true;
}

void write_cached(const std::string& path, const std::vector<char>& code) {
    if (path.empty()) return;
    const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    const bool ok = std::fwrite(code.data(), 1, code.size(), f) == code.size();
    if (std::fclose(f) == 0 && ok) std::rename(tmp.c_str(), path.c_str());
    else std::remove(tmp.c_str());
}
}  // namespace

bool compile_source(const std::string& src, const std::string& arch, std::vector<char>& code, std::string& err) {
    hiprtcProgram prog;
    const char* hdrs[] = {kDeviceHeader};
    const char* names[] = {"mono_device.hpp"};
    if (hiprtcCreateProgram(&prog, src.c_str(), "mpart_gen.hip", 1, hdrs, names) != HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return false;
    }
    const std::string archOpt = "--offload-arch=" + arch;
    const char* opts[7] = {archOpt.c_str()};
    for (int i = 0; i < 6; ++i) opts[i + 1] = kCompileOpts[i];
    if (hiprtcCompileProgram(prog, 7, opts) != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls + 1, '\0');
        hiprtcGetProgramLog(prog, &log[0]);
        err = "hipRTC compilation failed: " + log.substr(0, 4000);
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return true;
}

GenModule* get_module(const GenSpec& spec, std::string& err, int part) {
    const std::string src = generate_source(spec, part);
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(src);
    if (it != g_cache.end()) return it->second.get();
    auto t0 = std::chrono::steady_clock::now();
    int dev = 0;
    std::string arch = "gfx950";
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        const std::string a = prop.gcnArchName;
        arch = a.substr(0, a.find(':'));
    }
    // A code object compiled earlier (by this or another process) is reused
    // from MPART_CODEGEN_CACHE when that names a directory; a process then
    // runs exactly the binary the first one compiled, whatever compiler its
    // own hipRTC would load.
    std::vector<char> code;
    const std::string cpath = cache_path(src, arch);
    if (!read_cached(cpath, code)) {
        if (!compile_source(src, arch, code, err)) return nullptr;
        write_cached(cpath, code);
    }
    auto m = std::make_unique<GenModule>();
    if (hipModuleLoadData(&m->mod, code.data()) != hipSuccess) {
        err = "hipModuleLoadData failed";
        return nullptr;
    }
    const bool ok =
        part == 0 ? (hipModuleGetFunction(&m->fused, m->mod, "mpart_gen_fused") == hipSuccess &&
                     hipModuleGetFunction(&m->fused3, m->mod, "mpart_gen_fused_w3") == hipSuccess &&
                     hipModuleGetFunction(&m->eval, m->mod, "mpart_gen_eval") == hipSuccess &&
                     hipModuleGetFunction(&m->coef, m->mod, "mpart_gen_coef") == hipSuccess &&
                     hipModuleGetFunction(&m->ldcoef, m->mod, "mpart_gen_ldcoef") == hipSuccess)
                  : (hipModuleGetFunction(&m->ivjp, m->mod, "mpart_gen_ivjp") == hipSuccess &&
                     hipModuleGetFunction(&m->klx, m->mod, "mpart_gen_klx") == hipSuccess &&
                     hipModuleGetFunction(&m->fusedIn, m->mod, "mpart_gen_fused_in") == hipSuccess);
    if (!ok) {
        err = "hipModuleGetFunction failed";
        return nullptr;
    }
    m->compileSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    GenModule* out = m.get();
    g_cache.emplace(src, std::move(m));
    return out;
}

}  // namespace mpg
