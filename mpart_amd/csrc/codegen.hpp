// codegen.hpp -- runtime specialization of the term passes of a map.
//
// For a given map structure (multi-index sets, basis / bijector / quadrature
// kinds, degree bound) the generator writes HIP source in which every term of
// every component is a straight-line expression over basis values held in
// registers, with constant coefficient offsets, and compiles it with hipRTC
// for the current device.  The adaptive quadrature and the component core are
// the same device functions the precompiled kernels use (mono_device.hpp is
// compiled into the module from an embedded copy).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace mpg {

struct GenComp {
    unsigned dim;
    unsigned coeffOff;
    unsigned p;                       // largest last-coordinate order
    std::vector<uint32_t> starts, dims, orders;  // compressed multi-index set
};

struct GenSpec {
    int bk, pk, qk, pm;
    int lin = 0;  // LinearizedBasis bounds in use
    // Input-gradient kernels form d phi_o / dx as M.dfac[o] * phi_{o-1} instead
    // of holding a derivative per basis value (orthogonal polynomials, no bounds).
    int derivFromValues = 0;
    unsigned D;
    unsigned numCoeffs;
    std::vector<int> dimDeg;  // largest off-diagonal order per coordinate
    std::vector<GenComp> comps;
};

struct GenModule {
    hipModule_t mod = nullptr;
    hipFunction_t fused = nullptr, fused3 = nullptr, eval = nullptr, coef = nullptr, ldcoef = nullptr, ivjp = nullptr, klx = nullptr,
                  fusedIn = nullptr;
    double compileSeconds = 0.0;
};

// part 0: the fused pass (KL / VJP), Evaluate / LogDeterminant, CoeffGrad,
// LogDeterminantCoeffGrad; part 1: the kernels with input gradients (input
// VJP, KL input gradient, fused pass with input gradients).  Part 1 is
// compiled only when one of them is first used.
std::string generate_source(const GenSpec& spec, int part);

// hipRTC compilation of generated source for `arch` (e.g. "gfx950") into a code
// object; false with `err` set (including the compiler log) on failure.
bool compile_source(const std::string& src, const std::string& arch, std::vector<char>& code, std::string& err);

// Returns the compiled module for this structure (cached process-wide), or
// nullptr with `err` set when compilation or loading fails.
GenModule* get_module(const GenSpec& spec, std::string& err, int part);

}  // namespace mpg
