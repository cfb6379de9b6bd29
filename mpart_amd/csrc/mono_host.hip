// mono_host.hip -- C ABI (include/mpart_amd.h): map construction, device plans
// and kernel launches.  Host code only; the kernels live in k_*.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpart_amd.h"
#include "batch_io.hpp"
#include "codegen.hpp"
#include "launchers.hpp"

namespace {

using mpd::CompDev;
using mpd::GroupDev;
using mpd::MapDev;

thread_local std::string g_err;

struct ArgError : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct RtError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NotImpl : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return MPART_OK;
    } catch (ArgError& e) {
        g_err = e.what();
        return MPART_INVALID_ARGUMENT;
    } catch (std::invalid_argument& e) {
        g_err = e.what();
        return MPART_INVALID_ARGUMENT;
    } catch (HipError& e) {
        g_err = e.what();
        return MPART_HIP_ERROR;
    } catch (NotImpl& e) {
        g_err = e.what();
        return MPART_NOT_IMPLEMENTED;
    } catch (std::exception& e) {
        g_err = e.what();
        return MPART_RUNTIME_ERROR;
    }
}

// ---------------------------------------------------------------------------
// Compressed multi-index set (same layout as FixedMultiIndexSet).
struct MultiSet {
    unsigned dim = 0;
    std::vector<uint32_t> starts, dims, orders, maxDeg;
    unsigned size() const { return unsigned(starts.size()) - 1; }
    void max_degrees() {
        maxDeg.assign(dim, 0u);
        for (size_t i = 0; i < dims.size(); ++i) maxDeg[dims[i]] = std::max(maxDeg[dims[i]], orders[i]);
    }
};

// Total-order set in the enumeration order of FixedMultiIndexSet's total-order
// constructor (src/MultiIndices/FixedMultiIndexSet.cpp:507-549): multi-indices
// in lexicographic order of (alpha_0, ..., alpha_{dim-1}).
MultiSet total_order(unsigned dim, unsigned order) {
    MultiSet s;
    s.dim = dim;
    std::vector<unsigned> a(dim, 0u);
    // Odometer over the last coordinate first; each state is a multi-index with
    // |a| <= order visited in lexicographic order.
    while (true) {
        s.starts.push_back(uint32_t(s.dims.size()));
        for (unsigned i = 0; i < dim; ++i)
            if (a[i] > 0) {
                s.dims.push_back(i);
                s.orders.push_back(a[i]);
            }
        // advance
        unsigned sum = 0;
        for (unsigned v : a) sum += v;
        int i = int(dim) - 1;
        while (i >= 0) {
            if (sum < order) {
                a[i]++;
                break;
            }
            sum -= a[i];
            a[i] = 0;
            --i;
            if (i < 0) break;
            // carry into a[i]; recompute sum condition on next loop
        }
        if (i < 0) break;
    }
    s.starts.push_back(uint32_t(s.dims.size()));
    s.max_degrees();
    return s;
}

MultiSet csr_set(unsigned dim, unsigned nterms, const uint32_t* starts, const uint32_t* dims, const uint32_t* orders) {
    if (dim == 0) throw ArgError("CreateComponent: dimension must be positive");
    MultiSet s;
    s.dim = dim;
    s.starts.assign(starts, starts + nterms + 1);
    const unsigned nnz = starts[nterms];
    s.dims.assign(dims, dims + nnz);
    s.orders.assign(orders, orders + nnz);
    for (unsigned t = 0; t < nterms; ++t) {
        if (s.starts[t] > s.starts[t + 1]) throw ArgError("CreateComponent: nzStarts must be non-decreasing");
        std::vector<std::pair<uint32_t, uint32_t>> e;
        for (unsigned i = s.starts[t]; i < s.starts[t + 1]; ++i) {
            if (s.dims[i] >= dim) throw ArgError("CreateComponent: nzDims entry out of range");
            if (s.orders[i] == 0) throw ArgError("CreateComponent: nzOrders entries must be positive");
            e.push_back({s.dims[i], s.orders[i]});
        }
        std::stable_sort(e.begin(), e.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < e.size(); ++k) {
            s.dims[s.starts[t] + k] = e[k].first;
            s.orders[s.starts[t] + k] = e[k].second;
        }
    }
    s.max_degrees();
    return s;
}

// ---------------------------------------------------------------------------
// Device buffer.  Grown on demand.  Once `keep` is set (a call of the owner
// has been captured in a graph, whose launches hold this buffer's address),
// growing retires the old allocation instead of freeing it: it stays valid
// until the owner is destroyed, so the graph never reads freed memory.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool keep = false;
    std::vector<void*> retired;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
        for (void* q : retired) (void)hipFree(q);
    }
    void ensure(size_t b) {
        if (b <= bytes) return;
        if (p) {
            if (keep) retired.push_back(p);
            else hip_check(hipFree(p), "hipFree");
        }
        p = nullptr;
        bytes = 0;
        hip_check(hipMalloc(&p, b), "hipMalloc");
        bytes = b;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

template <class T>
void upload(DBuf& b, const std::vector<T>& v) {
    b.ensure(std::max<size_t>(v.size() * sizeof(T), 8));
    if (!v.empty()) hip_check(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

unsigned factorial_u32(unsigned d) {
    unsigned out = 1;
    for (unsigned i = 2; i <= d; ++i) out *= i;
    return out;
}

int g_numCU = 0;

// Compute units of the current device (queried once).
int num_cu() {
    if (g_numCU == 0) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&g_numCU, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    }
    return g_numCU;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host memory at the boundary.  The reference's ConditionalMapBase methods take
// host (Kokkos::HostSpace) views as well as device ones and copy what they need
// (src/ConditionalMapBase.cpp:7-81, src/ParameterizedFunctionBase.cpp:24-33).
// Here a pointer HIP does not know as device memory (pageable or pinned host
// memory) is staged: inputs are packed and copied to the device before the
// call, outputs are computed into device buffers and copied back, after which
// the call has completed (it synchronizes the stream).  Device pointers pass
// through untouched, so device calls stay asynchronous.
namespace {

// Only memory the kernels can read in place stays as it is: device and
// managed allocations.  Everything else is staged: pageable memory (reported
// as an error, or as hipMemoryTypeUnregistered, depending on the runtime),
// pinned host memory and any type this code does not know.
bool host_memory(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory HIP never saw: clear the error it recorded
        return true;
    }
    return !(a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeArray || a.type == hipMemoryTypeManaged);
}

class Staging {
   public:
    explicit Staging(hipStream_t s) : s_(s) {}
    mpart_cview in(const mpart_cview& v) {
        if (v.rows <= 0 || v.cols <= 0 || !host_memory(v.ptr)) return v;  // the callee checks the shape
        std::vector<double> h((size_t)(v.rows * v.cols));
        for (int64_t i = 0; i < v.rows; ++i)
            for (int64_t j = 0; j < v.cols; ++j) h[(size_t)(i * v.cols + j)] = v.ptr[i * v.row_stride + j * v.col_stride];
        double* d = upload_vec(h);
        return {d, v.rows, v.cols, v.cols, 1};
    }
    const double* in_vec(const double* p, int64_t n, int64_t stride) {
        if (n <= 0 || !host_memory(p)) return p;
        std::vector<double> h((size_t)n);
        for (int64_t i = 0; i < n; ++i) h[(size_t)i] = p[i * stride];
        return upload_vec(h);
    }
    mpart_view out(const mpart_view& v) {
        if (v.rows <= 0 || v.cols <= 0 || !host_memory(v.ptr)) return v;
        double* d = device((size_t)(v.rows * v.cols));
        const mpart_view hv = v;
        back_.push_back([this, hv, d] {
            std::vector<double> h((size_t)(hv.rows * hv.cols));
            hip_check(hipMemcpy(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
            for (int64_t i = 0; i < hv.rows; ++i)
                for (int64_t j = 0; j < hv.cols; ++j) hv.ptr[i * hv.row_stride + j * hv.col_stride] = h[(size_t)(i * hv.cols + j)];
        });
        return {d, v.rows, v.cols, v.cols, 1};
    }
    // a vector output (n entries with a stride); `read` copies its current
    // values in first (for outputs the call accumulates into)
    double* out_vec(double* p, int64_t n, int64_t stride, bool read = false) {
        if (n <= 0 || !host_memory(p)) return p;
        double* d = device((size_t)n);
        if (read) hip_check(hipMemcpy(d, in_vec(p, n, stride), n * sizeof(double), hipMemcpyDeviceToDevice), "hipMemcpy");
        back_.push_back([p, n, stride, d] {
            std::vector<double> h((size_t)n);
            hip_check(hipMemcpy(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
            for (int64_t i = 0; i < n; ++i) p[i * stride] = h[(size_t)i];
        });
        return d;
    }
    bool staged() const { return !bufs_.empty(); }
    // after the call: wait for it and copy host outputs back
    void finish() {
        if (bufs_.empty()) return;
        hip_check(hipStreamSynchronize(s_), "hipStreamSynchronize");
        for (auto& f : back_) f();
    }

   private:
    double* device(size_t n) {
        if (bufs_.empty()) {
            // staging copies synchronously and frees its buffers when the call
            // returns, which a graph being captured on the stream cannot take
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            hip_check(hipStreamIsCapturing(s_, &cs), "hipStreamIsCapturing");
            if (cs != hipStreamCaptureStatusNone)
                throw ArgError("host memory cannot be passed while the stream is being captured; pass device memory");
        }
        bufs_.push_back(std::make_unique<DBuf>());
        bufs_.back()->ensure(std::max<size_t>(8, n * sizeof(double)));
        return bufs_.back()->as<double>();
    }
    double* upload_vec(const std::vector<double>& h) {
        double* d = device(h.size());
        hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        return d;
    }
    hipStream_t s_;
    std::vector<std::unique_ptr<DBuf>> bufs_;
    std::vector<std::function<void()>> back_;
};

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of mpart_kl_value_and_grad_noalloc: a fixed device array, one copy
// per GPU, held until the process ends.  Host code never passes `scratch`
// itself (in host code the name stands for a host-side shadow, not a device
// address): each device's address comes from hipGetSymbolAddress, called once
// per device with that device current, and is kept in g_scratchDev.
static __device__ __align__(256) double scratch[MPART_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratchMu;
double* g_scratchDev[SCRATCH_MAX_DEVICES] = {};

// Device address of the current device's scratch (set up on first use).
double* noalloc_scratch() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) throw RtError("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchDev[dev]) {
        void* ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) throw RtError("hipGetSymbolAddress(scratch) returned no address");
        g_scratchDev[dev] = static_cast<double*>(ptr);
    }
    return g_scratchDev[dev];
}
}  // namespace

// ---------------------------------------------------------------------------
// Schedules of the specialized fused pass the tuner times (gen_fused_launch).
constexpr int GEN_FUSED_SCHEDULES = 4;
// Inverse hand-over defaults and sizes (heavy_args).
constexpr int HEAVY_SOLVE_STEPS = 1024;
constexpr long HEAVY_MAX_BLOCKS = 2560;  // ten 15.5 KB one-wavefront blocks per CU of 256

struct mpart_map_s {
    mpart_options opts{};
    bool triangular = false;
    unsigned inDim = 0, outDim = 0, numCoeffs = 0;
    std::vector<MultiSet> comps;  // per component
    std::vector<unsigned> coeffOff;
    int variant = 0;
    bool plain = false;

    // device plan
    MapDev dev{};
    DBuf dComps, dGroups, dTerms, dSlotDim, dDimSlot, dDimDeg;
    DBuf dStatus;  // MapDev::status
    DBuf dRule;    // quadrature rule of more than MAXNF points (MapDev::quad.big)
    // Component-split fused schedule: [waveStart[0..4] | component order]
    DBuf dSplit;
    int splitOcc[2] = {0, 0};   // resident blocks per CU of the two builds
    int splitMinB = 0;          // build the tuner chose (1 or 2; 0: not timed, 1 is used)
    double splitTrialMs[2] = {0, 0};

    // coefficients
    DBuf ownCoeffs;
    double* coeffs = nullptr;
    long coeffStride = 1;

    // work buffers
    DBuf partialGrad, partialLoss, gtab, wT, wLd, wG, wLG, wFull, wCoeffs;
    DBuf wInv;  // inverse: per-point records + the point counter and the hand-over count
    DBuf wHeavy;  // inverse hand-over: list of handed-over solves + per-wavefront workspaces
    int heavySolveSteps = HEAVY_SOLVE_STEPS;
    // A call on this map has been captured in a graph: from then on the work
    // buffers are never freed while the map lives (a larger eager call
    // allocates new ones and retires the old ones, which the graph reads).
    bool captured = false;
    void note_capture(hipStream_t s) {
        if (captured) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs == hipStreamCaptureStatusNone) return;
        captured = true;
        for (DBuf* b : {&ownCoeffs, &partialGrad, &partialLoss, &gtab, &wT, &wLd, &wG, &wLG, &wFull, &wCoeffs, &wInv,
                        &wHeavy})
            b->keep = true;
    }

    // runtime-specialized kernels (codegen.hip)
    bool genEnabled = true;
    bool genTried = false;
    bool genInTried = false;
    mpg::GenModule* genIn = nullptr;
    mpg::GenModule* gen = nullptr;
    double genSeconds = 0.0;
    std::string genErr;
    int genOccK[2] = {0, 0};
    int genOccIn = 0;  // resident blocks per CU of the fused kernel with point gradients
    size_t genOccSmem = 0;
    bool genW3Spills = false;  // the three-waves kernel needs more private memory (register spills)
    int fusedChoice = -1;  // tuned schedule of the fused kernel (-1: not yet)
    // Point schedule of the generic Evaluate / LogDeterminant: lanes per point
    // (1, 4, 8 or 64) for [Evaluate (with or without log-det), log-det only];
    // 0: not tuned yet.  pointGroupForced > 0 overrides the tuner.
    int evalGroup[2] = {0, 0};
    double evalTrialMs[2][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};
    int evalSplitOcc = 0;
    int pointGroupForced = 0;
    int fusedGroup = 0;  // the same for the generic fused pass
    // trials of the generic fused pass: lanes per point 1, 4, 8, 64, then the
    // component-split schedule
    double fusedGroupTrialMs[5] = {0, 0, 0, 0, 0};
    double fusedTrialMs[GEN_FUSED_SCHEDULES] = {0, 0, 0, 0};
};

namespace {

void build_plan(mpart_map_s& m) {
    const mpart_options& o = m.opts;
    if (o.basisType != MPART_BASIS_PROBABILIST_HERMITE && o.basisType != MPART_BASIS_PHYSICIST_HERMITE &&
        o.basisType != MPART_BASIS_HERMITE_FUNCTIONS)
        throw ArgError("unknown basisType");
    if (o.quadType != MPART_QUAD_CLENSHAW_CURTIS && o.quadType != MPART_QUAD_ADAPTIVE_SIMPSON &&
        o.quadType != MPART_QUAD_ADAPTIVE_CLENSHAW_CURTIS)
        throw ArgError("unknown quadType");
    // isLinearized = !isinf(basisLB) || !isinf(basisUB) (MParT/MapFactory.h:226)
    const bool lin = !std::isinf(o.basisLB) || !std::isinf(o.basisUB);
    if (lin && !(o.basisLB < o.basisUB)) throw ArgError("LinearizedBasis needs basisLB < basisUB");
    if (o.posFuncType != MPART_POSFUNC_EXP && o.posFuncType != MPART_POSFUNC_SOFTPLUS)
        throw ArgError("unknown posFuncType");
    if (o.quadMaxSub < 1 || o.quadMaxSub > (unsigned)mpd::MAXSUB)
        throw ArgError("quadMaxSub must be in [1, " + std::to_string(mpd::MAXSUB) +
                       "] (deeper trees overflow the reference's 32-bit subdivision path, MParT/Quadrature.h:566)");
    if (o.quadMinSub >= o.quadMaxSub)
        throw ArgError("quadMinSub must be smaller than quadMaxSub (the subdivision depth is bounded by quadMaxSub)");
    if (o.nugget < 0) throw ArgError("nugget must be non-negative");

    const int bk = (o.basisType == MPART_BASIS_HERMITE_FUNCTIONS)   ? mpd::BK_HF
                   : (o.basisType == MPART_BASIS_PHYSICIST_HERMITE) ? mpd::BK_PHYS
                                                                     : mpd::BK_PH;
    const int pk = (o.posFuncType == MPART_POSFUNC_EXP) ? mpd::PK_EXP : mpd::PK_SP;
    MapDev& M = m.dev;
    std::memset(&M, 0, sizeof(M));

    // Quadrature.
    M.quad.absTol = o.quadAbsTol;
    M.quad.relTol = o.quadRelTol;
    M.quad.maxSub = (int)o.quadMaxSub;
    M.quad.minSub = (int)o.quadMinSub;
    int qk = mpd::QK_AS;
    // Clenshaw-Curtis rule (MParT/Quadrature.h:212-258).
    auto rule = [](unsigned n, double* w, double* p) {
        if (n == 1) {
            p[0] = 0.0;
            w[0] = 2.0;
            return;
        }
        for (unsigned i = 0; i < n; ++i) p[i] = std::cos((n - (i + 1)) * M_PI / (n - 1));
        p[0] = -1.0;
        if (n % 2 == 1) p[(n - 1) / 2] = 0.0;
        p[n - 1] = 1.0;
        for (unsigned i = 0; i < n; ++i) {
            w[i] = 1.0;
            double theta = i * M_PI / (n - 1);
            for (unsigned j = 0; j < (n - 1) / 2; ++j) {
                double b = (2 * (j + 1) == (n - 1)) ? 1.0 : 2.0;
                w[i] = w[i] - b * std::cos(2.0 * (j + 1) * theta) / (4 * (j + 1) * (j + 1) - 1);
            }
        }
        w[0] = w[0] / (n - 1.0);
        for (unsigned i = 1; i < (n - 1); ++i) w[i] = 2.0 * w[i] / (n - 1.0);
        w[n - 1] = w[n - 1] / (n - 1.0);
    };
    // Rules of more than MAXNF points go to device memory (QuadDev::big):
    // cw, cp (nc each), then fw, fp (nf each).
    std::vector<double> bigRule;
    if (o.quadType == MPART_QUAD_CLENSHAW_CURTIS) {
        // ClenshawCurtisQuadrature(quadPts) (src/MapFactoryImpl5.cpp:17)
        if (o.quadPts < 1) throw ArgError("ClenshawCurtis needs quadPts >= 1");
        const unsigned n = o.quadPts;
        if (n > (unsigned)mpd::MAXNF) {
            bigRule.resize(2 * (size_t)n);
            rule(n, bigRule.data(), bigRule.data() + n);
        } else {
            rule(n, M.quad.fw, M.quad.fp);
        }
        M.quad.nf = (int)n;
        qk = mpd::QK_CC;
    }
    if (o.quadType == MPART_QUAD_ADAPTIVE_CLENSHAW_CURTIS) {
        if (o.quadPts < 3) throw ArgError("AdaptiveClenshawCurtis needs quadPts >= 4");
        // level = floor(log2(quadPts - 2)) (src/MapFactoryImpl4.cpp:18)
        const unsigned level = (unsigned)std::log2(o.quadPts - 2);
        if (level < 1) throw ArgError("AdaptiveClenshawCurtis needs a nested level >= 1 (quadPts >= 4)");
        if (level > 28) throw ArgError("AdaptiveClenshawCurtis: quadPts too large for a nested rule");
        const unsigned nc = (unsigned)std::pow(2, level) + 1, nf = (unsigned)std::pow(2, level + 1) + 1;
        if (nf > (unsigned)mpd::MAXNF) {
            bigRule.resize(2 * (size_t)nc + 2 * (size_t)nf);
            double* b = bigRule.data();
            rule(nc, b, b + nc);
            rule(nf, b + 2 * (size_t)nc, b + 2 * (size_t)nc + nf);
        } else {
            rule(nc, M.quad.cw, M.quad.cp);
            rule(nf, M.quad.fw, M.quad.fp);
        }
        M.quad.nc = (int)nc;
        M.quad.nf = (int)nf;
        qk = (nf == 5) ? mpd::QK_ACC5 : mpd::QK_ACC;
    }
    m.variant = bk * 100 + pk * 10 + qk;

    // Basis constants.
    M.normalize = (bk != mpd::BK_HF && o.basisNorm) ? 1 : 0;
    // Normalization sqrt(sqrt(2 pi) k!) (OrthogonalPolynomial.h:297) or
    // sqrt(sqrt(pi) 2^k k!) (313-315), with the 32-bit factorial.
    for (int k = 0; k <= mpd::MAXORD; ++k)
        M.norm[k] = (bk == mpd::BK_PHYS)
                        ? std::sqrt(std::sqrt(M_PI) * std::pow(2.0, static_cast<double>(k)) * factorial_u32(k))
                        : std::sqrt(std::sqrt(2.0 * M_PI) * factorial_u32(k));
    // p_o' = A o p_{o-1} (OrthogonalPolynomial.h:87-118 in closed form): with
    // phi_o = p_o / n_o, phi_o' = A o (n_{o-1} / n_o) phi_{o-1}, and phi_1' = A / n_1.
    if (bk != mpd::BK_HF) {
        const double A = (bk == mpd::BK_PHYS) ? 2.0 : 1.0;
        M.dfac[0] = 0.0;
        for (int k = 1; k <= mpd::MAXORD; ++k) {
            const double raw = A * k;
            if (M.normalize) M.dfac[k] = (k == 1) ? A / M.norm[1] : raw * M.norm[k - 1] / M.norm[k];
            else M.dfac[k] = raw;
        }
    }
    M.discrete = o.contDeriv ? 0 : 1;
    M.lin = lin ? 1 : 0;
    M.lb = lin ? o.basisLB : -std::numeric_limits<double>::infinity();
    M.ub = lin ? o.basisUB : std::numeric_limits<double>::infinity();
    M.hfPi = std::pow(M_PI, -0.25);
    M.hfSqrt2 = std::sqrt(2.0);
    {
        double fact = 1;
        for (int i = 0; i <= mpd::MAXORD; ++i) {
            M.hfA[i] = (i >= 1) ? std::sqrt(0.5 * (i - 1)) : 0.0;
            M.hfB[i] = std::sqrt(0.5 * i);
            if (i >= 1) fact *= i;
            M.hfS[i] = (i == 0) ? 1.0 : std::pow(std::pow(2, i) * fact, -0.5);
        }
    }
    // phi~_q(0) and the derivative folding rho_q.
    M.phi0[0] = 1.0;
    M.rho[0] = 0.0;
    if (bk != mpd::BK_HF) {
        const double A = (bk == mpd::BK_PHYS) ? 2.0 : 1.0;
        double hm2 = 1.0, hm1 = 0.0;  // p_0(0), p_1(0)
        for (int q = 1; q <= mpd::MAXP; ++q) {
            double h = (q == 1) ? 0.0 : (A * 0.0 + 0.0) * hm1 - (A * (q - 1.0)) * hm2;
            if (q >= 2) {
                hm2 = hm1;
                hm1 = h;
            }
            M.phi0[q] = M.normalize ? h / M.norm[q] : h;
            M.rho[q] = M.normalize ? 1.0 / M.norm[q] : 1.0;
        }
    } else {
        // EvaluateAll at 0: phi_1 = 0, phi_2 = pi^-1/4, phi_3 = 0, ...
        double v[mpd::MAXP + 1];
        v[0] = 1.0;
        v[1] = 0.0;
        if (mpd::MAXP >= 2) v[2] = std::pow(M_PI, -0.25) * std::exp(-0.5 * 0.0 * 0.0);
        if (mpd::MAXP >= 3) v[3] = std::sqrt(2.0) * 0.0 * v[2];
        for (int i = 2; i + 2 <= mpd::MAXP; ++i)
            v[i + 2] = (0.0 * v[i + 1] - std::sqrt(0.5 * (i - 1)) * v[i]) / std::sqrt(0.5 * i);
        for (int q = 1; q <= mpd::MAXP; ++q) {
            M.phi0[q] = v[q];
            M.rho[q] = (q == 1) ? 1.0 : M.hfPi * M.hfS[q - 2];
        }
    }
    // A linearized basis whose bounds exclude 0 continues linearly there
    // (LinearizedBasis::EvaluateAll at 0).
    if (lin && (0.0 < M.lb || 0.0 > M.ub)) {
        const double c = (0.0 < M.lb) ? M.lb : M.ub;
        for (int q = 1; q <= mpd::MAXP; ++q) {
            double v, d;
            if (bk == mpd::BK_HF) mpd::basis_vd<mpd::BK_HF>(M, q, c, v, d);
            else if (bk == mpd::BK_PHYS) mpd::basis_vd<mpd::BK_PHYS>(M, q, c, v, d);
            else mpd::basis_vd<mpd::BK_PH>(M, q, c, v, d);
            M.phi0[q] = v + d * (0.0 - c);
        }
    }
    M.nugget = o.nugget;
    // status word of the integrand check (0: no infinite g met since the last check)
    m.dStatus.ensure(sizeof(int));
    hip_check(hipMemset(m.dStatus.p, 0, sizeof(int)), "hipMemset");
    M.status = m.dStatus.as<int>();
    if (!bigRule.empty()) {
        m.dRule.ensure(bigRule.size() * sizeof(double));
        hip_check(hipMemcpy(m.dRule.p, bigRule.data(), bigRule.size() * sizeof(double), hipMemcpyHostToDevice),
                  "hipMemcpy");
        M.quad.big = m.dRule.as<double>();
    }

    // Off-diagonal table slots: max off-diagonal degree of every coordinate.
    const unsigned D = m.inDim;
    std::vector<int> dimDeg(D, 0), dimSlot(D, 0), slotDim(1, -1);
    for (auto& s : m.comps)
        for (unsigned d = 0; d + 1 < s.dim; ++d) dimDeg[d] = std::max<int>(dimDeg[d], (int)s.maxDeg[d]);
    int nslots = 1;
    for (unsigned d = 0; d < D; ++d) {
        if (dimDeg[d] > mpd::MAXORD)
            throw NotImpl("off-diagonal degree above " + std::to_string(mpd::MAXORD) + " is not provided");
        dimSlot[d] = nslots;
        for (int k = 0; k < dimDeg[d]; ++k) slotDim.push_back((int)d);
        nslots += dimDeg[d];
    }

    // Term records grouped by the order of the last coordinate.
    int nd = 1;
    for (auto& s : m.comps)
        for (unsigned t = 0; t < s.size(); ++t) {
            int off = 0;
            for (unsigned i = s.starts[t]; i < s.starts[t + 1]; ++i)
                if (s.dims[i] != s.dim - 1) off++;
            nd = std::max(nd, off);
        }
    const int trec = nd + 1;
    std::vector<CompDev> comps;
    std::vector<GroupDev> groups;
    std::vector<uint32_t> terms;
    int recCount = 0;
    unsigned row = 0;
    for (size_t c = 0; c < m.comps.size(); ++c) {
        const MultiSet& s = m.comps[c];
        CompDev cd{};
        cd.dim = (int)s.dim;
        cd.nterms = (int)s.size();
        cd.coeffOff = (int)m.coeffOff[c];
        cd.outRow = (int)row;
        row += 1;
        cd.p = (int)s.maxDeg[s.dim - 1];
        if (cd.p > mpd::MAXP) throw NotImpl("last-coordinate degree above " + std::to_string(mpd::MAXP) + " is not provided");
        cd.groupOff = (int)groups.size();
        std::vector<std::vector<unsigned>> byq(cd.p + 1);
        for (unsigned t = 0; t < s.size(); ++t) {
            unsigned q = 0;
            if (s.starts[t + 1] > s.starts[t] && s.dims[s.starts[t + 1] - 1] == s.dim - 1) q = s.orders[s.starts[t + 1] - 1];
            byq[q].push_back(t);
        }
        for (int q = 0; q <= cd.p; ++q) {
            GroupDev g{};
            g.start = recCount;
            g.count = (int)byq[q].size();
            g.countPad = (g.count + 15) / 16 * 16;
            for (int r = 0; r < g.countPad; ++r) {
                std::vector<uint32_t> rec(trec, 0u);
                if (r < g.count) {
                    unsigned t = byq[q][r];
                    rec[0] = t;
                    int e = 1;
                    for (unsigned i = s.starts[t]; i < s.starts[t + 1]; ++i) {
                        if (s.dims[i] == s.dim - 1) continue;
                        rec[e++] = uint32_t(dimSlot[s.dims[i]] + (int)s.orders[i] - 1);
                    }
                } else {
                    rec[0] = mpd::DUMMY;
                }
                terms.insert(terms.end(), rec.begin(), rec.end());
            }
            recCount += g.countPad;
            groups.push_back(g);
        }
        comps.push_back(cd);
    }
    // Component-split schedule: the components over the eight wavefronts of a
    // block, longest first to the least loaded wavefront (cost ~ terms, plus
    // the quadrature).
    {
        std::vector<int> order(comps.size());
        for (size_t c = 0; c < comps.size(); ++c) order[c] = (int)c;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return comps[a].nterms > comps[b].nterms; });
        std::vector<std::vector<int>> per(mpd::FUSED_SPLIT_WAVES);
        std::vector<double> load(mpd::FUSED_SPLIT_WAVES, 0.0);
        for (int c : order) {
            const int wv = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            per[wv].push_back(c);
            load[wv] += 2.0 * comps[c].nterms + 60.0;
        }
        std::vector<int> plan(mpd::FUSED_SPLIT_WAVES + 1 + comps.size());
        int k = mpd::FUSED_SPLIT_WAVES + 1;
        for (int wv = 0; wv < mpd::FUSED_SPLIT_WAVES; ++wv) {
            plan[wv] = k - (mpd::FUSED_SPLIT_WAVES + 1);
            std::sort(per[wv].begin(), per[wv].end());
            for (int c : per[wv]) plan[k++] = c;
        }
        plan[mpd::FUSED_SPLIT_WAVES] = (int)comps.size();
        upload(m.dSplit, plan);
    }
    upload(m.dComps, comps);
    upload(m.dGroups, groups);
    upload(m.dTerms, terms);
    upload(m.dSlotDim, slotDim);
    upload(m.dDimSlot, dimSlot);
    upload(m.dDimDeg, dimDeg);
    M.comps = m.dComps.as<CompDev>();
    M.groups = m.dGroups.as<GroupDev>();
    M.terms = m.dTerms.as<uint32_t>();
    M.slotDim = m.dSlotDim.as<int>();
    M.dimSlot = m.dDimSlot.as<int>();
    M.dimDeg = m.dDimDeg.as<int>();
    M.trec = trec;
    M.nd = nd;
    M.nslots = nslots;
    M.D = (int)D;
    M.ncomp = (int)comps.size();
    M.numCoeffs = (int)m.numCoeffs;
    int pmax = 0;
    for (auto& cd : comps) pmax = std::max(pmax, cd.p);
    M.pmax = pmax;
    // Register arrays of the kernels are sized by a compile-time degree bound.
    if (pmax > 4) m.variant += 1000;
}

void init_coeffs_zero(mpart_map_s& m) {
    m.ownCoeffs.ensure(std::max<size_t>(8, m.numCoeffs * sizeof(double)));
    hip_check(hipMemset(m.ownCoeffs.p, 0, std::max<size_t>(8, m.numCoeffs * sizeof(double))), "hipMemset");
    m.coeffs = m.ownCoeffs.as<double>();
    m.coeffStride = 1;
}

mpart_map_s* make_map(const mpart_options& o, std::vector<MultiSet> comps, bool triangular) {
    auto m = std::make_unique<mpart_map_s>();
    m->opts = o;
    m->triangular = triangular;
    m->comps = std::move(comps);
    m->inDim = m->comps.back().dim;
    m->outDim = (unsigned)m->comps.size();
    unsigned off = 0;
    for (size_t i = 0; i < m->comps.size(); ++i) {
        if (i > 0 && m->comps[i].dim != m->comps[i - 1].dim + 1)
            throw ArgError("TriangularMap: the input dimension of component " + std::to_string(i) +
                           " must be one more than that of component " + std::to_string(i - 1));
        m->coeffOff.push_back(off);
        off += m->comps[i].size();
    }
    m->numCoeffs = off;
    build_plan(*m);
    init_coeffs_zero(*m);
    return m.release();
}

void check_map(mpart_map m) {
    if (!m) throw ArgError("null map handle");
}

// Waits for s, then fails with the reference's runtime error
// (MonotoneIntegrand.h:183-190) when a kernel on the map met an infinite g at
// a quadrature node since the last check, and clears the status word.
const char* const kInfiniteIntegrand =
    "MonotoneIntegrand: nan was encountered in value of g(df(...)). Use MonotoneIntegrand::setFailOnNaN for "
    "enabling NaN propagation.";
bool take_integrand_status(mpart_map_s& m, hipStream_t s) {
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    int v = 0;
    hip_check(hipMemcpy(&v, m.dev.status, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
    if (v) hip_check(hipMemset(m.dev.status, 0, sizeof(int)), "hipMemset");
    return v != 0;
}
void check_integrand(mpart_map_s& m, hipStream_t s) {
    if (take_integrand_status(m, s)) throw RtError(kInfiniteIntegrand);
}
void check_coeffs(mpart_map m) {
    if (!m->coeffs) throw RtError("The coefficients have not been set for this map");
}
hipStream_t S(mpart_stream s) { return static_cast<hipStream_t>(s); }

constexpr size_t LDS_LIMIT = 64 * 1024;

// Block size and table placement of a lane-per-point launch.
struct Shape {
    mpd::Launch L;
    double* gtab;
};

Shape shape_for(mpart_map_s& m, long N, hipStream_t s, int tables = 1) {
    Shape sh{};
    const size_t perLane = (size_t)m.dev.nslots * tables * sizeof(double);
    int block = 256;
    while (block > 64 && perLane * block > LDS_LIMIT) block /= 2;
    sh.L.block = dim3(block);
    const long grid = std::max<long>(1, (N + block - 1) / block);
    sh.L.grid = dim3((unsigned)grid);
    sh.L.stream = s;
    if (perLane * block <= LDS_LIMIT) {
        sh.L.smem = perLane * block;
        sh.gtab = nullptr;
    } else {
        sh.L.smem = 0;
        m.gtab.ensure(perLane * (size_t)grid * block);
        sh.gtab = m.gtab.as<double>();
    }
    return sh;
}

void check_view(const char* what, int64_t rows, int64_t cols, int64_t erows, int64_t ecols) {
    if (rows != erows || cols != ecols)
        throw ArgError(std::string(what) + ": expected a " + std::to_string(erows) + " x " + std::to_string(ecols) +
                       " matrix, got " + std::to_string(rows) + " x " + std::to_string(cols));
}

// ---------------------------------------------------------------------------
// Runtime-specialized kernels.
constexpr unsigned GEN_MAX_COEFFS = 4096;
constexpr long FUSED_TUNE_MIN_POINTS = 65536;

bool codegen_env_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("MPART_CODEGEN");
        return !(e && e[0] == '0');
    }();
    return on;
}

mpg::GenSpec gen_spec(const mpart_map_s& m) {
    mpg::GenSpec spec;
    spec.bk = (m.variant / 100) % 10;
    spec.pk = (m.variant / 10) % 10;
    spec.qk = m.variant % 10;
    spec.lin = m.dev.lin;
    // Off-diagonal derivatives from the values (phi_o' = dfac[o] phi_{o-1}):
    // orthogonal polynomials without LinearizedBasis bounds.
    spec.derivFromValues = (!m.dev.lin && spec.bk != mpd::BK_HF) ? 1 : 0;
    spec.pm = std::max(1, m.dev.pmax);
    spec.D = m.inDim;
    spec.numCoeffs = m.numCoeffs;
    spec.dimDeg.assign(m.inDim, 0);
    for (auto& c : m.comps)
        for (unsigned d = 0; d + 1 < c.dim; ++d) spec.dimDeg[d] = std::max<int>(spec.dimDeg[d], (int)c.maxDeg[d]);
    for (size_t c = 0; c < m.comps.size(); ++c) {
        const MultiSet& s = m.comps[c];
        mpg::GenComp g;
        g.dim = s.dim;
        g.coeffOff = m.coeffOff[c];
        g.p = s.maxDeg[s.dim - 1];
        g.starts = s.starts;
        g.dims = s.dims;
        g.orders = s.orders;
        spec.comps.push_back(std::move(g));
    }
    return spec;
}

mpg::GenModule* gen_for(mpart_map_s& m) {
    if (!m.genEnabled || !codegen_env_enabled()) return nullptr;
    if (m.genTried) return m.gen;
    m.genTried = true;
    if (m.numCoeffs > GEN_MAX_COEFFS) {
        m.genErr = "more than " + std::to_string(GEN_MAX_COEFFS) + " coefficients: generic kernels";
        return nullptr;
    }
    const auto t0 = std::chrono::steady_clock::now();
    m.gen = mpg::get_module(gen_spec(m), m.genErr, 0);
    m.genSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return m.gen;
}

// The specialized kernels with input gradients, a second module compiled on
// their first use (nullptr: generic kernels).
mpg::GenModule* gen_in_for(mpart_map_s& m) {
    if (!gen_for(m)) return nullptr;
    if (m.genInTried) return m.genIn;
    m.genInTried = true;
    std::string err;
    m.genIn = mpg::get_module(gen_spec(m), err, 1);
    if (!m.genIn) m.genErr = err;
    return m.genIn;
}

// Coefficients as a contiguous vector (the specialized kernels index w[i]).
const double* gen_coeffs(mpart_map_s& m, hipStream_t s) {
    if (m.coeffStride == 1) return m.coeffs;
    m.wCoeffs.ensure(std::max<size_t>(8, m.numCoeffs * sizeof(double)));
    hip_check(mpd::launch_copy(m.coeffs, 1, m.numCoeffs, 0, m.coeffStride, m.wCoeffs.as<double>(), 0, 1, s),
              "launch copy");
    return m.wCoeffs.as<double>();
}

void gen_launch(hipFunction_t f, long grid, int block, size_t smem, hipStream_t s, void** args, const char* what) {
    hip_check(hipModuleLaunchKernel(f, (unsigned)grid, 1, 1, (unsigned)block, 1, 1, (unsigned)smem, s, args, nullptr),
              what);
}

// Evaluate / LogDeterminant through the specialized module for components [cb, ce).
void gen_eval(mpart_map_s& m, mpg::GenModule* g, const double* w, const mpart_cview& p, double* out, long ors,
              long ocs, double* ld, long lds, int ldAcc, int32_t* nevals, int cb, int ce, hipStream_t s) {
    MapDev M = m.dev;
    const double* pts = p.ptr;
    long prs = p.row_stride, pcs = p.col_stride, N = p.cols;
    int* ne = nevals;
    void* args[] = {&M, &w, &pts, &prs, &pcs, &N, &out, &ors, &ocs, &ld, &lds, &ldAcc, &ne, &cb, &ce};
    gen_launch(g->eval, (N + 255) / 256, 256, 0, s, args, "launch specialized evaluate");
}

constexpr long EVAL_TUNE_MIN_POINTS = 16384;
// Point-schedule code of the component-split fused pass (mono_split.hpp).
constexpr int FUSED_SPLIT = 2;

// Generic Evaluate and/or LogDeterminant of components [cb, ce) with the
// map's point schedule: a lane per point (k_eval) or a group of 4, 8 or 64
// lanes per point (k_eval_grp).  The first call with enough points (outputs
// overwritten, not under stream capture) times the four schedules with HIP
// events and keeps the fastest for this kind of call.
void generic_eval(mpart_map_s& m, const mpart_cview& p, double* out, long ors, long ocs, double* ld, long lds,
                  int ldAcc, int32_t* nevals, int cb, int ce, hipStream_t s) {
    const long N = p.cols;
    auto run = [&](int G) {
        if (G == FUSED_SPLIT) {
            // components split over the wavefronts of a block: every component
            // at once only, and only when the tile's table fits
            const size_t smem = (size_t)m.dev.nslots * mpd::FUSED_SPLIT_TILE * sizeof(double);
            if (cb == 0 && ce == (int)m.comps.size() && m.comps.size() >= 2 && smem <= LDS_LIMIT &&
                m.dev.nslots >= mpd::FUSED_SPLIT_WAVES) {
                if (m.evalSplitOcc == 0) m.evalSplitOcc = mpd::occupancy_eval_split(m.variant, smem);
                const long ntiles = std::max<long>(1, (N + mpd::FUSED_SPLIT_TILE - 1) / mpd::FUSED_SPLIT_TILE);
                const long grid = std::min<long>(ntiles, (long)num_cu() * m.evalSplitOcc);
                const int* plan = m.dSplit.as<int>();
                mpd::Launch L{dim3((unsigned)grid), dim3(mpd::FUSED_SPLIT_WAVES * 64), smem, s};
                hip_check(mpd::launch_eval_split(m.variant, L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                                 p.col_stride, N, out, ors, ocs, ld, lds, ldAcc, nevals,
                                                 plan + mpd::FUSED_SPLIT_WAVES + 1, plan),
                          "launch evaluate (components split over wavefronts)");
                return;
            }
            G = 1;
        }
        if (G > 1) {
            const size_t tabBytes = (size_t)m.dev.nslots * sizeof(double);
            int block = 256;
            while (block > G && (size_t)(block / G) * tabBytes > LDS_LIMIT) block /= 2;
            if ((size_t)(block / G) * tabBytes <= LDS_LIMIT) {
                const long per = block / G;
                mpd::Launch L{dim3((unsigned)((N + per - 1) / per)), dim3(block), (size_t)per * tabBytes, s};
                hip_check(mpd::launch_eval_grp(m.variant, G, L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                               p.col_stride, N, out, ors, ocs, ld, lds, ldAcc, nevals, cb, ce),
                          "launch evaluate (group per point)");
                return;
            }
        }
        Shape sh = shape_for(m, N, s);
        hip_check(mpd::launch_eval(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                   p.col_stride, N, out, ors, ocs, ld, lds, ldAcc, nevals, cb, ce, sh.gtab),
                  "launch evaluate");
    };
    const int kind = out ? 0 : 1;
    static const int kGroups[5] = {1, 4, 8, 64, FUSED_SPLIT};
    if (m.pointGroupForced > 0) {
        run(m.pointGroupForced);
        return;
    }
    if (m.evalGroup[kind] == 0 && N >= EVAL_TUNE_MIN_POINTS && !ldAcc && !m.plain) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
        if (cs == hipStreamCaptureStatusNone) {
            hipEvent_t e0, e1;
            hip_check(hipEventCreate(&e0), "hipEventCreate");
            hip_check(hipEventCreate(&e1), "hipEventCreate");
            int best = 0;
            for (int i = 0; i < 5; ++i) {
                run(kGroups[i]);
                hip_check(hipEventRecord(e0, s), "hipEventRecord");
                run(kGroups[i]);
                hip_check(hipEventRecord(e1, s), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0.f;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                m.evalTrialMs[kind][i] = ms;
                if (ms < m.evalTrialMs[kind][best]) best = i;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            m.evalGroup[kind] = kGroups[best];
        }
    }
    run(m.evalGroup[kind] > 0 && !m.plain ? m.evalGroup[kind] : 1);
}

void do_evaluate(mpart_map_s& m, const mpart_cview& p, const mpart_view& o, int32_t* nevals, hipStream_t s) {
    check_coeffs(&m);
    check_view("Evaluate: pts", p.rows, p.cols, m.inDim, p.cols);
    check_view("Evaluate: output", o.rows, o.cols, m.outDim, p.cols);
    if (p.cols == 0) return;
    if (mpg::GenModule* g = gen_for(m)) {
        const double* w = gen_coeffs(m, s);
        if (m.plain) {
            for (int c = 0; c < m.dev.ncomp; ++c)
                gen_eval(m, g, w, p, o.ptr, o.row_stride, o.col_stride, nullptr, 0, 0, nevals, c, c + 1, s);
        } else {
            gen_eval(m, g, w, p, o.ptr, o.row_stride, o.col_stride, nullptr, 0, 0, nevals, 0, m.dev.ncomp, s);
        }
        return;
    }
    if (m.plain) {
        Shape sh = shape_for(m, p.cols, s);
        for (int c = 0; c < m.dev.ncomp; ++c)
            hip_check(mpd::launch_eval(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                       p.col_stride, p.cols, o.ptr, o.row_stride, o.col_stride, nullptr, 0, 0,
                                       nevals, c, c + 1, sh.gtab),
                      "launch evaluate");
    } else {
        generic_eval(m, p, o.ptr, o.row_stride, o.col_stride, nullptr, 0, 0, nevals, 0, m.dev.ncomp, s);
    }
}

void do_logdet(mpart_map_s& m, const mpart_cview& p, double* out, long os, hipStream_t s) {
    check_coeffs(&m);
    check_view("LogDeterminant: pts", p.rows, p.cols, m.inDim, p.cols);
    if (p.cols == 0) return;
    if (!m.opts.contDeriv) {
        // DiscreteDerivative (MonotoneComponent.h:586-638), generic kernels.
        Shape sh = shape_for(m, p.cols, s);
        if (m.plain) {
            for (int c = 0; c < m.dev.ncomp; ++c)
                hip_check(mpd::launch_logdet_discrete(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr,
                                                      p.row_stride, p.col_stride, p.cols, out, os, c > 0 ? 1 : 0, c,
                                                      c + 1, sh.gtab),
                          "launch logdet (discrete)");
        } else {
            hip_check(mpd::launch_logdet_discrete(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                                  p.col_stride, p.cols, out, os, 0, 0, m.dev.ncomp, sh.gtab),
                      "launch logdet (discrete)");
        }
        return;
    }
    if (mpg::GenModule* g = gen_for(m)) {
        const double* w = gen_coeffs(m, s);
        if (m.plain) {
            for (int c = 0; c < m.dev.ncomp; ++c)
                gen_eval(m, g, w, p, nullptr, 0, 0, out, os, c > 0 ? 1 : 0, nullptr, c, c + 1, s);
        } else {
            gen_eval(m, g, w, p, nullptr, 0, 0, out, os, 0, nullptr, 0, m.dev.ncomp, s);
        }
        return;
    }
    if (m.plain) {
        Shape sh = shape_for(m, p.cols, s);
        for (int c = 0; c < m.dev.ncomp; ++c)
            hip_check(mpd::launch_eval(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                       p.col_stride, p.cols, nullptr, 0, 0, out, os, c > 0 ? 1 : 0, nullptr, c, c + 1,
                                       sh.gtab),
                      "launch logdet");
    } else {
        generic_eval(m, p, nullptr, 0, 0, out, os, 0, nullptr, 0, m.dev.ncomp, s);
    }
}

void do_coeff_grad(mpart_map_s& m, const mpart_cview& p, const mpart_cview& sens, const mpart_view& o,
                   hipStream_t s) {
    check_coeffs(&m);
    check_view("CoeffGrad: pts", p.rows, p.cols, m.inDim, p.cols);
    check_view("CoeffGrad: sens", sens.rows, sens.cols, m.outDim, p.cols);
    check_view("CoeffGrad: output", o.rows, o.cols, m.numCoeffs, p.cols);
    if (p.cols == 0) return;
    if (mpg::GenModule* g = gen_for(m)) {
        const double* w = gen_coeffs(m, s);
        MapDev M = m.dev;
        const double* pts = p.ptr;
        long prs = p.row_stride, pcs = p.col_stride, N = p.cols;
        const double* sp = sens.ptr;
        long srs = sens.row_stride, scs = sens.col_stride;
        double* op = o.ptr;
        long ors = o.row_stride, ocs = o.col_stride;
        for (int c = 0; c < m.dev.ncomp; ++c) {
            void* args[] = {&M, &w, &pts, &prs, &pcs, &N, &sp, &srs, &scs, &op, &ors, &ocs, &c};
            gen_launch(g->coef, (N + 255) / 256, 256, 0, s, args, "launch specialized coeff_grad");
        }
        return;
    }
    Shape sh = shape_for(m, p.cols, s);
    for (int c = 0; c < m.dev.ncomp; ++c)
        hip_check(mpd::launch_coeff_grad(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                         p.col_stride, p.cols, sens.ptr, sens.row_stride, sens.col_stride, o.ptr,
                                         o.row_stride, o.col_stride, c, sh.gtab),
                  "launch coeff_grad");
}

void do_logdet_coeff_grad(mpart_map_s& m, const mpart_cview& p, const mpart_view& o, hipStream_t s) {
    check_coeffs(&m);
    check_view("LogDeterminantCoeffGrad: pts", p.rows, p.cols, m.inDim, p.cols);
    check_view("LogDeterminantCoeffGrad: output", o.rows, o.cols, m.numCoeffs, p.cols);
    if (p.cols == 0) return;
    if (!m.opts.contDeriv) {
        // DiscreteMixedJacobian / DiscreteDerivative (MonotoneComponent.h:185-206, 928-980).
        Shape sh = shape_for(m, p.cols, s);
        for (int c = 0; c < m.dev.ncomp; ++c)
            hip_check(mpd::launch_logdet_coeff_grad_discrete(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr,
                                                             p.row_stride, p.col_stride, p.cols, o.ptr, o.row_stride,
                                                             o.col_stride, c, sh.gtab),
                      "launch logdet_coeff_grad (discrete)");
        return;
    }
    if (mpg::GenModule* g = gen_for(m)) {
        const double* w = gen_coeffs(m, s);
        MapDev M = m.dev;
        const double* pts = p.ptr;
        long prs = p.row_stride, pcs = p.col_stride, N = p.cols;
        double* op = o.ptr;
        long ors = o.row_stride, ocs = o.col_stride;
        for (int c = 0; c < m.dev.ncomp; ++c) {
            void* args[] = {&M, &w, &pts, &prs, &pcs, &N, &op, &ors, &ocs, &c};
            gen_launch(g->ldcoef, (N + 255) / 256, 256, 0, s, args, "launch specialized logdet_coeff_grad");
        }
        return;
    }
    Shape sh = shape_for(m, p.cols, s);
    for (int c = 0; c < m.dev.ncomp; ++c)
        hip_check(mpd::launch_logdet_coeff_grad(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                                p.col_stride, p.cols, o.ptr, o.row_stride, o.col_stride, c, sh.gtab),
                  "launch logdet_coeff_grad");
}

void do_input_vjp(mpart_map_s& m, const mpart_cview& p, int klMode, const mpart_cview* sens, const double* lsens,
                  long lss, double lval, double scale, const mpart_view& o, hipStream_t s, const char* nm);

void do_input_grad(mpart_map_s& m, const mpart_cview& p, const mpart_cview* sens, const mpart_view& o, int mode,
                   hipStream_t s) {
    check_coeffs(&m);
    const char* nm = mode == 0 ? "Gradient" : "LogDeterminantInputGrad";
    check_view((std::string(nm) + ": pts").c_str(), p.rows, p.cols, m.inDim, p.cols);
    if (sens) check_view((std::string(nm) + ": sens").c_str(), sens->rows, sens->cols, m.outDim, p.cols);
    check_view((std::string(nm) + ": output").c_str(), o.rows, o.cols, m.inDim, p.cols);
    if (mode == 1 && !m.opts.contDeriv)
        throw ArgError("Discrete derivative version is not implemented yet (To Do)");
    if (p.cols == 0) return;
    if (!m.plain && gen_for(m)) {
        // all components in one specialized launch
        do_input_vjp(m, p, 0, sens, nullptr, 0, mode == 1 ? 1.0 : 0.0, 1.0, o, s, nm);
        return;
    }
    hip_check(mpd::launch_fill(o.ptr, o.rows, o.cols, o.row_stride, o.col_stride, 0.0, s), "launch fill");
    Shape sh = shape_for(m, p.cols, s, 2);
    for (int c = 0; c < m.dev.ncomp; ++c)
        hip_check(mpd::launch_input_grad(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                         p.col_stride, p.cols, sens ? sens->ptr : nullptr, sens ? sens->row_stride : 0,
                                         sens ? sens->col_stride : 0, o.ptr, o.row_stride, o.col_stride, c, mode,
                                         sh.gtab),
                  "launch input_grad");
}

// Input VJP of every component in one launch: out = scale * grad_x [sum_c
// sens_c T_c + l logdet] with l = lsens[n] (or lval when lsens is null); in KL
// mode sens = T and l = -1.
void do_input_vjp(mpart_map_s& m, const mpart_cview& p, int klMode, const mpart_cview* sens, const double* lsens,
                  long lss, double lval, double scale, const mpart_view& o, hipStream_t s, const char* nm) {
    check_coeffs(&m);
    check_view((std::string(nm) + ": pts").c_str(), p.rows, p.cols, m.inDim, p.cols);
    if (sens) check_view((std::string(nm) + ": sens").c_str(), sens->rows, sens->cols, m.outDim, p.cols);
    check_view((std::string(nm) + ": output").c_str(), o.rows, o.cols, m.inDim, p.cols);
    if ((klMode || lsens || lval != 0.0) && !m.opts.contDeriv)
        throw ArgError("Discrete derivative version is not implemented yet (To Do)");
    if (p.cols == 0) return;
    const double* sp = sens ? sens->ptr : nullptr;
    long srs = sens ? sens->row_stride : 0, scs = sens ? sens->col_stride : 0;
    if (mpg::GenModule* g = gen_in_for(m)) {
        const double* w = gen_coeffs(m, s);
        MapDev M = m.dev;
        const double* pts = p.ptr;
        long prs = p.row_stride, pcs = p.col_stride, N = p.cols;
        double* op = o.ptr;
        long ors = o.row_stride, ocs = o.col_stride;
        void* args[] = {&M, &w, &pts, &prs, &pcs, &N, &sp, &srs, &scs, &lsens, &lss, &lval, &scale, &op, &ors, &ocs};
        gen_launch(klMode ? g->klx : g->ivjp, (N + 255) / 256, 256, 0, s, args, "launch specialized input_vjp");
        return;
    }
    Shape sh = shape_for(m, p.cols, s, 2);
    hip_check(mpd::launch_input_vjp(m.variant, sh.L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                    p.col_stride, p.cols, klMode, sp, srs, scs, lsens, lss, lval, scale, o.ptr,
                                    o.row_stride, o.col_stride, sh.gtab),
              "launch input_vjp");
}

// Hand-over of the longest inverse solves (adaptive rules): once no point is
// left to take, a queue lane hands a solve to the one-wavefront-per-solve
// pass if the solve has taken HEAVY_SOLVE_STEPS interval steps (the median
// solve takes a few hundred); the
// list holds up to N / 16 (at least 1 024) solves, taken one at a time by
// heavyBlocks wavefronts (one per 256 points, at most HEAVY_MAX_BLOCKS), each with an
// LDS workspace of COOP_NODES intervals per quadrature (larger trees are
// integrated depth-first by all lanes).
long heavy_cap(long N) { return std::max<long>(1024, N / 16); }
int heavy_blocks(long N) { return (int)std::min<long>(HEAVY_MAX_BLOCKS, std::max<long>(1, N / 256)); }
// Maps whose Evaluate in a batch, and whose Inverse setup, take 4 lanes per
// point: above the specialization limit, or with a per-lane basis table that
// does not fit 256 lanes in LDS.
bool many_terms(const mpart_map_s& m) {
    return m.numCoeffs >= GEN_MAX_COEFFS || (size_t)m.dev.nslots * 256 * sizeof(double) > LDS_LIMIT;
}

// Layout of the hand-over buffer: [statistics (256 B) | the arguments as the
// solve pass reads them (128 B), the take counter (128 B) | list].
constexpr size_t HEAVY_STATS_BYTES = 256;
mpd::InvHeavy heavy_args(DBuf& buf, unsigned long long* count, long N, int solveSteps, int& blocks,
                         mpd::InvHeavy*& Hd) {
    static_assert(mpd::HS_COUNT * sizeof(unsigned long long) <= HEAVY_STATS_BYTES, "statistics block");
    static_assert(sizeof(mpd::InvHeavy) <= 128, "argument block");
    mpd::InvHeavy H{};
    H.count = count;
    blocks = 0;
    Hd = nullptr;
    if (solveSteps <= 0) return H;
    const long cap = heavy_cap(N);
    blocks = heavy_blocks(N);
    buf.ensure(HEAVY_STATS_BYTES + 256 + (size_t)cap * sizeof(long));
    H.stats = buf.as<unsigned long long>();
    Hd = reinterpret_cast<mpd::InvHeavy*>(buf.as<char>() + HEAVY_STATS_BYTES);
    H.take = reinterpret_cast<unsigned long long*>(buf.as<char>() + HEAVY_STATS_BYTES + 128);
    H.list = reinterpret_cast<long*>(buf.as<char>() + HEAVY_STATS_BYTES + 256);
    H.cap = cap;
    H.solveSteps = solveSteps;
    return H;
}

// Statistics of the last Inverse call that used the hand-over (zeros if none).
void heavy_stats(const DBuf& buf, uint64_t* out, hipStream_t s) {
    for (int k = 0; k < mpd::HS_COUNT; ++k) out[k] = 0;
    if (!buf.p) return;
    hip_check(hipMemcpyAsync(out, buf.p, mpd::HS_COUNT * sizeof(uint64_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// Inverse of component c: a lane-per-point setup pass writes each point's
// record, then a persistent pass solves with per-lane scheduling (lanes take
// new points as their solves end).
void inverse_component(mpart_map_s& m, const Shape& sh, const double* xs, long xrows, long xcols, long xrs, long xcs,
                       const double* ys, long yss, long N, double* out, long os, int c, double xtol, double ytol) {
    if (g_numCU == 0) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&g_numCU, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    }
    const size_t recBytes = (size_t)mpd::inverse_record_rows(m.variant) * N * sizeof(double);
    m.wInv.ensure(recBytes + 256);
    double* rec = m.wInv.as<double>();
    auto* counter = reinterpret_cast<unsigned long long*>(m.wInv.as<char>() + recBytes);
    const int block = 256;
    const long grid = std::min<long>((N + block - 1) / block,
                                     (long)g_numCU * mpd::occupancy_inverse(m.variant, block));
    int heavyBlocks = 0;
    mpd::InvHeavy* Hd = nullptr;
    const mpd::InvHeavy H = heavy_args(m.wHeavy, counter + 1, N, m.heavySolveSteps, heavyBlocks, Hd);
    // maps with many terms: the 4-lanes-per-point setup of the batched Inverse (the same records)
    const size_t grpSmem = many_terms(m) ? (size_t)64 * m.dev.nslots * sizeof(double) : 0;
    hip_check(mpd::launch_inverse(m.variant, sh.L, dim3((unsigned)std::max<long>(1, grid)), dim3(block), m.dev,
                                  m.coeffs, m.coeffStride, xs, xrows, xcols, xrs, xcs, ys, yss, N, out, os, c, xtol,
                                  ytol, sh.gtab, rec, counter, H, Hd, heavyBlocks,
                                  grpSmem <= LDS_LIMIT ? grpSmem : 0),
              "launch inverse");
}

void do_inverse(mpart_map_s& m, const mpart_cview& x1, const mpart_cview& r, const mpart_view& o, double xtol,
                double ytol, hipStream_t s) {
    check_coeffs(&m);
    if (xtol < 0) throw ArgError("Invalid tolerance \"xtol\" given to MonotoneComponent::Inverse.  Value must be non-negative");
    if (ytol < 0) throw ArgError("Invalid tolerance \"ytol\" given to MonotoneComponent::Inverse.  Value must be non-negative");
    if (ytol <= 2.220446049250313e-16 && xtol <= 2.220446049250313e-16)
        throw ArgError("Invalid tolerances given to MonotoneComponent::Inverse.  Either \"xtol\" or \"ytol\" must be nonzero");
    const long N = r.cols;
    check_view("Inverse: r", r.rows, r.cols, m.outDim, N);
    check_view("Inverse: output", o.rows, o.cols, m.outDim, N);
    if (N == 0) return;
    if (!m.triangular) {
        if (x1.rows != m.inDim)
            throw ArgError("Inverse: x1 must have " + std::to_string(m.inDim) + " rows");
        if (x1.cols != 1 && x1.cols != N)
            throw ArgError("Invalid argument sizes given to MonotoneComponent::Inverse. The number of x points is " +
                           std::to_string(x1.cols) + ", but the number of y points is " + std::to_string(N) +
                           ".  If the number of xs is not 1 then it must match the number of ys.");
        Shape sh = shape_for(m, N, s);
        inverse_component(m, sh, x1.ptr, x1.rows, x1.cols, x1.row_stride, x1.col_stride, r.ptr, r.col_stride, N,
                          o.ptr, o.col_stride, 0, xtol, ytol);
        return;
    }
    // TriangularMap::InverseImpl (src/TriangularMap.cpp:147-184): components
    // run in order on a zero-initialized full point buffer.
    const long ip = m.inDim, op = m.outDim, extra = ip - op;
    check_view("Inverse: x1", x1.rows, x1.rows ? x1.cols : N, extra, N);
    m.wFull.ensure((size_t)ip * N * sizeof(double));
    double* full = m.wFull.as<double>();
    hip_check(mpd::launch_fill(full, ip, N, N, 1, 0.0, s), "launch fill");
    if (extra > 0)
        hip_check(mpd::launch_copy(x1.ptr, extra, N, x1.row_stride, x1.col_stride, full, N, 1, s), "launch copy");
    Shape sh = shape_for(m, N, s);
    for (int c = 0; c < m.dev.ncomp; ++c) {
        const long dimc = m.comps[c].dim;
        inverse_component(m, sh, full, dimc, N, N, 1, r.ptr + (long)c * r.row_stride, r.col_stride, N,
                          full + (extra + c) * N, 1, c, xtol, ytol);
    }
    hip_check(mpd::launch_copy(full + extra * N, op, N, N, 1, o.ptr, o.row_stride, o.col_stride, s), "launch copy");
}

// Block size, LDS bytes and resident blocks per CU of the specialized fused
// kernels (per-wave LDS accumulators of K doubles; K <= 4096 keeps them within
// 64 KiB at 128 lanes).  Queries only: nothing is allocated.
void gen_fused_shape(mpart_map_s& m, mpg::GenModule* g, int& gblock, size_t& gsmem) {
    const long K = m.numCoeffs;
    gblock = 256;
    while (gblock > 64 && (size_t)(gblock / 64) * K * sizeof(double) > LDS_LIMIT) gblock /= 2;
    gsmem = (size_t)(gblock / 64) * K * sizeof(double);
    if (m.genOccSmem == gsmem) return;
    hipFunction_t fns[2] = {g->fused, g->fused3};
    for (int k = 0; k < 2; ++k) {
        int nb = 0;
        hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fns[k], gblock, gsmem),
                  "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
        m.genOccK[k] = std::max(1, nb);
    }
    int priv[2] = {0, 0};
    for (int k = 0; k < 2; ++k)
        hip_check(hipFuncGetAttribute(&priv[k], HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, fns[k]), "hipFuncGetAttribute");
    m.genW3Spills = priv[1] > priv[0];
    m.genOccSmem = gsmem;
}

// One schedule of the specialized fused pass: mpart_gen_fused (cfg & 1: the
// three-waves build) on a persistent grid of one or two blocks per resident
// slot ((cfg >> 1) + 1).
struct GenFusedLaunch {
    hipFunction_t fn;
    int block;
    size_t smem;
    long grid;
};
GenFusedLaunch gen_fused_launch(const mpart_map_s& m, const mpg::GenModule* g, int gblock, size_t gsmem, int cfg,
                                long N) {
    GenFusedLaunch L{(cfg & 1) ? g->fused3 : g->fused, gblock, gsmem, 1};
    const long ntiles = std::max<long>(1, (N + gblock - 1) / gblock);
    L.grid = std::min<long>(ntiles, (long)g_numCU * m.genOccK[cfg & 1] * ((cfg >> 1) + 1));
    return L;
}

// Launch shape of the generic lane-per-point fused kernel: block size, LDS
// bytes, whether the per-wave gradient accumulators (ldsAcc) and the basis
// table (tabLds) fit in LDS, and the persistent grid.
struct LaneFusedShape {
    int block;
    size_t smem, tabLane;
    bool ldsAcc, tabLds;
    long grid;
};
LaneFusedShape lane_fused_shape(const mpart_map_s& m, long N) {
    const long K = m.numCoeffs;
    LaneFusedShape f{};
    f.tabLane = (size_t)m.dev.nslots * sizeof(double);
    auto smem_for = [&](int b, bool ldsAcc) {
        return (ldsAcc ? (size_t)(b / 64) * K * sizeof(double) : 0) + f.tabLane * b;
    };
    int block = 256;
    f.ldsAcc = true;
    while (block > 64 && smem_for(block, true) > LDS_LIMIT) block /= 2;
    if (smem_for(block, true) > LDS_LIMIT) f.ldsAcc = false;
    block = 256;
    while (block > 64 && smem_for(block, f.ldsAcc) > LDS_LIMIT) block /= 2;
    f.block = block;
    f.tabLds = smem_for(block, f.ldsAcc) <= LDS_LIMIT;
    f.smem = f.tabLds ? smem_for(block, f.ldsAcc) : (f.ldsAcc ? (size_t)(block / 64) * K * sizeof(double) : 0);
    const long ntiles = std::max<long>(1, (N + block - 1) / block);
    const int occ = mpd::occupancy_fused(m.variant, block, f.smem);
    f.grid = std::min<long>(ntiles, (long)g_numCU * occ);
    return f;
}

// Launch shape of the fused kernel: persistent grid, LDS accumulators when
// they fit, otherwise global atomics.
void do_input_vjp(mpart_map_s& m, const mpart_cview& p, int klMode, const mpart_cview* sens, const double* lsens,
                  long lss, double lval, double scale, const mpart_view& o, hipStream_t s, const char* nm);

// Fused forward + adjoint: the coefficient gradient (and in KL mode the loss)
// summed over the points, scaled by `scale`; with gin, also the per-point
// input gradient of the same sensitivities times gscale, from the same
// quadrature (specialized kernels) or by a second launch (generic kernels).
void do_fused(mpart_map_s& m, const mpart_cview& p, const double* sens, long srs, long scs, const double* lsens,
              long lss, int klMode, double* loss, double* grad, double scale, hipStream_t s,
              const mpart_view* gin = nullptr, double gscale = 1.0) {
    check_coeffs(&m);
    if (gin) check_view("input gradient output", gin->rows, gin->cols, m.inDim, p.cols);
    const long N = p.cols;
    const long K = m.numCoeffs;
    if (g_numCU == 0) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&g_numCU, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    }
    // the specialized kernels implement the continuous derivative only
    if (mpg::GenModule* g = m.dev.discrete ? nullptr : gen_for(m)) {
        int gblock = 256;
        size_t gsmem = 0;
        gen_fused_shape(m, g, gblock, gsmem);
        const long ntiles = std::max<long>(1, (N + gblock - 1) / gblock);
        auto grid_of = [&](int cfg) { return gen_fused_launch(m, g, gblock, gsmem, cfg, N).grid; };
        long maxGrid = 1;
        for (int cfg = 0; cfg < GEN_FUSED_SCHEDULES; ++cfg) maxGrid = std::max(maxGrid, grid_of(cfg));
        // Partial rows of the widest schedule, then the reduction's chunk rows.
        m.partialGrad.ensure((size_t)(maxGrid + mpd::COLSUM_MAX_CHUNKS) * (K + 1) * sizeof(double));
        const double* w = gen_coeffs(m, s);
        MapDev M = m.dev;
        const double* pts = p.ptr;
        long prs = p.row_stride, pcs = p.col_stride, NN = N;
        const double* sp = sens;
        long srs_ = srs, scs_ = scs;
        const double* lp = lsens;
        long lss_ = lss;
        int kl = klMode;
        double* part = m.partialGrad.as<double>();
        void* args[] = {&M, &w, &pts, &prs, &pcs, &NN, &sp, &srs_, &scs_, &lp, &lss_, &kl, &part};
        auto launch_cfg = [&](int cfg) {
            const GenFusedLaunch L = gen_fused_launch(m, g, gblock, gsmem, cfg, N);
            gen_launch(L.fn, L.grid, L.block, L.smem, s, args, "launch specialized fused");
        };
        mpg::GenModule* gi = gin ? gen_in_for(m) : nullptr;
        if (gi) {
            // One kernel for both gradients; persistent grid of two blocks per
            // resident slot (the schedule the tuner picks for the fused kernel).
            if (m.genOccIn == 0) {
                int nb = 0;
                hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gi->fusedIn, gblock, gsmem),
                          "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
                m.genOccIn = std::max(1, nb);
            }
            const long gridIn = std::min<long>(ntiles, (long)g_numCU * m.genOccIn * 2);
            m.partialGrad.ensure((size_t)(std::max(maxGrid, gridIn) + mpd::COLSUM_MAX_CHUNKS) * (K + 1) *
                                 sizeof(double));
            part = m.partialGrad.as<double>();
            double gs = gscale;
            double* gp = gin->ptr;
            long gors = gin->row_stride, gocs = gin->col_stride;
            void* argsIn[] = {&M, &w, &pts, &prs, &pcs, &NN, &sp, &srs_, &scs_, &lp, &lss_, &kl, &part,
                              &gs, &gp, &gors, &gocs};
            const long rows = (N > 0) ? gridIn : 0;
            if (N > 0) gen_launch(gi->fusedIn, gridIn, gblock, gsmem, s, argsIn, "launch specialized fused (inputs)");
            hip_check(mpd::launch_colsum(part, rows, K + 1, loss ? K + 1 : K, scale, grad, loss,
                                         part + std::max(maxGrid, gridIn) * (K + 1), s),
                      "launch reduce");
            return;
        }
        // Autotuning: on the first call with enough points (and not under
        // stream capture) every schedule runs twice and the faster one is kept.
        // The three-waves kernel is left out when its register cap makes it
        // spill (the spill traffic goes to memory); its trials report -1.
        if (m.fusedChoice < 0 && N >= FUSED_TUNE_MIN_POINTS) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
            if (cs == hipStreamCaptureStatusNone) {
                hipEvent_t e0, e1;
                hip_check(hipEventCreate(&e0), "hipEventCreate");
                hip_check(hipEventCreate(&e1), "hipEventCreate");
                int best = 0;
                for (int cfg = 0; cfg < GEN_FUSED_SCHEDULES; ++cfg) {
                    if ((cfg & 1) && m.genW3Spills) {
                        m.fusedTrialMs[cfg] = -1.0;
                        continue;
                    }
                    launch_cfg(cfg);
                    hip_check(hipEventRecord(e0, s), "hipEventRecord");
                    launch_cfg(cfg);
                    hip_check(hipEventRecord(e1, s), "hipEventRecord");
                    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                    float ms = 0.f;
                    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                    m.fusedTrialMs[cfg] = ms;
                    if (ms < m.fusedTrialMs[best]) best = cfg;
                }
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                m.fusedChoice = best;
            }
        }
        const int cfg = m.fusedChoice < 0 ? 0 : m.fusedChoice;
        if (N > 0) launch_cfg(cfg);
        // Column sums of the partial rows: K gradient columns, then the loss column.
        const long rows = (N > 0) ? grid_of(cfg) : 0;
        hip_check(mpd::launch_colsum(part, rows, K + 1, loss ? K + 1 : K, scale, grad, loss, part + maxGrid * (K + 1),
                                     s),
                  "launch reduce");
        if (gin) {
            // the input-gradient module did not build: a second (generic) launch
            mpart_cview sv{sens, m.outDim, N, srs, scs};
            do_input_vjp(m, p, klMode, sens ? &sv : nullptr, lsens, lss, 0.0, gscale, *gin, s, "input gradient");
        }
        return;
    }
    // Generic kernels: a lane per point (k_fused) or a group of 4, 8 or 64 lanes
    // per point (k_fused_grp), tuned like the generic Evaluate.
    auto run_lane = [&]() {
        const LaneFusedShape f = lane_fused_shape(m, N);
        const long grid = f.grid;
        const bool ldsAcc = f.ldsAcc;
        mpd::Launch L{dim3((unsigned)grid), dim3(f.block), f.smem, s};
        double* gtab = nullptr;
        if (!f.tabLds) {
            m.gtab.ensure(f.tabLane * (size_t)grid * f.block);
            gtab = m.gtab.as<double>();
        }
        m.partialGrad.ensure(std::max<size_t>(8, (size_t)grid * K * sizeof(double)));
        m.partialLoss.ensure(std::max<size_t>(8, (size_t)grid * sizeof(double)));
        double* gradGlobal = nullptr;
        if (!ldsAcc) {
            gradGlobal = m.partialGrad.as<double>();
            hip_check(mpd::launch_fill(gradGlobal, 1, K, K, 1, 0.0, s), "launch fill");
        }
        if (N > 0)
            hip_check(mpd::launch_fused(m.variant, L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                        p.col_stride, N, sens, srs, scs, lsens, lss, klMode,
                                        m.partialGrad.as<double>(), m.partialLoss.as<double>(), gtab, gradGlobal),
                      "launch fused");
        const long rows = (N > 0) ? (ldsAcc ? grid : 1) : 0;
        hip_check(mpd::launch_reduce_rows(m.partialGrad.as<double>(), rows, K, scale, grad, s), "launch reduce");
        if (loss)
            hip_check(mpd::launch_reduce_rows(m.partialLoss.as<double>(), N > 0 ? grid : 0, 1, scale, loss, s),
                      "launch reduce");
    };
    // false when the group kernel does not fit (the caller then runs run_lane)
    auto run_grp = [&](int G) -> bool {
        const size_t tabBytes = (size_t)m.dev.nslots * sizeof(double);
        const int block = 256;
        const size_t smem = (size_t)(block / G) * tabBytes;
        if (smem > LDS_LIMIT) return false;
        const long pointsPerBlock = block / G;
        const long nblocks = std::max<long>(1, (N + pointsPerBlock - 1) / pointsPerBlock);
        const int occ = mpd::occupancy_fused_grp(m.variant, G, block, smem);
        // one partial row of K + 1 doubles per wavefront, at most 2 GiB of rows
        const long rowBytes = (long)(K + 1) * (long)sizeof(double) * (block / 64);
        const long maxBlocks = std::max<long>(1, (2L << 30) / rowBytes);
        const long grid = std::min<long>(std::min<long>(nblocks, (long)g_numCU * occ), maxBlocks);
        const long nrows = grid * (block / 64);
        m.partialGrad.ensure((size_t)(nrows + mpd::COLSUM_MAX_CHUNKS) * (K + 1) * sizeof(double));
        double* rowsp = m.partialGrad.as<double>();
        mpd::Launch L{dim3((unsigned)grid), dim3(block), smem, s};
        if (N > 0)
            hip_check(mpd::launch_fused_grp(m.variant, G, L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                            p.col_stride, N, sens, srs, scs, lsens, lss, klMode, rowsp),
                      "launch fused (group per point)");
        hip_check(mpd::launch_colsum(rowsp, N > 0 ? nrows : 0, K + 1, loss ? K + 1 : K, scale, grad, loss,
                                     rowsp + nrows * (K + 1), s),
                  "launch reduce");
        return true;
    };
    // Components split over the wavefronts of a block (k_fused_split): false
    // when the tile's table does not fit or the map has one component.
    auto run_split = [&](int minb) -> bool {
        const size_t smem = (size_t)m.dev.nslots * mpd::FUSED_SPLIT_TILE * sizeof(double);
        // (the kernel's static LDS is the eight wavefronts' loss sums)
        if (smem + mpd::FUSED_SPLIT_WAVES * sizeof(double) > LDS_LIMIT || m.comps.size() < 2) return false;
        if (minb != 2) minb = 1;
        int& occ = m.splitOcc[minb - 1];
        if (occ == 0) occ = mpd::occupancy_fused_split(m.variant, minb, smem);
        const long ntiles = std::max<long>(1, (N + mpd::FUSED_SPLIT_TILE - 1) / mpd::FUSED_SPLIT_TILE);
        const long grid = std::min<long>(ntiles, (long)num_cu() * occ);
        m.partialGrad.ensure((size_t)(grid + mpd::COLSUM_MAX_CHUNKS) * (K + 1) * sizeof(double));
        double* rowsp = m.partialGrad.as<double>();
        const int* plan = m.dSplit.as<int>();
        mpd::Launch L{dim3((unsigned)grid), dim3(mpd::FUSED_SPLIT_WAVES * 64), smem, s};
        if (N > 0)
            hip_check(mpd::launch_fused_split(m.variant, minb, L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride,
                                              p.col_stride, N, sens, srs, scs, lsens, lss, klMode,
                                              plan + mpd::FUSED_SPLIT_WAVES + 1, plan, rowsp),
                      "launch fused (components split over wavefronts)");
        hip_check(mpd::launch_colsum(rowsp, N > 0 ? grid : 0, K + 1, loss ? K + 1 : K, scale, grad, loss,
                                     rowsp + grid * (K + 1), s),
                  "launch reduce");
        return true;
    };
    auto run = [&](int G) {
        // the discrete derivative: the lane-per-point and split kernels only
        if (m.dev.discrete && G != FUSED_SPLIT) G = 1;
        if (G == FUSED_SPLIT) {
            if (run_split(m.splitMinB)) return;
            G = 1;
        }
        if (G > 1 && run_grp(G)) return;
        run_lane();
    };
    static const int kGroups[5] = {1, 4, 8, 64, FUSED_SPLIT};
    if (m.pointGroupForced > 0) {
        run(m.pointGroupForced);
    } else {
        if (m.fusedGroup == 0 && N >= EVAL_TUNE_MIN_POINTS) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
            if (cs == hipStreamCaptureStatusNone) {
                hipEvent_t e0, e1;
                hip_check(hipEventCreate(&e0), "hipEventCreate");
                hip_check(hipEventCreate(&e1), "hipEventCreate");
                // the split schedule's two builds (one or two wavefronts per
                // SIMD) first; the faster represents it below
                for (int b = 0; b < 2; ++b) {
                    if (!run_split(b + 1)) break;
                    hip_check(hipEventRecord(e0, s), "hipEventRecord");
                    run_split(b + 1);
                    hip_check(hipEventRecord(e1, s), "hipEventRecord");
                    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                    float ms = 0.f;
                    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                    m.splitTrialMs[b] = ms;
                    m.splitMinB = (b == 0 || ms < m.splitTrialMs[0]) ? b + 1 : 1;
                }
                int best = 0;
                for (int i = 0; i < 5; ++i) {
                    run(kGroups[i]);
                    hip_check(hipEventRecord(e0, s), "hipEventRecord");
                    run(kGroups[i]);
                    hip_check(hipEventRecord(e1, s), "hipEventRecord");
                    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                    float ms = 0.f;
                    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                    m.fusedGroupTrialMs[i] = ms;
                    if (ms < m.fusedGroupTrialMs[best]) best = i;
                }
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                m.fusedGroup = kGroups[best];
            }
        }
        run(m.fusedGroup > 0 ? m.fusedGroup : 1);
    }
    if (gin) {
        mpart_cview sv{sens, m.outDim, N, srs, scs};
        do_input_vjp(m, p, klMode, sens ? &sv : nullptr, lsens, lss, 0.0, gscale, *gin, s, "input gradient");
    }
}

// The fused KL value and gradient (do_fused in KL mode) with every work
// buffer in `scr` (cap doubles, the no-allocation scratch): no device memory
// is allocated or freed and nothing is compiled.  The specialized kernels run
// when an earlier call (or mpart_codegen_prepare) compiled them, on the schedule
// the autotuner picked then (the default schedule if none was tuned); otherwise
// the generic lane-per-point kernel runs.  The persistent grid is reduced when
// its partial rows would not fit in `cap`.
void do_fused_noalloc(mpart_map_s& m, const mpart_cview& p, double* loss, double* grad, double scale, hipStream_t s,
                      double* scr, long cap) {
    check_coeffs(&m);
    const long N = p.cols, K = m.numCoeffs;
    if (g_numCU == 0) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&g_numCU, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    }
    const std::string tooBig = "KL without allocation: the map's " + std::to_string(K) +
                               " coefficients need more than the " + std::to_string(MPART_NOALLOC_SCRATCH_BYTES) +
                               "-byte scratch";
    // (the specialized kernels implement the continuous derivative only)
    mpg::GenModule* g = (m.genEnabled && codegen_env_enabled() && m.genTried && !m.dev.discrete) ? m.gen : nullptr;
    if (g) {
        int gblock = 256;
        size_t gsmem = 0;
        gen_fused_shape(m, g, gblock, gsmem);
        const int cfg = (m.fusedChoice < 0) ? 0 : m.fusedChoice;
        const GenFusedLaunch GL = gen_fused_launch(m, g, gblock, gsmem, cfg, N);
        long grid = GL.grid;
        // Layout: [partial rows: grid x (K + 1)][reduction chunks: <= 64 x (K + 1)] ...
        // [contiguous copy of strided coefficients: K, at the end].
        const long extra = (m.coeffStride == 1) ? 0 : K;
        long rmax = (cap - extra) / (K + 1) - mpd::COLSUM_MAX_CHUNKS;
        if (rmax < 1) {
            if (2 * (K + 1) + extra > cap) throw ArgError(tooBig);
            rmax = 1;
        }
        grid = std::min(grid, rmax);
        const double* w = m.coeffs;
        if (extra) {
            double* wc = scr + (cap - K);
            hip_check(mpd::launch_copy(m.coeffs, 1, K, 0, m.coeffStride, wc, 0, 1, s), "launch copy");
            w = wc;
        }
        MapDev M = m.dev;
        const double* pts = p.ptr;
        long prs = p.row_stride, pcs = p.col_stride, NN = N;
        const double* sp = nullptr;
        long srs = 0, scs = 0;
        const double* lp = nullptr;
        long lss = 0;
        int kl = 1;
        double* part = scr;
        void* args[] = {&M, &w, &pts, &prs, &pcs, &NN, &sp, &srs, &scs, &lp, &lss, &kl, &part};
        if (N > 0) gen_launch(GL.fn, grid, GL.block, GL.smem, s, args, "launch specialized fused");
        hip_check(mpd::launch_colsum(part, N > 0 ? grid : 0, K + 1, K + 1, scale, grad, loss, part + grid * (K + 1), s),
                  "launch reduce");
        return;
    }
    const LaneFusedShape f = lane_fused_shape(m, N);
    if (!f.tabLds)
        throw NotImpl("KL without allocation: the map's per-lane basis table does not fit in LDS (it would need a "
                      "global table)");
    // Layout: [gradient rows: grid x K (LDS accumulators) or one row of K
    // (global atomics)][loss partials: grid].
    long grid = f.grid;
    if (f.ldsAcc) grid = std::min(grid, cap / (K + 1));
    else grid = std::min(grid, cap - K);
    if (grid < 1) throw ArgError(tooBig);
    double* pg = scr;
    double* pl = scr + (f.ldsAcc ? grid * K : K);
    double* gradGlobal = nullptr;
    if (!f.ldsAcc) {
        gradGlobal = pg;
        hip_check(mpd::launch_fill(gradGlobal, 1, K, K, 1, 0.0, s), "launch fill");
    }
    mpd::Launch L{dim3((unsigned)grid), dim3(f.block), f.smem, s};
    if (N > 0)
        hip_check(mpd::launch_fused(m.variant, L, m.dev, m.coeffs, m.coeffStride, p.ptr, p.row_stride, p.col_stride,
                                    N, nullptr, 0, 0, nullptr, 0, 1, pg, pl, nullptr, gradGlobal),
                  "launch fused");
    hip_check(mpd::launch_reduce_rows(pg, N > 0 ? (f.ldsAcc ? grid : 1) : 0, K, scale, grad, s), "launch reduce");
    hip_check(mpd::launch_reduce_rows(pl, N > 0 ? grid : 0, 1, scale, loss, s), "launch reduce");
}

// The KL objective with the reference's launch structure
// (src/MapObjective.cpp:107-215): per-component Evaluate, LogDeterminant,
// CoeffGrad(sens = T) and LogDeterminantCoeffGrad with stored Jacobians,
// then row reductions.
void do_kl_plain(mpart_map_s& m, const mpart_cview& pts, double* loss, double* grad, hipStream_t st) {
    check_coeffs(&m);
    const long N = pts.cols, K = m.numCoeffs, Mo = m.outDim;
    check_view("KL: pts", pts.rows, pts.cols, m.inDim, N);
    if (N == 0) throw ArgError("KL: no points");
    m.wT.ensure((size_t)Mo * N * sizeof(double));
    m.wLd.ensure((size_t)N * sizeof(double));
    m.wG.ensure((size_t)K * N * sizeof(double));
    m.wLG.ensure((size_t)K * N * sizeof(double));
    const bool plainSave = m.plain;
    m.plain = true;
    mpart_view T{m.wT.as<double>(), Mo, N, N, 1};
    mpart_cview Tc{m.wT.as<double>(), Mo, N, N, 1};
    mpart_view G{m.wG.as<double>(), K, N, N, 1};
    mpart_view LG{m.wLG.as<double>(), K, N, N, 1};
    try {
        do_evaluate(m, pts, T, nullptr, st);
        do_logdet(m, pts, m.wLd.as<double>(), 1, st);
        do_coeff_grad(m, pts, Tc, G, st);
        do_logdet_coeff_grad(m, pts, LG, st);
    } catch (...) {
        m.plain = plainSave;
        throw;
    }
    m.plain = plainSave;
    hip_check(mpd::launch_kl_rows(m.wG.as<double>(), m.wLG.as<double>(), N, K, grad, st), "launch kl_rows");
    const long nb = (N + 255) / 256;
    m.partialLoss.ensure(std::max<size_t>(8, (size_t)nb * sizeof(double)));
    hip_check(mpd::launch_kl_loss(m.wT.as<double>(), Mo, m.wLd.as<double>(), N, m.partialLoss.as<double>(), st),
              "launch kl_loss");
    hip_check(mpd::launch_reduce_rows(m.partialLoss.as<double>(), nb, 1, 1.0 / (double)N, loss, st),
              "launch reduce");
}

// ---------------------------------------------------------------------------
// Batches of independent maps (one persistent launch per method).
}  // namespace

// Device-side arguments of one batched method (each method keeps its own, so a
// graph can hold a sequence of different batched calls).
struct BatchSlot {
    DBuf dIO, dTileMap, dTileStart, dTileLanes, dCounter;
    bool valid = false;
    bool captured = false;  // a call of this method has been captured: its arguments are frozen
    std::vector<mpd::BatchIO> io;
    std::vector<int> tileMap, tileLanes;
    std::vector<long> tileStart;
};

// Batched Inverse: per component index c, one pass over the maps that have a
// component c (segments), with its setup tiles and queue chunks; built and
// uploaded by a call whose arguments differ from the previous one.
struct InvPlan {
    struct Pass {
        long seg0, nseg, tile0, ntiles, chunk0, Ntot;
        long ntilesGrp;  // the last ntilesGrp tiles: maps with 4 lanes per point in the setup
    };
    bool valid = false;
    std::vector<mpd::BatchIO> io;
    std::vector<Pass> passes;
    std::vector<long> fullOff;  // per map: offset of its point buffer (-1: none)
    long fullElems = 0;
    long maxNtot = 0;
    DBuf dSegs, dTileSeg, dTileStart, dChunkSeg, dCounters, wFull, wRec, wHeavy;
};

struct mpart_batch_s {
    std::vector<mpart_map> maps;
    int variant = 0;
    int block = 256;
    size_t smem = 0;
    DBuf dMaps;
    BatchSlot slot[3];
    int occ = 0;
    // Mixed schedule of the batched Evaluate / LogDeterminant: lanes per point
    // of each map (4 for maps with many terms, else 1), the largest table of
    // the 4-lane maps, the LDS bytes (0: the layout does not fit), occupancy,
    // and the tuner's choice (1: a lane per point everywhere, 4: the mixed
    // schedule; 0: not tuned; never changed once a call has been captured,
    // since a captured graph reads the tile list of its schedule).
    InvPlan inv;
    // Batched Inverse schedule: 1 per-lane solve scheduling (component passes),
    // 2 a lane per point; 0 autotune (timed on the first uncaptured call with
    // enough points; untuned calls take 2).  Frozen once a call is captured.
    int invSchedule = 0;
    int heavySolveSteps = HEAVY_SOLVE_STEPS;  // hand-over (inverse_component)
    double invTrialMs[2] = {0, 0};
    bool invCaptured = false;  // an Inverse call has been captured (its schedule is frozen)
    std::vector<int> mapLanes;
    size_t grpSetupSmem = 0;  // LDS of the 4-lanes-per-point Inverse setup (0: not used)
    size_t laneSetupSmem = 0; // with it, the LDS of the other maps' lane-per-point setup
    int grpSlots = 1;
    size_t smemGrp = 0;
    int occGrp = 0;
    int evalGroup = 0;
    double evalTrialMs[2] = {0, 0};
    bool captured = false;
};

namespace {

mpart_batch_s* make_batch(const mpart_map* maps, uint32_t n) {
    if (!maps || n == 0) throw ArgError("CreateBatch: no maps");
    auto b = std::make_unique<mpart_batch_s>();
    int pmClass = 0;
    int maxSlots = 1;
    std::vector<MapDev> devs;
    for (uint32_t i = 0; i < n; ++i) {
        check_map(maps[i]);
        if (std::memcmp(&maps[i]->opts, &maps[0]->opts, sizeof(mpart_options)) != 0)
            throw NotImpl("a batch needs maps with identical MapOptions");
        pmClass = std::max(pmClass, maps[i]->variant / 1000);
        maxSlots = std::max(maxSlots, maps[i]->dev.nslots);
        if (maps[i]->inDim > (unsigned)mpd::BATCH_MAX_DIM)
            throw NotImpl("batched maps are limited to input dimension " + std::to_string(mpd::BATCH_MAX_DIM));
        b->maps.push_back(maps[i]);
        devs.push_back(maps[i]->dev);
    }
    b->variant = pmClass * 1000 + maps[0]->variant % 1000;
    int block = 256;
    while (block > 64 && (size_t)maxSlots * block * sizeof(double) > LDS_LIMIT) block /= 2;
    if ((size_t)maxSlots * block * sizeof(double) > LDS_LIMIT)
        throw NotImpl("batched maps need a basis table that fits in LDS (at most " +
                      std::to_string(LDS_LIMIT / (64 * sizeof(double))) + " table slots)");
    b->block = block;
    b->smem = (size_t)maxSlots * block * sizeof(double);
    // maps above the specialization limit, or whose per-lane table does not
    // fit 256 lanes, take 4 lanes per point in the mixed schedule
    int laneSlots = 1, grpSlots = 1;
    bool anyGroup = false;
    for (uint32_t i = 0; i < n; ++i) {
        const bool big = many_terms(*maps[i]);
        b->mapLanes.push_back(big ? 4 : 1);
        anyGroup = anyGroup || big;
        if (big) grpSlots = std::max(grpSlots, maps[i]->dev.nslots);
        else laneSlots = std::max(laneSlots, maps[i]->dev.nslots);
    }
    b->grpSlots = grpSlots;
    const size_t mixedSmem =
        std::max((size_t)256 * laneSlots, (size_t)(256 / 4) * grpSlots) * sizeof(double);
    b->smemGrp = (anyGroup && mixedSmem <= LDS_LIMIT) ? mixedSmem : 0;
    // the batched Inverse's setup of those maps: 64 points (4 lanes each) per 256-lane block
    const size_t grpSetup = (size_t)64 * grpSlots * sizeof(double);
    b->grpSetupSmem = (anyGroup && grpSetup <= LDS_LIMIT) ? grpSetup : 0;
    b->laneSetupSmem = (size_t)laneSlots * block * sizeof(double);
    upload(b->dMaps, devs);
    for (auto& sl : b->slot) sl.dCounter.ensure(256);
    return b.release();
}

// Uploads the call's descriptors and tile list when they differ from the ones
// on the device, then launches.  Under stream capture the descriptors must
// already be on the device (run the same call once before capturing).
void run_batch(mpart_batch_s& b, int method, std::vector<mpd::BatchIO>& io, double xtol, double ytol, hipStream_t s,
               int lanes = 1) {
    const size_t n = b.maps.size();
    // mixed schedule: Evaluate / LogDeterminant only, and only when its layout fits
    const bool mixed = lanes > 1 && method == mpd::BM_EVAL && b.smemGrp > 0;
    // Tiles, most expensive map first (cost ~ terms + quadrature per component);
    // a tile has 256 / lanes points in the mixed schedule, else block points.
    std::vector<std::pair<double, std::pair<int, long>>> tiles;
    for (size_t i = 0; i < n; ++i) {
        const double cost = (double)b.maps[i]->numCoeffs + 40.0 * (double)b.maps[i]->comps.size();
        const long tilePoints = mixed ? 256 / b.mapLanes[i] : b.block;
        for (long t0 = 0; t0 < io[i].N; t0 += tilePoints) tiles.push_back({cost, {(int)i, t0}});
    }
    std::stable_sort(tiles.begin(), tiles.end(), [](const auto& a, const auto& c) { return a.first > c.first; });
    std::vector<int> tileMap(tiles.size()), tileLanes(tiles.size());
    std::vector<long> tileStart(tiles.size());
    for (size_t t = 0; t < tiles.size(); ++t) {
        tileMap[t] = tiles[t].second.first;
        tileStart[t] = tiles[t].second.second;
        tileLanes[t] = mixed ? b.mapLanes[tileMap[t]] : 1;
    }
    BatchSlot& sl = b.slot[method];
    const bool same = sl.valid && tileMap == sl.tileMap && tileStart == sl.tileStart && tileLanes == sl.tileLanes &&
                      io.size() == sl.io.size() &&
                      std::memcmp(io.data(), sl.io.data(), io.size() * sizeof(mpd::BatchIO)) == 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
    if (!same) {
        if (cs != hipStreamCaptureStatusNone)
            throw ArgError("batched call: its arguments differ from the previous call of this method on this batch; "
                           "run the call once before capturing it in a graph");
        // A captured launch holds this method's tables, tile count and grid:
        // new arguments would re-upload (and possibly reallocate) the tables
        // under it, so they are refused before anything is launched.
        if (sl.captured)
            throw ArgError("batched call: this method has been captured in a graph on this batch, so its arguments "
                           "can no longer change; use another batch for other arguments");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        upload(sl.dIO, io);
        upload(sl.dTileMap, tileMap);
        upload(sl.dTileStart, tileStart);
        upload(sl.dTileLanes, tileLanes);
        sl.io = io;
        sl.tileMap = tileMap;
        sl.tileStart = tileStart;
        sl.tileLanes = tileLanes;
        sl.valid = true;
    }
    if (cs != hipStreamCaptureStatusNone) {
        b.captured = true;
        sl.captured = true;
    }
    if (tiles.empty()) return;
    if (g_numCU == 0) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&g_numCU, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    }
    if (mixed) {
        if (b.occGrp == 0) b.occGrp = mpd::occupancy_batch_eval_grp(b.variant, 256, b.smemGrp);
        const long grid = std::min<long>((long)tiles.size(), (long)g_numCU * b.occGrp);
        mpd::Launch L{dim3((unsigned)grid), dim3(256), b.smemGrp, s};
        hip_check(mpd::launch_batch_eval_mixed(b.variant, L, b.dMaps.as<MapDev>(), sl.dIO.as<mpd::BatchIO>(),
                                               sl.dTileMap.as<int>(), sl.dTileStart.as<long>(),
                                               sl.dTileLanes.as<int>(), (int)tiles.size(),
                                               sl.dCounter.as<unsigned>(), b.grpSlots),
                  "launch batch (mixed lanes per point)");
        return;
    }
    if (b.occ == 0) b.occ = mpd::occupancy_batch(b.variant, b.block, b.smem);
    const long grid = std::min<long>((long)tiles.size(), (long)g_numCU * b.occ);
    mpd::Launch L{dim3((unsigned)grid), dim3(b.block), b.smem, s};
    hip_check(mpd::launch_batch(b.variant, L, b.dMaps.as<MapDev>(), sl.dIO.as<mpd::BatchIO>(), sl.dTileMap.as<int>(),
                                sl.dTileStart.as<long>(), (int)tiles.size(), sl.dCounter.as<unsigned>(), method, xtol,
                                ytol),
              "launch batch");
}

// Batched Inverse with per-lane solve scheduling (the single-map queue solver
// run over the points of all maps): component by component, a setup pass and
// a persistent solve pass over every map that has that component.  A
// triangular map solves into a zero-initialized point buffer (its x1 rows
// copied in) as TriangularMap::InverseImpl does (src/TriangularMap.cpp:147-184);
// each root is written there and to the caller's output.  The per-point
// arithmetic is that of the single-map path, so the results are equal bit for
// bit.
void batch_inverse(mpart_batch_s& b, const std::vector<mpd::BatchIO>& io, double xtol, double ytol, hipStream_t s) {
    InvPlan& P = b.inv;
    const size_t nm = b.maps.size();
    const bool same = P.valid && io.size() == P.io.size() &&
                      std::memcmp(io.data(), P.io.data(), io.size() * sizeof(mpd::BatchIO)) == 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
    if (cs != hipStreamCaptureStatusNone) b.captured = true;
    if (!same) {
        if (cs != hipStreamCaptureStatusNone)
            throw ArgError("batched call: its arguments differ from the previous call of this method on this batch; "
                           "run the call once before capturing it in a graph");
        // a captured graph holds the pass offsets and counts of this plan as
        // launch arguments: a new plan would leave it indexing the new tables
        if (b.invCaptured)
            throw ArgError("batched Inverse: this batch's Inverse has been captured in a graph, so its arguments can "
                           "no longer change; use another batch for other arguments");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        P.valid = false;
        P.passes.clear();
        P.fullOff.assign(nm, -1);
        P.fullElems = 0;
        P.maxNtot = 0;
        size_t ncompMax = 0;
        for (size_t i = 0; i < nm; ++i) {
            ncompMax = std::max(ncompMax, b.maps[i]->comps.size());
            if (b.maps[i]->triangular && io[i].N > 0) {
                P.fullOff[i] = P.fullElems;
                P.fullElems += (long)b.maps[i]->inDim * io[i].N;
            }
        }
        std::vector<mpd::InvSeg> segs;
        std::vector<long> xOff, oOff;  // per segment: offsets in the point buffers (-1: caller memory)
        std::vector<int> tileSeg, chunkSeg;
        std::vector<long> tileStart;
        const long tileN = b.block;
        for (size_t c = 0; c < ncompMax; ++c) {
            InvPlan::Pass ps{(long)segs.size(), 0, (long)tileSeg.size(), 0, (long)chunkSeg.size(), 0, 0};
            std::vector<int> grpSeg;  // tiles of maps with 4 lanes per point, appended after the others
            std::vector<long> grpStart;
            for (size_t i = 0; i < nm; ++i) {
                mpart_map m = b.maps[i];
                const mpd::BatchIO& d = io[i];
                if (c >= m->comps.size() || d.N <= 0) continue;
                const MultiSet& ms = m->comps[c];
                mpd::InvSeg g{};
                g.map = (int)i;
                g.p = (int)ms.maxDeg[ms.dim - 1];
                g.status = m->dev.status;
                g.base = ps.Ntot;
                g.N = d.N;
                g.w = d.w;
                g.ws = d.ws;
                g.y = d.s + (long)c * d.srs;
                g.yss = d.scs;
                if (m->triangular) {
                    // the solved coordinates follow the x1 rows in the point buffer
                    xOff.push_back(P.fullOff[i]);
                    oOff.push_back(P.fullOff[i] + (d.xrows + (long)c) * d.N);
                    g.xrows = (long)ms.dim;
                    g.xcols = d.N;
                    g.xrs = d.N;
                    g.xcs = 1;
                    g.os = 1;
                    g.o2 = d.o + (long)c * d.ors;
                    g.os2 = d.ocs;
                } else {
                    xOff.push_back(-1);
                    oOff.push_back(-1);
                    g.x = d.p;
                    g.xrows = d.xrows;
                    g.xcols = d.pcols;
                    g.xrs = d.prs;
                    g.xcs = d.pcs;
                    g.o = d.o;
                    g.os = d.ocs;
                    g.o2 = nullptr;
                    g.os2 = 0;
                }
                const int si = (int)segs.size() - (int)ps.seg0;
                const bool grp = b.grpSetupSmem > 0 && b.mapLanes[i] == 4;
                for (long t0 = 0; t0 < d.N; t0 += tileN) {
                    (grp ? grpSeg : tileSeg).push_back(si);
                    (grp ? grpStart : tileStart).push_back(t0);
                }
                const long padded = (d.N + mpd::INV_CHUNK - 1) / mpd::INV_CHUNK * mpd::INV_CHUNK;
                for (long k = 0; k < padded / mpd::INV_CHUNK; ++k) chunkSeg.push_back(si);
                ps.Ntot += padded;
                segs.push_back(g);
            }
            ps.ntilesGrp = (long)grpSeg.size();
            tileSeg.insert(tileSeg.end(), grpSeg.begin(), grpSeg.end());
            tileStart.insert(tileStart.end(), grpStart.begin(), grpStart.end());
            ps.nseg = (long)segs.size() - ps.seg0;
            ps.ntiles = (long)tileSeg.size() - ps.tile0;
            P.maxNtot = std::max(P.maxNtot, ps.Ntot);
            P.passes.push_back(ps);
        }
        // device buffers, then the point-buffer offsets become addresses
        P.wFull.ensure(std::max<long>(P.fullElems, 1) * sizeof(double));
        const int rows = mpd::inverse_record_rows(b.variant);
        P.wRec.ensure((size_t)rows * std::max<long>(P.maxNtot, 1) * sizeof(double));
        // a point counter per pass, then the hand-over count
        P.dCounters.ensure((std::max<size_t>(P.passes.size(), 1) + 1) * sizeof(unsigned long long));
        double* fullBuf = P.wFull.as<double>();
        for (size_t k = 0; k < segs.size(); ++k) {
            if (xOff[k] >= 0) segs[k].x = fullBuf + xOff[k];
            if (oOff[k] >= 0) segs[k].o = fullBuf + oOff[k];
        }
        // host-side checks of the tables the kernels index
        for (const auto& ps : P.passes) {
            if (ps.Ntot % mpd::INV_CHUNK != 0) throw std::runtime_error("batched Inverse: unpadded pass");
            for (long k = 0; k < ps.Ntot / mpd::INV_CHUNK; ++k) {
                const int si = chunkSeg[ps.chunk0 + k];
                if (si < 0 || si >= ps.nseg) throw std::runtime_error("batched Inverse: chunk table out of range");
                const mpd::InvSeg& g = segs[ps.seg0 + si];
                const long padded = (g.N + mpd::INV_CHUNK - 1) / mpd::INV_CHUNK * mpd::INV_CHUNK;
                if (k * mpd::INV_CHUNK < g.base || k * mpd::INV_CHUNK >= g.base + padded)
                    throw std::runtime_error("batched Inverse: chunk outside its segment");
            }
            for (long t = 0; t < ps.ntiles; ++t) {
                const int si = tileSeg[ps.tile0 + t];
                if (si < 0 || si >= ps.nseg || tileStart[ps.tile0 + t] < 0 ||
                    tileStart[ps.tile0 + t] >= segs[ps.seg0 + si].N)
                    throw std::runtime_error("batched Inverse: tile table out of range");
            }
        }
        upload(P.dSegs, segs);
        upload(P.dTileSeg, tileSeg);
        upload(P.dTileStart, tileStart);
        upload(P.dChunkSeg, chunkSeg);
        P.io = io;
        P.valid = true;
    }
    if (g_numCU == 0) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&g_numCU, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    }
    // zeroed point buffers with the x1 rows copied in
    double* full = P.wFull.as<double>();
    if (P.fullElems > 0) hip_check(mpd::launch_fill(full, 1, P.fullElems, P.fullElems, 1, 0.0, s), "launch fill");
    for (size_t i = 0; i < nm; ++i) {
        if (P.fullOff[i] < 0 || io[i].xrows == 0) continue;
        hip_check(mpd::launch_copy(io[i].p, io[i].xrows, io[i].N, io[i].prs, io[i].pcs, full + P.fullOff[i], io[i].N, 1,
                                   s),
                  "launch copy");
    }
    const int qblock = 256;
    const long qcap = (long)g_numCU * mpd::occupancy_inverse_batch(b.variant, qblock);
    const auto* segs = P.dSegs.as<mpd::InvSeg>();
    int heavyBlocks = 0;
    mpd::InvHeavy* Hd = nullptr;
    const mpd::InvHeavy H = heavy_args(P.wHeavy, P.dCounters.as<unsigned long long>() + std::max<size_t>(P.passes.size(), 1),
                                       P.maxNtot, b.heavySolveSteps, heavyBlocks, Hd);
    for (size_t c = 0; c < P.passes.size(); ++c) {
        const InvPlan::Pass& ps = P.passes[c];
        const long qgrid = std::max<long>(1, std::min<long>((ps.Ntot + qblock - 1) / qblock, qcap));
        // the lane-per-point setup holds only the tables of the maps without 4 lanes per point
        mpd::Launch L{dim3(1), dim3(b.block), b.grpSetupSmem > 0 ? b.laneSetupSmem : b.smem, s};
        hip_check(mpd::launch_inverse_batch(b.variant, L, (int)ps.ntiles, dim3((unsigned)qgrid), dim3(qblock),
                                            b.maps[0]->dev, b.dMaps.as<MapDev>(), segs + ps.seg0,
                                            P.dTileSeg.as<int>() + ps.tile0, P.dTileStart.as<long>() + ps.tile0,
                                            P.dChunkSeg.as<int>() + ps.chunk0, ps.Ntot, (int)c, xtol, ytol,
                                            P.wRec.as<double>(), P.dCounters.as<unsigned long long>() + c, H,
                                            Hd, heavyBlocks, (int)ps.ntilesGrp, b.block, b.grpSetupSmem),
                  "launch batched inverse");
    }
}

mpd::BatchIO base_io(mpart_map m) {
    check_coeffs(m);
    mpd::BatchIO io{};
    io.w = m->coeffs;
    io.ws = m->coeffStride;
    return io;
}

// ---------------------------------------------------------------------------
// Training.  TrainMap (src/TrainMap.cpp:168-216) trains a triangular map under
// the standard-Gaussian KL objective one component at a time
// (TrainMap_Triangular_StandardNormalDensity, 141-166): component i sees rows
// [0, inputDim_i) of the samples through FastGaussianReverseKLObjective
// (src/MapObjective.cpp:93-234), L_i = mean(T_i^2 / 2 - log dT_i/dx_i), and the
// returned error is the sum of the L_i at the end.  Here every objective
// evaluation is one fused KL pass (value + coefficient gradient).  The
// optimizer is L-BFGS (Nocedal & Wright, Numerical Optimization, 2nd ed.,
// Algorithm 7.4, memory `history`) with a strong-Wolfe line search
// (Algorithms 3.5 / 3.6, c1 = 1e-4, c2 = 0.9, cubic interpolation) in place of
// NLopt, stopped by NLopt's rules (TrainOptions, MParT/TrainMap.h:19-40):
// f <= stopval; |df| < ftol_abs or |df| < ftol_rel |f|; every |dx_i| < xtol_abs
// or < xtol_rel |x_i|; maxeval objective evaluations; maxtime seconds.

// Value and coefficient gradient of the KL objective of one map at host
// coefficients w (written into the map's coefficient storage).
struct KLObjectiveEval {
    mpart_map_s& m;
    mpart_cview pts;
    hipStream_t s;
    DBuf dev;  // [w (K) | loss | grad (K)]
    std::vector<double> host;
    int evals = 0;
    double offset = 0.0;  // constant added to the objective value (the Gaussian normalization, if any)
    KLObjectiveEval(mpart_map_s& m_, const mpart_cview& p, hipStream_t st, double off = 0.0)
        : m(m_), pts(p), s(st), offset(off) {
        dev.ensure((2 * (size_t)m.numCoeffs + 1) * sizeof(double));
        host.resize((size_t)m.numCoeffs + 1);
    }
    double operator()(const std::vector<double>& w, std::vector<double>& g) {
        const long K = m.numCoeffs;
        double* d = dev.as<double>();
        hip_check(hipMemcpyAsync(d, w.data(), K * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        hip_check(mpd::launch_copy(d, 1, K, 0, 1, m.coeffs, 0, m.coeffStride, s), "launch copy");
        do_fused(m, pts, nullptr, 0, 0, nullptr, 0, 1, d + K, d + K + 1, 1.0 / (double)pts.cols, s);
        hip_check(hipMemcpyAsync(host.data(), d + K, (K + 1) * sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        // an infinite integrand at these coefficients makes the trial point
        // infeasible: the line search backtracks from it
        const bool infG = take_integrand_status(m, s);
        ++evals;
        g.assign(host.begin() + 1, host.end());
        const double f = host[0] + offset;
        return (infG || std::isnan(f)) ? std::numeric_limits<double>::infinity() : f;
    }
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double t = 0.0;
    for (size_t i = 0; i < a.size(); ++i) t += a[i] * b[i];
    return t;
}

// Minimizer of the cubic interpolating (a, fa, da) and (b, fb, db), kept in
// [lo + 0.1 (hi - lo), hi - 0.1 (hi - lo)] of the interval they span.
double cubic_step(double a, double fa, double da, double b, double fb, double db) {
    const double lo = std::min(a, b), hi = std::max(a, b);
    const double d1 = da + db - 3.0 * (fa - fb) / (a - b);
    const double r = d1 * d1 - da * db;
    double t = 0.5 * (lo + hi);
    if (r >= 0.0 && std::isfinite(r)) {
        const double d2 = std::copysign(std::sqrt(r), b - a);
        t = b - (b - a) * (db + d2 - d1) / (db - da + 2.0 * d2);
    }
    if (!std::isfinite(t)) t = 0.5 * (lo + hi);
    return std::min(std::max(t, lo + 0.1 * (hi - lo)), hi - 0.1 * (hi - lo));
}

struct TrainResult {
    double f;
    int evals;
    const char* status;
};

TrainResult lbfgs_minimize(KLObjectiveEval& fg, std::vector<double>& x, const mpart_train_options& o) {
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    const size_t n = x.size();
    const int hist = std::max(1, (int)o.history);
    std::vector<double> g(n), d(n), xn(n), gn(n);
    double f = fg(x, g);
    if (!std::isfinite(f)) return {f, fg.evals, "objective not finite at the initial coefficients"};
    if (f <= o.stopval) return {f, fg.evals, "stopval reached"};
    std::vector<std::vector<double>> S, Y;
    std::vector<double> rho;
    auto budget = [&] { return fg.evals < o.maxeval && elapsed() < o.maxtime; };
    for (int iter = 0;; ++iter) {
        if (!budget()) return {f, fg.evals, fg.evals >= o.maxeval ? "maxeval reached" : "maxtime reached"};
        // Two-loop recursion: d = -H g.
        for (size_t i = 0; i < n; ++i) d[i] = -g[i];
        const int mcur = (int)S.size();
        std::vector<double> alpha(mcur);
        for (int j = mcur - 1; j >= 0; --j) {
            alpha[j] = rho[j] * dot(S[j], d);
            for (size_t i = 0; i < n; ++i) d[i] -= alpha[j] * Y[j][i];
        }
        if (mcur > 0) {
            const double gamma = dot(S.back(), Y.back()) / dot(Y.back(), Y.back());
            for (size_t i = 0; i < n; ++i) d[i] *= gamma;
        }
        for (int j = 0; j < mcur; ++j) {
            const double b = rho[j] * dot(Y[j], d);
            for (size_t i = 0; i < n; ++i) d[i] += (alpha[j] - b) * S[j][i];
        }
        double gtd = dot(g, d);
        if (!(gtd < 0.0)) {  // not a descent direction: restart from steepest descent
            S.clear();
            Y.clear();
            rho.clear();
            for (size_t i = 0; i < n; ++i) d[i] = -g[i];
            gtd = dot(g, d);
            if (!(gtd < 0.0)) return {f, fg.evals, "zero gradient"};
        }
        double gnorm1 = 0.0;
        for (double v : g) gnorm1 += std::fabs(v);
        double step = (mcur == 0) ? std::min(1.0, 1.0 / gnorm1) : 1.0;
        // Strong-Wolfe line search (bracketing, then zoom).
        const double c1 = 1e-4, c2 = 0.9;
        double aPrev = 0.0, fPrev = f, dPrev = gtd;
        double aLo = 0.0, fLo = f, dLo = gtd, aHi = 0.0, fHi = 0.0, dHi = 0.0;
        bool zoom = false, found = false;
        double fNew = f, aNew = 0.0;
        std::vector<double> gLo = g;
        auto eval_at = [&](double a, double& fa, double& da) {
            for (size_t i = 0; i < n; ++i) xn[i] = x[i] + a * d[i];
            fa = fg(xn, gn);
            da = dot(gn, d);
        };
        for (int ls = 0; ls < 25 && budget(); ++ls) {
            double fa, da;
            eval_at(step, fa, da);
            if (fa > f + c1 * step * gtd || (ls > 0 && fa >= fPrev)) {
                aLo = aPrev; fLo = fPrev; dLo = dPrev;
                aHi = step; fHi = fa; dHi = da;
                zoom = true;
                break;
            }
            if (std::fabs(da) <= -c2 * gtd) {
                found = true; aNew = step; fNew = fa;
                break;
            }
            if (da >= 0.0) {
                aLo = step; fLo = fa; dLo = da; gLo = gn;
                aHi = aPrev; fHi = fPrev; dHi = dPrev;
                zoom = true;
                break;
            }
            aPrev = step; fPrev = fa; dPrev = da;
            aLo = step; fLo = fa; dLo = da; gLo = gn;
            step *= 2.0;
        }
        if (zoom) {
            for (int z = 0; z < 25 && budget(); ++z) {
                if (std::fabs(aHi - aLo) * std::sqrt(dot(d, d)) < 1e-12) break;
                const double a = cubic_step(aLo, fLo, dLo, aHi, std::isfinite(fHi) ? fHi : fLo + 1e300, dHi);
                double fa, da;
                eval_at(a, fa, da);
                if (fa > f + c1 * a * gtd || fa >= fLo) {
                    aHi = a; fHi = fa; dHi = da;
                } else {
                    if (std::fabs(da) <= -c2 * gtd) {
                        found = true; aNew = a; fNew = fa;
                        break;
                    }
                    if (da * (aHi - aLo) >= 0.0) {
                        aHi = aLo; fHi = fLo; dHi = dLo;
                    }
                    aLo = a; fLo = fa; dLo = da; gLo = gn;
                }
            }
        }
        if (!found) {
            // Best point with sufficient decrease seen by the search, if any.
            if (aLo > 0.0 && fLo < f) {
                aNew = aLo;
                fNew = fLo;
                gn = gLo;
            } else {
                return {f, fg.evals, budget() ? "line search made no progress" : "maxeval reached"};
            }
        }
        // Accept the step; gn holds the gradient at x + aNew d.
        for (size_t i = 0; i < n; ++i) xn[i] = x[i] + aNew * d[i];
        bool xconv = true;
        std::vector<double> sv(n), yv(n);
        for (size_t i = 0; i < n; ++i) {
            sv[i] = xn[i] - x[i];
            yv[i] = gn[i] - g[i];
            const double dx = std::fabs(sv[i]);
            if (!(dx < o.xtol_abs || dx < o.xtol_rel * std::fabs(xn[i]))) xconv = false;
        }
        const double df = std::fabs(fNew - f);
        const bool fconv = df < o.ftol_abs || df < o.ftol_rel * std::fabs(fNew);
        x = xn;
        g = gn;
        f = fNew;
        if (o.verbose > 1) printf("  iteration %d: f = %.12g, step %.3g, evaluations %d\n", iter, f, aNew, fg.evals);
        if (f <= o.stopval) return {f, fg.evals, "stopval reached"};
        if (fconv) return {f, fg.evals, "ftol reached"};
        if (xconv) return {f, fg.evals, "xtol reached"};
        const double sy = dot(sv, yv);
        if (sy > 1e-10 * std::sqrt(dot(sv, sv) * dot(yv, yv))) {
            if ((int)S.size() == hist) {
                S.erase(S.begin());
                Y.erase(Y.begin());
                rho.erase(rho.begin());
            }
            S.push_back(sv);
            Y.push_back(yv);
            rho.push_back(1.0 / sy);
        }
    }
}

// One map (a component, or any map trained as a whole): returns the final
// objective value and leaves the trained coefficients in the map.
TrainResult train_one(mpart_map_s& m, const mpart_cview& pts, const mpart_train_options& o, hipStream_t s,
                      double offset = 0.0) {
    check_coeffs(&m);
    const long K = m.numCoeffs;
    std::vector<double> x(K);
    DBuf tmp;
    tmp.ensure(std::max<long>(1, K) * sizeof(double));
    hip_check(mpd::launch_copy(m.coeffs, 1, K, 0, m.coeffStride, tmp.as<double>(), 0, 1, s), "launch copy");
    hip_check(hipMemcpyAsync(x.data(), tmp.as<double>(), K * sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    KLObjectiveEval fg(m, pts, s, offset);
    const TrainResult r = lbfgs_minimize(fg, x, o);
    // The map keeps the accepted iterate (the last evaluation may have been a
    // rejected trial point of the line search).
    double* d = fg.dev.as<double>();
    hip_check(hipMemcpyAsync(d, x.data(), K * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(mpd::launch_copy(d, 1, K, 0, 1, m.coeffs, 0, m.coeffStride, s), "launch copy");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    return r;
}

void do_train(mpart_map_s& m, const mpart_cview& pts, const mpart_train_options& o, double* error, int32_t* evals,
              hipStream_t s) {
    if (o.algorithm != MPART_TRAIN_LBFGS) throw NotImpl("TrainMap: only the L-BFGS optimizer is available");
    if (o.maxeval <= 0) throw ArgError("TrainMap: opt_maxeval must be positive");
    if (o.history <= 0) throw ArgError("TrainMap: history must be positive");
    check_view("TrainMap: samples", pts.rows, pts.cols, m.inDim, pts.cols);
    if (pts.cols == 0) throw ArgError("TrainMap: no samples");
    check_coeffs(&m);
    double total = 0.0;
    int nev = 0;
    if (!m.triangular) {
        // A lone component goes through the generic NLopt path with the
        // KLObjective (src/TrainMap.cpp:184-199, src/MapObjective.cpp:55-91),
        // whose value carries the Gaussian normalization 0.5 d log(2 pi)
        // (GaussianSamplerDensity, d = the output dimension), so that the
        // optimizer sees the same values as NLopt would (stopval and
        // ftol_rel compare against them).
        const TrainResult r = train_one(m, pts, o, s, 0.5 * (double)m.outDim * std::log(2.0 * M_PI));
        total = r.f;
        if (o.verbose) printf("Optimization result: %s\nOptimization error: %.12g\nOptimization evaluations: %d\n", r.status, total, r.evals);
        nev = r.evals;
    } else {
        if (o.verbose) printf("Detected standard normal reference density with Triangular Map\n");
        for (size_t i = 0; i < m.comps.size(); ++i) {
            std::unique_ptr<mpart_map_s> comp(make_map(m.opts, std::vector<MultiSet>{m.comps[i]}, false));
            const long nt = comp->numCoeffs;
            double* block = m.coeffs + (long)m.coeffOff[i] * m.coeffStride;
            hip_check(mpd::launch_copy(block, 1, nt, 0, m.coeffStride, comp->ownCoeffs.as<double>(), 0, 1, s), "launch copy");
            mpart_cview sub = pts;
            sub.rows = comp->inDim;
            if (o.verbose) printf("Training component %zu...\n", i);
            const TrainResult r = train_one(*comp, sub, o, s);
            hip_check(mpd::launch_copy(comp->coeffs, 1, nt, 0, comp->coeffStride, block, 0, m.coeffStride, s),
                      "launch copy");
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // before the component is freed
            if (o.verbose)
                printf("Optimization result: %s\nOptimization error: %.12g\nOptimization evaluations: %d\nComponent %zu trained.\n\n",
                       r.status, r.f, r.evals, i);
            total += r.f;
            nev += r.evals;
        }
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    if (error) *error = total;
    if (evals) *evals = nev;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* mpart_last_error(void) { return g_err.c_str(); }

int mpart_is_host_memory(const void* p) { return host_memory(p) ? 1 : 0; }

int mpart_check_integrand(mpart_map m, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        check_integrand(*m, S(s));
    });
}

int mpart_init(void) {
    return guarded([&] {
        static std::once_flag once;
        std::call_once(once, [] {
            int n = 0;
            if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return;
        });
    });
}

void mpart_default_options(mpart_options* o) {
    o->basisType = MPART_BASIS_PROBABILIST_HERMITE;
    o->posFuncType = MPART_POSFUNC_SOFTPLUS;
    o->quadType = MPART_QUAD_ADAPTIVE_SIMPSON;
    o->contDeriv = 1;
    o->basisNorm = 1;
    o->quadMaxSub = 30;
    o->quadMinSub = 0;
    o->quadPts = 5;
    o->quadAbsTol = 1e-6;
    o->quadRelTol = 1e-6;
    o->nugget = 0.0;
    o->basisLB = -std::numeric_limits<double>::infinity();
    o->basisUB = std::numeric_limits<double>::infinity();
}

int mpart_total_order_set(uint32_t dim, uint32_t order, uint32_t* nterms, uint32_t* nnz, uint32_t* nzStarts,
                          uint32_t* nzDims, uint32_t* nzOrders) {
    return guarded([&] {
        if (dim == 0) throw ArgError("FixedMultiIndexSet: dimension must be positive");
        MultiSet s = total_order(dim, order);
        if (nterms) *nterms = s.size();
        if (nnz) *nnz = (uint32_t)s.dims.size();
        if (nzStarts) std::copy(s.starts.begin(), s.starts.end(), nzStarts);
        if (nzDims) std::copy(s.dims.begin(), s.dims.end(), nzDims);
        if (nzOrders) std::copy(s.orders.begin(), s.orders.end(), nzOrders);
    });
}

int mpart_create_component(const mpart_options* opts, uint32_t dim, uint32_t nterms, const uint32_t* nzStarts,
                           const uint32_t* nzDims, const uint32_t* nzOrders, mpart_map* out) {
    return guarded([&] {
        if (!opts || !out || !nzStarts) throw ArgError("null argument");
        if (nterms == 0) throw ArgError("CreateComponent: empty multi-index set");
        std::vector<MultiSet> c{csr_set(dim, nterms, nzStarts, nzDims, nzOrders)};
        *out = make_map(*opts, std::move(c), false);
    });
}

int mpart_create_total_order_component(const mpart_options* opts, uint32_t dim, uint32_t order, mpart_map* out) {
    return guarded([&] {
        if (!opts || !out) throw ArgError("null argument");
        if (dim == 0) throw ArgError("CreateComponent: dimension must be positive");
        std::vector<MultiSet> c{total_order(dim, order)};
        *out = make_map(*opts, std::move(c), false);
    });
}

int mpart_create_triangular(uint32_t inputDim, uint32_t outputDim, uint32_t totalOrder, const mpart_options* opts,
                            mpart_map* out) {
    return guarded([&] {
        if (!opts || !out) throw ArgError("null argument");
        if (outputDim == 0 || outputDim > inputDim)
            throw ArgError("CreateTriangular: need 0 < outputDim <= inputDim");
        const unsigned extra = inputDim - outputDim;
        std::vector<MultiSet> c;
        for (unsigned i = 0; i < outputDim; ++i) c.push_back(total_order(i + extra + 1, totalOrder));
        *out = make_map(*opts, std::move(c), true);
    });
}

int mpart_create_triangular_from(const mpart_map* comps, uint32_t n, mpart_map* out) {
    return guarded([&] {
        if (!comps || !out || n == 0) throw ArgError("null argument");
        std::vector<MultiSet> c;
        for (uint32_t i = 0; i < n; ++i) {
            check_map(comps[i]);
            if (comps[i]->triangular) throw NotImpl("nested triangular maps are not provided");
            if (std::memcmp(&comps[i]->opts, &comps[0]->opts, sizeof(mpart_options)) != 0)
                throw NotImpl("components with different MapOptions are not provided");
            c.push_back(comps[i]->comps[0]);
        }
        mpart_map m = make_map(comps[0]->opts, std::move(c), true);
        // moveCoeffs semantics (src/TriangularMap.cpp:40-60): copy each component's coefficients.
        for (uint32_t i = 0; i < n; ++i)
            hip_check(mpd::launch_copy(comps[i]->coeffs, 1, comps[i]->numCoeffs, 0, comps[i]->coeffStride,
                                       m->coeffs + m->coeffOff[i], 0, 1, nullptr),
                      "launch copy");
        hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        *out = m;
    });
}

int mpart_destroy(mpart_map map) {
    return guarded([&] { delete map; });
}

int mpart_map_dims(mpart_map m, uint32_t* i, uint32_t* o, uint32_t* c) {
    return guarded([&] {
        check_map(m);
        if (i) *i = m->inDim;
        if (o) *o = m->outDim;
        if (c) *c = m->numCoeffs;
    });
}

int mpart_map_options(mpart_map m, mpart_options* o, int32_t* triangular) {
    return guarded([&] {
        check_map(m);
        if (o) *o = m->opts;
        if (triangular) *triangular = m->triangular ? 1 : 0;
    });
}

int mpart_map_num_components(mpart_map m, uint32_t* n) {
    return guarded([&] {
        check_map(m);
        *n = (uint32_t)m->comps.size();
    });
}

int mpart_component_multiset(mpart_map m, uint32_t i, uint32_t* dim, uint32_t* nterms, uint32_t* nnz,
                             uint32_t* nzStarts, uint32_t* nzDims, uint32_t* nzOrders) {
    return guarded([&] {
        check_map(m);
        if (i >= m->comps.size()) throw ArgError("component index out of range");
        const MultiSet& s = m->comps[i];
        if (dim) *dim = s.dim;
        if (nterms) *nterms = s.size();
        if (nnz) *nnz = (uint32_t)s.dims.size();
        if (nzStarts) std::copy(s.starts.begin(), s.starts.end(), nzStarts);
        if (nzDims) std::copy(s.dims.begin(), s.dims.end(), nzDims);
        if (nzOrders) std::copy(s.orders.begin(), s.orders.end(), nzOrders);
    });
}

int mpart_set_coeffs(mpart_map m, const double* c, int64_t len, int64_t stride, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        if (len != m->numCoeffs)
            throw ArgError("SetCoeffs: expected " + std::to_string(m->numCoeffs) + " coefficients, got " +
                           std::to_string(len));
        m->ownCoeffs.ensure(std::max<size_t>(8, m->numCoeffs * sizeof(double)));
        hip_check(mpd::launch_copy(c, 1, len, 0, stride, m->ownCoeffs.as<double>(), 0, 1, S(s)), "launch copy");
        m->coeffs = m->ownCoeffs.as<double>();
        m->coeffStride = 1;
    });
}

int mpart_wrap_coeffs(mpart_map m, double* c, int64_t len, int64_t stride) {
    return guarded([&] {
        check_map(m);
        if (len != m->numCoeffs)
            throw ArgError("WrapCoeffs: expected " + std::to_string(m->numCoeffs) + " coefficients, got " +
                           std::to_string(len));
        m->coeffs = c;
        m->coeffStride = stride;
    });
}

int mpart_get_coeffs(mpart_map m, double* out, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        check_coeffs(m);
        hip_check(mpd::launch_copy(m->coeffs, 1, m->numCoeffs, 0, m->coeffStride, out, 0, 1, S(s)), "launch copy");
    });
}

int mpart_evaluate(mpart_map m, mpart_cview pts, mpart_view out, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        Staging st(S(s));
        do_evaluate(*m, st.in(pts), st.out(out), nullptr, S(s));
        st.finish();
        if (st.staged()) check_integrand(*m, S(s));
    });
}

int mpart_evaluate_counted(mpart_map m, mpart_cview pts, mpart_view out, int32_t* nevals, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        do_evaluate(*m, pts, out, nevals, S(s));
    });
}

int mpart_log_determinant(mpart_map m, mpart_cview pts, double* out, int64_t stride, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        Staging st(S(s));
        const mpart_cview p = st.in(pts);
        double* o = st.out_vec(out, pts.cols, stride);
        do_logdet(*m, p, o, o != out ? 1 : stride, S(s));
        st.finish();
        if (st.staged()) check_integrand(*m, S(s));
    });
}

int mpart_gradient(mpart_map m, mpart_cview pts, mpart_cview sens, mpart_view out, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        Staging st(S(s));
        const mpart_cview sv = st.in(sens);
        do_input_grad(*m, st.in(pts), &sv, st.out(out), 0, S(s));
        st.finish();
        if (st.staged()) check_integrand(*m, S(s));
    });
}

int mpart_coeff_grad(mpart_map m, mpart_cview pts, mpart_cview sens, mpart_view out, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        Staging st(S(s));
        do_coeff_grad(*m, st.in(pts), st.in(sens), st.out(out), S(s));
        st.finish();
        if (st.staged()) check_integrand(*m, S(s));
    });
}

int mpart_log_determinant_coeff_grad(mpart_map m, mpart_cview pts, mpart_view out, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        Staging st(S(s));
        do_logdet_coeff_grad(*m, st.in(pts), st.out(out), S(s));
        st.finish();
        if (st.staged()) check_integrand(*m, S(s));
    });
}

int mpart_log_determinant_input_grad(mpart_map m, mpart_cview pts, mpart_view out, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        Staging st(S(s));
        do_input_grad(*m, st.in(pts), nullptr, st.out(out), 1, S(s));
        st.finish();
        if (st.staged()) check_integrand(*m, S(s));
    });
}

int mpart_inverse(mpart_map m, mpart_cview x1, mpart_cview r, mpart_view out, double xtol, double ytol,
                  mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        Staging st(S(s));
        do_inverse(*m, st.in(x1), st.in(r), st.out(out), xtol, ytol, S(s));
        st.finish();
        if (st.staged()) check_integrand(*m, S(s));
    });
}

int mpart_kl_value_and_grad(mpart_map m, mpart_cview pts, double* loss, double* grad, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        check_view("KL: pts", pts.rows, pts.cols, m->inDim, pts.cols);
        if (pts.cols == 0) throw ArgError("KL: no points");
        Staging st(S(s));
        do_fused(*m, st.in(pts), nullptr, 0, 0, nullptr, 0, 1, st.out_vec(loss, 1, 1), st.out_vec(grad, m->numCoeffs, 1),
                 1.0 / (double)pts.cols, S(s));
        st.finish();
        if (st.staged()) check_integrand(*m, S(s));
    });
}

int mpart_kl_value_and_grad_noalloc(mpart_map m, mpart_cview pts, double* loss, double* grad, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        check_view("KL: pts", pts.rows, pts.cols, m->inDim, pts.cols);
        if (pts.cols == 0) throw ArgError("KL: no points");
        if (!loss || !grad) throw ArgError("KL: null output");
        double* scr = noalloc_scratch();
        do_fused_noalloc(*m, pts, loss, grad, 1.0 / (double)pts.cols, S(s), scr,
                         (long)(MPART_NOALLOC_SCRATCH_BYTES / sizeof(double)));
    });
}

int mpart_coeff_vjp(mpart_map m, mpart_cview pts, mpart_cview sens, const double* lsens, int64_t lss, double* grad,
                    mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        check_view("vjp: pts", pts.rows, pts.cols, m->inDim, pts.cols);
        if (sens.ptr) check_view("vjp: sens", sens.rows, sens.cols, m->outDim, pts.cols);
        do_fused(*m, pts, sens.ptr, sens.row_stride, sens.col_stride, lsens, lss, 0, nullptr, grad, 1.0, S(s));
    });
}

int mpart_kl_value_and_grads(mpart_map m, mpart_cview pts, double* loss, double* grad, mpart_view gradInputs,
                             mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        check_view("KL: pts", pts.rows, pts.cols, m->inDim, pts.cols);
        if (pts.cols == 0) throw ArgError("KL: no points");
        if (!gradInputs.ptr) {
            do_fused(*m, pts, nullptr, 0, 0, nullptr, 0, 1, loss, grad, 1.0 / (double)pts.cols, S(s));
            return;
        }
        if (!m->opts.contDeriv) throw ArgError("Discrete derivative version is not implemented yet (To Do)");
        do_fused(*m, pts, nullptr, 0, 0, nullptr, 0, 1, loss, grad, 1.0 / (double)pts.cols, S(s), &gradInputs,
                 1.0 / (double)pts.cols);
    });
}

int mpart_vjp(mpart_map m, mpart_cview pts, mpart_cview sens, const double* lsens, int64_t lss, double* grad,
              mpart_view gradInputs, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        check_view("vjp: pts", pts.rows, pts.cols, m->inDim, pts.cols);
        if (sens.ptr) check_view("vjp: sens", sens.rows, sens.cols, m->outDim, pts.cols);
        if (!m->opts.contDeriv && gradInputs.ptr)
            throw ArgError("Discrete derivative version is not implemented yet (To Do)");
        do_fused(*m, pts, sens.ptr, sens.row_stride, sens.col_stride, lsens, lss, 0, nullptr, grad, 1.0, S(s),
                 gradInputs.ptr ? &gradInputs : nullptr, 1.0);
    });
}

int mpart_kl_input_grad(mpart_map m, mpart_cview pts, double scale, mpart_view out, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        do_input_vjp(*m, pts, 1, nullptr, nullptr, 0, 0.0, scale, out, S(s), "KL input gradient");
    });
}

int mpart_input_vjp(mpart_map m, mpart_cview pts, mpart_cview sens, const double* logdetSens, int64_t lstride,
                    mpart_view out, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        do_input_vjp(*m, pts, 0, sens.ptr ? &sens : nullptr, logdetSens, (long)lstride, 0.0, 1.0, out, S(s),
                     "InputVJP");
    });
}

void mpart_default_train_options(mpart_train_options* o) {
    if (!o) return;
    o->algorithm = MPART_TRAIN_LBFGS;
    o->stopval = -std::numeric_limits<double>::infinity();
    o->ftol_rel = 1e-3;
    o->ftol_abs = 1e-3;
    o->xtol_rel = 1e-4;
    o->xtol_abs = 1e-4;
    o->maxeval = 1000;
    o->maxtime = std::numeric_limits<double>::infinity();
    o->verbose = 0;
    o->history = 10;
}

int mpart_train_map(mpart_map m, mpart_cview train, const mpart_train_options* opts, double* error, int32_t* evals,
                    mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        if (!opts) throw ArgError("null argument");
        do_train(*m, train, *opts, error, evals, S(s));
    });
}

int mpart_kl_value_and_grad_plain(mpart_map m, mpart_cview pts, double* loss, double* grad, mpart_stream s) {
    return guarded([&] {
        check_map(m);
        m->note_capture(S(s));
        do_kl_plain(*m, pts, loss, grad, S(s));
    });
}

int mpart_set_plain_launches(mpart_map m, int plain) {
    return guarded([&] {
        check_map(m);
        m->plain = plain != 0;
    });
}

int mpart_set_codegen(mpart_map m, int enable) {
    return guarded([&] {
        check_map(m);
        m->genEnabled = enable != 0;
    });
}

int mpart_codegen_prepare(mpart_map m, int32_t* specialized, double* compileSeconds) {
    return guarded([&] {
        check_map(m);
        const bool fresh = !m->genTried;
        mpg::GenModule* g = gen_for(*m);
        if (specialized) *specialized = g ? 1 : 0;
        if (compileSeconds) *compileSeconds = (g && fresh) ? m->genSeconds : 0.0;
    });
}

const char* mpart_codegen_error(mpart_map m) { return m ? m->genErr.c_str() : "null map handle"; }

int mpart_set_point_group(mpart_map m, int32_t lanes) {
    return guarded([&] {
        check_map(m);
        if (lanes != 0 && lanes != 1 && lanes != FUSED_SPLIT && lanes != 4 && lanes != 8 && lanes != 64)
            throw ArgError("lanes per point must be 0, 1, 2 (fused pass split over wavefronts), 4, 8 or 64");
        m->pointGroupForced = lanes;
    });
}

int mpart_eval_schedule(mpart_map m, int32_t* lanes, double* trialMs) {
    return guarded([&] {
        check_map(m);
        for (int k = 0; k < 2; ++k) {
            if (lanes) lanes[k] = m->evalGroup[k];
            if (trialMs)
                for (int i = 0; i < 5; ++i) trialMs[k * 5 + i] = m->evalTrialMs[k][i];
        }
        if (lanes) lanes[2] = m->fusedGroup;
        if (trialMs)
            for (int i = 0; i < 5; ++i) trialMs[10 + i] = m->fusedGroupTrialMs[i];
    });
}

int mpart_fused_schedule(mpart_map m, int32_t* choice, double* trialMs) {
    return guarded([&] {
        check_map(m);
        if (choice) *choice = m->fusedChoice;
        if (trialMs)
            for (int i = 0; i < GEN_FUSED_SCHEDULES; ++i) trialMs[i] = m->fusedTrialMs[i];
    });
}

int mpart_set_fused_schedule(mpart_map m, int32_t choice) {
    return guarded([&] {
        check_map(m);
        if (choice < -1 || choice >= GEN_FUSED_SCHEDULES)
            throw ArgError("fused schedule must be -1 (autotune) or 0.." + std::to_string(GEN_FUSED_SCHEDULES - 1));
        m->fusedChoice = choice;
        if (choice < 0)
            for (double& t : m->fusedTrialMs) t = 0.0;
    });
}

int mpart_codegen_kernel_info(mpart_map m, const char* kernel, int32_t* vgprs, int32_t* scratchBytes,
                              int32_t* ldsBytes) {
    return guarded([&] {
        check_map(m);
        if (!kernel) throw ArgError("null kernel name");
        const std::string k(kernel);
        mpg::GenModule* g = m->gen;
        mpg::GenModule* gi = m->genIn;
        hipFunction_t f = nullptr;
        if (k == "fused") f = g ? g->fused : nullptr;
        else if (k == "fused_w3") f = g ? g->fused3 : nullptr;
        else if (k == "eval") f = g ? g->eval : nullptr;
        else if (k == "coef") f = g ? g->coef : nullptr;
        else if (k == "ldcoef") f = g ? g->ldcoef : nullptr;
        else if (k == "fused_in") f = gi ? gi->fusedIn : nullptr;
        else if (k == "ivjp") f = gi ? gi->ivjp : nullptr;
        else if (k == "klx") f = gi ? gi->klx : nullptr;
        else throw ArgError("unknown specialized kernel '" + k + "'");
        if (!f) throw ArgError("specialized kernel '" + k + "' is not compiled for this map");
        int v = 0, sc = 0, lds = 0;
        hip_check(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, f), "hipFuncGetAttribute");
        hip_check(hipFuncGetAttribute(&sc, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f), "hipFuncGetAttribute");
        hip_check(hipFuncGetAttribute(&lds, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, f), "hipFuncGetAttribute");
        if (vgprs) *vgprs = v;
        if (scratchBytes) *scratchBytes = sc;
        if (ldsBytes) *ldsBytes = lds;
    });
}

int mpart_create_batch(const mpart_map* maps, uint32_t n, mpart_batch* out) {
    return guarded([&] {
        if (!out) throw ArgError("null argument");
        *out = make_batch(maps, n);
    });
}

int mpart_destroy_batch(mpart_batch b) {
    return guarded([&] { delete b; });
}

int mpart_batch_size(mpart_batch b, uint32_t* n) {
    return guarded([&] {
        if (!b || !n) throw ArgError("null argument");
        *n = (uint32_t)b->maps.size();
    });
}

constexpr long BATCH_TUNE_MIN_POINTS = 65536;

int mpart_batch_set_inverse_schedule(mpart_batch b, int32_t schedule) {
    return guarded([&] {
        if (!b) throw ArgError("null argument");
        if (schedule < 0 || schedule > 2) throw ArgError("the inverse schedule must be 0, 1 or 2");
        if (b->invCaptured && schedule != b->invSchedule)
            throw ArgError("the batch's Inverse has been captured in a graph: its schedule can no longer change");
        b->invSchedule = schedule;
    });
}

int mpart_set_inverse_handover(mpart_map m, int32_t solveSteps) {
    return guarded([&] {
        check_map(m);
        if (solveSteps < 0) throw ArgError("the hand-over budget must be non-negative");
        m->heavySolveSteps = solveSteps;
    });
}

int mpart_batch_set_inverse_handover(mpart_batch b, int32_t solveSteps) {
    return guarded([&] {
        if (!b) throw ArgError("null argument");
        if (solveSteps < 0) throw ArgError("the hand-over budget must be non-negative");
        b->heavySolveSteps = solveSteps;
    });
}

int mpart_inverse_handover_stats(mpart_map m, uint64_t* stats, void* stream) {
    return guarded([&] {
        check_map(m);
        if (!stats) throw ArgError("null argument");
        heavy_stats(m->wHeavy, stats, (hipStream_t)stream);
    });
}

int mpart_batch_inverse_handover_stats(mpart_batch b, uint64_t* stats, void* stream) {
    return guarded([&] {
        if (!b || !stats) throw ArgError("null argument");
        heavy_stats(b->inv.wHeavy, stats, (hipStream_t)stream);
    });
}

int mpart_batch_inverse_schedule(mpart_batch b, int32_t* schedule, double* trialMs) {
    return guarded([&] {
        if (!b) throw ArgError("null argument");
        if (schedule) *schedule = b->invSchedule;
        if (trialMs)
            for (int i = 0; i < 2; ++i) trialMs[i] = b->invTrialMs[i];
    });
}

int mpart_batch_set_point_group(mpart_batch b, int32_t lanes) {
    return guarded([&] {
        if (!b) throw ArgError("null argument");
        if (lanes != 0 && lanes != 1 && lanes != 4) throw ArgError("lanes per point must be 0, 1 or 4");
        if (b->captured && lanes != b->evalGroup)
            throw ArgError("the batch has been captured in a graph: its schedule can no longer change");
        b->evalGroup = lanes;
    });
}

int mpart_batch_eval_schedule(mpart_batch b, int32_t* lanes, double* trialMs) {
    return guarded([&] {
        if (!b) throw ArgError("null argument");
        if (lanes) *lanes = b->evalGroup;
        if (trialMs)
            for (int i = 0; i < 2; ++i) trialMs[i] = b->evalTrialMs[i];
    });
}

int mpart_batch_evaluate(mpart_batch b, const mpart_cview* pts, const mpart_view* out, double* const* logdet,
                         const int64_t* logdetStride, mpart_stream s) {
    return guarded([&] {
        if (!b || !pts) throw ArgError("null argument");
        std::vector<mpd::BatchIO> io;
        for (size_t i = 0; i < b->maps.size(); ++i) {
            mpart_map m = b->maps[i];
            mpd::BatchIO d = base_io(m);
            const std::string w = "batched Evaluate, map " + std::to_string(i);
            check_view((w + ": pts").c_str(), pts[i].rows, pts[i].cols, m->inDim, pts[i].cols);
            d.p = pts[i].ptr;
            d.prs = pts[i].row_stride;
            d.pcs = pts[i].col_stride;
            d.N = pts[i].cols;
            if (out && out[i].ptr) {
                check_view((w + ": output").c_str(), out[i].rows, out[i].cols, m->outDim, pts[i].cols);
                d.o = out[i].ptr;
                d.ors = out[i].row_stride;
                d.ocs = out[i].col_stride;
            }
            if (logdet && logdet[i]) {
                if (!m->opts.contDeriv) throw NotImpl("LogDeterminant with contDeriv = false is not provided yet");
                d.ld = logdet[i];
                d.lds = logdetStride ? logdetStride[i] : 1;
            }
            io.push_back(d);
        }
        // Lanes per point (1 or 8): timed on the first uncaptured call with
        // enough points, before any capture, then kept.
        long total = 0;
        for (auto& d : io) total += d.N;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(S(s), &cs), "hipStreamIsCapturing");
        if (b->evalGroup == 0 && !b->captured && cs == hipStreamCaptureStatusNone && b->smemGrp > 0 &&
            total >= BATCH_TUNE_MIN_POINTS) {
            hipEvent_t e0, e1;
            hip_check(hipEventCreate(&e0), "hipEventCreate");
            hip_check(hipEventCreate(&e1), "hipEventCreate");
            const int lanes[2] = {1, 4};
            for (int i = 0; i < 2; ++i) {
                run_batch(*b, mpd::BM_EVAL, io, 0.0, 0.0, S(s), lanes[i]);
                hip_check(hipEventRecord(e0, S(s)), "hipEventRecord");
                run_batch(*b, mpd::BM_EVAL, io, 0.0, 0.0, S(s), lanes[i]);
                hip_check(hipEventRecord(e1, S(s)), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0.f;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                b->evalTrialMs[i] = ms;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            b->evalGroup = b->evalTrialMs[1] < b->evalTrialMs[0] ? 4 : 1;
        }
        run_batch(*b, mpd::BM_EVAL, io, 0.0, 0.0, S(s), b->evalGroup > 0 ? b->evalGroup : 1);
    });
}

int mpart_batch_coeff_grad(mpart_batch b, const mpart_cview* pts, const mpart_cview* sens, const mpart_view* out,
                           mpart_stream s) {
    return guarded([&] {
        if (!b || !pts || !sens || !out) throw ArgError("null argument");
        std::vector<mpd::BatchIO> io;
        for (size_t i = 0; i < b->maps.size(); ++i) {
            mpart_map m = b->maps[i];
            mpd::BatchIO d = base_io(m);
            const std::string w = "batched CoeffGrad, map " + std::to_string(i);
            check_view((w + ": pts").c_str(), pts[i].rows, pts[i].cols, m->inDim, pts[i].cols);
            check_view((w + ": sens").c_str(), sens[i].rows, sens[i].cols, m->outDim, pts[i].cols);
            check_view((w + ": output").c_str(), out[i].rows, out[i].cols, m->numCoeffs, pts[i].cols);
            d.p = pts[i].ptr;
            d.prs = pts[i].row_stride;
            d.pcs = pts[i].col_stride;
            d.N = pts[i].cols;
            d.s = sens[i].ptr;
            d.srs = sens[i].row_stride;
            d.scs = sens[i].col_stride;
            d.o = out[i].ptr;
            d.ors = out[i].row_stride;
            d.ocs = out[i].col_stride;
            io.push_back(d);
        }
        run_batch(*b, mpd::BM_COEFF_GRAD, io, 0.0, 0.0, S(s));
    });
}

int mpart_batch_inverse(mpart_batch b, const mpart_cview* x1, const mpart_cview* r, const mpart_view* out, double xtol,
                        double ytol, mpart_stream s) {
    return guarded([&] {
        if (!b || !x1 || !r || !out) throw ArgError("null argument");
        if (xtol < 0 || ytol < 0) throw ArgError("Invalid tolerance given to Inverse.  Value must be non-negative");
        if (ytol <= 2.220446049250313e-16 && xtol <= 2.220446049250313e-16)
            throw ArgError("Invalid tolerances given to Inverse.  Either \"xtol\" or \"ytol\" must be nonzero");
        std::vector<mpd::BatchIO> io;
        for (size_t i = 0; i < b->maps.size(); ++i) {
            mpart_map m = b->maps[i];
            mpd::BatchIO d = base_io(m);
            const std::string w = "batched Inverse, map " + std::to_string(i);
            const long N = r[i].cols;
            check_view((w + ": r").c_str(), r[i].rows, r[i].cols, m->outDim, N);
            check_view((w + ": output").c_str(), out[i].rows, out[i].cols, m->outDim, N);
            const long xrows = m->triangular ? (long)(m->inDim - m->outDim) : (long)m->inDim;
            if (x1[i].rows != xrows) throw ArgError(w + ": x1 must have " + std::to_string(xrows) + " rows");
            if (xrows > 0 && x1[i].cols != N && !(x1[i].cols == 1 && !m->triangular))
                throw ArgError(w + ": x1 must have one column per target");
            d.p = x1[i].ptr;
            d.prs = x1[i].row_stride;
            d.pcs = x1[i].col_stride;
            d.pcols = x1[i].cols;
            d.xrows = xrows;
            d.N = N;
            d.s = r[i].ptr;
            d.srs = r[i].row_stride;
            d.scs = r[i].col_stride;
            d.o = out[i].ptr;
            d.ors = out[i].row_stride;
            d.ocs = out[i].col_stride;
            io.push_back(d);
        }
        auto run = [&](int sched) {
            if (sched == 1) batch_inverse(*b, io, xtol, ytol, S(s));
            else run_batch(*b, mpd::BM_INVERSE, io, xtol, ytol, S(s));
        };
        long total = 0;
        for (auto& d : io) total += d.N;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(S(s), &cs), "hipStreamIsCapturing");
        if (cs != hipStreamCaptureStatusNone) b->invCaptured = true;
        if (b->invSchedule == 0 && !b->invCaptured && total >= BATCH_TUNE_MIN_POINTS) {
            hipEvent_t e0, e1;
            hip_check(hipEventCreate(&e0), "hipEventCreate");
            hip_check(hipEventCreate(&e1), "hipEventCreate");
            for (int i = 0; i < 2; ++i) {
                run(i + 1);  // untimed: builds and uploads the schedule's tables
                hip_check(hipEventRecord(e0, S(s)), "hipEventRecord");
                run(i + 1);
                hip_check(hipEventRecord(e1, S(s)), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0.f;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                b->invTrialMs[i] = ms;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            b->invSchedule = b->invTrialMs[0] < b->invTrialMs[1] ? 1 : 2;
        }
        run(b->invSchedule == 1 ? 1 : 2);
    });
}

}  // extern "C"
