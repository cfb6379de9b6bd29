/* mpart_amd.h -- C ABI of the MI355X monotone-component library (libmpart_amd.so).
 *
 * Drop-in boundary for MParT's monotone-component path.  Every entry point
 * takes device pointers (HBM) and a HIP stream, returns 0 on success or a
 * negative status, and never throws.  The message of the last failure on the
 * calling thread is returned by mpart_last_error().
 *
 * Status codes:
 *   MPART_OK               0
 *   MPART_INVALID_ARGUMENT -1  (reference: std::invalid_argument)
 *   MPART_RUNTIME_ERROR    -2  (reference: std::runtime_error)
 *   MPART_HIP_ERROR        -3  (a HIP runtime call failed)
 *   MPART_NOT_IMPLEMENTED  -4
 *
 * Matrix arguments are strided views {ptr, rows, cols, row_stride,
 * col_stride} in units of doubles: element (i, j) is ptr[i*row_stride +
 * j*col_stride].  Point matrices are inputDim x N with a column per point,
 * exactly like the reference's StridedMatrix views built from
 * (ptr, (rows, cols), (rowStride, colStride)) tuples
 * (MParT/Utilities/ArrayConversions.h:102-130); the fields are 64-bit here.
 * Calls are asynchronous on `stream` (a hipStream_t; NULL is the null stream).
 *
 * Host memory.  The reference's ConditionalMapBase methods take HostSpace
 * views as well as device views (src/ConditionalMapBase.cpp:7-81).  Here the
 * matrix and vector arguments of mpart_evaluate, mpart_log_determinant,
 * mpart_gradient, mpart_coeff_grad, mpart_log_determinant_coeff_grad,
 * mpart_log_determinant_input_grad, mpart_inverse and mpart_kl_value_and_grad
 * may each lie in host memory (pageable or pinned; any strides): such an
 * input is packed and copied to a temporary device buffer, such an output is
 * computed into one and copied back, and the call then returns only once the
 * stream has finished it (a synchronous call).  Results equal the device-path
 * results bitwise.  Host memory is refused while the stream is being captured
 * in a graph.  Every other entry point takes device memory only.
 *
 * Graph capture.  A call on a map may be captured in a HIP graph (run it once
 * eagerly first, so its kernels are compiled, tuned and its work buffers
 * sized).  The captured launches hold the addresses of the map's work
 * buffers, so once a call on a map has been captured the map never frees a
 * work buffer while it lives: a later call that needs larger buffers
 * allocates new ones and keeps the old ones until mpart_destroy.  Replays
 * and eager calls on one map must not overlap in time (they share the work
 * buffers); keep them on one stream or order them with events.
 */
#ifndef MPART_AMD_H
#define MPART_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPART_OK 0
#define MPART_INVALID_ARGUMENT (-1)
#define MPART_RUNTIME_ERROR (-2)
#define MPART_HIP_ERROR (-3)
#define MPART_NOT_IMPLEMENTED (-4)

/* Enumerations of MParT/MapOptions.h:11-29. */
#define MPART_BASIS_PROBABILIST_HERMITE 0
#define MPART_BASIS_PHYSICIST_HERMITE 1
#define MPART_BASIS_HERMITE_FUNCTIONS 2
#define MPART_POSFUNC_EXP 0
#define MPART_POSFUNC_SOFTPLUS 1
#define MPART_QUAD_CLENSHAW_CURTIS 0
#define MPART_QUAD_ADAPTIVE_SIMPSON 1
#define MPART_QUAD_ADAPTIVE_CLENSHAW_CURTIS 2

/* Replaces mpart::MapOptions (MParT/MapOptions.h:58-123); defaults from
 * mpart_default_options() are the reference defaults.  A finite basisLB or
 * basisUB selects the LinearizedBasis of the factory (MParT/MapFactory.h:226,
 * MParT/LinearizedBasis.h); quadType ClenshawCurtis is the fixed rule of
 * quadPts points.  quadMaxSub is at most 32: the reference keeps the path of
 * its subdivision tree in a 32-bit mask (MParT/Quadrature.h:566,636-647),
 * whose shifts are undefined past level 31.  The last coordinate's degree is
 * at most 12 (the degree bound of the kernels' register arrays; the
 * reference's 32-bit Factorial, MParT/Utilities/MathFunctions.h:16, wraps at
 * 13! in the default normalization), other coordinates' degrees at most 32. */
typedef struct mpart_options {
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
    double basisLB;
    double basisUB;
} mpart_options;

typedef struct mpart_map_s* mpart_map;

typedef struct mpart_cview {
    const double* ptr;
    int64_t rows, cols, row_stride, col_stride;
} mpart_cview;

typedef struct mpart_view {
    double* ptr;
    int64_t rows, cols, row_stride, col_stride;
} mpart_view;

typedef void* mpart_stream; /* hipStream_t */

const char* mpart_last_error(void);

/* 1 when the library stages `p` through a device buffer (host memory: pageable,
 * pinned or unknown to HIP), 0 when kernels read it in place (device or managed
 * memory, or NULL).  No counterpart in the reference, whose Kokkos views carry
 * their memory space in their type. */
int mpart_is_host_memory(const void* p);

/* The reference fails when g(df) is infinite at a quadrature node
 * (MParT/MonotoneIntegrand.h:183-190: std::runtime_error on the host,
 * Kokkos::abort in device code).  Here every kernel that evaluates the
 * integrand (Evaluate, Gradient, CoeffGrad, Inverse, the KL passes and VJPs,
 * the discrete-derivative LogDeterminant and its gradient, the batches) sets
 * the map's status word instead, and no kernel stops.  mpart_check_integrand
 * waits for `stream`, returns MPART_RUNTIME_ERROR with the reference's message
 * when any call on the map met an infinite g since the last check, and clears
 * the word.  The host-memory calls above check it themselves (they are
 * synchronous): they fail the same way, after writing their outputs.
 * mpart_train_map treats coefficients at which g is infinite as an infeasible
 * trial point (objective +inf, the line search backtracks). */
int mpart_check_integrand(mpart_map map, mpart_stream stream);

/* Replaces mpart::Initialize (MParT/Initialization.h:78-116): idempotent. */
int mpart_init(void);

void mpart_default_options(mpart_options* opts);

/* FixedMultiIndexSet(dim, maxOrder) (src/MultiIndices/FixedMultiIndexSet.cpp:208-233,
 * 487-549): the total-order set in compressed form, host arrays sized by a
 * first call with NULL pointers.  Needs no GPU. */
int mpart_total_order_set(uint32_t dim, uint32_t order, uint32_t* nterms, uint32_t* nnz, uint32_t* nzStarts,
                          uint32_t* nzDims, uint32_t* nzOrders);

/* MapFactory::CreateComponent(FixedMultiIndexSet, MapOptions)
 * (MParT/MapFactory.h:49-60, src/MapFactory.cpp:22-27).  The multi-index set
 * is given in the compressed form of FixedMultiIndexSet
 * (MParT/MultiIndices/FixedMultiIndexSet.h: nzStarts[nterms+1], nzDims, nzOrders).
 * Coefficients start at zero (src/MapFactoryImpl6.cpp:24-25). */
int mpart_create_component(const mpart_options* opts, uint32_t dim, uint32_t nterms, const uint32_t* nzStarts,
                           const uint32_t* nzDims, const uint32_t* nzOrders, mpart_map* out);

/* CreateComponent with FixedMultiIndexSet(dim, order) (total order,
 * src/MultiIndices/FixedMultiIndexSet.cpp:208-233). */
int mpart_create_total_order_component(const mpart_options* opts, uint32_t dim, uint32_t order, mpart_map* out);

/* MapFactory::CreateTriangular(inputDim, outputDim, totalOrder, options)
 * (MParT/MapFactory.h:62-87, src/MapFactory.cpp:105-125). */
int mpart_create_triangular(uint32_t inputDim, uint32_t outputDim, uint32_t totalOrder, const mpart_options* opts,
                            mpart_map* out);

/* TriangularMap(components) (src/TriangularMap.cpp:10-61).  The components
 * must share one MapOptions; their coefficients are copied into the new map. */
int mpart_create_triangular_from(const mpart_map* comps, uint32_t ncomps, mpart_map* out);

int mpart_destroy(mpart_map map);

/* inputDim / outputDim / numCoeffs of ConditionalMapBase. */
int mpart_map_dims(mpart_map map, uint32_t* inputDim, uint32_t* outputDim, uint32_t* numCoeffs);

/* The MapOptions the map was built with and whether it is a TriangularMap
 * (1) or a single component (0); what a serialized map needs besides its
 * multi-index sets and coefficients (the reference keeps them in the objects
 * its cereal archives save, MParT/MonotoneComponent.h and TriangularMap.h). */
int mpart_map_options(mpart_map map, mpart_options* opts, int32_t* triangular);

/* Number of components and the multi-index set of component i (host arrays
 * sized by a first call with NULL pointers). */
int mpart_map_num_components(mpart_map map, uint32_t* ncomps);
int mpart_component_multiset(mpart_map map, uint32_t i, uint32_t* dim, uint32_t* nterms, uint32_t* nnz,
                             uint32_t* nzStarts, uint32_t* nzDims, uint32_t* nzOrders);

/* ParameterizedFunctionBase::SetCoeffs (deep copy, src/ParameterizedFunctionBase.cpp:184-214)
 * and WrapCoeffs (alias, 219-228).  coeffs is a device vector with `stride`. */
int mpart_set_coeffs(mpart_map map, const double* coeffs, int64_t len, int64_t stride, mpart_stream stream);
int mpart_wrap_coeffs(mpart_map map, double* coeffs, int64_t len, int64_t stride);
/* Copies the current coefficients into a device vector. */
int mpart_get_coeffs(mpart_map map, double* out, mpart_stream stream);

/* ConditionalMapBase::EvaluateImpl (MParT/ParameterizedFunctionBase.h:101-110):
 * out (outputDim x N) = T(pts). */
int mpart_evaluate(mpart_map map, mpart_cview pts, mpart_view out, mpart_stream stream);

/* ConditionalMapBase::LogDeterminantImpl (MParT/ConditionalMapBase.h:79-90):
 * out[n*stride] = log det dT/dx at pts(:, n). */
int mpart_log_determinant(mpart_map map, mpart_cview pts, double* out, int64_t stride, mpart_stream stream);

/* ParameterizedFunctionBase::GradientImpl (MParT/ParameterizedFunctionBase.h:140-160):
 * out (inputDim x N) = sum_i sens(i, n) d T_i / dx. */
int mpart_gradient(mpart_map map, mpart_cview pts, mpart_cview sens, mpart_view out, mpart_stream stream);

/* ParameterizedFunctionBase::CoeffGradImpl (MParT/ParameterizedFunctionBase.h:162-182):
 * out (numCoeffs x N) = sum_i sens(i, n) d T_i / dw. */
int mpart_coeff_grad(mpart_map map, mpart_cview pts, mpart_cview sens, mpart_view out, mpart_stream stream);

/* ConditionalMapBase::LogDeterminantCoeffGradImpl (MParT/ConditionalMapBase.h:130-140). */
int mpart_log_determinant_coeff_grad(mpart_map map, mpart_cview pts, mpart_view out, mpart_stream stream);

/* ConditionalMapBase::LogDeterminantInputGradImpl (MParT/ConditionalMapBase.h:148-159). */
int mpart_log_determinant_input_grad(mpart_map map, mpart_cview pts, mpart_view out, mpart_stream stream);

/* ConditionalMapBase::InverseImpl (MParT/ConditionalMapBase.h:99-110) with the
 * Bracket method and tolerances xtol, ytol (MParT/MonotoneComponent.h:331-443).
 * Component: x1 is dim x (1 or N) with the initial guess in its last row, r is
 * 1 x N.  Triangular map: x1 is (inputDim-outputDim) x N, r is outputDim x N. */
int mpart_inverse(mpart_map map, mpart_cview x1, mpart_cview r, mpart_view out, double xtol, double ytol,
                  mpart_stream stream);

/* Fused forward + adjoint of the Gaussian reverse-KL objective
 * (FastGaussianReverseKLObjective::ObjectivePlusCoeffGradImpl, src/MapObjective.cpp:93-234,
 * generalized to every output row):  *loss = mean_n(0.5 |T(x_n)|^2 - logdet(x_n)),
 * grad = mean_n(sum_i T_i dT_i/dw - dlogdet/dw).  One launch over every
 * component; no per-point Jacobian is stored.  loss and grad are device pointers. */
int mpart_kl_value_and_grad(mpart_map map, mpart_cview pts, double* loss, double* grad, mpart_stream stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mpart_kl_value_and_grad with every work buffer (partial
 * rows, reduction chunks, a contiguous copy of strided coefficients) in a
 * fixed device array of MPART_NOALLOC_SCRATCH_BYTES that the library holds
 * per GPU until the process ends.  The call allocates, frees and compiles
 * nothing: the map's specialized kernels run only if an earlier call (or
 * mpart_codegen_prepare) compiled them, on the schedule the autotuner picked
 * then, otherwise the generic kernel runs; the persistent grid shrinks when
 * its partial rows would not fit (results then differ from
 * mpart_kl_value_and_grad's in the last bits, by summation order).  Maps
 * whose work does not fit are refused with MPART_INVALID_ARGUMENT (more than
 * about 2 million coefficients, far beyond the specialization limit), maps
 * whose basis table needs a global buffer with MPART_NOT_IMPLEMENTED.  Every call on a GPU shares that GPU's scratch:
 * calls that overlap in time on one GPU (other streams, other threads) race
 * and give wrong results; serialize them. */
#define MPART_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
int mpart_kl_value_and_grad_noalloc(mpart_map map, mpart_cview pts, double* loss, double* grad, mpart_stream stream);

/* The same objective computed the way the reference structures it: per
 * component Evaluate, LogDeterminant, CoeffGrad(sens = T) and
 * LogDeterminantCoeffGrad launches with the numCoeffs x N Jacobians stored,
 * then a row reduction (src/MapObjective.cpp:107-215). */
int mpart_kl_value_and_grad_plain(mpart_map map, mpart_cview pts, double* loss, double* grad, mpart_stream stream);

/* Input gradient of the same objective, per point and fused over the
 * components: out (inputDim x N) = scale * grad_x [ |T(x_n)|^2 / 2 -
 * log det grad T(x_n) ], the point gradient the torch autograd path forms as
 * GradientImpl(pts, sens = T) - LogDeterminantInputGradImpl(pts)
 * (MParT/ParameterizedFunctionBase.h:140-160, MParT/ConditionalMapBase.h:148-159;
 * bindings/python/package/torch_helpers.py:91-107), in one launch without
 * storing T.  scale = 1/N gives the gradient of the mean. */
int mpart_kl_input_grad(mpart_map map, mpart_cview pts, double scale, mpart_view out, mpart_stream stream);

/* KL value, coefficient gradient and per-point input gradient in one pass:
 * as mpart_kl_value_and_grad, and gradInputs (inputDim x N; ptr may be NULL)
 * receives grad_x of the mean objective, (Gradient(x, T) -
 * LogDeterminantInputGrad(x)) / N per point, computed from the same
 * quadrature as the coefficient gradient (reference pieces as for
 * mpart_kl_value_and_grad and mpart_kl_input_grad). */
int mpart_kl_value_and_grads(mpart_map map, mpart_cview pts, double* loss, double* gradCoeffs,
                             mpart_view gradInputs, mpart_stream stream);

/* Both sides of the torch backward (bindings/python/package/torch_helpers.py:78-107)
 * in one pass: grad as mpart_coeff_vjp, and gradInputs (inputDim x N; ptr may
 * be NULL) as mpart_input_vjp, for the same sens / logdetSens. */
int mpart_vjp(mpart_map map, mpart_cview pts, mpart_cview sens, const double* logdetSens, int64_t lstride,
              double* grad, mpart_view gradInputs, mpart_stream stream);

/* Input-side vector-Jacobian product used by the torch backward
 * (bindings/python/package/torch_helpers.py:91-107), per point and fused over
 * the components: out(:, n) = Gradient(sens)(:, n) + logdetSens[n]
 * LogDeterminantInputGrad(:, n) (MParT/ParameterizedFunctionBase.h:140-160,
 * MParT/ConditionalMapBase.h:148-159) in one launch.  sens.ptr or logdetSens
 * may be NULL (that term is left out). */
int mpart_input_vjp(mpart_map map, mpart_cview pts, mpart_cview sens, const double* logdetSens, int64_t lstride,
                    mpart_view out, mpart_stream stream);

/* TrainOptions (MParT/TrainMap.h:19-40): NLopt-style stopping rules.  The
 * reference's opt_alg names an NLopt algorithm (default "LD_SLSQP"); here the
 * one optimizer is L-BFGS (MPART_TRAIN_LBFGS) with a strong-Wolfe line search;
 * `history` is its memory.  mpart_default_train_options() gives the reference
 * defaults (ftol_rel = ftol_abs = 1e-3, xtol_rel = xtol_abs = 1e-4,
 * maxeval = 1000, stopval = -inf, maxtime = inf) and history = 10. */
#define MPART_TRAIN_LBFGS 0
typedef struct mpart_train_options {
    int32_t algorithm;
    double stopval;
    double ftol_rel, ftol_abs;
    double xtol_rel, xtol_abs;
    int32_t maxeval;
    double maxtime;
    int32_t verbose;
    int32_t history;
} mpart_train_options;
void mpart_default_train_options(mpart_train_options* opts);

/* TrainMap(map, CreateGaussianKLObjective(train), options) (src/TrainMap.cpp:168-216):
 * minimizes the standard-Gaussian KL objective over the map's coefficients,
 * a triangular map one component at a time on its rows of `train`
 * (TrainMap_Triangular_StandardNormalDensity, 141-166).  Every objective
 * evaluation is one fused KL pass on the device; the optimizer runs on the
 * host.  For a triangular map *error receives the sum of the components'
 * final objective values (mean(T_i^2 / 2 - log dT_i/dx_i), without the
 * Gaussian normalization constant, as FastGaussianReverseKLObjective returns
 * them); a lone component is trained as a whole under the KLObjective
 * (184-199), whose value includes 0.5 d log(2 pi) with d its output
 * dimension, and *error is that value.  *evals is the number of objective
 * evaluations; the trained coefficients stay in the map.  train is a device
 * view (inputDim x N). */
int mpart_train_map(mpart_map map, mpart_cview train, const mpart_train_options* opts, double* error, int32_t* evals,
                    mpart_stream stream);

/* Fused vector-Jacobian product used by the torch backward
 * (bindings/python/package/torch_helpers.py:91-107):
 * grad = sum_n [ CoeffGrad(sens)(:, n) + logdetSens[n] LogDeterminantCoeffGrad(:, n) ].
 * sens.ptr or logdetSens may be NULL. */
int mpart_coeff_vjp(mpart_map map, mpart_cview pts, mpart_cview sens, const double* logdetSens, int64_t lstride,
                    double* grad, mpart_stream stream);

/* Evaluate that also returns the number of integrand evaluations per point and
 * component (ncomps x N, int32) -- exposes the adaptive trees for parity checks. */
int mpart_evaluate_counted(mpart_map map, mpart_cview pts, mpart_view out, int32_t* nevals, mpart_stream stream);

/* Selects per-component launches ("plain", 1) or one launch over all
 * components (0, default) for Evaluate and LogDeterminant. */
int mpart_set_plain_launches(mpart_map map, int plain);

/* Runtime specialization.  By default the first Evaluate / LogDeterminant /
 * CoeffGrad / LogDeterminantCoeffGrad / KL / VJP call on a map compiles, with
 * hipRTC, kernels specialized to the map's multi-index sets (every term a
 * straight-line expression over register-resident basis values) and uses them
 * from then on; maps with more than 4096 coefficients, or a failed
 * compilation, use the precompiled generic kernels.  Equal structures share
 * one compiled module per process.  No reference counterpart: the reference
 * evaluates the multi-index set through its Kokkos kernels
 * (MParT/MultivariateExpansionWorker.h:140-420).
 *
 * mpart_set_codegen(map, 0) selects the generic kernels for this map (1
 * restores the default).  mpart_codegen_prepare compiles now (instead of on
 * the first call) and reports *specialized = 1 when the specialized kernels
 * are in use, 0 otherwise, and the compile time of the module in seconds (0
 * when it came from the cache).  mpart_codegen_error returns the reason of a
 * failed compilation ("" otherwise).  When the environment variable
 * MPART_CODEGEN_CACHE names a writable directory, every compiled module's
 * code object is stored there (keyed by hashes of the generated source, the
 * device header, the target and the options) and later processes load it
 * instead of compiling: they skip the compile time and run the identical
 * binary, whichever compiler their own hipRTC would load. */
/* Batches of independent maps.  A batch runs one method over many maps (each
 * with its own points, outputs and coefficients) in one persistent launch:
 * every map's points are split into tiles, the tiles are ordered by
 * estimated cost and the GPU's blocks take them from a shared counter.  No
 * reference counterpart (the reference loops over maps and components on the
 * host, src/TriangularMap.cpp:100-310); it is the batched form of the
 * per-map EvaluateImpl / LogDeterminantImpl / CoeffGradImpl / InverseImpl.
 *
 * The maps must share MapOptions, must outlive the batch, and keep their own
 * coefficients (read at call time).  Arrays hold one entry per map, in batch
 * order.  Each method keeps its arguments on the device and uploads a call's
 * arguments when they differ from the previous call of that method; under
 * stream capture they must not differ (make the same calls once before
 * capturing), so a captured graph -- e.g. Evaluate, CoeffGrad and Inverse in
 * sequence -- replays with no host work.  A captured launch holds the tables,
 * tile counts and grid of its call, so once a method has been captured on a
 * batch its arguments are frozen: a later call of that method with other
 * arguments (other pointers, sizes or strides, or a map whose coefficients
 * were re-wrapped elsewhere) returns MPART_INVALID_ARGUMENT before launching
 * anything, and the graph keeps replaying the captured call; use another
 * batch for other arguments.  Calls with the captured arguments stay allowed
 * between replays.  Batched
 * Inverse follows ConditionalMapBase::Inverse per map: x1 has
 * inDim - outDim rows for a triangular map, and all inDim rows (the last one
 * the initial guess) for a single component. */
typedef struct mpart_batch_s* mpart_batch;
int mpart_create_batch(const mpart_map* maps, uint32_t n, mpart_batch* out);
int mpart_destroy_batch(mpart_batch batch);
int mpart_batch_size(mpart_batch batch, uint32_t* n);
/* Schedule of the batched Evaluate / LogDeterminant: 1 (a lane per point for
 * every map) or 4 (mixed: maps with many terms -- at least 4096 coefficients,
 * or a per-lane basis table that does not fit 256 lanes of LDS -- take a
 * group of 4 lanes per point sharing its table, the others a lane per point,
 * in the same launch).  mpart_batch_set_point_group forces it (0: autotune, the
 * default: the first uncaptured call with >= 65536 points in all times both
 * and keeps the faster; the choice is frozen once a call has been captured).
 * mpart_batch_eval_schedule reports it (0: not tuned yet) and the two trial
 * times in ms.  No counterpart in the reference. */
int mpart_batch_set_point_group(mpart_batch batch, int32_t lanes);
int mpart_batch_eval_schedule(mpart_batch batch, int32_t* lanes, double* trialMs);

/* out and logdet may be NULL (not computed); so may out[i].ptr / logdet[i]. */
int mpart_batch_evaluate(mpart_batch batch, const mpart_cview* pts, const mpart_view* out, double* const* logdet,
                         const int64_t* logdetStride, mpart_stream stream);
int mpart_batch_coeff_grad(mpart_batch batch, const mpart_cview* pts, const mpart_cview* sens, const mpart_view* out,
                           mpart_stream stream);
int mpart_batch_inverse(mpart_batch batch, const mpart_cview* x1, const mpart_cview* r, const mpart_view* out,
                        double xtol, double ytol, mpart_stream stream);
/* Schedule of the batched Inverse: 1 runs, component by component, one setup
 * pass and one persistent solve pass over the points of every map, lanes
 * taking new points as their solves end; 2 solves each point to the end in
 * its lane (a lane per point, all components in one launch).  0 (the
 * default) autotunes: the first uncaptured call with >= 65536 points in all
 * times both and keeps the faster (until then 2; frozen once an Inverse call
 * of the batch has been captured).  Both take the reference's steps (MonotoneComponent::InverseImpl,
 * MParT/MonotoneComponent.h:331-443); their roots can differ in the last bits
 * where the compiled arithmetic differs.  mpart_batch_inverse_schedule reports
 * the schedule (0: not tuned yet) and the trial times in ms of 1 and 2.
 * No counterpart in the reference. */
int mpart_batch_set_inverse_schedule(mpart_batch batch, int32_t schedule);

/* Hand-over of the longest inverse solves (adaptive quadrature rules; the
 * per-lane solve scheduling of mpart_inverse and of batch schedule 1): once
 * a wavefront of the solve pass finds no point left to take, a solve of it
 * that has taken more than solveSteps quadrature interval steps is solved
 * again by a pass that gives each such solve a whole wavefront and spreads
 * each of its quadratures over the 64 lanes, level by level of the
 * subdivision tree (adaptive Simpson and adaptive Clenshaw-Curtis; the fixed
 * rule has no such tail).  The trees, sums and roots are the same as the
 * one-lane solve's, bit for bit; only the time differs.  solveSteps = 0
 * turns the hand-over off.  Default 1024 (the median solve takes a few
 * hundred steps).  No counterpart in the reference, whose solves run a
 * thread each (MParT/MonotoneComponent.h:331-443). */
int mpart_set_inverse_handover(mpart_map map, int32_t solveSteps);
int mpart_batch_set_inverse_handover(mpart_batch batch, int32_t solveSteps);
/* Totals of the last Inverse call with the hand-over on (waits for stream):
 * stats[0] solves handed over, [1] their T evaluations, [2] intervals
 * evaluated level by level, [3] T evaluations integrated depth-first because
 * the tree outgrew the wavefront's workspace, [4] ... because an interval
 * error was not finite, [5] the largest tree (intervals), [6] the deepest
 * level, [7] the most T evaluations of one solve, [8] the longest solve in
 * ticks of the 100 MHz device clock.  No counterpart in the reference. */
int mpart_inverse_handover_stats(mpart_map map, uint64_t stats[9], void* stream);
int mpart_batch_inverse_handover_stats(mpart_batch batch, uint64_t stats[9], void* stream);
int mpart_batch_inverse_schedule(mpart_batch batch, int32_t* schedule, double* trialMs);

int mpart_set_codegen(mpart_map map, int enable);
int mpart_codegen_prepare(mpart_map map, int32_t* specialized, double* compileSeconds);
const char* mpart_codegen_error(mpart_map map);

/* Point schedule of the generic (non-specialized) Evaluate / LogDeterminant
 * kernels: lanes per point, 1 (a lane per point, the basis table per lane),
 * 4, 8 or 64 (a group of lanes per point sharing one table and splitting the
 * terms).  mpart_set_point_group forces it (0: autotune, the default: the
 * first call with >= 16384 points times the four and keeps the fastest).
 * A fifth schedule, code 2, splits the components over the four wavefronts
 * of a block, which share one basis table for a tile of 64 points and run a
 * lane per point over their own components (maps with several components,
 * all components in one call).  The generic fused KL / VJP pass has the same
 * five schedules and its own tuning.  mpart_eval_schedule reports the
 * choice for [Evaluate, LogDeterminant only, fused pass] in lanes[3] (0: not
 * tuned yet) and the trial times in ms in trialMs[15], five per kind (lanes
 * 1, 4, 8, 64, then the split schedule).  No counterpart in the reference,
 * which runs a thread per point (MParT/Utilities/KokkosHelpers.h:17-28). */
int mpart_set_point_group(mpart_map map, int32_t lanes);
int mpart_eval_schedule(mpart_map map, int32_t* lanes, double* trialMs);

/* Schedule of the specialized fused KL / VJP kernel.  The first such call on
 * a map with at least 65536 points (outside stream capture) times four
 * schedules -- {compiler-chosen occupancy, three waves per SIMD} x
 * {persistent grid of one or two blocks per resident slot} -- twice each on
 * the caller's stream and keeps the fastest for the map.  *choice is the
 * schedule index (bit 0: three-waves kernel, bit 1: doubled grid) or -1
 * before tuning; trialMs (4 doubles, may be NULL) receives the trial times in
 * milliseconds, -1 for a schedule left out because its kernel would spill
 * registers.  No reference counterpart. */
int mpart_fused_schedule(mpart_map map, int32_t* choice, double* trialMs);
/* Forces schedule `choice` (0-3 as above) for the map's specialized fused
 * calls, or -1 to tune again on the next call (schedules differ in the order
 * of the sums over points, so forcing one makes results repeatable across
 * processes).  No reference counterpart. */
int mpart_set_fused_schedule(mpart_map map, int32_t choice);

/* Resources of one of the map's specialized (hipRTC-compiled) kernels, as the
 * HIP runtime reports them (hipFuncGetAttribute): registers per lane,
 * private (scratch) bytes per lane and static LDS bytes.  kernel is one of
 * "fused", "fused_w3", "eval", "coef", "ldcoef" (first module) or
 * "fused_in", "ivjp", "klx" (module with input gradients);
 * MPART_INVALID_ARGUMENT when that kernel is not compiled.  The compiler is
 * the one the process's hipRTC loads, so a map's kernels can differ between
 * processes that load different ROCm runtimes.  No reference counterpart. */
int mpart_codegen_kernel_info(mpart_map map, const char* kernel, int32_t* vgprs, int32_t* scratchBytes,
                              int32_t* ldsBytes);

#ifdef __cplusplus
}
#endif

#endif /* MPART_AMD_H */
