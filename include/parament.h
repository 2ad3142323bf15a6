/*
 * parament.h -- C ABI of the MI355X-native equiprop library (libparament.so).
 *
 * Every function, type and enum value of the reference header
 * (parament-integrator/parament src/cuda/parament.h) is re-declared here with the
 * same name, argument order and binary layout, so existing callers (the pyparament
 * ctypes bindings, src/python/pyparament/parament/paramentlib.py:57-72, and C
 * programs) link against this library unchanged.  The complex types are plain
 * {float, float} / {double, double} structs, layout-identical to the reference's
 * cuComplex / cuDoubleComplex.
 *
 * Entry points after the "Additions" marker have no counterpart in the reference;
 * each comment says which caller they are for.
 */
#ifndef PARAMENT_H_
#define PARAMENT_H_

#ifdef __cplusplus
extern "C" {
#else
#include <stdbool.h>
#include <stddef.h>
#endif

#if defined(__GNUC__)
#define LIBSPEC __attribute__((visibility("default")))
#else
#define LIBSPEC
#endif

/* Layout-identical stand-ins for cuComplex / cuDoubleComplex (reference parament.h:52-55). */
typedef struct cuComplex { float x, y; } cuComplex;
typedef struct cuDoubleComplex { double x, y; } cuDoubleComplex;

struct Parament_Context_f32;
struct Parament_Context_f64;
typedef struct Parament_Context_f32 Parament_Context_f32;
typedef struct Parament_Context_f64 Parament_Context_f64;

/* Quadrature rule (reference parament.h:64-69; same values). */
typedef enum Parament_QuadratureSpec {
    PARAMENT_QUADRATURE_NONE = 0x00000000,
    PARAMENT_QUADRATURE_MIDPOINT = 0x01000000,
    PARAMENT_QUADRATURE_SIMPSON = 0x02000000
} Parament_QuadratureSpec;

/* Error codes (reference parament.h:74-130; same values).  CUBLAS_INIT_FAILED and
 * CUBLAS_FAILED keep their numbers; here they mean "device runtime set-up failed"
 * and "a device kernel or copy failed". */
typedef enum Parament_ErrorCode {
    PARAMENT_STATUS_SUCCESS = 0,
    PARAMENT_STATUS_HOST_ALLOC_FAILED = 10,
    PARAMENT_STATUS_DEVICE_ALLOC_FAILED = 20,
    PARAMENT_STATUS_CUBLAS_INIT_FAILED = 30,
    PARAMENT_STATUS_INVALID_VALUE = 50,
    PARAMENT_STATUS_CUBLAS_FAILED = 60,
    PARAMENT_STATUS_SELECT_SMALLER_DT = 70,
    PARAMENT_STATUS_NO_HAMILTONIAN = 80,
    PARAMENT_STATUS_INVALID_QUADRATURE_SELECTION = 90,
    /* addition: the call needs more device memory than a fixed workspace holds */
    PARAMENT_STATUS_WORKSPACE_TOO_SMALL = 100,
    PARAMENT_FAIL = 1000
} Parament_ErrorCode;

/* ---- single precision (complex64) API: reference parament.h:155-357 ---- */

/* Reference parament.h:155 / parament.cpp:51-128.  Allocates the host context only;
 * device resources are bound to the device current at the first setHamiltonian. */
LIBSPEC Parament_ErrorCode Parament_create(Parament_Context_f32 **handle_p);

/* Reference parament.h:177 / parament.cpp:185-205.  NULL is a no-op. */
LIBSPEC Parament_ErrorCode Parament_destroy(Parament_Context_f32 *handle);

/* Reference parament.h:219-220 / parament.cpp:210-368.  H0 is dim x dim, H1 is
 * amps consecutive dim x dim matrices, all row-major as passed by pyparament. */
LIBSPEC Parament_ErrorCode Parament_setHamiltonian(Parament_Context_f32 *handle, cuComplex *H0, cuComplex *H1,
    unsigned int dim, unsigned int amps, bool use_magnus, enum Parament_QuadratureSpec quadrature_mode);

/* Reference parament.h:261-262 / parament.cpp:791-851.  carr holds amps arrays of
 * pts samples (carr[k*pts + j]); out receives the dim x dim propagator.  Blocking. */
LIBSPEC Parament_ErrorCode Parament_equiprop(Parament_Context_f32 *handle, cuComplex *carr, double dt,
    unsigned int pts, unsigned int amps, cuComplex *out);

/* Reference parament.h:286 / parament.cpp:723-739. */
LIBSPEC int Parament_selectIterationCycles_fp32(double H_norm, double dt);

/* Reference parament.h:308 / parament.cpp:771-776.  cycles must be odd and >= 3
 * (the reference's Clenshaw loop reads out of bounds otherwise); other values
 * return PARAMENT_STATUS_INVALID_VALUE and leave the setting unchanged. */
LIBSPEC Parament_ErrorCode Parament_setIterationCyclesManually(Parament_Context_f32 *handle, unsigned int cycles);

/* Reference parament.h:327 / parament.cpp:781-786. */
LIBSPEC Parament_ErrorCode Parament_automaticIterationCycles(Parament_Context_f32 *handle);

/* Reference parament.h:347 / parament.cpp:854-857. */
LIBSPEC Parament_ErrorCode Parament_peekAtLastError(Parament_Context_f32 *handle);

/* Reference parament.h:357 / parament.cpp:859-882 (same strings). */
LIBSPEC const char *Parament_errorMessage(Parament_ErrorCode errorCode);

/* ---- double precision (complex128) API: reference parament.h:363-400 ---- */

LIBSPEC Parament_ErrorCode Parament_create_fp64(Parament_Context_f64 **handle_p);
LIBSPEC Parament_ErrorCode Parament_destroy_fp64(Parament_Context_f64 *handle);
LIBSPEC Parament_ErrorCode Parament_setHamiltonian_fp64(Parament_Context_f64 *handle, cuDoubleComplex *H0,
    cuDoubleComplex *H1, unsigned int dim, unsigned int amps, bool use_magnus, Parament_QuadratureSpec quadrature_mode);
LIBSPEC Parament_ErrorCode Parament_equiprop_fp64(Parament_Context_f64 *handle, cuDoubleComplex *carr, double dt,
    unsigned int pts, unsigned int amps, cuDoubleComplex *out);
LIBSPEC Parament_ErrorCode Parament_setIterationCyclesManually_fp64(Parament_Context_f64 *handle, unsigned int cycles);
LIBSPEC Parament_ErrorCode Parament_automaticIterationCycles_fp64(Parament_Context_f64 *handle);
LIBSPEC Parament_ErrorCode Parament_peekAtLastError_fp64(Parament_Context_f64 *handle);
LIBSPEC int Parament_selectIterationCycles_fp64(double H_norm, double dt);

/* Reference deviceInfo.c:30-59: prints the visible devices to stdout. */
LIBSPEC void device_info(void);

/* Reference mathhelper.h:67-68: max row sum of moduli of a row-major dim x dim matrix. */
LIBSPEC double OneNorm(cuComplex *mat, unsigned int dim);
LIBSPEC double OneNorm_fp64(cuDoubleComplex *mat, unsigned int dim);

/* ======================== Additions (no reference counterpart) ======================== */

/* Arithmetic of a context's equiprop:
 *   PARAMENT_PRECISION_NATIVE: every step in the context's precision: FP64 for complex128
 *       (the reference algorithm), FP32 for complex64.  Default of complex128 contexts.  Where
 *       tau Hnorm <= 1 and M >= 5, complex64 slices (dims above 16) evaluate the same degree-M
 *       polynomial in powers of tau X (Horner's rule in (tau X)^2: (M + 1) / 2 products instead
 *       of the recurrence's M - 1); complex128 keeps the reference's recurrence.
 *   PARAMENT_PRECISION_MIXED:  FP64 for Hnorm, the Bessel coefficients, the slice
 *       Hamiltonians, the first-order Chebyshev terms and the upper levels of the product
 *       tree; the higher-order Chebyshev remainder and the lowest three tree levels run as
 *       split products with FP32 accumulation on the matrix cores (FP32-accurate products):
 *       scaled fp16 two-way splits (three fp16 products per product) in the power form
 *       (tau Hnorm <= 1: the fused kernel at dim <= 64, the batched GEMMs above from M = 5 on),
 *       bf16 three-way splits (six bf16 products) otherwise.  The environment variable
 *       PARAMENT_MIXED_SPLIT=bf16x3 keeps the bf16 three-way split everywhere (f16x4: all four
 *       fp16 products in the fused kernel), for precision studies; it is read once per
 *       process.  Its accuracy (within 1e-5 of FP64 at the
 *       benchmark workloads) depends on the slices' actual norm ||tau X_j||, which the order M
 *       and the power form do not see: as in the reference, Hnorm sums the operators' norms
 *       without the control amplitudes.  Measured at the dim-64, 65,536-step workload with the
 *       controls scaled so that max ||tau X_j||_2 = 0.22 / 0.71 / 1.38: 6.5e-6 / 1.2e-5 / 1.7e-4;
 *       controls much larger than 1 in modulus call for NATIVE (FP64) or a smaller dt.
 *       Default of complex64 contexts, which keep the FP32
 *       iteration table's M and, for dims <= 16 on the automatic kernel path, the FP32
 *       wavefront-resident kernels.  A complex128 context in this mode takes M from the FP32
 *       table raised until the accumulated truncation error stays below 2e-6.
 *   PARAMENT_PRECISION_MIXED_FP32:   the same with FP32 matrix-core products.
 *   PARAMENT_PRECISION_MIXED_BF16X2: the same with bf16 hi+lo (two-way split) products
 *       (about 16 significant bits; for precision studies).
 * MIXED runs the fused on-chip kernel for dim <= 64 and batched GEMMs above; the other two
 * always run the batched GEMMs.  getPrecisionMode returns the current mode (-1 for NULL). */
typedef enum Parament_PrecisionMode {
    PARAMENT_PRECISION_NATIVE = 0,
    PARAMENT_PRECISION_MIXED = 1,
    PARAMENT_PRECISION_MIXED_FP32 = 2,
    PARAMENT_PRECISION_MIXED_BF16X2 = 3
} Parament_PrecisionMode;

LIBSPEC Parament_ErrorCode Parament_setPrecisionMode(Parament_Context_f32 *handle, Parament_PrecisionMode mode);
LIBSPEC Parament_ErrorCode Parament_setPrecisionMode_fp64(Parament_Context_f64 *handle, Parament_PrecisionMode mode);
LIBSPEC int Parament_getPrecisionMode(Parament_Context_f32 *handle);
LIBSPEC int Parament_getPrecisionMode_fp64(Parament_Context_f64 *handle);

/* Chebyshev order M used by the last successful equiprop (-1 before any). */
LIBSPEC int Parament_getIterationCycles(Parament_Context_f32 *handle);
LIBSPEC int Parament_getIterationCycles_fp64(Parament_Context_f64 *handle);

/* Device-resident variant of equiprop for pipelines and benchmarks that keep the
 * controls and the result in device memory: carr_dev and out_dev are device
 * pointers (same layouts as Parament_equiprop).  The call is asynchronous on the
 * context's stream; Parament_synchronize waits for it. */
LIBSPEC Parament_ErrorCode Parament_equipropDevice(Parament_Context_f32 *handle, const cuComplex *carr_dev, double dt,
    unsigned int pts, unsigned int amps, cuComplex *out_dev);
LIBSPEC Parament_ErrorCode Parament_equipropDevice_fp64(Parament_Context_f64 *handle, const cuDoubleComplex *carr_dev,
    double dt, unsigned int pts, unsigned int amps, cuDoubleComplex *out_dev);
/* Allocation-free variant of Parament_equipropDevice_fp64.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, arithmetic and result as
 * Parament_equipropDevice_fp64 (every precision mode and kernel path), but the call's working
 * memory comes from a fixed 256 MiB device workspace that the library holds on each GPU it runs
 * on until the process ends, and it allocates nothing.  A problem whose working set exceeds the
 * workspace returns PARAMENT_STATUS_WORKSPACE_TOO_SMALL (e.g. dim 64 at 65,536 steps fits in
 * the mixed mode, about 3,000 slices in FP64).  All calls on one GPU share the workspace: calls
 * that overlap on one GPU (other streams, other contexts, other threads) race.  Asynchronous on
 * the context's stream, like equipropDevice; setHamiltonian must have run (it allocates). */
LIBSPEC Parament_ErrorCode Parament_equipropDeviceNoAlloc_fp64(Parament_Context_f64 *handle,
    const cuDoubleComplex *carr_dev, double dt, unsigned int pts, unsigned int amps, cuDoubleComplex *out_dev);
LIBSPEC Parament_ErrorCode Parament_synchronize(Parament_Context_f32 *handle);
LIBSPEC Parament_ErrorCode Parament_synchronize_fp64(Parament_Context_f64 *handle);

/* Selects the kernel family for testing and benchmarking:
 *   0 = automatic (fused on-chip kernels for dim <= 64, step-wise batched kernels above),
 *   1 = always the step-wise batched kernels (the reference's launch structure). */
LIBSPEC Parament_ErrorCode Parament_setKernelPath(Parament_Context_f32 *handle, int path);
LIBSPEC Parament_ErrorCode Parament_setKernelPath_fp64(Parament_Context_f64 *handle, int path);

/* Stream control, for callers that order equiprop against their own device work (e.g. a
 * PyTorch stream).  A context works on its own non-blocking stream (created by the first
 * setHamiltonian; getStream returns it) until setStream is called; setStream(NULL) selects
 * the device's null stream, and passing back the handle getStream returned earlier restores
 * the context's own stream. */
LIBSPEC Parament_ErrorCode Parament_setStream(Parament_Context_f32 *handle, void *hip_stream);
LIBSPEC Parament_ErrorCode Parament_setStream_fp64(Parament_Context_f64 *handle, void *hip_stream);
LIBSPEC void *Parament_getStream(Parament_Context_f32 *handle);
LIBSPEC void *Parament_getStream_fp64(Parament_Context_f64 *handle);

/* Kernel timing for benchmarks: when enabled, each equiprop brackets its Chebyshev slice
 * kernel and its product-tree launches with HIP events on the context's stream;
 * getKernelTimes waits for them and returns the two durations of the last call in ms. */
LIBSPEC Parament_ErrorCode Parament_setTiming(Parament_Context_f32 *handle, int enable);
LIBSPEC Parament_ErrorCode Parament_setTiming_fp64(Parament_Context_f64 *handle, int enable);
LIBSPEC Parament_ErrorCode Parament_getKernelTimes(Parament_Context_f32 *handle, double *slice_ms, double *tree_ms);
LIBSPEC Parament_ErrorCode Parament_getKernelTimes_fp64(Parament_Context_f64 *handle, double *slice_ms, double *tree_ms);

/* Adjoint (GRAPE) gradient of the gate fidelity F = |Tr(target^H U)|^2 / dim^2 of the
 * propagator U that Parament_equiprop[_fp64] computes, with respect to every control sample
 * (carr layout, real controls: the imaginary parts must be zero).  grad receives amps x pts
 * values dF/dc[k*pts + j] (quadrature chain rule included), fidelity (optional) receives F.
 * For callers optimising controls (GRAPE loops, the PyTorch autograd wrapper).  Any dim (dims
 * above 32 run on batched GEMMs and keep the whole product tree), every quadrature mode
 * including Magnus.  A complex64 context in PARAMENT_PRECISION_MIXED (its default) runs the
 * batched GEMMs of the power form (tau Hnorm <= 1, M >= 5) on scaled fp16 two-way split
 * products, with the costates formed from target^H / ||target||_F and the factor applied to
 * the traces; its other modes use FP32 products.  The Device variants take device pointers and
 * are asynchronous on the context's stream; they read only the real parts of carr_dev. */
LIBSPEC Parament_ErrorCode Parament_gateFidelityGradient(Parament_Context_f32 *handle, const cuComplex *carr,
    double dt, unsigned int pts, unsigned int amps, const cuComplex *target, float *fidelity, float *grad);
LIBSPEC Parament_ErrorCode Parament_gateFidelityGradient_fp64(Parament_Context_f64 *handle, const cuDoubleComplex *carr,
    double dt, unsigned int pts, unsigned int amps, const cuDoubleComplex *target, double *fidelity, double *grad);
LIBSPEC Parament_ErrorCode Parament_gateFidelityGradientDevice(Parament_Context_f32 *handle,
    const cuComplex *carr_dev, double dt, unsigned int pts, unsigned int amps, const cuComplex *target_dev,
    float *fidelity_dev, float *grad_dev);
LIBSPEC Parament_ErrorCode Parament_gateFidelityGradientDevice_fp64(Parament_Context_f64 *handle,
    const cuDoubleComplex *carr_dev, double dt, unsigned int pts, unsigned int amps, const cuDoubleComplex *target_dev,
    double *fidelity_dev, double *grad_dev);

/* Memory policy of the gradient (no reference counterpart; for long control sequences).
 *   segment_slices = 0: keep the whole product tree of the N slices (default; fastest).
 *   segment_slices = 1: checkpoint at segments of 2^ceil(log2 sqrt(N)) slices.
 *   segment_slices >= 2: checkpoint at segments of that many slices (rounded up to a power of two).
 * With checkpointing only the segment products and the tree above them stay in memory; the
 * slices are processed in chunks of chunk_slices (0: max(segment, 2048); rounded up to a
 * multiple of the segment) whose propagators are computed again in the backward pass, so
 * the working set is O(N / segment + chunk) matrices for the cost of one more forward pass.
 * Results are identical with and without checkpointing.  Dims above 32 always keep the whole
 * tree and process their per-slice arrays in chunks of chunk_slices slices (0: as many as fit
 * in 4 GiB); with segment_slices = 0 they also keep the power form's S and Horner brackets of
 * every slice for the derivative ((M - 1) / 2 more matrices per slice, when that takes at most
 * 8 GiB), and any other segment_slices forms them again instead.  getGradientWorkspace returns
 * the device bytes the last gradient call worked in. */
LIBSPEC Parament_ErrorCode Parament_setGradientCheckpointing(Parament_Context_f32 *handle, unsigned int segment_slices,
    unsigned int chunk_slices);
LIBSPEC Parament_ErrorCode Parament_setGradientCheckpointing_fp64(Parament_Context_f64 *handle,
    unsigned int segment_slices, unsigned int chunk_slices);
LIBSPEC size_t Parament_getGradientWorkspace(Parament_Context_f32 *handle);
LIBSPEC size_t Parament_getGradientWorkspace_fp64(Parament_Context_f64 *handle);

/* Ensembles of independent complex64 problems (no reference counterpart; for batched studies
 * and optimisations over many systems of different sizes).  Instance i has dimension dims[i],
 * amps[i] control Hamiltonians, pts[i] control samples and time step dt[i]; one quadrature
 * mode (NONE, MIDPOINT or SIMPSON) applies to all.  Packed layouts, instance after instance:
 *   H:   (1 + amps[i]) row-major dims[i] x dims[i] matrices (H0, then H_1 ..),
 *   carr: amps[i] x pts[i] samples (carr layout of Parament_equiprop),
 *   out: dims[i] x dims[i] propagators.
 * Each instance computes what Parament_equiprop computes for it in PARAMENT_PRECISION_NATIVE
 * mode (FP32 throughout, FP32 iteration-count table, the power form where that mode takes it,
 * same quadrature and product order), except that instances above dim 64 run the power form's
 * products and their product tree on scaled fp16 two-way splits (FP32-accurate, as the
 * default mode's step-wise path).  setExecution(0) (default) runs the whole ensemble as
 * one captured HIP graph: the instances of dim <= 64 in multi-instance kernels per dim class
 * (16, 32, 48, 64), the larger ones on batched GEMMs, on concurrent branches (up to four for the
 * larger ones); the tree block size and grid of each class are tuned on first use (or by
 * ensembleTune).  setExecution(1) runs one instance
 * after the other (the per-instance loop).  getTuning fills config[10]: (T, log2 tree block,
 * grid multiple) for classes 16, 32, 64, then the number of branches (the 48 class is tuned
 * alike and not reported). */
typedef struct Parament_Ensemble Parament_Ensemble;
LIBSPEC Parament_ErrorCode Parament_ensembleCreate(Parament_Ensemble **ensemble_p, unsigned int count,
    const unsigned int *dims, const unsigned int *amps, const unsigned int *pts, const double *dt,
    Parament_QuadratureSpec quadrature_mode);
LIBSPEC Parament_ErrorCode Parament_ensembleDestroy(Parament_Ensemble *ensemble);
LIBSPEC Parament_ErrorCode Parament_ensembleSetHamiltonians(Parament_Ensemble *ensemble, const cuComplex *H);
LIBSPEC Parament_ErrorCode Parament_ensembleSetExecution(Parament_Ensemble *ensemble, int mode);
LIBSPEC Parament_ErrorCode Parament_ensembleTune(Parament_Ensemble *ensemble);
LIBSPEC Parament_ErrorCode Parament_ensembleGetTuning(Parament_Ensemble *ensemble, int *config);
LIBSPEC Parament_ErrorCode Parament_ensembleEquiprop(Parament_Ensemble *ensemble, const cuComplex *carr, cuComplex *out);
LIBSPEC Parament_ErrorCode Parament_ensembleEquipropDevice(Parament_Ensemble *ensemble, const cuComplex *carr_dev,
    cuComplex *out_dev);
LIBSPEC Parament_ErrorCode Parament_ensembleSynchronize(Parament_Ensemble *ensemble);
LIBSPEC Parament_ErrorCode Parament_ensemblePeekAtLastError(Parament_Ensemble *ensemble);
/* Gate-fidelity gradients of every instance (as Parament_gateFidelityGradient computes them
 * for it in a complex64 context's default mode; real controls).  targets packed like out (dims[i] x dims[i] row-major
 * each), fidelity[i] = F_i (optional), grad packed like carr: amps[i] x pts[i] real values
 * dF_i/dc per instance.  setExecution(0): the instances on concurrent branches (fused kernels
 * for dims <= 32, batched GEMMs above), captured as one HIP graph per (targets, fidelity,
 * grad) pointer triple and replayed; setExecution(1): one instance after the other.  The
 * Device variant takes device pointers and is asynchronous on the ensemble's stream. */
LIBSPEC Parament_ErrorCode Parament_ensembleGateFidelityGradient(Parament_Ensemble *ensemble, const cuComplex *carr,
    const cuComplex *targets, float *fidelity, float *grad);
LIBSPEC Parament_ErrorCode Parament_ensembleGateFidelityGradientDevice(Parament_Ensemble *ensemble,
    const cuComplex *carr_dev, const cuComplex *targets_dev, float *fidelity_dev, float *grad_dev);

#ifdef __cplusplus
}
#endif

#endif  /* PARAMENT_H_ */
