// Launchers for the equiprop device kernels (defined in kernels.hip / mixed.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "device_common.hpp"

namespace pa {

constexpr int kMaxCheb = 48;        // Chebyshev coefficients carried in kernel arguments (M <= 41)
constexpr int kMaxCoeffs = 256;     // operator coefficients per slice (controls + Magnus terms)
constexpr int kGradCoeffs = 64;     // controls per slice of the gradient kernels

template <class R> struct ChebArgs {
    const cplx<R>* ops;    // (1 + ncoef) row-major d x d operators: H0, H_1.., Magnus terms
    int ncoef;             // operators after H0
    int d;
    Controls<R> ctl;
    int N;                 // slices computed by this launch: j0 .. j0 + N - 1
    int M;                 // Chebyshev order (odd, >= 3)
    R s;                   // 2 / Hnorm
    R asm_scale = R(1);    // step-wise assembly only: X_j is written multiplied by this
    R are[kMaxCheb], aim[kMaxCheb];   // a_k = (-i)^k J_k(tau Hnorm); are[0] holds a_0 - 1
    cplx<R>* out;          // N x d x d, E_j = U_j - I (out[0] is slice j0)
    int j0 = 0;            // first slice
    // gradient only: the same polynomial in powers of Y = tau X (tau Hnorm <= 1), m_k = pre/pim[k]
    // with pre[0] = m_0 - 1; the fused slice kernel and the derivative kernel then use the power form
    int ps = 0;
    R pre[kMaxCheb], pim[kMaxCheb];
    const cplx<R>* opsT = nullptr;   // gradient only: the ncoef operators after H0, each transposed
};

// Mixed precision: FP64 operators / coefficients; output: the level-F nodes E = U - I of the
// pairwise product tree (F = mixed_fused_levels(N)) in complex64.
struct MixedArgs {
    const double2* ops;
    int ncoef;
    int d;
    Controls<double> ctl;
    int N;
    int M;
    double s;
    double are[kMaxCheb], aim[kMaxCheb];   // are[0] holds a_0 - 1 (ps: m_0 - 1)
    float2* out;           // ceil(N / 2^F) x d x d
    int ps = 0;            // 1: are/aim hold the power-form coefficients m_k of Y = tau X
    int ps_levels = -1;    // power form: bits 2j, 2j+1 = split level (0..2) of the product Y^2 Q_{j+1}
};

// Fused Chebyshev slice kernel, dims <= 64 (padded to 16 / 32 / 64).
template <class R> hipError_t launch_cheb_fused(const ChebArgs<R>& a, hipStream_t st);
hipError_t launch_cheb_mixed(const MixedArgs& a, hipStream_t st);
// The power form of the mixed kernel on scaled fp16 two-way split products (mixed_h2.hip): full =
// 3 or 4 fp16 products per FP32-accurate product; a.ps must be set.
hipError_t launch_cheb_h2(const MixedArgs& a, int full, hipStream_t st);

// Small dims (d <= 16, ncoef <= 64, operators within 48 KB): the slices of a.j0 .. a.j0 + N - 1
// with the lowest small_fold_levels(N) tree levels folded in; a.out receives the level-F nodes
// (E = U - I), ceil(N / 2^F) of them.
bool small_fold_supported(int d, int ncoef, size_t elem);
int small_fold_levels(int N);
template <class R> hipError_t launch_small_fold(const ChebArgs<R>& a, hipStream_t st);
bool mixed_supported(int d);
int mixed_fused_levels(int N);   // tree levels the mixed slice kernel folds in

// Pairwise product tree (dims <= 64) on deviations from identity, one launch per
// block-reduction level.  leaves: N d x d matrices E_j = U_j - I, of cplx<R> or
// (leaves_are_c64) complex64; scratch0 / scratch1: tree_scratch_nodes<R>(N) matrices each;
// root receives the product U = I + E (not E).
template <class R> int tree_scratch_nodes(int N);
template <class R> hipError_t launch_tree(const void* leaves, bool leaves_are_c64, int N, int d, cplx<R>* scratch0,
                                          cplx<R>* scratch1, cplx<R>* root, hipStream_t st);

// Generic step-wise kernels (any dim).
template <class R> hipError_t launch_assemble(const ChebArgs<R>& a, cplx<R>* X, hipStream_t st);
// out_b = alpha * X_b + diag * I  (X may be null when alpha == 0)
template <class R> hipError_t launch_fill(const cplx<R>* X, cplx<R> alpha, cplx<R> diag, int d, int batch,
                                          cplx<R>* out, hipStream_t st);
// X_j assembled in FP64 and rounded to complex64 (operand of the low-precision GEMMs).
hipError_t launch_assemble_c64(const ChebArgs<double>& a, float2* X, hipStream_t st);
// n d x d complex64 matrices -> complex128.
hipError_t launch_widen(const float2* in, int d, int n, double2* out, hipStream_t st);
// n complex128 values -> complex64 (rounded to nearest).
hipError_t launch_narrow(const double2* in, long n, float2* out, hipStream_t st);
// E_j = (a_0 - 1) I + a_1 s X_j + R_j in FP64 (are[0] = a_0 - 1), stored as complex64.
hipError_t launch_first_order(const ChebArgs<double>& a, const float2* R, float2* E, hipStream_t st);

// Batched complex GEMM on the matrix cores (gemm.hip), any dim:
//     C_b = alpha A_b B_b + beta Cin_b + beta2 Cin2_b + diag I (+ A_b + B_b when add_ab),
// strides in matrices of d x d (0 broadcasts one matrix).  Element types: complex128 where the
// flag is set, complex64 otherwise (Cin and Cin2 have C's type; C may be Cin or Cin2 itself).  GEMM_F64 needs a complex128 output and
// takes complex128 or complex64 operands; the other modes run on complex64 throughout.
// GEMM_F16X2: scaled fp16 two-way split products (22 bits), for operands below 4 in modulus.
enum : int { GEMM_F64 = 0, GEMM_F32 = 1, GEMM_BF16X2 = 2, GEMM_BF16X3 = 3, GEMM_F16X2 = 4 };
struct GemmArgs {
    const void* A; long sa;
    const void* B; long sb;
    const void* Cin; long sc;
    void* C; long so;
    double alpha_re, alpha_im, beta_re, beta_im, diag_re, diag_im;
    const void* Cin2 = nullptr; long sc2 = 0;
    double beta2_re = 0.0, beta2_im = 0.0;
    // a second product added before the epilogue, C = alpha (A B + A2 B2) + ..., operands of A's
    // and B's types; FP64, FP32 and fp16 split modes (A2 and B2 both set, no add_ab)
    const void* A2 = nullptr; long sa2 = 0;
    const void* B2 = nullptr; long sb2 = 0;
    int d, batch, add_ab;
    int a_c128, b_c128, out_c128;
    int tile = 0;          // output tile 32 or 64, 16 = whole product per work-group (FP32, dims 33-128); 0 = automatic
};
hipError_t launch_zgemm(const GemmArgs& g, int mode, hipStream_t st);

// Ensemble kernels (ensemble.hip): per-instance parameters in a device table, task tables
// of (instance, slice) and (level descriptor, block); complex64.  Dim classes D = 16, 32, 48, 64
// (instances padded up), wave split T (TileShape<D, T>), tree block size 2^logb (2 or 3).
struct EnsInstance {
    const float2* ops;     // (1 + ncoef) d x d: H0, H_1..
    const float2* c;       // amps x pts controls
    float2* leaves;        // N x d x d, E_j = U_j - I
    int d, ncoef, amps, pts, mode, M;
    float s, tau;
    float are[kMaxCheb], aim[kMaxCheb];   // are[0] = a_0 - 1
    int ps;                               // slices in the power form (power_slice_tile)
    float pre[kMaxCheb], pim[kMaxCheb];   // its coefficients, pre[0] = m_0 - 1
};
struct EnsLevel {
    const float2* src;     // n_in nodes of one instance's tree level (E form)
    float2* dst;           // ceil(n_in / 2^logb) nodes; the instance's U (= I + E) when final_level
    int n_in, d, final_level;
};
bool ens_shape_valid(int D, int T);
hipError_t launch_ens_slices(int D, int T, const EnsInstance* inst, const int2* tasks, int ntasks, int grid,
                             hipStream_t st);
hipError_t launch_ens_tree(int D, int T, int logb, const EnsLevel* lv, const int2* tasks, int ntasks, int grid,
                           hipStream_t st);

// Adjoint gate-fidelity gradient (grad.hip), dims <= 32.  Tree levels are stored as
// consecutive d x d matrices; see grad.hip for the scan.
bool grad_supported(int d);
// Intermediate levels of a block reduction that the caller keeps: lev[k], n[k] = the array and
// node count of level k above the input (k = 1 .. LOGB - 1; level LOGB is the block's output).
template <class R> struct TreeStore {
    cplx<R>* lev[4];
    int n[4];
};
// up-sweep of logb (1..3) levels in one launch, every node kept as E = V - I (st: arrays and
// counts of the levels above the input, the last one included)
template <class R> hipError_t launch_up_block(const cplx<R>* in, int n_in, int d, int logb, const TreeStore<R>& st,
                                              hipStream_t s);
template <class R> hipError_t launch_down_level(const cplx<R>* V, int n_child, int d, const cplx<R>* Wp, cplx<R>* Wc,
                                                hipStream_t st);
// The root's two children (V: their deviations; W(root) = t^H, staged from the target t) and
// g = Tr(t^H (I + u)) from the root's deviation u, into g[0]: the root launch and the first
// down-sweep level in one.
template <class R> hipError_t launch_down_root(const cplx<R>* V, const cplx<R>* t, const cplx<R>* u, int d,
                                               cplx<R>* Wc, cplx<R>* g, hipStream_t st);
// With gd.grad set (no quadrature, no Magnus terms): the slice kernel maps its traces to the gradient
// itself, grad[t * pts + j] = 2 Re(conj(g) Tr(Q_j O_t)) / d^2 for t < amps with g = gd.g[0], and
// writes F = |g|^2 / d^2 to gd.fid (if set); dg is then unused.
template <class R> struct GradDirect {
    const cplx<R>* g = nullptr;
    R* grad = nullptr;
    R* fid = nullptr;
    int pts = 0, amps = 0;
};
template <class R> hipError_t launch_grad_slices(const ChebArgs<R>& a, const cplx<R>* E, const cplx<R>* W1,
                                                 cplx<R>* dg, int dg_stride, hipStream_t st,
                                                 GradDirect<R> gd = GradDirect<R>());
// out[t] = in[t]^T for t < count d x d matrices
template <class R> hipError_t launch_transpose(const cplx<R>* in, int count, int d, cplx<R>* out, hipStream_t st);
// dg[t * dg_stride + j0 + j] = Tr(Q_j O_t), O_t^T = opsT[t], for the n matrices Q_j (any d), times
// scale[0].x when scale (a device value) is given
template <class R> hipError_t launch_grad_trace(const cplx<R>* Q, int n, int d, const cplx<R>* opsT, int ncoef,
                                                cplx<R>* dg, int dg_stride, int j0, hipStream_t st,
                                                const cplx<R>* scale = nullptr);
// root costate wroot = t^H (times wscale[0].y when given) and g = Tr(t^H u), one launch
template <class R> hipError_t launch_grad_root(const cplx<R>* t, const cplx<R>* u, int d, cplx<R>* wroot, cplx<R>* g,
                                               hipStream_t st, const cplx<R>* wscale = nullptr);
// out[0] = (c, 1 / c), c = ||t||_F (1 for t = 0): the costate scale of the split-product gradient
template <class R> hipError_t launch_target_scale(const cplx<R>* t, int d, cplx<R>* out, hipStream_t st);
template <class R> hipError_t launch_grad_finish(const cplx<R>* dg, const cplx<R>* g, int N, int pts, int amps,
                                                 int mode, int d, R* grad, R* fid, hipStream_t st);
// the chain rule of the Magnus coefficients (commutator terms included) to the raw samples
template <class R> hipError_t launch_grad_finish_magnus(const cplx<R>* dg, const cplx<R>* g, const Controls<R>& q, int N,
                                                       int d, R* grad, R* fid, hipStream_t st);

}  // namespace pa
