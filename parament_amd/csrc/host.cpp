// Host orchestration and C ABI of libparament.so (MI355X-native equiprop).
//
// Mirrors the reference's context lifecycle and pipeline (parament.cpp:51-882):
// create -> setHamiltonian -> equiprop* -> destroy, with the same error codes,
// messages, iteration tables, Bessel coefficients and slice counts.  The device work is
// done by the kernels in kernels.hip / mixed.hip on one HIP stream per context.
#include <math.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <new>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/parament.h"
#include "kernels.hpp"

using pa::cplx;

namespace {

template <class R> struct ApiCplx;
template <> struct ApiCplx<float> { typedef cuComplex type; };
template <> struct ApiCplx<double> { typedef cuDoubleComplex type; };

// ----- host math (mathhelper.cpp:35-114, parament.cpp:723-766) -----

int select_fp32(double h, double dt) {
    static const double b[] = {0.032516793, 0.219062571, 0.619625593, 1.218059203, 1.979888284,
                               2.873301187, 3.872963682, 4.959398466, 6.117657121, 7.336154907,
                               8.605792444, 9.919320831, 11.27088616, 12.65570085};
    const double x = h * dt;
    for (int i = 0; i < 14; ++i)
        if (x <= b[i]) return 3 + 2 * i;
    return -1;
}

int select_fp64(double h, double dt) {
    static const double b[] = {0.000213616, 0.00768149, 0.0501474, 0.162592, 0.368382, 0.676861, 1.08784,
                               1.59605, 2.19402, 2.87366, 3.62716, 4.44725, 5.3274, 6.26179,
                               7.2453, 8.27338, 9.34206, 10.4478, 11.5875, 12.7584};
    const double x = h * dt;
    for (int i = 0; i < 20; ++i)
        if (x <= b[i]) return 3 + 2 * i;
    return -1;
}

// Max over rows of the summed moduli; single precision takes each modulus in float and
// sums in double, as the reference does (mathhelper.cpp:82-114).
double one_norm(const cuComplex* m, unsigned d) {
    double best = 0;
    for (unsigned i = 0; i < d; ++i) {
        double s = 0;
        for (unsigned j = 0; j < d; ++j) s += (double)hypotf(m[(size_t)d * i + j].x, m[(size_t)d * i + j].y);
        best = std::max(best, s);
    }
    return best;
}
double one_norm(const cuDoubleComplex* m, unsigned d) {
    double best = 0;
    for (unsigned i = 0; i < d; ++i) {
        double s = 0;
        for (unsigned j = 0; j < d; ++j) s += hypot(m[(size_t)d * i + j].x, m[(size_t)d * i + j].y);
        best = std::max(best, s);
    }
    return best;
}

// a_k = (-i)^k J_k(x) in FP64 (glibc jn, as the reference).
void bessel(int m, double x, double* re, double* im) {
    for (int k = 0; k <= m; ++k) {
        const double j = jn(k, x);
        switch (k % 4) {
            case 0: re[k] = j; im[k] = 0; break;
            case 1: re[k] = 0; im[k] = -j; break;
            case 2: re[k] = -j; im[k] = 0; break;
            default: re[k] = 0; im[k] = j; break;
        }
    }
}

// Power-form coefficients of the Chebyshev approximant: with y = X / Hnorm and the series
// p(y) = a_0 + 2 sum_{k>=1} a_k T_k(y) that the Clenshaw recurrence evaluates, returns m_k with
// p = sum_k m_k Y^k for Y = tau X = x y (x = tau Hnorm), in FP64; mr[0] holds m_0 - 1.
void power_coefficients(int m, double x, const double* are, const double* aim, double* mr, double* mi) {
    // monomial coefficients of T_{k-1}, T_k, T_{k+1} (m < kMaxCheb)
    double t0[pa::kMaxCheb + 1] = {}, t1[pa::kMaxCheb + 1] = {}, t2[pa::kMaxCheb + 1] = {};
    double cr[pa::kMaxCheb + 1] = {}, ci[pa::kMaxCheb + 1] = {};
    double *tprev = t0, *tcur = t1, *tnext = t2;
    tprev[0] = 1.0;   // T_0
    tcur[1] = 1.0;    // T_1
    cr[0] += are[0]; ci[0] += aim[0];
    for (int k = 1; k <= m; ++k) {
        for (int i = 0; i <= k; ++i) { cr[i] += 2.0 * are[k] * tcur[i]; ci[i] += 2.0 * aim[k] * tcur[i]; }
        if (k == m) break;
        for (int i = 0; i <= k + 1; ++i) tnext[i] = 0.0;   // T_{k+1} = 2 y T_k - T_{k-1}
        for (int i = 0; i <= k; ++i) tnext[i + 1] += 2.0 * tcur[i];
        for (int i = 0; i <= k - 1; ++i) tnext[i] -= tprev[i];
        double* t = tprev;
        tprev = tcur;
        tcur = tnext;
        tnext = t;
    }
    double xp = 1.0;
    for (int k = 0; k <= m; ++k, xp *= x) { mr[k] = cr[k] / xp; mi[k] = ci[k] / xp; }
    mr[0] -= 1.0;
}

// Split level of each inner product Y^2 Q_{j+1} of the power form (j = 1 .. (M-1)/2 - 1), two
// bits per j.  The product enters U with weight <= x^(2j+2) / (2j+2)! (|Y| <= x = tau Hnorm
// on the series' domain, Q_{j+1} ~ m_{2j+2}); a level dropping the slice pairs p + q > L has
// relative error ~2^-8(L+1), and the level is the lowest whose error stays below 1e-8 -- an
// unbiased rounding-size error, far under the FP32 rounding of the stored slice itself.
int power_split_levels(int m, double x) {
    int bits = 0;
    double w = x * x / 2.0;   // x^(2j+2) / (2j+2)! for j = 0
    for (int j = 1; j < (m - 1) / 2; ++j) {
        w *= x * x / ((2.0 * j + 1.0) * (2.0 * j + 2.0));
        const int lv = (w * 0x1p-8 <= 1e-8) ? 0 : (w * 0x1p-16 <= 1e-8 ? 1 : 2);
        bits |= lv << (2 * j);
    }
    return bits;
}

// The same for the fp16 two-way split form: a product on the leading parts only has relative error
// ~2^-11; bits 2j = 0 select it, 1 the full product.
int power_split_levels_h2(int m, double x) {
    int bits = 0;
    double w = x * x / 2.0;
    for (int j = 1; j < (m - 1) / 2; ++j) {
        w *= x * x / ((2.0 * j + 1.0) * (2.0 * j + 2.0));
        bits |= (w * 0x1p-11 <= 1e-8 ? 0 : 1) << (2 * j);
    }
    return bits;
}

// Split scheme of the mixed power form: 3 (default) or 4 = scaled fp16 two-way split with three or
// four products per FP32-accurate product (mixed_h2.hip; the batched GEMMs take three in both), 0 =
// bf16 three-way split with six (mixed.hip, and the batched GEMMs' bf16 mode).
// PARAMENT_MIXED_SPLIT = f16x3 | f16x4 | bf16x3 selects one, for comparisons.
int mixed_split_scheme() {
    static const int scheme = [] {
        const char* v = getenv("PARAMENT_MIXED_SPLIT");
        if (v && !strcmp(v, "bf16x3")) return 0;
        if (v && !strcmp(v, "f16x4")) return 4;
        return 3;
    }();
    return scheme;
}

// Chebyshev order of the mixed-precision path.  The FP32 table bounds the per-slice tail
// 2 sum_{k>M} |J_k(x)| at ~0.05 eps32; that truncation error is the same for every slice and
// adds up over the product, so the mixed path (whose target is 1e-5 against FP64) also
// requires N * tail(M) <= 2e-6, raising M by 2 until it holds (never below the FP32 choice).
double cheb_tail(int m, double x) {
    double t = 0;
    for (int k = m + 1; k <= m + 40; ++k) t += fabs(jn(k, x));
    return 2 * t;
}
int mixed_order(int m_fp32, double x, int n_slices) {
    int m = m_fp32;
    while (m + 2 < 48 && (double)n_slices * cheb_tail(m, x) > 2e-6) m += 2;
    return m;
}

// ----- context -----

// Everything a gradient call passes to its kernels, pointers and values; zero-filled before
// use so that byte comparison is exact.
struct GradKey {
    const void *carr, *target, *fid, *grad, *ops, *E, *V, *W, *dg, *aux;
    double dt, hnorm;
    unsigned pts, amps, amps_set, dim, seg, chunk;
    int quad, ncoef, M, device, precision;
    GradKey() { memset(this, 0, sizeof(*this)); }
    bool operator==(const GradKey& o) const { return memcmp(this, &o, sizeof(*this)) == 0; }
};

// Everything an equiprop call passes to its kernels: the arguments, the settings that choose M,
// the kernels and the coefficients, and the addresses of the operators and of every working
// buffer.  Zero-filled before use so that byte comparison is exact.
struct PropKey {
    const void *carr, *out, *ops, *ops64, *leaves, *nodes, *gen0, *gen1, *gen2, *c64, *out64, *stream;
    double dt, hnorm;
    unsigned pts, amps, amps_set, dim;
    int quad, magnus, ncoef, precision, path, manual, mmax, device;
    PropKey() { memset(this, 0, sizeof(*this)); }
    bool operator==(const PropKey& o) const { return memcmp(this, &o, sizeof(*this)) == 0; }
};

template <class R> struct Context {
    typedef cplx<R> C;
    Parament_ErrorCode last_error = PARAMENT_STATUS_SUCCESS;
    // options (defaults of parament.cpp:72-75)
    int mmax = 11;
    bool manual = false;
    int last_m = -1;
    // complex128 contexts compute in FP64 unless a mixed mode is selected; complex64 contexts
    // default to the mixed pipeline (FP64 first-order terms and upper tree levels)
    int precision = sizeof(R) == 4 ? PARAMENT_PRECISION_MIXED : PARAMENT_PRECISION_NATIVE;
    int path = 0;
    // Hamiltonian
    bool has_h = false;
    unsigned dim = 0, amps = 0;
    bool magnus = false;
    int quad = PARAMENT_QUADRATURE_SIMPSON;
    double hnorm = 0;
    int ncoef = 0;
    // device
    bool bound = false;
    int device = -1;
    hipStream_t stream = nullptr;      // stream in use (own_stream unless the caller set one)
    hipStream_t own_stream = nullptr;
    bool user_stream = false;          // setStream was called: `stream` is the caller's (NULL = null stream)
    bool timing = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    C* d_ops = nullptr;
    // complex64 contexts: the operators (Magnus commutators included) in FP64, for the mixed pipeline
    double2* d_ops64 = nullptr;
    // working memory (grow-only, like parament.cpp:418)
    C* d_c = nullptr; size_t cap_c = 0;
    void* d_leaves = nullptr; size_t cap_leaves = 0;
    void* d_nodes = nullptr; size_t cap_nodes = 0;
    C* d_gen[4] = {nullptr, nullptr, nullptr, nullptr}; size_t cap_gen = 0; int n_gen = 0;
    void* d_root = nullptr; size_t cap_root = 0;  // d x d result of the last equiprop
    void* d_c64 = nullptr; size_t cap_c64 = 0;    // complex64 contexts: controls widened to FP64
    void* d_out64 = nullptr; size_t cap_out64 = 0;  // complex64 contexts: FP64 result before rounding
    // gradient: leaves E, tree levels V, costates W, per-slice traces, trace g; tree layout policy
    void* d_gE = nullptr; size_t cap_gE = 0;
    void* d_gV = nullptr; size_t cap_gV = 0;
    void* d_gW = nullptr; size_t cap_gW = 0;
    void* d_gdg = nullptr; size_t cap_gdg = 0;
    void* d_gaux = nullptr; size_t cap_gaux = 0;
    unsigned ckpt_seg = 0, ckpt_chunk = 0;
    size_t grad_workspace = 0;  // device bytes the last gradient call worked in
    // gradient launch sequence captured as a graph once two consecutive calls agree on every
    // argument that reaches a kernel (GradKey), replayed while they keep agreeing
    GradKey grad_key, grad_prev;
    hipGraphExec_t grad_exec = nullptr;
    // the same for equiprop: repeated identical calls replay a captured graph of the launches
    PropKey prop_key, prop_prev;
    hipGraphExec_t prop_exec = nullptr;
    hipStream_t cap_stream = nullptr;
};

template <class R> void free_hamiltonian(Context<R>* h) {
    if (h->d_ops) hipFree(h->d_ops);
    h->d_ops = nullptr;
    if (h->d_ops64) hipFree(h->d_ops64);
    h->d_ops64 = nullptr;
    h->has_h = false;
}

template <class R> void free_working(Context<R>* h) {
    hipFree(h->d_c); h->d_c = nullptr; h->cap_c = 0;
    hipFree(h->d_leaves); h->d_leaves = nullptr; h->cap_leaves = 0;
    hipFree(h->d_nodes); h->d_nodes = nullptr; h->cap_nodes = 0;
    for (auto& p : h->d_gen) { hipFree(p); p = nullptr; }
    h->cap_gen = 0; h->n_gen = 0;
    hipFree(h->d_root); h->d_root = nullptr; h->cap_root = 0;
    hipFree(h->d_c64); h->d_c64 = nullptr; h->cap_c64 = 0;
    hipFree(h->d_out64); h->d_out64 = nullptr; h->cap_out64 = 0;
    for (void** p : {&h->d_gE, &h->d_gV, &h->d_gW, &h->d_gdg, &h->d_gaux}) { hipFree(*p); *p = nullptr; }
    h->cap_gE = h->cap_gV = h->cap_gW = h->cap_gdg = h->cap_gaux = 0;
}

// Grow a device buffer to at least `bytes` (contents are not preserved).
bool grow(void** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes && *p) return true;
    if (*p) hipFree(*p);
    *p = nullptr; *cap = 0;
    if (hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; hipGetLastError(); return false; }
    *cap = bytes;
    return true;
}

// Device working memory of one equiprop call.  By default every buffer is one of the context's
// grow-only allocations; with `fixed` set, buffers are consecutive 256-byte aligned pieces of
// that caller-provided region, and a request beyond its end fails instead of allocating.
struct Workspace {
    char* fixed = nullptr;
    size_t cap = 0, used = 0;
    // `bytes` of device memory: the context slot (grown when needed) or the next fixed piece
    void* get(void** slot, size_t* slot_cap, size_t bytes) {
        if (!fixed) return grow(slot, slot_cap, bytes) ? *slot : nullptr;
        const size_t at = (used + 255) & ~(size_t)255;
        if (bytes > cap || at > cap - bytes) return nullptr;
        used = at + bytes;
        return fixed + at;
    }
    Parament_ErrorCode failure() const {
        return fixed ? PARAMENT_STATUS_WORKSPACE_TOO_SMALL : PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
    }
};

// Batched complex GEMM on the matrix cores (gemm.hip): C = alpha A B + beta Cin + beta2 Cin2 +
// diag I (+ A + B with add_ab); pointers with their element type (true = complex128), strides in
// matrices.
struct GemmOp { const void* p; bool c128; long stride; };
hipError_t zgemm(int mode, GemmOp A, GemmOp B, GemmOp Cin, void* C, bool c_c128, long so, double2 alpha, double2 beta,
                 double2 diag, int d, int batch, hipStream_t st, bool add_ab = false, int tile = 0,
                 GemmOp Cin2 = GemmOp{nullptr, false, 0}, double2 beta2 = double2{0.0, 0.0}) {
    pa::GemmArgs g;
    g.tile = tile;
    g.A = A.p; g.sa = A.stride; g.B = B.p; g.sb = B.stride; g.Cin = Cin.p; g.sc = Cin.stride; g.C = C; g.so = so;
    g.alpha_re = alpha.x; g.alpha_im = alpha.y; g.beta_re = beta.x; g.beta_im = beta.y;
    g.Cin2 = Cin2.p; g.sc2 = Cin2.stride; g.beta2_re = beta2.x; g.beta2_im = beta2.y;
    g.diag_re = diag.x; g.diag_im = diag.y;
    g.d = d; g.batch = batch; g.add_ab = add_ab ? 1 : 0;
    g.a_c128 = A.c128; g.b_c128 = B.c128; g.out_c128 = c_c128;
    return pa::launch_zgemm(g, mode, st);
}
// The same in the context's own precision (FP64 products for complex128, FP32 for complex64).
template <class R>
hipError_t zgemm_native(const cplx<R>* A, long sa, const cplx<R>* B, long sb, const cplx<R>* Cin, long sc, cplx<R>* C,
                        long so, double2 alpha, double2 beta, double2 diag, int d, int batch, hipStream_t st,
                        bool add_ab = false, int tile = 0) {
    const bool w = sizeof(R) == 8;
    return zgemm(w ? pa::GEMM_F64 : pa::GEMM_F32, GemmOp{A, w, sa}, GemmOp{B, w, sb}, GemmOp{Cin, w, sc}, C, w, so, alpha,
                 beta, diag, d, batch, st, add_ab, tile);
}
// C = alpha (A B + A2 B2) + beta Cin + beta2 Cin2 + diag I: two products summed in one launch
// (FP64, FP32 and fp16 split modes; the context's element type, unit matrix strides over the batch).
template <class R>
hipError_t zgemm_sum(int mode, const cplx<R>* A, const cplx<R>* B, const cplx<R>* A2, const cplx<R>* B2,
                     const cplx<R>* Cin, double2 beta, const cplx<R>* Cin2, double2 beta2, cplx<R>* C, double2 alpha,
                     double2 diag, int d, int batch, hipStream_t st) {
    const int w = sizeof(R) == 8;
    pa::GemmArgs g;
    g.A = A; g.sa = 1; g.B = B; g.sb = 1; g.A2 = A2; g.sa2 = 1; g.B2 = B2; g.sb2 = 1;
    g.Cin = Cin; g.sc = Cin ? 1 : 0; g.Cin2 = Cin2; g.sc2 = Cin2 ? 1 : 0; g.C = C; g.so = 1;
    g.alpha_re = alpha.x; g.alpha_im = alpha.y; g.beta_re = beta.x; g.beta_im = beta.y;
    g.beta2_re = beta2.x; g.beta2_im = beta2.y; g.diag_re = diag.x; g.diag_im = diag.y;
    g.d = d; g.batch = batch; g.add_ab = 0; g.a_c128 = g.b_c128 = g.out_c128 = w;
    return pa::launch_zgemm(g, mode, st);
}
inline double2 c2(double re, double im = 0.0) { return pa::mk<double>(re, im); }
inline double2 cmul2(double2 x, double2 y) { return c2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x); }
inline double2 cscale(double f, double2 x) { return c2(f * x.x, f * x.y); }

// Slices E_j = p(X_j) - I of a batch of cnt slices by the Clenshaw recurrence on batched GEMMs
// (parament.cpp:569-652): b_k = s X b_{k+1} - b_{k+2} + a_k I, E = s X b_1 - 2 b_2 + (a_0 - 1) I,
// every step's diagonal and scaling in the GEMM epilogue.  The start values b_M = a_M I and
// b_{M-1} = s a_M X + a_{M-1} I are affine in X, so the first two steps take them through X as
// C_in instead of materialising them:
//   b_{M-2} = s^2 a_M X X + s a_{M-1} X + (a_{M-2} - a_M) I,
//   b_{M-3} = s X b_{M-2} + beta s a_M X + (a_{M-3} + beta a_{M-1}) I   (beta = -2 when M = 3).
// a[k] = a_k (a[0] = a_0 - 1); M >= 3.  w0, w1: cnt slices each.  The result goes to out, or,
// when out is null, stays in w0 or w1 (*result says which).
template <class R>
hipError_t clenshaw_batched(const cplx<R>* X, int d, int cnt, int M, double s, const double2* a, cplx<R>* w0,
                            cplx<R>* w1, cplx<R>* out, cplx<R>** result, hipStream_t st, int tile = 0) {
    const int k3 = M - 3;
    const double beta3 = k3 == 0 ? -2.0 : -1.0;
    hipError_t r = zgemm_native<R>(X, 1, X, 1, X, 1, w0, 1, cscale(s * s, a[M]), cscale(s, a[M - 1]),
                                   c2(a[M - 2].x - a[M].x, a[M - 2].y - a[M].y), d, cnt, st, false, tile);
    cplx<R>* dst = (k3 == 0 && out) ? out : w1;
    if (r == hipSuccess)
        r = zgemm_native<R>(X, 1, w0, 1, X, 1, dst, 1, c2(s), cscale(beta3 * s, a[M]),
                            c2(a[k3].x + beta3 * a[M - 1].x, a[k3].y + beta3 * a[M - 1].y), d, cnt, st, false, tile);
    cplx<R>* b1 = dst;
    cplx<R>* b2 = w0;
    for (int k = M - 4; r == hipSuccess && k >= 0; --k) {
        cplx<R>* o = (k == 0 && out) ? out : b2;
        r = zgemm_native<R>(X, 1, b1, 1, b2, 1, o, 1, c2(s), c2(k == 0 ? -2.0 : -1.0), a[k], d, cnt, st, false, tile);
        b2 = b1;
        b1 = o;
    }
    if (result) *result = b1;
    return r;
}

// Slices E_j = p(X_j) - I of a batch of cnt slices in the power form on batched GEMMs, for
// tau Hnorm <= 1: p = sum_k m_k Y^k, Y = tau X (m from power_coefficients, m[0] = m_0 - 1), by
// Horner's rule in S = Y^2 with J = (M - 1) / 2:
//   B_{J-1} = m_{2J+1} S Y + m_{2J} S + m_{2J-1} Y + m_{2J-2} I,
//   B_j     = S B_{j+1} + m_{2j+1} Y + m_{2j} I   (j = J-2 .. 0),   E = B_0,
// (M + 1) / 2 GEMMs against the recurrence's M - 1, every Y term and diagonal in the FP64 GEMM
// epilogue.  The polynomial is the one the recurrence evaluates; only the rounding differs.
// With first_order false the last step leaves out m_1 Y + (m_0 - 1) I, which the caller adds in
// FP64.  Buffers of cnt slices: S and w0, w1 as well when M >= 7.  The last step writes out, or,
// when out is null, w0 (M = 3) or X (M >= 5, read in place).  *result: the buffer that holds E.
// keep (optional, J buffers of cnt slices): S goes to keep[0] and each B_j (j >= 1) to keep[j]
// instead, so that a caller can read them back (S, w0 and w1 are then unused).
template <class R>
hipError_t power_batched(int mode, cplx<R>* X, int d, int cnt, int M, double tau, const double2* m, bool first_order,
                         cplx<R>* S, cplx<R>* w0, cplx<R>* w1, cplx<R>* out, cplx<R>** result, hipStream_t st,
                         int tile = 0, cplx<R>* const* keep = nullptr) {
    const bool w = sizeof(R) == 8;
    const int J = (M - 1) / 2;
    if (J < 1 || (J >= 3 && !w1 && !keep)) return hipErrorInvalidValue;
    if (keep) S = keep[0];
    auto op = [&](const void* p) { return GemmOp{p, w, 1}; };
    const GemmOp none{nullptr, w, 0};
    hipError_t r = zgemm(mode, op(X), op(X), none, S, w, 1, c2(tau * tau), c2(0), c2(0), d, cnt, st, false, tile);
    cplx<R>* cur = (J == 1 && out) ? out : (keep && J > 1 ? keep[J - 1] : w0);
    {
        const int j = J - 1;
        const bool lin = j > 0 || first_order;   // the step's own Y term and diagonal
        if (r == hipSuccess)
            r = zgemm(mode, op(S), op(X), op(S), cur, w, 1, cscale(tau, m[2 * J + 1]), m[2 * J], lin ? m[2 * j] : c2(0), d,
                      cnt, st, false, tile, lin ? op(X) : none, lin ? cscale(tau, m[2 * j + 1]) : c2(0));
    }
    for (int j = J - 2; j >= 0 && r == hipSuccess; --j) {
        const bool lin = j > 0 || first_order;
        cplx<R>* o = j == 0 ? (out ? out : X) : (keep ? keep[j] : (cur == w0 ? w1 : w0));
        r = zgemm(mode, op(S), op(cur), lin ? op(X) : none, o, w, 1, c2(1), lin ? cscale(tau, m[2 * j + 1]) : c2(0),
                  lin ? m[2 * j] : c2(0), d, cnt, st, false, tile);
        cur = o;
    }
    if (result) *result = cur;
    return r;
}

// Whether the power form applies at x = tau Hnorm: x <= 1 (the power coefficients are then close
// to (-i)^k / k! and the form is well conditioned) and x^M a normal double (power_coefficients
// divides by x^k; a manual M far above the table's at a tiny x would underflow it).
inline bool power_form_ok(int M, double x) { return x <= 1.0 && std::pow(x, M) >= 1e-300; }
// The step-wise complex64 paths take it from M = 5 on (at M = 3 both forms take two GEMMs).
inline bool stepwise_power(int M, double x) { return M >= 5 && power_form_ok(M, x); }

template <class R> Parament_ErrorCode bind_device(Context<R>* h) {
    if (h->bound) {
        if (hipSetDevice(h->device) != hipSuccess) return PARAMENT_FAIL;
        return PARAMENT_STATUS_SUCCESS;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { hipGetLastError(); return PARAMENT_FAIL; }
    if (hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking) != hipSuccess) {
        hipGetLastError();
        return PARAMENT_STATUS_CUBLAS_INIT_FAILED;
    }
    if (!h->user_stream) h->stream = h->own_stream;
    h->device = dev;
    h->bound = true;
    return PARAMENT_STATUS_SUCCESS;
}

template <class R> Parament_ErrorCode create(Context<R>** hp) {
    if (!hp) return PARAMENT_STATUS_INVALID_VALUE;
    *hp = new (std::nothrow) Context<R>();
    if (!*hp) return PARAMENT_STATUS_HOST_ALLOC_FAILED;
    return PARAMENT_STATUS_SUCCESS;
}

template <class R> Parament_ErrorCode destroy(Context<R>* h) {
    if (!h) return PARAMENT_STATUS_SUCCESS;
    if (h->bound) {
        hipSetDevice(h->device);
        hipStreamSynchronize(h->stream);
        hipStreamSynchronize(h->own_stream);
        free_hamiltonian(h);
        free_working(h);
        for (auto& e : h->ev)
            if (e) hipEventDestroy(e);
        if (h->grad_exec) hipGraphExecDestroy(h->grad_exec);
        if (h->prop_exec) hipGraphExecDestroy(h->prop_exec);
        if (h->cap_stream) hipStreamDestroy(h->cap_stream);
        hipStreamDestroy(h->own_stream);
    }
    delete h;
    return PARAMENT_STATUS_SUCCESS;
}

// Magnus commutators on the device: [H0, H_k] then [H_a, H_b] (a < b, lexicographic),
// each as A B - B A with the generic GEMM (parament.cpp:289-359).
template <class R> hipError_t magnus_commutators(cplx<R>* ops, unsigned d, unsigned A, hipStream_t st) {
    typedef cplx<R> C;
    const size_t dd = (size_t)d * d;
    auto comm = [&](const C* X, const C* Y, C* out) -> hipError_t {
        hipError_t e = zgemm_native<R>(X, 0, Y, 0, nullptr, 0, out, 0, c2(1), c2(0), c2(0), d, 1, st);
        if (e != hipSuccess) return e;
        return zgemm_native<R>(Y, 0, X, 0, out, 0, out, 0, c2(-1), c2(1), c2(0), d, 1, st);
    };
    for (unsigned k = 0; k < A; ++k) {
        hipError_t e = comm(ops, ops + (1 + k) * dd, ops + (1 + A + k) * dd);
        if (e != hipSuccess) return e;
    }
    unsigned slot = 0;
    for (unsigned a = 0; a < A; ++a)
        for (unsigned b = a + 1; b < A; ++b, ++slot) {
            hipError_t e = comm(ops + (1 + a) * dd, ops + (1 + b) * dd, ops + (1 + 2 * A + slot) * dd);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

template <class R>
Parament_ErrorCode set_hamiltonian(Context<R>* h, const typename ApiCplx<R>::type* H0, const typename ApiCplx<R>::type* H1,
                                   unsigned dim, unsigned amps, bool use_magnus, int quad) {
    if (!h) return PARAMENT_STATUS_INVALID_VALUE;
    if (h->bound) hipSetDevice(h->device);
    free_hamiltonian(h);  // the old Hamiltonian is dropped even when this call fails (parament.cpp:216)
    if (use_magnus && quad != PARAMENT_QUADRATURE_SIMPSON) return h->last_error = PARAMENT_STATUS_INVALID_QUADRATURE_SELECTION;
    if (quad != PARAMENT_QUADRATURE_NONE && quad != PARAMENT_QUADRATURE_MIDPOINT && quad != PARAMENT_QUADRATURE_SIMPSON)
        return h->last_error = PARAMENT_STATUS_INVALID_VALUE;
    if (!H0 || !H1 || dim == 0 || amps == 0) return h->last_error = PARAMENT_STATUS_INVALID_VALUE;
    const int ncoef = pa::n_coeffs((int)amps, use_magnus ? 1 : 0);
    if (ncoef > pa::kMaxCoeffs) return h->last_error = PARAMENT_STATUS_INVALID_VALUE;

    Parament_ErrorCode err = bind_device(h);
    if (err != PARAMENT_STATUS_SUCCESS) return h->last_error = err;

    const size_t dd = (size_t)dim * dim;
    const size_t nops = 1 + (size_t)ncoef;
    // H0 and the ncoef operators, then the operators transposed (read by the gradient's traces)
    if (hipMalloc(&h->d_ops, (nops + ncoef) * dd * sizeof(cplx<R>)) != hipSuccess) {
        hipGetLastError();
        h->d_ops = nullptr;
        return h->last_error = PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
    }
    if (hipMemcpyAsync(h->d_ops, H0, dd * sizeof(cplx<R>), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(h->d_ops + dd, H1, amps * dd * sizeof(cplx<R>), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
        free_hamiltonian(h);
        return h->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    }
    h->dim = dim;
    h->amps = amps;
    h->magnus = use_magnus;
    h->quad = use_magnus ? PARAMENT_QUADRATURE_SIMPSON : quad;
    h->ncoef = ncoef;
    h->hnorm = one_norm(H0, dim);
    for (unsigned k = 0; k < amps; ++k) h->hnorm += one_norm(H1 + k * dd, dim);
    if (use_magnus && magnus_commutators<R>(h->d_ops, dim, amps, h->stream) != hipSuccess) {
        free_hamiltonian(h);
        return h->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    }
    if constexpr (sizeof(R) == 4) {
        // FP64 copy for the mixed pipeline: the complex64 inputs widen exactly; the Magnus
        // commutators are formed again in FP64
        std::vector<double2> wide((1 + (size_t)amps) * dd);
        for (size_t i = 0; i < dd; ++i) wide[i] = c2(H0[i].x, H0[i].y);
        for (size_t i = 0; i < amps * dd; ++i) wide[dd + i] = c2(H1[i].x, H1[i].y);
        if (hipMalloc(&h->d_ops64, nops * dd * sizeof(double2)) != hipSuccess) {
            hipGetLastError();
            h->d_ops64 = nullptr;
            free_hamiltonian(h);
            return h->last_error = PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
        }
        if (hipMemcpyAsync(h->d_ops64, wide.data(), wide.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream) !=
                hipSuccess ||
            (use_magnus && magnus_commutators<double>(h->d_ops64, dim, amps, h->stream) != hipSuccess) ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            free_hamiltonian(h);
            return h->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
        }
    }
    if (pa::launch_transpose<R>(h->d_ops + dd, ncoef, (int)dim, h->d_ops + nops * dd, h->stream) != hipSuccess) {
        free_hamiltonian(h);
        return h->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        free_hamiltonian(h);
        return h->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    }
    if (!grow(&h->d_root, &h->cap_root, dd * sizeof(cplx<R>))) {
        free_hamiltonian(h);
        return h->last_error = PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
    }
    h->has_h = true;
    return h->last_error = PARAMENT_STATUS_SUCCESS;
}

// Slice count and slice length per mode (parament.cpp:800-802, 820-831).
template <class R> void slices(const Context<R>* h, unsigned pts, double dt, int* N, double* tau) {
    const bool doubled = h->magnus || h->quad == PARAMENT_QUADRATURE_SIMPSON;
    *tau = doubled ? 2 * dt : dt;
    if (doubled) *N = pts >= 1 ? (int)((pts - 1) / 2) : 0;
    else if (h->quad == PARAMENT_QUADRATURE_MIDPOINT) *N = pts >= 1 ? (int)(pts - 1) : 0;
    else *N = (int)pts;
}

// Timing marks: 0 before the slice kernel, 1 between slice kernel and tree, 2 after the tree.
template <class R> void mark(Context<R>* h, int i) {
    if (!h->timing) return;
    if (!h->ev[i] && hipEventCreate(&h->ev[i]) != hipSuccess) { hipGetLastError(); h->ev[i] = nullptr; return; }
    hipEventRecord(h->ev[i], h->stream);
}

// Mixed-precision pipeline.  FP64 for Hnorm, the Bessel coefficients, the slice Hamiltonians,
// the first-order part (a_0 - 1) I + a_1 s X_j and the tree levels above the lowest three; the
// Chebyshev remainder and the lowest tree levels in FP32 arithmetic whose products are bf16
// three-way split products (FP32 accurate), FP32 products or bf16 hi+lo products.  Inputs are
// FP64 operators and controls: a complex64 context passes widened copies of its own.
struct MixedSrc {
    const double2* ops;
    const double2* carr;
    int N, M;
    double tau;
    unsigned pts, amps;
};

int gemm_mode_of(int precision) {
    switch (precision) {
        case PARAMENT_PRECISION_MIXED_FP32: return pa::GEMM_F32;
        case PARAMENT_PRECISION_MIXED_BF16X2: return pa::GEMM_BF16X2;
        default: return pa::GEMM_BF16X3;
    }
}

constexpr int kLowTreeLevels = 3;   // tree levels run in the low-precision products

// The three step-wise working arrays (X, b_k, b_{k+1}) of `bytes` each: the context's, grown
// together, or three pieces of a fixed workspace.
// n (3 or 4) step-wise working buffers of bytes each, grow-only.
template <class R> bool gen_buffers(Context<R>* h, Workspace& ws, size_t bytes, void* out[], int n = 3) {
    if (ws.fixed) {
        for (int i = 0; i < n; ++i)
            if (!(out[i] = ws.get(nullptr, nullptr, bytes))) return false;
        return true;
    }
    if (h->cap_gen < bytes || h->n_gen < n) {
        for (auto& p : h->d_gen) { hipFree(p); p = nullptr; }
        h->cap_gen = 0; h->n_gen = 0;
        for (int i = 0; i < n; ++i)
            if (hipMalloc(&h->d_gen[i], bytes) != hipSuccess) {
                hipGetLastError();
                for (auto& q : h->d_gen) { hipFree(q); q = nullptr; }
                return false;
            }
        h->cap_gen = bytes; h->n_gen = n;
    }
    for (int i = 0; i < n; ++i) out[i] = h->d_gen[i];
    return true;
}

// Step-wise form (any dim): the Clenshaw GEMMs (or the power form's) batched over all slices on
// complex64 data with the products of the context's mode, every GEMM epilogue in FP64.  In the
// default mode the power form runs on scaled fp16 two-way split products (its operands Y = tau X,
// S = Y^2, the Horner brackets and the tree's deviations stay below 4 in modulus while
// ||tau X_j||_2 <= 1, as in the fused kernel), the recurrence on bf16 three-way splits.
template <class R>
Parament_ErrorCode run_mixed_stepwise(Context<R>* h, const MixedSrc& m, double2* out_dev, Workspace& ws) {
    const int d = (int)h->dim, N = m.N, M = m.M;
    const size_t dd = (size_t)d * d;
    const bool power = stepwise_power(M, m.tau * h->hnorm);
    const bool f16 = power && h->precision == PARAMENT_PRECISION_MIXED && mixed_split_scheme() != 0;
    const int mode = f16 ? pa::GEMM_F16X2 : gemm_mode_of(h->precision);
    hipStream_t st = h->stream;
    void* gen[4] = {nullptr, nullptr, nullptr, nullptr};
    if (!gen_buffers(h, ws, (size_t)N * dd * sizeof(float2), gen, power && M >= 7 ? 4 : 3)) return ws.failure();
    double are[pa::kMaxCheb], aim[pa::kMaxCheb];
    bessel(M, m.tau * h->hnorm, are, aim);
    pa::ChebArgs<double> ca;
    ca.ops = m.ops; ca.ncoef = h->ncoef; ca.d = d; ca.N = N; ca.M = M; ca.s = 2.0 / h->hnorm;
    ca.ctl.c = m.carr; ca.ctl.pts = (int)m.pts; ca.ctl.amps_in = (int)m.amps; ca.ctl.amps_set = (int)h->amps;
    ca.ctl.mode = h->quad; ca.ctl.magnus = h->magnus ? 1 : 0; ca.ctl.tau = m.tau;
    for (int k = 0; k <= M; ++k) { ca.are[k] = are[k]; ca.aim[k] = aim[k]; }
    ca.are[0] = are[0] - 1.0;
    ca.out = nullptr;
    float2* X = reinterpret_cast<float2*>(gen[0]);
    float2* b1 = reinterpret_cast<float2*>(gen[1]);
    float2* b2 = reinterpret_cast<float2*>(gen[2]);
    float2* rd = b1;   // the slices E_j
    float2* wr = X;    // the other half of the tree's ping-pong
    mark(h, 0);
    // the fp16 split takes Y = tau X itself (entries below 1 while ||tau X||_2 <= 1), so the buffer
    // holds Y and the power form runs with tau = 1
    pa::ChebArgs<double> xa = ca;
    if (f16) xa.asm_scale = m.tau;
    if (pa::launch_assemble_c64(xa, X, st) != hipSuccess) return PARAMENT_STATUS_CUBLAS_FAILED;
    if (power) {
        // R = E - m_1 Y - (m_0 - 1) I by the power form (into X), then E = R + m_1 Y + (m_0 - 1) I
        // in FP64, in place
        double mr[pa::kMaxCheb], mi[pa::kMaxCheb];
        double2 mc[pa::kMaxCheb];
        power_coefficients(M, m.tau * h->hnorm, are, aim, mr, mi);
        for (int k = 0; k <= M; ++k) mc[k] = c2(mr[k], mi[k]);
        if (power_batched<float>(mode, X, d, N, M, f16 ? 1.0 : m.tau, mc, false, b1, b2, reinterpret_cast<float2*>(gen[3]),
                                 nullptr, &rd, st) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        pa::ChebArgs<double> fa = ca;
        fa.s = m.tau;
        fa.are[0] = mr[0]; fa.aim[0] = mi[0]; fa.are[1] = mr[1]; fa.aim[1] = mi[1];
        if (pa::launch_first_order(fa, rd, rd, st) != hipSuccess) return PARAMENT_STATUS_CUBLAS_FAILED;
        wr = (rd == X) ? b1 : X;
    } else {
        const float2 aM = make_float2((float)are[M], (float)aim[M]);
        const float2 saM = make_float2((float)(ca.s * are[M]), (float)(ca.s * aim[M]));
        if (pa::launch_fill<float>(nullptr, make_float2(0, 0), aM, d, N, b2, st) != hipSuccess ||
            pa::launch_fill<float>(X, saM, make_float2((float)are[M - 1], (float)aim[M - 1]), d, N, b1, st) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        // b_k = s X b_{k+1} - b_{k+2} + a_k I, k = M-2 .. 2, then b_1' (no a_1 I)
        for (int k = M - 2; k >= 1; --k) {
            const double2 diag = (k >= 2) ? c2(are[k], aim[k]) : c2(0);
            if (zgemm(mode, GemmOp{X, false, 1}, GemmOp{b1, false, 1}, GemmOp{b2, false, 1}, b2, false, 1, c2(ca.s), c2(-1),
                      diag, d, N, st) != hipSuccess)
                return PARAMENT_STATUS_CUBLAS_FAILED;
            std::swap(b1, b2);
        }
        // R = s X b_1' - 2 b_2 (into b2), then E = (a_0 - 1) I + a_1 s X + R in FP64 (into b1)
        if (zgemm(mode, GemmOp{X, false, 1}, GemmOp{b1, false, 1}, GemmOp{b2, false, 1}, b2, false, 1, c2(ca.s), c2(-2),
                  c2(0), d, N, st) != hipSuccess ||
            pa::launch_first_order(ca, b2, b1, st) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        rd = b1;
        wr = X;
    }
    mark(h, 1);
    // lowest tree levels on complex64 (E' = A + B + A B), ping-pong rd <-> wr
    int r = N;
    for (int l = 0; l < kLowTreeLevels && r > 1; ++l) {
        const int half = r / 2;
        if (zgemm(mode, GemmOp{rd + dd, false, 2}, GemmOp{rd, false, 2}, GemmOp{nullptr, false, 0}, wr, false, 1, c2(1), c2(0),
                  c2(0), d, half, st, true) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        if (r % 2 && hipMemcpyAsync(wr + half * dd, rd + (size_t)(r - 1) * dd, dd * sizeof(float2), hipMemcpyDeviceToDevice,
                                    st) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        r = half + (r % 2);
        std::swap(rd, wr);
    }
    // FP64 levels above: the first one reads the complex64 nodes
    const size_t n64 = (size_t)(r + 1) / 2;
    void* nodes = ws.get(&h->d_nodes, &h->cap_nodes, 2 * std::max<size_t>(n64, 1) * dd * sizeof(double2));
    if (!nodes) return ws.failure();
    double2* cur = reinterpret_cast<double2*>(nodes);
    double2* nxt = cur + std::max<size_t>(n64, 1) * dd;
    if (r == 1) {
        if (pa::launch_widen(rd, d, 1, cur, st) != hipSuccess) return PARAMENT_STATUS_CUBLAS_FAILED;
    } else {
        const int half = r / 2;
        if (zgemm(pa::GEMM_F64, GemmOp{rd + dd, false, 2}, GemmOp{rd, false, 2}, GemmOp{nullptr, true, 0}, cur, true, 1, c2(1),
                  c2(0), c2(0), d, half, st, true) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        if (r % 2 && pa::launch_widen(rd + (size_t)(r - 1) * dd, d, 1, cur + half * dd, st) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        r = half + (r % 2);
        while (r > 1) {
            const int hf = r / 2;
            if (zgemm(pa::GEMM_F64, GemmOp{cur + dd, true, 2}, GemmOp{cur, true, 2}, GemmOp{nullptr, true, 0}, nxt, true, 1,
                      c2(1), c2(0), c2(0), d, hf, st, true) != hipSuccess)
                return PARAMENT_STATUS_CUBLAS_FAILED;
            if (r % 2 && hipMemcpyAsync(nxt + hf * dd, cur + (size_t)(r - 1) * dd, dd * sizeof(double2),
                                        hipMemcpyDeviceToDevice, st) != hipSuccess)
                return PARAMENT_STATUS_CUBLAS_FAILED;
            r = hf + (r % 2);
            std::swap(cur, nxt);
        }
    }
    if (pa::launch_fill<double>(cur, c2(1), c2(1), d, 1, out_dev, st) != hipSuccess) return PARAMENT_STATUS_CUBLAS_FAILED;
    mark(h, 2);
    return PARAMENT_STATUS_SUCCESS;
}

// Fused form (dims <= 64): one kernel forms each slice on chip and folds the lowest tree levels;
// the FP64 tree takes its complex64 nodes.
template <class R>
Parament_ErrorCode run_mixed_fused(Context<R>* h, const MixedSrc& m, double2* out_dev, Workspace& ws) {
    const unsigned d = h->dim;
    const size_t dd = (size_t)d * d;
    const int N = m.N, M = m.M;
    pa::MixedArgs ma;
    ma.ops = m.ops;
    ma.ncoef = h->ncoef; ma.d = (int)d;
    ma.ctl.c = m.carr; ma.ctl.pts = (int)m.pts; ma.ctl.amps_in = (int)m.amps;
    ma.ctl.amps_set = (int)h->amps; ma.ctl.mode = h->quad; ma.ctl.magnus = h->magnus ? 1 : 0; ma.ctl.tau = m.tau;
    // The low-precision recurrence multiplies by s in FP32: take s = fl32(2 / Hnorm) and expand
    // with Hnorm' = 2 / s (within 6e-8 of Hnorm) so that every term uses the same, exact s.
    const float sf = (float)(2.0 / h->hnorm);
    const double hn2 = 2.0 / (double)sf;
    double bre[pa::kMaxCheb], bim[pa::kMaxCheb];
    bessel(M, m.tau * hn2, bre, bim);
    ma.N = N; ma.M = M; ma.s = (double)sf;
    for (int k = 0; k <= M; ++k) { ma.are[k] = bre[k]; ma.aim[k] = bim[k]; }
    ma.are[0] = bre[0] - 1.0;
    // Small tau Hnorm: evaluate the same degree-M polynomial in powers of Y = tau X, which
    // needs (M + 1) / 2 products instead of M - 1 (the power coefficients are then close to
    // (-i)^k / k! and the form is well conditioned).
    const double xarg = m.tau * hn2;
    ma.ps = power_form_ok(M, xarg) ? 1 : 0;
    if (ma.ps) {
        power_coefficients(M, xarg, bre, bim, ma.are, ma.aim);
        ma.ps_levels = power_split_levels(M, xarg);
    }
    // the slice kernel folds the lowest F tree levels; the FP64 tree takes its level-F nodes
    const int F = pa::mixed_fused_levels(N);
    const int n_nodes = (N + (1 << F) - 1) >> F;
    const size_t nodes = (size_t)pa::tree_scratch_nodes<double>(n_nodes);
    void* leaves = ws.get(&h->d_leaves, &h->cap_leaves, (size_t)n_nodes * dd * sizeof(float2));
    void* scratch_nodes = leaves ? ws.get(&h->d_nodes, &h->cap_nodes, 2 * nodes * dd * sizeof(double2)) : nullptr;
    if (!scratch_nodes) return ws.failure();
    ma.out = reinterpret_cast<float2*>(leaves);
    const int split = mixed_split_scheme();
    if (ma.ps && split) ma.ps_levels = power_split_levels_h2(M, xarg);
    mark(h, 0);
    if ((ma.ps && split ? pa::launch_cheb_h2(ma, split, h->stream) : pa::launch_cheb_mixed(ma, h->stream)) != hipSuccess)
        return PARAMENT_STATUS_CUBLAS_FAILED;
    mark(h, 1);
    double2* nd = reinterpret_cast<double2*>(scratch_nodes);
    if (pa::launch_tree<double>(leaves, true, n_nodes, (int)d, nd, nd + nodes * dd, out_dev, h->stream) != hipSuccess)
        return PARAMENT_STATUS_CUBLAS_FAILED;
    mark(h, 2);
    return PARAMENT_STATUS_SUCCESS;
}

// Whether a call runs the mixed pipeline: a mixed mode on a complex128 context; on a complex64
// context every mode but NATIVE, except that the default mode keeps the wavefront-resident FP32
// kernels for dims <= 16 on the automatic path.
template <class R> bool uses_mixed(const Context<R>* h) {
    if (h->precision == PARAMENT_PRECISION_NATIVE) return false;
    if (sizeof(R) == 4 && h->precision == PARAMENT_PRECISION_MIXED && h->dim <= 16 && h->path == 0) return false;
    return true;
}

// Runs the whole pipeline on the context's stream; carr_dev / out_dev are device pointers.
template <class R>
Parament_ErrorCode run(Context<R>* h, const cplx<R>* carr_dev, double dt, unsigned pts, unsigned amps, cplx<R>* out_dev,
                       Workspace& ws) {
    typedef cplx<R> C;
    int N = 0;
    double tau = 0;
    slices(h, pts, dt, &N, &tau);
    if (N <= 0) return PARAMENT_STATUS_INVALID_VALUE;

    const bool mixed = uses_mixed(h);
    int M = h->mmax;
    if (!h->manual) {
        M = (sizeof(R) == 4 || mixed) ? select_fp32(h->hnorm, tau) : select_fp64(h->hnorm, tau);
        if (M < 3) return PARAMENT_STATUS_SELECT_SMALLER_DT;
        // a complex128 context in a mixed mode targets 1e-5 against FP64: the truncation error,
        // the same in every slice, must stay below that over the whole product; a complex64
        // context keeps the FP32 table's order
        if (sizeof(R) == 8 && mixed) M = mixed_order(M, tau * h->hnorm, N);
    }
    if (M >= pa::kMaxCheb) return PARAMENT_STATUS_INVALID_VALUE;
    h->last_m = M;

    const unsigned d = h->dim;
    const size_t dd = (size_t)d * d;
    hipStream_t st = h->stream;

    if (mixed) {
        MixedSrc m;
        m.N = N; m.M = M; m.tau = tau; m.pts = pts; m.amps = amps;
        double2* out64 = nullptr;
        if constexpr (sizeof(R) == 8) {
            m.ops = h->d_ops;
            m.carr = carr_dev;
            out64 = out_dev;
        } else {
            const size_t n = (size_t)amps * pts;
            void* c64 = ws.get(&h->d_c64, &h->cap_c64, std::max<size_t>(n, 1) * sizeof(double2));
            void* o64 = c64 ? ws.get(&h->d_out64, &h->cap_out64, dd * sizeof(double2)) : nullptr;
            if (!o64) return ws.failure();
            if (pa::launch_widen(carr_dev, 1, (int)n, reinterpret_cast<double2*>(c64), st) != hipSuccess)
                return PARAMENT_STATUS_CUBLAS_FAILED;
            m.ops = h->d_ops64;
            m.carr = reinterpret_cast<const double2*>(c64);
            out64 = reinterpret_cast<double2*>(o64);
        }
        const bool fused = h->precision == PARAMENT_PRECISION_MIXED && d <= 64 && h->path == 0 && pa::mixed_supported((int)d);
        const Parament_ErrorCode err = fused ? run_mixed_fused(h, m, out64, ws) : run_mixed_stepwise(h, m, out64, ws);
        if (err != PARAMENT_STATUS_SUCCESS) return err;
        if constexpr (sizeof(R) == 4) {
            if (pa::launch_narrow(out64, (long)dd, out_dev, st) != hipSuccess) return PARAMENT_STATUS_CUBLAS_FAILED;
        }
        return PARAMENT_STATUS_SUCCESS;
    }

    double are[pa::kMaxCheb], aim[pa::kMaxCheb];
    bessel(M, tau * h->hnorm, are, aim);
    // Slices are carried as E_j = U_j - I: the last Clenshaw step adds (a_0 - 1) I, formed in FP64.
    const double a0m1 = are[0] - 1.0;
    const bool fused = (d <= 64) && h->path == 0;

    pa::ChebArgs<R> ca;
    ca.ops = h->d_ops; ca.ncoef = h->ncoef; ca.d = (int)d; ca.N = N; ca.M = M;
    ca.ctl.c = carr_dev; ca.ctl.pts = (int)pts; ca.ctl.amps_in = (int)amps; ca.ctl.amps_set = (int)h->amps;
    ca.ctl.mode = h->quad; ca.ctl.magnus = h->magnus ? 1 : 0; ca.ctl.tau = (R)tau;
    ca.s = (R)(2.0 / h->hnorm);  // the reference casts s and a_k to the working precision (parament.cpp:583-584, 406)
    for (int k = 0; k <= M; ++k) { ca.are[k] = (R)are[k]; ca.aim[k] = (R)aim[k]; }
    ca.are[0] = (R)a0m1;

    if (fused) {
        // dims <= 16: wavefront-resident slices with the lowest tree levels folded in
        const bool small = pa::small_fold_supported((int)d, h->ncoef, sizeof(C));
        // complex64 slices in the power form where the step-wise path takes it (complex128 keeps
        // the reference's recurrence)
        if (!small && sizeof(R) == 4 && stepwise_power(M, tau * h->hnorm)) {
            double mr[pa::kMaxCheb], mi[pa::kMaxCheb];
            power_coefficients(M, tau * h->hnorm, are, aim, mr, mi);
            ca.ps = 1;
            for (int k = 0; k <= M; ++k) { ca.pre[k] = (R)mr[k]; ca.pim[k] = (R)mi[k]; }
        }
        const int F = small ? pa::small_fold_levels(N) : 0;
        const int n_nodes = (N + (1 << F) - 1) >> F;
        const size_t nodes = (size_t)pa::tree_scratch_nodes<R>(n_nodes);
        void* leaves = ws.get(&h->d_leaves, &h->cap_leaves, (size_t)n_nodes * dd * sizeof(C));
        void* scratch_nodes = leaves ? ws.get(&h->d_nodes, &h->cap_nodes, 2 * nodes * dd * sizeof(C)) : nullptr;
        if (!scratch_nodes) return ws.failure();
        ca.out = reinterpret_cast<C*>(leaves);
        mark(h, 0);
        if ((small ? pa::launch_small_fold<R>(ca, st) : pa::launch_cheb_fused<R>(ca, st)) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        mark(h, 1);
        C* nd = reinterpret_cast<C*>(scratch_nodes);
        if (pa::launch_tree<R>(leaves, false, n_nodes, (int)d, nd, nd + nodes * dd, out_dev, st) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        mark(h, 2);
        return PARAMENT_STATUS_SUCCESS;
    }

    // Step-wise path: X, then M-1 batched GEMM launches (complex64: (M+1)/2 in the power form when
    // tau Hnorm <= 1; complex128 keeps the reference's recurrence), then ceil(log2 N) tree levels.
    const bool power = sizeof(R) == 4 && stepwise_power(M, tau * h->hnorm);
    void* gen[4] = {nullptr, nullptr, nullptr, nullptr};
    if (!gen_buffers(h, ws, (size_t)N * dd * sizeof(C), gen, power && M >= 7 ? 4 : 3)) return ws.failure();
    C* X = reinterpret_cast<C*>(gen[0]);
    C* w1 = reinterpret_cast<C*>(gen[1]);
    C* w2 = reinterpret_cast<C*>(gen[2]);
    mark(h, 0);
    if (pa::launch_assemble<R>(ca, X, st) != hipSuccess) return PARAMENT_STATUS_CUBLAS_FAILED;
    C* leaves = nullptr;
    if (power) {
        double mr[pa::kMaxCheb], mi[pa::kMaxCheb];
        double2 mc[pa::kMaxCheb];
        power_coefficients(M, tau * h->hnorm, are, aim, mr, mi);
        for (int k = 0; k <= M; ++k) mc[k] = c2(mr[k], mi[k]);
        if (power_batched<R>(pa::GEMM_F32, X, (int)d, N, M, tau, mc, true, w1, w2, reinterpret_cast<C*>(gen[3]), nullptr,
                             &leaves, st) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
    } else {
        // the slices as E = U - I (a_0 - 1 on the diagonal of the last step)
        double2 ak[pa::kMaxCheb];
        for (int k = 0; k <= M; ++k) ak[k] = c2(k == 0 ? a0m1 : (double)ca.are[k], (double)ca.aim[k]);
        if (clenshaw_batched<R>(X, (int)d, N, M, (double)ca.s, ak, w1, w2, nullptr, &leaves, st) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
    }
    mark(h, 1);
    // E_j = U_j - I for the N slices; reduce level by level (parament.cpp:675-714) on deviations
    // from identity: E' = A + B + A B.
    C* rd = leaves;
    C* wr = leaves == w1 ? w2 : w1;
    int r = N;
    while (r > 1) {
        const int half = r / 2;
        if (zgemm_native<R>(rd + dd, 2, rd, 2, nullptr, 0, wr, 1, c2(1), c2(0), c2(0), (int)d, half, st, true) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
        if (r % 2) {
            if (hipMemcpyAsync(wr + half * dd, rd + (size_t)(r - 1) * dd, dd * sizeof(C), hipMemcpyDeviceToDevice, st) !=
                hipSuccess)
                return PARAMENT_STATUS_CUBLAS_FAILED;
        }
        r = half + (r % 2);
        std::swap(rd, wr);
    }
    const C one = pa::mk<R>(1, 0);
    if (pa::launch_fill<R>(rd, one, one, (int)d, 1, out_dev, st) != hipSuccess)  // U = I + E
        return PARAMENT_STATUS_CUBLAS_FAILED;
    mark(h, 2);
    return PARAMENT_STATUS_SUCCESS;
}

// run() on the context's own buffers.  A call whose key (PropKey) equals the previous call's is
// captured as a HIP graph on a capture stream and replayed by later calls with that key: the
// launches and their arguments are then issued with one graph launch (small problems are bound by
// that issue cost).  The key holds every buffer address, so a call that grew a buffer in between
// never replays a graph with stale addresses.  Timing mode always launches directly.
template <class R>
Parament_ErrorCode run_replayable(Context<R>* h, const cplx<R>* carr_dev, double dt, unsigned pts, unsigned amps,
                                  cplx<R>* out_dev) {
    Workspace ws;
    if (h->timing) return run(h, carr_dev, dt, pts, amps, out_dev, ws);
    PropKey key;
    key.carr = carr_dev; key.out = out_dev; key.ops = h->d_ops; key.ops64 = h->d_ops64;
    key.leaves = h->d_leaves; key.nodes = h->d_nodes; key.gen0 = h->d_gen[0]; key.gen1 = h->d_gen[1];
    key.gen2 = h->d_gen[2]; key.c64 = h->d_c64; key.out64 = h->d_out64; key.stream = h->stream;
    key.dt = dt; key.hnorm = h->hnorm; key.pts = pts; key.amps = amps; key.amps_set = h->amps; key.dim = h->dim;
    key.quad = h->quad; key.magnus = h->magnus ? 1 : 0; key.ncoef = h->ncoef; key.precision = h->precision;
    key.path = h->path; key.manual = h->manual ? 1 : 0; key.mmax = h->mmax; key.device = h->device;
    if (h->prop_exec && key == h->prop_key)
        return hipGraphLaunch(h->prop_exec, h->stream) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
    if (!(key == h->prop_prev)) {   // first call with this key: launch directly
        h->prop_prev = key;
        return run(h, carr_dev, dt, pts, amps, out_dev, ws);
    }
    if (!h->cap_stream && hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking) != hipSuccess) {
        hipGetLastError();
        h->cap_stream = nullptr;
        return run(h, carr_dev, dt, pts, amps, out_dev, ws);
    }
    if (h->prop_exec) { hipGraphExecDestroy(h->prop_exec); h->prop_exec = nullptr; }
    hipGraph_t g = nullptr;
    hipStream_t own = h->stream;
    h->stream = h->cap_stream;
    Parament_ErrorCode err = PARAMENT_STATUS_CUBLAS_FAILED;
    if (hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        err = run(h, carr_dev, dt, pts, amps, out_dev, ws);
        const hipError_t r = hipStreamEndCapture(h->cap_stream, &g);
        if (err == PARAMENT_STATUS_SUCCESS && r == hipSuccess && g &&
            hipGraphInstantiate(&h->prop_exec, g, nullptr, nullptr, 0) != hipSuccess)
            h->prop_exec = nullptr;
        if (g) hipGraphDestroy(g);
    }
    hipGetLastError();
    h->stream = own;
    if (err != PARAMENT_STATUS_SUCCESS || !h->prop_exec) {   // capture unavailable: launch directly
        h->prop_exec = nullptr;
        Workspace direct;
        return run(h, carr_dev, dt, pts, amps, out_dev, direct);
    }
    h->prop_key = key;
    return hipGraphLaunch(h->prop_exec, h->stream) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
}

template <class R> Parament_ErrorCode ensure_common(Context<R>* h, unsigned amps) {
    if (!h->has_h) return PARAMENT_STATUS_NO_HAMILTONIAN;
    if (amps > h->amps || amps == 0) return PARAMENT_STATUS_INVALID_VALUE;
    if (hipSetDevice(h->device) != hipSuccess) return PARAMENT_FAIL;
    return PARAMENT_STATUS_SUCCESS;
}

template <class R>
Parament_ErrorCode equiprop(Context<R>* h, const typename ApiCplx<R>::type* carr, double dt, unsigned pts, unsigned amps,
                            typename ApiCplx<R>::type* out) {
    typedef cplx<R> C;
    if (!h) return PARAMENT_STATUS_INVALID_VALUE;
    if (!h->has_h) return h->last_error = PARAMENT_STATUS_NO_HAMILTONIAN;
    if (!carr || !out) return h->last_error = PARAMENT_STATUS_INVALID_VALUE;
    Parament_ErrorCode err = ensure_common(h, amps);
    if (err != PARAMENT_STATUS_SUCCESS) return h->last_error = err;
    const size_t dd = (size_t)h->dim * h->dim;
    const size_t cbytes = (size_t)amps * pts * sizeof(C);
    if (!grow(reinterpret_cast<void**>(&h->d_c), &h->cap_c, std::max<size_t>(cbytes, sizeof(C))))
        return h->last_error = PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
    if (hipMemcpyAsync(h->d_c, carr, cbytes, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        return h->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    err = run_replayable(h, h->d_c, dt, pts, amps, reinterpret_cast<C*>(h->d_root));
    if (err == PARAMENT_STATUS_SUCCESS) {
        if (hipMemcpyAsync(out, h->d_root, dd * sizeof(C), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess)
            err = PARAMENT_STATUS_CUBLAS_FAILED;
    } else {
        hipStreamSynchronize(h->stream);
    }
    return h->last_error = err;
}

template <class R>
Parament_ErrorCode equiprop_device(Context<R>* h, const typename ApiCplx<R>::type* carr_dev, double dt, unsigned pts,
                                   unsigned amps, typename ApiCplx<R>::type* out_dev) {
    if (!h) return PARAMENT_STATUS_INVALID_VALUE;
    if (!h->has_h) return h->last_error = PARAMENT_STATUS_NO_HAMILTONIAN;
    if (!carr_dev || !out_dev) return h->last_error = PARAMENT_STATUS_INVALID_VALUE;
    Parament_ErrorCode err = ensure_common(h, amps);
    if (err != PARAMENT_STATUS_SUCCESS) return h->last_error = err;
    err = run_replayable(h, reinterpret_cast<const cplx<R>*>(carr_dev), dt, pts, amps, reinterpret_cast<cplx<R>*>(out_dev));
    return h->last_error = err;
}


// ----- allocation-free equiprop (no counterpart in the reference) -----
// For callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc): the working memory of a call is carved from one fixed
// device array instead of the context's grow-only buffers.  Each GPU the library runs on holds
// its own copy of the array until the process ends; all calls on one GPU share it, so calls that
// overlap on one GPU race.  No kernel reads workspace memory it has not written in the same call,
// so nothing needs clearing between calls.
constexpr size_t kFixedWorkspaceBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kFixedWorkspaceBytes];

// Device address of `scratch` on device `dev`, which must be current: looked up on the first
// call on each device and kept (null when the lookup fails).
char* fixed_workspace(int dev) {
    static std::mutex mu;
    static char* addr[kMaxDevices] = {};
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            hipGetLastError();
            return nullptr;
        }
        addr[dev] = static_cast<char*>(p);
    }
    return addr[dev];
}

template <class R>
Parament_ErrorCode equiprop_device_no_alloc(Context<R>* h, const cplx<R>* carr_dev, double dt, unsigned pts, unsigned amps,
                                            cplx<R>* out_dev) {
    if (!h) return PARAMENT_STATUS_INVALID_VALUE;
    if (!h->has_h) return h->last_error = PARAMENT_STATUS_NO_HAMILTONIAN;
    if (!carr_dev || !out_dev) return h->last_error = PARAMENT_STATUS_INVALID_VALUE;
    Parament_ErrorCode err = ensure_common(h, amps);   // makes the context's device current
    if (err != PARAMENT_STATUS_SUCCESS) return h->last_error = err;
    char* base = fixed_workspace(h->device);
    if (!base) return h->last_error = PARAMENT_FAIL;
    Workspace ws;
    ws.fixed = base;
    ws.cap = kFixedWorkspaceBytes;
    // kernel timing creates its events on first use: none are recorded by this entry point
    const bool timing = h->timing;
    h->timing = false;
    err = run(h, carr_dev, dt, pts, amps, out_dev, ws);
    h->timing = timing;
    return h->last_error = err;
}

// ----- adjoint gate-fidelity gradient (grad.hip) -----
// Layout of one gradient call over the pairwise tree.  Level l has n[l] = ceil(N / 2^l) nodes,
// level `top` (>= 1) is the root.  Levels seg .. top are kept for the whole call (the "top"
// arrays: V = node products, W = costates); levels 1 .. seg-1 live in chunk arrays that are
// reused by consecutive chunks of `chunk` slices (a multiple of 2^seg).  seg = top with one
// chunk keeps the whole tree; seg ~ log2(sqrt N) is the checkpointed layout: the first pass
// stores only the level-seg products, the second pass recomputes each chunk's slices.
struct GradPlan {
    int N = 0, top = 1, seg = 1, chunk = 1, nchunks = 1;
    std::vector<int> n;
    std::vector<size_t> coff, toff;   // matrix offsets of chunk levels (1..seg-1) and top levels (seg..top)
    size_t chunk_mats = 0, top_mats = 0;
};

GradPlan grad_plan(int N, unsigned seg_req, unsigned chunk_req) {
    GradPlan p;
    p.N = N;
    while ((1L << p.top) < N) ++p.top;
    for (int l = 0; l <= p.top; ++l) p.n.push_back((int)(((long)N + (1L << l) - 1) >> l));
    int seg = p.top;
    if (seg_req == 1) {
        int b = 1;
        while ((1L << (2 * b)) < N) ++b;   // 2^b >= sqrt(N)
        seg = std::min(b, p.top);
    } else if (seg_req >= 2) {
        int b = 1;
        while ((1L << b) < (long)seg_req) ++b;
        seg = std::min(b, p.top);
    }
    const long B = 1L << seg;
    long chunk = chunk_req ? (long)chunk_req : std::max<long>(B, 2048);
    chunk = (chunk + B - 1) / B * B;
    if (chunk >= N) { seg = p.top; chunk = N; }
    p.seg = seg;
    p.chunk = (int)chunk;
    p.nchunks = (int)((N + chunk - 1) / chunk);
    p.coff.assign(p.top + 1, 0);
    p.toff.assign(p.top + 1, 0);
    for (int l = 1; l < seg; ++l) { p.coff[l] = p.chunk_mats; p.chunk_mats += (size_t)((chunk + (1L << l) - 1) >> l); }
    for (int l = seg; l <= p.top; ++l) { p.toff[l] = p.top_mats; p.top_mats += (size_t)p.n[l]; }
    return p;
}

// The chain rule from the per-slice coefficient derivatives dg[t][j] to the raw samples: the
// quadrature weights, or with Magnus terms the derivatives of their coefficients.
template <class R>
hipError_t grad_finish(const pa::ChebArgs<R>& ca, const cplx<R>* dg, const cplx<R>* gsum, int N, R* grad, R* fid,
                       hipStream_t st) {
    if (ca.ctl.magnus) return pa::launch_grad_finish_magnus<R>(dg, gsum, ca.ctl, N, ca.d, grad, fid, st);
    return pa::launch_grad_finish<R>(dg, gsum, N, ca.ctl.pts, ca.ctl.amps_in, ca.ctl.mode, ca.d, grad, fid, st);
}

// Step-wise gradient (any dim, batched GEMMs).  Level l of the pairwise tree has n[l] nodes
// (level 0: the N slices, kept as E = U - I in the leaf array); V holds levels 1 .. top and W the
// costates of levels 0 .. top, all kept.  The nwork per-slice arrays (X and the recurrence's
// or the power form's buffers) cover `chunk` slices at a time.
struct StepGradPlan {
    int N = 0, top = 0, chunk = 1, nwork = 5;
    int keep = 0;   // matrices per slice kept from the slices' pass for the derivative (S, Q_1 .. Q_{J-1})
    bool split = false;   // complex64 power form: every product on scaled fp16 two-way splits
    std::vector<int> n;
    std::vector<size_t> voff, woff;   // matrix offsets of level l in V (l >= 1) and in W
    size_t vmats = 0, wmats = 0;
};

// Per-slice arrays of the step-wise gradient: X and four buffers (S, dS, Q, dQ in the power form),
// a fifth for the power form's Horner levels when M >= 7.
inline int step_grad_work(int M, bool power) { return (power && M >= 7) ? 6 : 5; }
// The power form's S and Horner brackets Q_1 .. Q_{J-1} of every slice, kept from the slices' pass
// so that the derivative need not form them again (J of its products per slice fewer): J matrices
// per slice, when they take at most max_bytes.
inline int step_grad_keep(int M, bool power, int N, int d, size_t elem_bytes, size_t max_bytes) {
    const int J = (M - 1) / 2;
    if (!power || M < 5) return 0;
    return (size_t)J * N * d * d * elem_bytes <= max_bytes ? J : 0;
}

StepGradPlan step_grad_plan(int N, int d, size_t elem_bytes, unsigned chunk_req = 0, int nwork = 5) {
    StepGradPlan p;
    p.N = N;
    p.nwork = nwork;
    while ((1L << p.top) < N) ++p.top;
    p.voff.assign(p.top + 1, 0);
    p.woff.assign(p.top + 1, 0);
    for (int l = 0; l <= p.top; ++l) {
        p.n.push_back((int)(((long)N + (1L << l) - 1) >> l));
        if (l >= 1) { p.voff[l] = p.vmats; p.vmats += (size_t)p.n[l]; }
        p.woff[l] = p.wmats; p.wmats += (size_t)p.n[l];
    }
    // the per-slice arrays within 4 GiB
    const size_t per_slice = (size_t)nwork * d * d * elem_bytes;
    size_t chunk = std::max<size_t>(1, std::min<size_t>((size_t)N, ((size_t)4 << 30) / per_slice));
    if (chunk_req) chunk = std::min<size_t>(chunk, chunk_req);
    p.chunk = (int)chunk;
    return p;
}

// Buffers of one step-wise gradient: E (N slices), work (nwork x chunk slices), keep (plan.keep x N
// slices, when kept), V (vmats), W (wmats), dg (ncoef x N values), gsum (one value).
template <class R> struct StepGradBufs {
    cplx<R> *E = nullptr, *work = nullptr, *keep = nullptr, *V = nullptr, *W = nullptr, *dg = nullptr, *gsum = nullptr;
};

// The adjoint gradient on batched GEMMs: slices E_j by the Clenshaw recurrence (as equiprop's
// step-wise path), the up-sweep keeping every node (E' = A + B + A B), the root costate t^H, the
// down-sweep (earlier child W V(sibling), later child V(sibling) W, a lone child W), then per
// slice the derivative of the same recurrence in the direction of the slice's costate,
//   db_k = s (W b_{k+1} + X db_{k+1}) - db_{k+2},   Q = s (W b_1 + X db_1) - 2 db_2,
// so that Tr(W dU_j) = Tr(Q_j dX_j) (the polynomial's Frechet derivative is self-adjoint under
// the trace pairing), and dg[t][j] = Tr(Q_j O_t).  When ca.ps is set and M >= 5 the slices and the
// derivative take the power form instead (power_batched; the derivative of its Horner rule as in
// the fused gradient kernel).  ca carries the controls, M, s and the coefficients with
// are[0] = a_0 - 1 (pre/pim the power coefficients, pre[0] = m_0 - 1); ca.opsT the transposed
// control operators.
// With p.split (complex64, power form) every product runs on scaled fp16 two-way splits, whose
// operands must stay below 4 in modulus: the slices are assembled as Y = tau X (||Y||_2 <= 1, so
// S, the brackets and the deviations E = U - I stay below 4 as in the forward), the costates run
// from t^H / ||t||_F (entries of every costate below 1; the traces take the factor back), and
// the derivative is taken in the direction W itself (dS = Y W + W Y below 2, the dQ_j below e),
// its last level scaled by tau.
template <class R>
hipError_t grad_stepwise_enqueue(pa::ChebArgs<R> ca, const StepGradPlan& p, const StepGradBufs<R>& b,
                                 const cplx<R>* target, R* fid, R* grad, hipStream_t st) {
    typedef cplx<R> C;
    const int d = ca.d, M = ca.M, N = p.N;
    const size_t dd = (size_t)d * d;
    const C zero = pa::mk<R>(0, 0), one = pa::mk<R>(1, 0);
    const double s = (double)ca.s;
    const bool w64 = sizeof(R) == 8;
    const bool power = ca.ps && M >= 5;
    const bool split = !w64 && power && p.split;
    const int mode = w64 ? pa::GEMM_F64 : (split ? pa::GEMM_F16X2 : pa::GEMM_F32);
    if (power && p.nwork < step_grad_work(M, true)) return hipErrorInvalidValue;
    const double tau = (double)ca.ctl.tau;
    // Y = ty X_buf (the slice buffer), the derivative's direction tv W, its last level times fin
    const double ty = split ? 1.0 : tau, tv = split ? 1.0 : tau, fin = split ? tau : 1.0;
    if (split) ca.asm_scale = (R)tau;
    const C* wscale = split ? b.gsum + 1 : nullptr;
    // C = alpha A B + beta Cin + diag I (+ A + B) over cnt products with the given matrix strides
    auto gm = [&](const C* A, long sa, const C* B, long sb, const C* Cin, long sc, C* out, long so, double2 alpha,
                  double2 beta, double2 diag, int cnt, bool add_ab) {
        return zgemm(mode, GemmOp{A, w64, sa}, GemmOp{B, w64, sb}, GemmOp{Cin, w64, sc}, out, w64, so, alpha, beta, diag, d,
                     cnt, st, add_ab);
    };
    double2 a[pa::kMaxCheb], m[pa::kMaxCheb];
    for (int k = 0; k <= M; ++k) {
        a[k] = c2((double)ca.are[k], (double)ca.aim[k]);
        m[k] = c2((double)ca.pre[k], (double)ca.pim[k]);
    }
    C* X = b.work;
    C* w[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int i = 0; i + 1 < p.nwork; ++i) w[i] = b.work + (size_t)(1 + i) * p.chunk * dd;
    // the kept S and Q_j of slices j0 .. : keep[0] = S, keep[j] = Q_j
    const bool kept = power && p.keep == (M - 1) / 2 && b.keep;
    auto keep_at = [&](int j0, C** kq) {
        for (int k = 0; k < p.keep; ++k) kq[k] = b.keep + ((size_t)k * N + j0) * dd;
    };
    hipError_t r = hipSuccess;
    // slices
    for (int j0 = 0; r == hipSuccess && j0 < N; j0 += p.chunk) {
        const int cnt = std::min(p.chunk, N - j0);
        ca.j0 = j0; ca.N = cnt;
        r = pa::launch_assemble<R>(ca, X, st);
        if (r != hipSuccess) break;
        C* kq[pa::kMaxCheb / 2 + 1];
        if (kept) keep_at(j0, kq);
        if (power)
            r = power_batched<R>(mode, X, d, cnt, M, ty, m, true, w[0], w[1], w[2], b.E + (size_t)j0 * dd, nullptr, st, 0,
                                 kept ? kq : nullptr);
        else r = clenshaw_batched<R>(X, d, cnt, M, s, a, w[0], w[1], b.E + (size_t)j0 * dd, nullptr, st);
    }
    // up-sweep
    const C* rd = b.E;
    for (int l = 1; r == hipSuccess && l <= p.top; ++l) {
        const int n = p.n[l - 1], half = n / 2;
        C* wr = b.V + p.voff[l] * dd;
        if (half) r = gm(rd + dd, 2, rd, 2, nullptr, 0, wr, 1, c2(1), c2(0), c2(0), half, true);
        if (r == hipSuccess && (n & 1)) r = pa::launch_fill<R>(rd + (size_t)(n - 1) * dd, one, zero, d, 1, wr + half * dd, st);
        rd = wr;
    }
    if (r == hipSuccess && split) r = pa::launch_target_scale<R>(target, d, b.gsum + 1, st);
    if (r == hipSuccess) r = pa::launch_grad_root<R>(target, rd, d, b.W + p.woff[p.top] * dd, b.gsum, st, wscale);
    // down-sweep
    for (int l = p.top - 1; r == hipSuccess && l >= 0; --l) {
        const C* Vl = l == 0 ? b.E : b.V + p.voff[l] * dd;
        const C* Wp = b.W + p.woff[l + 1] * dd;
        C* Wc = b.W + p.woff[l] * dd;
        const int n = p.n[l], half = n / 2;
        if (half) {
            r = gm(Wp, 1, Vl + dd, 2, Wp, 1, Wc, 2, c2(1), c2(1), c2(0), half, false);
            if (r == hipSuccess) r = gm(Vl, 2, Wp, 1, Wp, 1, Wc + dd, 2, c2(1), c2(1), c2(0), half, false);
        }
        if (r == hipSuccess && (n & 1))
            r = pa::launch_fill<R>(Wp + (size_t)(n / 2) * dd, one, zero, d, 1, Wc + (size_t)(n - 1) * dd, st);
    }
    // per-slice derivatives and traces; the start values b_M, b_{M-1}, db_M = 0 and
    // db_{M-1} = s a_M W are affine in X and W and enter the first two steps as C_in:
    //   db_{M-2} = s^2 a_M (X W + W X) + s a_{M-1} W,
    //   db_{M-3} = s (X db_{M-2} + W b_{M-2}) + beta s a_M W   (beta = -2 when M = 3)
    const int k3 = M - 3;
    const double beta3 = k3 == 0 ? -2.0 : -1.0;
    for (int j0 = 0; r == hipSuccess && j0 < N; j0 += p.chunk) {
        const int cnt = std::min(p.chunk, N - j0);
        ca.j0 = j0; ca.N = cnt;
        const C* Wj = b.W + (p.woff[0] + (size_t)j0) * dd;
        if (p.chunk < N) r = pa::launch_assemble<R>(ca, X, st);   // one chunk: X is still in place
        auto gemm = [&](const C* A, const C* B, const C* Cin, C* out, double2 alpha, double2 beta, double2 diag) {
            if (r == hipSuccess) r = gm(A, 1, B, 1, Cin, 1, out, 1, alpha, beta, diag, cnt, false);
        };
        // alpha (A B + A2 B2) + beta Cin + beta2 Cin2 + diag I: the two products that share an output in
        // one launch, accumulated before one rounding
        auto gsum = [&](const C* A, const C* B, const C* A2, const C* B2, const C* Cin, double2 beta, const C* Cin2,
                        double2 beta2, C* out, double2 alpha, double2 diag) {
            if (r == hipSuccess) r = zgemm_sum<R>(mode, A, B, A2, B2, Cin, beta, Cin2, beta2, out, alpha, diag, d, cnt, st);
        };
        if (power) {
            // Horner's rule in S = Y^2 (Y = ty X_buf) differentiated in the direction V = tv W:
            //   dS = Y V + V Y,  Q_J = m_{2J} I + m_{2J+1} Y,  dQ_J = m_{2J+1} V,
            //   Q_j = S Q_{j+1} + m_{2j+1} Y + m_{2j} I,  dQ_j = dS Q_{j+1} + S dQ_{j+1} + m_{2j+1} V,
            // and Q = dQ_0.  Level J-1 expands Q_J and dQ_J into the epilogue terms; below it each
            // level forms dQ_j into a free buffer, then Q_j (j >= 1) over the dead dQ_{j+1}.
            const GemmOp none{nullptr, w64, 0};
            auto op = [&](const C* x) { return x ? GemmOp{x, w64, 1} : none; };
            auto gemm2 = [&](const C* A, const C* B, const C* Cin, double2 beta, const C* Cin2, double2 beta2, C* out,
                             double2 alpha, double2 diag) {
                if (r == hipSuccess)
                    r = zgemm(mode, op(A), op(B), op(Cin), out, w64, 1, alpha, beta, diag, d, cnt, st, false, 0, op(Cin2),
                              beta2);
            };
            const int J = (M - 1) / 2;
            C* S = w[0];
            C* dS = w[1];
            if (kept) {
                // S and every Q_j of these slices were kept by the slices' pass: per level two products
                C* kq[pa::kMaxCheb / 2 + 1];
                keep_at(j0, kq);
                S = kq[0];
                C* dq = w[3];
                gsum(X, Wj, Wj, X, nullptr, c2(0), nullptr, c2(0), dS, c2(ty * tv), c2(0));
                gsum(dS, X, S, Wj, dS, m[2 * J], Wj, cscale(tv, m[2 * J - 1]), dq, cscale(ty, m[2 * J + 1]), c2(0));
                for (int j = J - 2; j >= 0; --j) {
                    C* z = dq == w[3] ? w[2] : w[3];
                    const double f = j == 0 ? fin : 1.0;
                    gsum(dS, kq[j + 1], S, dq, Wj, cscale(f * tv, m[2 * j + 1]), nullptr, c2(0), z, c2(f), c2(0));
                    dq = z;
                }
                if (r == hipSuccess) r = pa::launch_grad_trace<R>(dq, cnt, d, ca.opsT, ca.ncoef, b.dg, N, j0, st, wscale);
                continue;
            }
            gemm(X, X, nullptr, S, c2(ty * ty), c2(0), c2(0));
            gsum(X, Wj, Wj, X, nullptr, c2(0), nullptr, c2(0), dS, c2(ty * tv), c2(0));
            C* q = w[2];
            C* dq = w[3];
            C* spare = w[4];
            gemm2(S, X, S, m[2 * J], X, cscale(ty, m[2 * J - 1]), q, cscale(ty, m[2 * J + 1]), m[2 * J - 2]);
            gsum(dS, X, S, Wj, dS, m[2 * J], Wj, cscale(tv, m[2 * J - 1]), dq, cscale(ty, m[2 * J + 1]), c2(0));
            for (int j = J - 2; j >= 0; --j) {
                C* z = j == 0 ? X : spare;   // X is dead once no Q_j with j >= 1 is left to form
                const double f = j == 0 ? fin : 1.0;
                gsum(dS, q, S, dq, Wj, cscale(f * tv, m[2 * j + 1]), nullptr, c2(0), z, c2(f), c2(0));
                if (j >= 1) {
                    gemm(S, q, X, dq, c2(1), cscale(ty, m[2 * j + 1]), m[2 * j]);
                    spare = q;
                    q = dq;
                }
                dq = z;
            }
            if (r == hipSuccess) r = pa::launch_grad_trace<R>(dq, cnt, d, ca.opsT, ca.ncoef, b.dg, N, j0, st, wscale);
            continue;
        }
        gemm(X, X, X, w[0], cscale(s * s, a[M]), cscale(s, a[M - 1]), c2(a[M - 2].x - a[M].x, a[M - 2].y - a[M].y));
        gsum(X, Wj, Wj, X, Wj, cscale(s, a[M - 1]), nullptr, c2(0), w[2], cscale(s * s, a[M]), c2(0));
        gsum(X, w[2], Wj, w[0], Wj, cscale(beta3 * s, a[M]), nullptr, c2(0), w[3], c2(s), c2(0));
        if (k3 > 0)
            gemm(X, w[0], X, w[1], c2(s), cscale(beta3 * s, a[M]),
                 c2(a[k3].x + beta3 * a[M - 1].x, a[k3].y + beta3 * a[M - 1].y));
        // b1 = b_{k+1}, b2 = b_{k+2}, d1 = db_{k+1}, d2 = db_{k+2}
        C* b1 = w[1];
        C* b2 = w[0];
        C* d1 = w[3];
        C* d2 = w[2];
        for (int k = M - 4; k >= 0; --k) {
            gsum(X, d1, Wj, b1, d2, c2(k == 0 ? -2.0 : -1.0), nullptr, c2(0), d2, c2(s), c2(0));
            if (k > 0) gemm(X, b1, b2, b2, c2(s), c2(-1.0), a[k]);
            std::swap(b1, b2);
            std::swap(d1, d2);
        }
        // Q = Dp(X)[W] is in d1 (M = 3: the second step's w[3])
        if (r == hipSuccess) r = pa::launch_grad_trace<R>(d1, cnt, d, ca.opsT, ca.ncoef, b.dg, N, j0, st);
    }
    if (r == hipSuccess) r = grad_finish<R>(ca, b.dg, b.gsum, N, grad, fid, st);
    return r;
}

// Buffers of one gradient on the fused kernels (layout of GradPlan): E (chunk slices), V and W
// (chunk levels, then top levels), dg (ncoef x N values), gsum (one value).
template <class R> struct FusedGradBufs { cplx<R> *E, *V, *W, *dg, *gsum; };

// The adjoint gradient on the fused kernels (dims <= 32): slices and up-sweep chunk by chunk,
// the levels above the segment, the root, the down-sweep, then per chunk (recomputed when
// checkpointed) the lower down-sweep levels and the per-slice derivatives; the chain rule to
// the samples last.  h (optional) receives timing marks 0 (start) and 1 (root done).
template <class R>
hipError_t grad_fused_enqueue(pa::ChebArgs<R> ca, const GradPlan& p, const FusedGradBufs<R>& b, const cplx<R>* target,
                              R* fid, R* grad, hipStream_t st, Context<R>* h) {
    typedef cplx<R> C;
    const int d = ca.d, N = p.N;
    const size_t dd = (size_t)d * d;
    auto vc = [&](int l) { return b.V + p.coff[l] * dd; };
    auto wc = [&](int l) { return b.W + p.coff[l] * dd; };
    auto vt = [&](int l) { return b.V + (p.chunk_mats + p.toff[l]) * dd; };
    auto wt = [&](int l) { return b.W + (p.chunk_mats + p.toff[l]) * dd; };
    auto count = [](int cnt, int l) { return (int)(((long)cnt + (1L << l) - 1) >> l); };
    // Up-sweep from level lo (n_lo nodes at `in`) to level hi; out(l) is the array of level l.
    // Every node is kept as E = V - I.  Three levels per launch while that still leaves a
    // work-group per CU (a block's products run one after the other), one level per launch above.
    auto up_levels = [&](const C* in, int lo, int hi, int n_lo, auto out) -> hipError_t {
        for (int l = lo, n = n_lo; l < hi;) {
            const int k = (n >= 8 * 256) ? std::min(3, hi - l) : 1;
            pa::TreeStore<R> ts;
            for (int i = 1; i <= k; ++i) { n = (n + 1) / 2; ts.lev[i] = out(l + i); ts.n[i] = n; }
            const hipError_t r = pa::launch_up_block<R>(in, n_lo, d, k, ts, st);
            if (r != hipSuccess) return r;
            in = ts.lev[k]; n_lo = n; l += k;
        }
        return hipSuccess;
    };
    // Slices of chunk c and their products up to level `upto` (level seg goes to the top arrays).
    auto forward_chunk = [&](int c, int upto) -> hipError_t {
        const int j0 = c * p.chunk, cnt = std::min(p.chunk, N - j0);
        ca.j0 = j0; ca.N = cnt; ca.out = b.E;
        const hipError_t r = pa::launch_cheb_fused<R>(ca, st);
        if (r != hipSuccess) return r;
        return up_levels(b.E, 0, upto, cnt,
                         [&](int l) { return (l == p.seg) ? vt(p.seg) + (size_t)(j0 >> p.seg) * dd : vc(l); });
    };
    // without quadrature or Magnus terms a sample is a slice: the slice kernel writes the gradient
    // itself and no chain-rule launch follows
    pa::GradDirect<R> direct;
    if (ca.ctl.mode == pa::QUAD_NONE && !ca.ctl.magnus && ca.ctl.amps_in <= ca.ncoef) {
        direct.g = b.gsum; direct.grad = grad; direct.fid = fid; direct.pts = ca.ctl.pts; direct.amps = ca.ctl.amps_in;
    }
    if (h) mark(h, 0);
    hipError_t r = hipSuccess;
    for (int c = 0; r == hipSuccess && c < p.nchunks; ++c) r = forward_chunk(c, p.seg);
    if (r == hipSuccess) r = up_levels(vt(p.seg), p.seg, p.top, p.n[p.seg], [&](int l) { return vt(l); });
    // the root's costate t^H is needed only by its two children: one launch forms them and g, from
    // the top arrays (checkpointed layout) or from the chunk arrays (whole tree, one chunk);
    // otherwise (one or two slices) the root launch writes W(root)
    const bool merged_top = p.top >= 1 && p.top - 1 >= p.seg && p.n[p.top - 1] == 2;
    const bool merged_chunk = !merged_top && p.top >= 2 && p.seg == p.top && p.nchunks == 1 && p.n[p.top - 1] == 2;
    if (r == hipSuccess) {
        if (merged_top) r = pa::launch_down_root<R>(vt(p.top - 1), target, vt(p.top), d, wt(p.top - 1), b.gsum, st);
        else if (!merged_chunk) r = pa::launch_grad_root<R>(target, vt(p.top), d, wt(p.top), b.gsum, st);
    }
    if (h) mark(h, 1);
    for (int l = p.top - (merged_top ? 2 : 1); r == hipSuccess && l >= p.seg; --l)
        r = pa::launch_down_level<R>(vt(l), p.n[l], d, wt(l + 1), wt(l), st);
    for (int c = 0; r == hipSuccess && c < p.nchunks; ++c) {
        const int j0 = c * p.chunk, cnt = std::min(p.chunk, N - j0);
        if (p.nchunks > 1) r = forward_chunk(c, p.seg - 1);
        for (int l = p.seg - 1; r == hipSuccess && l >= 1; --l) {
            if (merged_chunk && l == p.top - 1) {
                r = pa::launch_down_root<R>(vc(l), target, vt(p.top), d, wc(l), b.gsum, st);
                continue;
            }
            const C* parent = (l + 1 == p.seg) ? wt(p.seg) + (size_t)(j0 >> p.seg) * dd : wc(l + 1);
            r = pa::launch_down_level<R>(vc(l), count(cnt, l), d, parent, wc(l), st);
        }
        const C* W1 = (p.seg == 1) ? wt(1) + (size_t)(j0 >> 1) * dd : wc(1);
        ca.j0 = j0; ca.N = cnt;
        if (r == hipSuccess) r = pa::launch_grad_slices<R>(ca, b.E, W1, b.dg, N, st, direct);
    }
    if (r == hipSuccess && !direct.grad) r = grad_finish<R>(ca, b.dg, b.gsum, N, grad, fid, st);
    return r;
}

// All device pointers; target is d x d row-major, grad receives amps x pts values of dF/dc
// (real controls: only the real parts of carr are read), fid (optional) F = |Tr(U_t^H U)|^2 / d^2.
template <class R>
Parament_ErrorCode gradient(Context<R>* h, const cplx<R>* carr_dev, double dt, unsigned pts, unsigned amps,
                            const cplx<R>* target_dev, R* fid_dev, R* grad_dev) {
    typedef cplx<R> C;
    if (!h->has_h) return PARAMENT_STATUS_NO_HAMILTONIAN;
    // dims <= 32 with at most kGradCoeffs coefficient operators: the fused kernels (on-chip slices
    // and tree levels); larger dims or more operators: batched GEMMs
    const bool stepwise = !pa::grad_supported((int)h->dim) || h->ncoef > pa::kGradCoeffs;
    if (amps == 0 || amps > h->amps) return PARAMENT_STATUS_INVALID_VALUE;
    if (hipSetDevice(h->device) != hipSuccess) return PARAMENT_FAIL;
    int N = 0;
    double tau = 0;
    slices(h, pts, dt, &N, &tau);
    if (N <= 0) return PARAMENT_STATUS_INVALID_VALUE;
    int M = h->mmax;
    if (!h->manual) {
        M = sizeof(R) == 4 ? select_fp32(h->hnorm, tau) : select_fp64(h->hnorm, tau);
        if (M < 3) return PARAMENT_STATUS_SELECT_SMALLER_DT;
    }
    if (M >= pa::kMaxCheb) return PARAMENT_STATUS_INVALID_VALUE;
    h->last_m = M;
    double are[pa::kMaxCheb], aim[pa::kMaxCheb];
    bessel(M, tau * h->hnorm, are, aim);
    const int d = (int)h->dim;
    const size_t dd = (size_t)d * d;

    pa::ChebArgs<R> ca;
    ca.ops = h->d_ops; ca.ncoef = h->ncoef; ca.d = d; ca.M = M; ca.s = (R)(2.0 / h->hnorm);
    ca.opsT = h->d_ops + (size_t)(1 + h->ncoef) * dd;
    ca.ctl.c = carr_dev; ca.ctl.pts = (int)pts; ca.ctl.amps_in = (int)amps; ca.ctl.amps_set = (int)h->amps;
    ca.ctl.mode = h->quad; ca.ctl.magnus = h->magnus ? 1 : 0; ca.ctl.tau = (R)tau;
    for (int k = 0; k <= M; ++k) { ca.are[k] = (R)are[k]; ca.aim[k] = (R)aim[k]; }
    ca.are[0] = (R)(are[0] - 1.0);  // slices are written as E_j = U_j - I
    // The derivative kernel evaluates the same polynomial's derivative in powers of tau X when
    // tau Hnorm <= 1 ((M-1)/2 + 1 product passes instead of M - 1).
    if (power_form_ok(M, tau * h->hnorm)) {
        double mr[pa::kMaxCheb], mi[pa::kMaxCheb];
        power_coefficients(M, tau * h->hnorm, are, aim, mr, mi);
        ca.ps = 1;
        for (int k = 0; k <= M; ++k) { ca.pre[k] = (R)mr[k]; ca.pim[k] = (R)mi[k]; }
    }

    // the fused path's tree layout (checkpointed or not); the step-wise path keeps the whole tree
    const GradPlan p = grad_plan(stepwise ? 1 : N, h->ckpt_seg, h->ckpt_chunk);
    StepGradPlan sp = step_grad_plan(stepwise ? N : 1, d, sizeof(C), h->ckpt_chunk, step_grad_work(M, ca.ps != 0));
    // the slices' S and Q_j kept for the derivative within 8 GiB (checkpointing asks for less memory: not kept)
    if (stepwise && !h->ckpt_seg) sp.keep = step_grad_keep(M, ca.ps != 0, N, d, sizeof(C), (size_t)8 << 30);
    // complex64 in the default mode: the power form's products, the sweeps and the derivative on
    // the scaled fp16 split, as the forward's step-wise path
    sp.split = sizeof(R) == 4 && ca.ps && M >= 5 && h->precision == PARAMENT_PRECISION_MIXED && mixed_split_scheme() != 0;
    const size_t mats_vw = p.chunk_mats + p.top_mats;
    const size_t bytes_e =
        stepwise ? ((size_t)N * (1 + sp.keep) + (size_t)sp.nwork * sp.chunk) * dd * sizeof(C) : (size_t)p.chunk * dd * sizeof(C);
    const size_t bytes_v = (stepwise ? std::max<size_t>(sp.vmats, 1) : mats_vw) * dd * sizeof(C);
    const size_t bytes_w = (stepwise ? sp.wmats : mats_vw) * dd * sizeof(C);
    const size_t bytes_dg = (size_t)h->ncoef * N * sizeof(C);
    if (!grow(&h->d_gE, &h->cap_gE, bytes_e) || !grow(&h->d_gV, &h->cap_gV, bytes_v) ||
        !grow(&h->d_gW, &h->cap_gW, bytes_w) || !grow(&h->d_gdg, &h->cap_gdg, bytes_dg) ||
        !grow(&h->d_gaux, &h->cap_gaux, 4 * sizeof(C)))
        return PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
    h->grad_workspace = bytes_e + bytes_v + bytes_w + bytes_dg + 4 * sizeof(C);
    C* E = reinterpret_cast<C*>(h->d_gE);
    C* V = reinterpret_cast<C*>(h->d_gV);
    C* W = reinterpret_cast<C*>(h->d_gW);
    C* dg = reinterpret_cast<C*>(h->d_gdg);
    C* gsum = reinterpret_cast<C*>(h->d_gaux);
    hipStream_t st = h->stream;   // the stream the launches below go to (a capture stream when recording)
    auto enqueue = [&]() -> Parament_ErrorCode {
        hipError_t r;
        if (stepwise) {
            StepGradBufs<R> sb;
            sb.E = E; sb.work = E + (size_t)N * dd; sb.V = V; sb.W = W; sb.dg = dg; sb.gsum = gsum;
            sb.keep = sp.keep ? sb.work + (size_t)sp.nwork * sp.chunk * dd : nullptr;
            mark(h, 0);
            r = grad_stepwise_enqueue<R>(ca, sp, sb, target_dev, fid_dev, grad_dev, st);
            mark(h, 1);
        } else {
            r = grad_fused_enqueue<R>(ca, p, FusedGradBufs<R>{E, V, W, dg, gsum}, target_dev, fid_dev, grad_dev, st, h);
        }
        mark(h, 2);
        return r == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
    };

    // Some thirty launches, most of them short: a repeated call replays them as one graph.
    if (h->timing) return enqueue();
    GradKey key;
    key.carr = carr_dev; key.target = target_dev; key.fid = fid_dev; key.grad = grad_dev; key.ops = h->d_ops;
    key.E = E; key.V = V; key.W = W; key.dg = dg; key.aux = gsum;
    key.dt = dt; key.hnorm = h->hnorm; key.pts = pts; key.amps = amps; key.amps_set = h->amps; key.dim = h->dim;
    key.seg = h->ckpt_seg; key.chunk = h->ckpt_chunk; key.quad = h->quad; key.ncoef = h->ncoef; key.M = M;
    key.device = h->device; key.precision = h->precision;
    if (h->grad_exec && key == h->grad_key)
        return hipGraphLaunch(h->grad_exec, h->stream) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
    if (!(key == h->grad_prev)) {   // first call with these arguments: launch directly
        h->grad_prev = key;
        return enqueue();
    }
    if (!h->cap_stream && hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking) != hipSuccess) {
        hipGetLastError(); h->cap_stream = nullptr;
        return enqueue();
    }
    if (h->grad_exec) { hipGraphExecDestroy(h->grad_exec); h->grad_exec = nullptr; }
    hipGraph_t g = nullptr;
    st = h->cap_stream;
    Parament_ErrorCode err = PARAMENT_STATUS_CUBLAS_FAILED;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        err = enqueue();
        const hipError_t r = hipStreamEndCapture(st, &g);
        if (err == PARAMENT_STATUS_SUCCESS && r == hipSuccess && g &&
            hipGraphInstantiate(&h->grad_exec, g, nullptr, nullptr, 0) != hipSuccess)
            h->grad_exec = nullptr;
        if (g) hipGraphDestroy(g);
    }
    hipGetLastError();
    st = h->stream;
    if (err != PARAMENT_STATUS_SUCCESS || !h->grad_exec) {   // capture unavailable: launch directly
        h->grad_exec = nullptr;
        return enqueue();
    }
    h->grad_key = key;
    return hipGraphLaunch(h->grad_exec, h->stream) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
}

template <class R>
Parament_ErrorCode gradient_host(Context<R>* h, const typename ApiCplx<R>::type* carr, double dt, unsigned pts,
                                 unsigned amps, const typename ApiCplx<R>::type* target, R* fidelity, R* grad) {
    typedef cplx<R> C;
    if (!h) return PARAMENT_STATUS_INVALID_VALUE;
    if (!h->has_h) return h->last_error = PARAMENT_STATUS_NO_HAMILTONIAN;
    if (!carr || !target || !grad) return h->last_error = PARAMENT_STATUS_INVALID_VALUE;
    for (size_t i = 0; i < (size_t)amps * pts; ++i)
        if (carr[i].y != 0) return h->last_error = PARAMENT_STATUS_INVALID_VALUE;  // real controls only
    Parament_ErrorCode err = ensure_common(h, amps);
    if (err != PARAMENT_STATUS_SUCCESS) return h->last_error = err;
    const size_t dd = (size_t)h->dim * h->dim;
    const size_t cbytes = (size_t)amps * pts * sizeof(C);
    // staging: controls, target, grad (amps * pts values), fidelity
    const size_t need = cbytes + dd * sizeof(C) + ((size_t)amps * pts + 2) * sizeof(R);
    if (!grow(reinterpret_cast<void**>(&h->d_c), &h->cap_c, need)) return h->last_error = PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
    char* base = reinterpret_cast<char*>(h->d_c);
    C* dc = reinterpret_cast<C*>(base);
    C* dtgt = reinterpret_cast<C*>(base + cbytes);
    R* dgrad = reinterpret_cast<R*>(base + cbytes + dd * sizeof(C));
    R* dfid = dgrad + (size_t)amps * pts;
    if (hipMemcpyAsync(dc, carr, cbytes, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(dtgt, target, dd * sizeof(C), hipMemcpyHostToDevice, h->stream) != hipSuccess)
        return h->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    err = gradient(h, dc, dt, pts, amps, dtgt, dfid, dgrad);
    if (err == PARAMENT_STATUS_SUCCESS) {
        if (hipMemcpyAsync(grad, dgrad, (size_t)amps * pts * sizeof(R), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            (fidelity && hipMemcpyAsync(fidelity, dfid, sizeof(R), hipMemcpyDeviceToHost, h->stream) != hipSuccess) ||
            hipStreamSynchronize(h->stream) != hipSuccess)
            err = PARAMENT_STATUS_CUBLAS_FAILED;
    } else {
        hipStreamSynchronize(h->stream);
    }
    return h->last_error = err;
}

template <class R>
Parament_ErrorCode gradient_device(Context<R>* h, const typename ApiCplx<R>::type* carr_dev, double dt, unsigned pts,
                                   unsigned amps, const typename ApiCplx<R>::type* target_dev, R* fid_dev, R* grad_dev) {
    if (!h) return PARAMENT_STATUS_INVALID_VALUE;
    if (!h->has_h) return h->last_error = PARAMENT_STATUS_NO_HAMILTONIAN;
    if (!carr_dev || !target_dev || !grad_dev) return h->last_error = PARAMENT_STATUS_INVALID_VALUE;
    Parament_ErrorCode err = ensure_common(h, amps);
    if (err != PARAMENT_STATUS_SUCCESS) return h->last_error = err;
    return h->last_error = gradient(h, reinterpret_cast<const cplx<R>*>(carr_dev), dt, pts, amps,
                                    reinterpret_cast<const cplx<R>*>(target_dev), fid_dev, grad_dev);
}

template <class R> Parament_ErrorCode set_precision(Context<R>* h, int mode) {
    if (mode != PARAMENT_PRECISION_NATIVE && mode != PARAMENT_PRECISION_MIXED && mode != PARAMENT_PRECISION_MIXED_FP32 &&
        mode != PARAMENT_PRECISION_MIXED_BF16X2)
        return h->last_error = PARAMENT_STATUS_INVALID_VALUE;
    h->precision = mode;
    return h->last_error = PARAMENT_STATUS_SUCCESS;
}

template <class R> Parament_ErrorCode set_checkpointing(Context<R>* h, unsigned seg, unsigned chunk) {
    if (!h) return PARAMENT_STATUS_INVALID_VALUE;
    h->ckpt_seg = seg;
    h->ckpt_chunk = chunk;
    return h->last_error = PARAMENT_STATUS_SUCCESS;
}

template <class R> Parament_ErrorCode kernel_times(Context<R>* h, double* a, double* b) {
    if (!h->timing || !h->ev[0] || !h->ev[1] || !h->ev[2]) return h->last_error = PARAMENT_STATUS_INVALID_VALUE;
    if (hipEventSynchronize(h->ev[2]) != hipSuccess) return h->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    float t0 = 0, t1 = 0;
    if (hipEventElapsedTime(&t0, h->ev[0], h->ev[1]) != hipSuccess ||
        hipEventElapsedTime(&t1, h->ev[1], h->ev[2]) != hipSuccess)
        return h->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    if (a) *a = t0;
    if (b) *b = t1;
    return h->last_error = PARAMENT_STATUS_SUCCESS;
}
// ----- ensemble of independent complex64 problems (ensemble.hip + gemm.hip) -----
// Instances of dim <= 64 are padded to a dim class (16, 32, 48, 64) and run by the multi-instance
// kernels, one slice kernel and one tree-level kernel per class and level; larger instances run
// the step-wise path on batched GEMMs.  The classes and the large instances go on separate
// branches (streams forked from and joined to the ensemble's stream), the whole solve is
// captured once as a hipGraph and replayed.  The tuner picks, per class, the tree block size and
// the persistent grid, and per band of 16 dims the output tile of the large instances' GEMMs, by
// timing each candidate on the ensemble itself; large instances run on up to four branches.
constexpr int kEnsSide = 8;        // side streams: 4 classes + up to 4 large-instance branches
constexpr int kEnsClasses = 4;     // dim classes 16, 32, 48, 64
constexpr int kEnsClassDim[kEnsClasses] = {16, 32, 48, 64};
inline int ens_class_of(unsigned d) { return d <= 16 ? 0 : (d <= 32 ? 1 : (d <= 48 ? 2 : (d <= 64 ? 3 : -1))); }

// nodes of one tree scratch buffer of an ensemble instance with N slices
size_t ens_tree_scratch(int N) {
    return std::max((size_t)((N + 3) / 4), (size_t)pa::tree_scratch_nodes<float>(N));
}

struct EnsInst {
    unsigned d = 0, amps = 0, pts = 0;
    double dt = 0, tau = 0, hnorm = 0;
    int N = 0, M = 0, cls = -1;                // cls: class index, -1 for the batched-GEMM path
    int tile = 0;                               // batched-GEMM tile (0 = automatic)
    size_t ops_off = 0, c_off = 0, out_off = 0; // element offsets in the packed arrays
    size_t work_off = 0;                        // element offset of its working memory
    size_t opsT_off = 0;                        // element offset of its transposed controls (gradient)
    double are[pa::kMaxCheb], aim[pa::kMaxCheb];
};

// Working buffers of N slices that a large instance's forward takes: X, the recurrence's or the
// power form's two, and a fourth for the power form at M >= 7.
inline int ens_large_buffers(const EnsInst& x) { return (stepwise_power(x.M, x.tau * x.hnorm) && x.M >= 7) ? 4 : 3; }

struct EnsClassPlan {
    int D = 0;
    int2* d_slice = nullptr; int n_slice = 0;
    // per tree block size (index logb - 2): levels, each a descriptor table and a task table
    std::vector<pa::EnsLevel*> d_lv[2];
    std::vector<int2*> d_lt[2];
    std::vector<int> n_lt[2];
};

struct EnsCfg { int T = 1, logb = 3, grid_mult = 0; };

struct Ensemble {
    Parament_ErrorCode last_error = PARAMENT_STATUS_SUCCESS;
    std::vector<EnsInst> in;
    int quad = 0;
    bool has_h = false;
    int device = 0, cus = 256;
    hipStream_t stream = nullptr, side[kEnsSide] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kEnsSide] = {}, t0 = nullptr, t1 = nullptr;
    float2 *d_ops = nullptr, *d_c = nullptr, *d_out = nullptr, *d_work = nullptr;
    size_t n_ops = 0, n_c = 0, n_out = 0, n_work = 0;
    pa::EnsInstance* d_inst = nullptr;
    EnsClassPlan cls[kEnsClasses];
    std::vector<int> large;
    EnsCfg cfg[kEnsClasses];
    int branches = 2;
    bool tuned = false;
    int exec_mode = 0;              // 0: captured graph, 1: per-instance loop
    hipGraphExec_t gexec = nullptr;
    std::vector<void*> allocs;      // small device tables
    // gradient: transposed control operators, instance -> branch, per-branch workspaces, the
    // captured graph and the pointers it was captured with, staging for the host-array call
    float2* d_opsT = nullptr;
    std::vector<int> gbranch;
    float2* d_gws[kEnsSide] = {};
    bool grad_ready = false;
    hipGraphExec_t gexec_grad = nullptr;
    const void* gkey[3] = {};
    float2* d_tgt = nullptr;
    float *d_fid = nullptr, *d_grad = nullptr;
};

void ens_free_plan(Ensemble* e) {
    for (void* p : e->allocs) hipFree(p);
    e->allocs.clear();
    for (auto& c : e->cls) {
        c.d_slice = nullptr; c.n_slice = 0;
        for (int b = 0; b < 2; ++b) { c.d_lv[b].clear(); c.d_lt[b].clear(); c.n_lt[b].clear(); }
    }
    if (e->gexec) { hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
    if (e->gexec_grad) { hipGraphExecDestroy(e->gexec_grad); e->gexec_grad = nullptr; }
    for (auto& w : e->d_gws) { hipFree(w); w = nullptr; }
    hipFree(e->d_opsT); e->d_opsT = nullptr;
    e->gbranch.clear();
    e->grad_ready = false;
    hipFree(e->d_work); e->d_work = nullptr; e->n_work = 0;
    e->d_inst = nullptr;   // one of allocs
    e->large.clear();
    e->tuned = false;
}

template <class T> T* ens_upload(Ensemble* e, const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    void* p = nullptr;
    if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess) { hipGetLastError(); return nullptr; }
    e->allocs.push_back(p);
    if (hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return static_cast<T*>(p);
}

Parament_ErrorCode ens_create(Ensemble** ep, unsigned count, const unsigned* dims, const unsigned* amps,
                              const unsigned* pts, const double* dts, int quad) {
    if (!ep || count == 0 || !dims || !amps || !pts || !dts) return PARAMENT_STATUS_INVALID_VALUE;
    if (quad != PARAMENT_QUADRATURE_NONE && quad != PARAMENT_QUADRATURE_MIDPOINT && quad != PARAMENT_QUADRATURE_SIMPSON)
        return PARAMENT_STATUS_INVALID_VALUE;
    Ensemble* e = new (std::nothrow) Ensemble();
    if (!e) return PARAMENT_STATUS_HOST_ALLOC_FAILED;
    e->quad = quad;
    e->in.resize(count);
    const bool doubled = quad == PARAMENT_QUADRATURE_SIMPSON;
    for (unsigned i = 0; i < count; ++i) {
        EnsInst& x = e->in[i];
        x.d = dims[i]; x.amps = amps[i]; x.pts = pts[i]; x.dt = dts[i];
        if (x.d == 0 || x.amps == 0 || x.amps > (unsigned)pa::kMaxCoeffs || !(x.dt > 0)) { delete e; return PARAMENT_STATUS_INVALID_VALUE; }
        x.tau = doubled ? 2 * x.dt : x.dt;
        x.N = doubled ? (int)((x.pts - 1) / 2) : (quad == PARAMENT_QUADRATURE_MIDPOINT ? (int)x.pts - 1 : (int)x.pts);
        if (x.pts == 0 || x.N <= 0) { delete e; return PARAMENT_STATUS_INVALID_VALUE; }
        x.ops_off = e->n_ops; e->n_ops += (size_t)(1 + x.amps) * x.d * x.d;
        x.c_off = e->n_c; e->n_c += (size_t)x.amps * x.pts;
        x.out_off = e->n_out; e->n_out += (size_t)x.d * x.d;
    }
    if (hipGetDevice(&e->device) != hipSuccess) { hipGetLastError(); delete e; return PARAMENT_FAIL; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, e->device) == hipSuccess) e->cus = prop.multiProcessorCount;
    bool ok = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&e->fork_ev, hipEventDisableTiming) == hipSuccess &&
              hipEventCreate(&e->t0) == hipSuccess && hipEventCreate(&e->t1) == hipSuccess;
    for (int k = 0; ok && k < kEnsSide; ++k)
        ok = hipStreamCreateWithFlags(&e->side[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&e->join_ev[k], hipEventDisableTiming) == hipSuccess;
    ok = ok && hipMalloc(&e->d_ops, e->n_ops * sizeof(float2)) == hipSuccess &&
         hipMalloc(&e->d_c, std::max<size_t>(e->n_c, 1) * sizeof(float2)) == hipSuccess &&
         hipMalloc(&e->d_out, e->n_out * sizeof(float2)) == hipSuccess;
    *ep = e;
    if (!ok) { hipGetLastError(); return PARAMENT_STATUS_DEVICE_ALLOC_FAILED; }
    return PARAMENT_STATUS_SUCCESS;
}

Parament_ErrorCode ens_destroy(Ensemble* e) {
    if (!e) return PARAMENT_STATUS_SUCCESS;
    hipSetDevice(e->device);
    if (e->stream) hipStreamSynchronize(e->stream);
    ens_free_plan(e);
    hipFree(e->d_ops); hipFree(e->d_c); hipFree(e->d_out);
    hipFree(e->d_tgt); hipFree(e->d_fid); hipFree(e->d_grad);
    for (int k = 0; k < kEnsSide; ++k) {
        if (e->side[k]) hipStreamDestroy(e->side[k]);
        if (e->join_ev[k]) hipEventDestroy(e->join_ev[k]);
    }
    if (e->fork_ev) hipEventDestroy(e->fork_ev);
    if (e->t0) hipEventDestroy(e->t0);
    if (e->t1) hipEventDestroy(e->t1);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return PARAMENT_STATUS_SUCCESS;
}

// H: packed instance by instance, (1 + amps_i) row-major d_i x d_i matrices (H0 then H_k).
Parament_ErrorCode ens_set_hamiltonians(Ensemble* e, const cuComplex* H) {
    if (!H) return PARAMENT_STATUS_INVALID_VALUE;
    if (hipSetDevice(e->device) != hipSuccess) return PARAMENT_FAIL;
    ens_free_plan(e);
    e->has_h = false;
    if (hipMemcpy(e->d_ops, H, e->n_ops * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
        return PARAMENT_STATUS_CUBLAS_FAILED;
    std::vector<pa::EnsInstance> tab(e->in.size());
    std::vector<std::vector<int2>> slices(kEnsClasses);
    for (size_t i = 0; i < e->in.size(); ++i) {
        EnsInst& x = e->in[i];
        const size_t dd = (size_t)x.d * x.d;
        x.hnorm = 0;
        for (unsigned k = 0; k <= x.amps; ++k) x.hnorm += one_norm(H + x.ops_off + k * dd, x.d);
        x.M = select_fp32(x.hnorm, x.tau);
        if (x.M < 3) return PARAMENT_STATUS_SELECT_SMALLER_DT;
        bessel(x.M, x.tau * x.hnorm, x.are, x.aim);
        x.cls = ens_class_of(x.d);
        // working memory: classes: leaves N + two tree scratch buffers (the multi-instance tree
        // writes ceil(N / 4) nodes, launch_tree tree_scratch_nodes(N)); large: X, b1, b2
        x.work_off = e->n_work;
        e->n_work += x.cls >= 0 ? ((size_t)x.N + 2 * ens_tree_scratch(x.N)) * dd : ens_large_buffers(x) * (size_t)x.N * dd;
        if (x.cls < 0) e->large.push_back((int)i);
    }
    if (hipMalloc(&e->d_work, e->n_work * sizeof(float2)) != hipSuccess) { hipGetLastError(); e->d_work = nullptr; return PARAMENT_STATUS_DEVICE_ALLOC_FAILED; }
    for (size_t i = 0; i < e->in.size(); ++i) {
        const EnsInst& x = e->in[i];
        pa::EnsInstance& t = tab[i];
        t.ops = e->d_ops + x.ops_off; t.c = e->d_c + x.c_off; t.leaves = e->d_work + x.work_off;
        t.d = (int)x.d; t.ncoef = (int)x.amps; t.amps = (int)x.amps; t.pts = (int)x.pts; t.mode = e->quad; t.M = x.M;
        t.s = (float)(2.0 / x.hnorm); t.tau = (float)x.tau;
        for (int k = 0; k < pa::kMaxCheb; ++k) { t.are[k] = k <= x.M ? (float)x.are[k] : 0.f; t.aim[k] = k <= x.M ? (float)x.aim[k] : 0.f; }
        t.are[0] = (float)(x.are[0] - 1.0);
        // the class kernels take the power form where the step-wise path does
        t.ps = stepwise_power(x.M, x.tau * x.hnorm) ? 1 : 0;
        double mr[pa::kMaxCheb] = {}, mi[pa::kMaxCheb] = {};
        if (t.ps) power_coefficients(x.M, x.tau * x.hnorm, x.are, x.aim, mr, mi);
        for (int k = 0; k < pa::kMaxCheb; ++k) { t.pre[k] = (float)mr[k]; t.pim[k] = (float)mi[k]; }
        if (x.cls >= 0)
            for (int j = 0; j < x.N; ++j) slices[x.cls].push_back(make_int2((int)i, j));
    }
    e->d_inst = ens_upload(e, tab);
    if (!e->d_inst) return PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
    for (int c = 0; c < kEnsClasses; ++c) {
        EnsClassPlan& P = e->cls[c];
        P.D = kEnsClassDim[c];
        P.n_slice = (int)slices[c].size();
        if (P.n_slice) {
            P.d_slice = ens_upload(e, slices[c]);
            if (!P.d_slice) return PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
        }
        for (int b = 0; b < 2; ++b) {
            const int logb = b + 2, B = 1 << logb;
            // level l of every member: src -> dst, until one node remains (written as U = I + E)
            std::vector<std::vector<pa::EnsLevel>> lv;
            std::vector<std::vector<int2>> lt;
            for (size_t i = 0; i < e->in.size(); ++i) {
                const EnsInst& x = e->in[i];
                if (x.cls != c) continue;
                const size_t dd = (size_t)x.d * x.d;
                float2* leaves = e->d_work + x.work_off;
                float2* scr[2] = {leaves + (size_t)x.N * dd, leaves + ((size_t)x.N + ens_tree_scratch(x.N)) * dd};
                int n = x.N, l = 0;
                const float2* src = leaves;
                do {
                    const int n_out = (n + B - 1) / B;
                    if ((int)lv.size() <= l) { lv.emplace_back(); lt.emplace_back(); }
                    pa::EnsLevel L;
                    L.src = src; L.n_in = n; L.d = (int)x.d; L.final_level = n_out == 1;
                    L.dst = n_out == 1 ? e->d_out + x.out_off : scr[l & 1];
                    const int li = (int)lv[l].size();
                    lv[l].push_back(L);
                    for (int blk = 0; blk < n_out; ++blk) lt[l].push_back(make_int2(li, blk));
                    src = L.dst; n = n_out; ++l;
                } while (n > 1);
            }
            for (size_t l = 0; l < lv.size(); ++l) {
                P.d_lv[b].push_back(ens_upload(e, lv[l]));
                P.d_lt[b].push_back(ens_upload(e, lt[l]));
                P.n_lt[b].push_back((int)lt[l].size());
                if (!P.d_lv[b].back() || !P.d_lt[b].back()) return PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
            }
        }
    }
    e->has_h = true;
    return PARAMENT_STATUS_SUCCESS;
}

int ens_grid(const Ensemble* e, int ntasks, int mult) {
    return mult > 0 ? std::max(1, std::min(ntasks, e->cus * mult)) : ntasks;
}

hipError_t ens_enqueue_class(Ensemble* e, int c, const EnsCfg& cf, hipStream_t st) {
    const EnsClassPlan& P = e->cls[c];
    if (!P.n_slice) return hipSuccess;
    hipError_t r = pa::launch_ens_slices(P.D, cf.T, e->d_inst, P.d_slice, P.n_slice, ens_grid(e, P.n_slice, cf.grid_mult), st);
    const int b = cf.logb - 2;
    for (size_t l = 0; r == hipSuccess && l < P.d_lv[b].size(); ++l)
        r = pa::launch_ens_tree(P.D, cf.T, cf.logb, P.d_lv[b][l], P.d_lt[b][l], P.n_lt[b][l],
                                ens_grid(e, P.n_lt[b][l], cf.grid_mult), st);
    return r;
}

// One large instance: the step-wise path on batched matrix-core GEMMs.  Where the power form
// applies (tau Hnorm <= 1) its products and the tree's run on scaled fp16 two-way splits, as in a
// complex64 context's default mode: the slices are assembled as Y = tau X (entries below 1), and
// S = Y^2, the Horner brackets and the deviations E = U - I stay below 4 in modulus.  Otherwise
// the recurrence and the tree run on FP32 products.
hipError_t ens_enqueue_large(Ensemble* e, int i, hipStream_t st) {
    const EnsInst& x = e->in[i];
    const int d = (int)x.d, N = x.N, M = x.M;
    const size_t dd = (size_t)d * d;
    float2* X = e->d_work + x.work_off;
    float2* b1 = X + (size_t)N * dd;
    float2* b2 = b1 + (size_t)N * dd;
    float2* b3 = ens_large_buffers(x) > 3 ? b2 + (size_t)N * dd : nullptr;
    pa::ChebArgs<float> ca;
    ca.ops = e->d_ops + x.ops_off; ca.ncoef = (int)x.amps; ca.d = d; ca.N = N; ca.M = M; ca.s = (float)(2.0 / x.hnorm);
    ca.ctl.c = e->d_c + x.c_off; ca.ctl.pts = (int)x.pts; ca.ctl.amps_in = (int)x.amps; ca.ctl.amps_set = (int)x.amps;
    ca.ctl.mode = e->quad; ca.ctl.magnus = 0; ca.ctl.tau = (float)x.tau;
    for (int k = 0; k <= M; ++k) { ca.are[k] = (float)x.are[k]; ca.aim[k] = (float)x.aim[k]; }
    ca.out = nullptr;
    double2 ak[pa::kMaxCheb];
    for (int k = 0; k <= M; ++k) ak[k] = c2(k == 0 ? x.are[0] - 1.0 : (double)ca.are[k], (double)ca.aim[k]);
    float2* leaves = nullptr;
    const bool power = stepwise_power(M, x.tau * x.hnorm);
    const int mode = power ? pa::GEMM_F16X2 : pa::GEMM_F32;
    if (power) ca.asm_scale = (float)x.tau;
    hipError_t r = pa::launch_assemble<float>(ca, X, st);
    if (power) {
        double mr[pa::kMaxCheb], mi[pa::kMaxCheb];
        double2 mc[pa::kMaxCheb];
        power_coefficients(M, x.tau * x.hnorm, x.are, x.aim, mr, mi);
        for (int k = 0; k <= M; ++k) mc[k] = c2(mr[k], mi[k]);
        if (r == hipSuccess)
            r = power_batched<float>(mode, X, d, N, M, 1.0, mc, true, b1, b2, b3, nullptr, &leaves, st, x.tile);
    } else if (r == hipSuccess) {
        r = clenshaw_batched<float>(X, d, N, M, (double)ca.s, ak, b1, b2, nullptr, &leaves, st, x.tile);
    }
    // the leaves E_j in one buffer, the other free for the other half of the ping-pong
    float2* rd = leaves;
    float2* wr = leaves == b1 ? b2 : b1;   // in the power form the leaves are in X and b1 is free
    int n = N;
    while (r == hipSuccess && n > 1) {
        const int half = n / 2;
        r = zgemm(mode, GemmOp{rd + dd, false, 2}, GemmOp{rd, false, 2}, GemmOp{nullptr, false, 0}, wr, false, 1, c2(1),
                  c2(0), c2(0), d, half, st, true, x.tile);
        if (r == hipSuccess && n % 2)
            r = pa::launch_fill<float>(rd + (size_t)(n - 1) * dd, make_float2(1, 0), make_float2(0, 0), d, 1, wr + half * dd, st);
        n = half + (n % 2);
        std::swap(rd, wr);
    }
    if (r == hipSuccess) r = pa::launch_fill<float>(rd, make_float2(1, 0), make_float2(1, 0), d, 1, e->d_out + x.out_off, st);
    return r;
}

// The whole solve on e->stream: branches for the classes and for the large instances.
hipError_t ens_enqueue(Ensemble* e) {
    hipError_t r = hipEventRecord(e->fork_ev, e->stream);
    const int nb = std::max(1, std::min(e->branches, kEnsSide - kEnsClasses));
    int used[kEnsSide] = {};
    for (int c = 0; r == hipSuccess && c < kEnsClasses; ++c) {
        if (!e->cls[c].n_slice) continue;
        r = hipStreamWaitEvent(e->side[c], e->fork_ev, 0);
        if (r == hipSuccess) r = ens_enqueue_class(e, c, e->cfg[c], e->side[c]);
        used[c] = 1;
    }
    for (size_t q = 0; r == hipSuccess && q < e->large.size(); ++q) {
        const int k = kEnsClasses + (int)(q % nb);
        if (!used[k]) { r = hipStreamWaitEvent(e->side[k], e->fork_ev, 0); used[k] = 1; }
        if (r == hipSuccess) r = ens_enqueue_large(e, e->large[q], e->side[k]);
    }
    for (int k = 0; r == hipSuccess && k < kEnsSide; ++k) {
        if (!used[k]) continue;
        r = hipEventRecord(e->join_ev[k], e->side[k]);
        if (r == hipSuccess) r = hipStreamWaitEvent(e->stream, e->join_ev[k], 0);
    }
    return r;
}

// Per-instance loop (the baseline the ensemble replaces): one problem after the other on one
// stream with the single-problem kernels.
hipError_t ens_enqueue_loop(Ensemble* e) {
    hipError_t r = hipSuccess;
    for (size_t i = 0; r == hipSuccess && i < e->in.size(); ++i) {
        const EnsInst& x = e->in[i];
        if (x.cls < 0) { r = ens_enqueue_large(e, (int)i, e->stream); continue; }
        const size_t dd = (size_t)x.d * x.d;
        pa::ChebArgs<float> ca;
        ca.ops = e->d_ops + x.ops_off; ca.ncoef = (int)x.amps; ca.d = (int)x.d; ca.N = x.N; ca.M = x.M;
        ca.s = (float)(2.0 / x.hnorm);
        ca.ctl.c = e->d_c + x.c_off; ca.ctl.pts = (int)x.pts; ca.ctl.amps_in = (int)x.amps; ca.ctl.amps_set = (int)x.amps;
        ca.ctl.mode = e->quad; ca.ctl.magnus = 0; ca.ctl.tau = (float)x.tau;
        for (int k = 0; k <= x.M; ++k) { ca.are[k] = (float)x.are[k]; ca.aim[k] = (float)x.aim[k]; }
        ca.are[0] = (float)(x.are[0] - 1.0);
        if (stepwise_power(x.M, x.tau * x.hnorm)) {   // as the class kernels
            double mr[pa::kMaxCheb], mi[pa::kMaxCheb];
            power_coefficients(x.M, x.tau * x.hnorm, x.are, x.aim, mr, mi);
            ca.ps = 1;
            for (int k = 0; k <= x.M; ++k) { ca.pre[k] = (float)mr[k]; ca.pim[k] = (float)mi[k]; }
        }
        float2* leaves = e->d_work + x.work_off;
        ca.out = leaves;
        r = pa::launch_cheb_fused<float>(ca, e->stream);
        float2* s0 = leaves + (size_t)x.N * dd;
        float2* s1 = s0 + ens_tree_scratch(x.N) * dd;
        if (r == hipSuccess) r = pa::launch_tree<float>(leaves, false, x.N, (int)x.d, s0, s1, e->d_out + x.out_off, e->stream);
    }
    return r;
}

float ens_time(Ensemble* e, const std::function<hipError_t()>& f, int reps, bool warm = true) {
    // a candidate whose launches or timing fail is never the fastest
    if (warm && f() != hipSuccess) return 1e30f;   // warm-up
    if (hipEventRecord(e->t0, e->stream) != hipSuccess) return 1e30f;
    for (int k = 0; k < reps; ++k)
        if (f() != hipSuccess) return 1e30f;
    float ms = 0;
    if (hipEventRecord(e->t1, e->stream) != hipSuccess || hipEventSynchronize(e->t1) != hipSuccess ||
        hipEventElapsedTime(&ms, e->t0, e->t1) != hipSuccess)
        return 1e30f;
    return ms / reps;
}

Parament_ErrorCode ens_tune(Ensemble* e) {
    if (!e->has_h) return PARAMENT_STATUS_NO_HAMILTONIAN;
    if (hipSetDevice(e->device) != hipSuccess) return PARAMENT_FAIL;
    if (e->gexec) { hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
    for (int c = 0; c < kEnsClasses; ++c) {
        if (!e->cls[c].n_slice) continue;
        const int D = kEnsClassDim[c];
        // the tree block size on the default grid first, then the persistent grid size for the
        // better of those (the grid size has not changed the block size chosen).  One wave split
        // (T = 1): the split of a slice over two waves' tiles was slower in every class whenever
        // it was timed (512-system ensemble, every run).
        float best = 1e30f;
        auto trial = [&](const EnsCfg& cf) {
            const float t = ens_time(e, [&] { return ens_enqueue_class(e, c, cf, e->stream); }, 1);
            if (t < best) { best = t; e->cfg[c] = cf; }
        };
        if (!pa::ens_shape_valid(D, 1)) continue;
        for (int logb = 2; logb <= 3; ++logb) {
            EnsCfg cf; cf.T = 1; cf.logb = logb; cf.grid_mult = 0;
            trial(cf);
        }
        const EnsCfg base = e->cfg[c];
        for (int gm : {2, 8}) {
            EnsCfg cf = base; cf.grid_mult = gm;
            trial(cf);
        }
    }
    // the large instances' batched GEMMs take the automatic output tile (the smaller padded
    // product volume).  Timing each tile on the first 512 slices of one instance per band of 16
    // dims chose tiles that made the 512-system ensemble's 106 large systems 10% slower (179
    // against 162 ms as a graph) once their products ran on the fp16 split.
    for (int i : e->large) e->in[i].tile = 0;
    // large-instance branches: four (or one per large instance when fewer).  Timing the branch
    // counts costs whole solves of the large instances, and four has been fastest wherever it
    // was timed (512-system ensemble: 1, 2, 4 branches)
    e->branches = std::max(1, std::min(4, (int)e->large.size()));
    if (hipStreamSynchronize(e->stream) != hipSuccess) return PARAMENT_STATUS_CUBLAS_FAILED;
    e->tuned = true;
    return PARAMENT_STATUS_SUCCESS;
}

Parament_ErrorCode ens_run(Ensemble* e) {
    if (!e->has_h) return PARAMENT_STATUS_NO_HAMILTONIAN;
    if (e->exec_mode == 1) return ens_enqueue_loop(e) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
    if (!e->tuned) {
        const Parament_ErrorCode err = ens_tune(e);
        if (err != PARAMENT_STATUS_SUCCESS) return err;
    }
    if (!e->gexec) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return PARAMENT_STATUS_CUBLAS_FAILED;
        const hipError_t r = ens_enqueue(e);
        const hipError_t r2 = hipStreamEndCapture(e->stream, &g);
        if (r != hipSuccess || r2 != hipSuccess || !g) { hipGetLastError(); if (g) hipGraphDestroy(g); return PARAMENT_STATUS_CUBLAS_FAILED; }
        const hipError_t r3 = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (r3 != hipSuccess) { hipGetLastError(); e->gexec = nullptr; return PARAMENT_STATUS_CUBLAS_FAILED; }
    }
    return hipGraphLaunch(e->gexec, e->stream) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
}

// ----- ensemble gradients -----
// Each instance's gate-fidelity gradient runs the single-problem gradient: the fused kernels
// for dims <= 32 (whole tree kept), batched GEMMs above.  Instances are spread over the side
// streams by longest-first greedy balancing of their arithmetic; one branch reuses one
// workspace for its instances, one after the other.  The whole solve is captured as a graph.
struct EnsGradLayout {
    bool fused = false;
    GradPlan p;
    StepGradPlan sp;
    size_t off[5] = {};   // element offsets: E (+ per-slice arrays), V, W, dg, gsum
    size_t total = 0;
};

EnsGradLayout ens_grad_layout(const EnsInst& x) {
    EnsGradLayout L;
    const size_t dd = (size_t)x.d * x.d;
    auto al = [](size_t n) { return (n + 63) / 64 * 64; };
    L.fused = pa::grad_supported((int)x.d) && (int)x.amps <= pa::kGradCoeffs;
    size_t sz[5];
    if (L.fused) {
        L.p = grad_plan(x.N, 0, 0);
        const size_t vw = L.p.chunk_mats + L.p.top_mats;
        sz[0] = (size_t)L.p.chunk * dd; sz[1] = vw * dd; sz[2] = vw * dd;
    } else {
        // per-slice arrays within 1 GiB per instance
        const int nwork = step_grad_work(x.M, power_form_ok(x.M, x.tau * x.hnorm));
        const unsigned cap = (unsigned)std::max<size_t>(1, ((size_t)1 << 30) / (nwork * dd * sizeof(float2)));
        L.sp = step_grad_plan(x.N, (int)x.d, sizeof(float2), cap, nwork);
        // the slices' S and Q_j kept for the derivative within 4 GiB per instance
        L.sp.keep = step_grad_keep(x.M, power_form_ok(x.M, x.tau * x.hnorm), x.N, (int)x.d, sizeof(float2), (size_t)4 << 30);
        // the power form on the scaled fp16 split, as the forward's large instances
        L.sp.split = x.M >= 5 && power_form_ok(x.M, x.tau * x.hnorm);
        sz[0] = ((size_t)x.N * (1 + L.sp.keep) + (size_t)nwork * L.sp.chunk) * dd;
        sz[1] = std::max<size_t>(L.sp.vmats, 1) * dd;
        sz[2] = L.sp.wmats * dd;
    }
    sz[3] = (size_t)x.amps * x.N;
    sz[4] = 4;
    for (int k = 0; k < 5; ++k) { L.off[k] = L.total; L.total += al(sz[k]); }
    return L;
}

Parament_ErrorCode ens_grad_setup(Ensemble* e) {
    if (e->grad_ready) return PARAMENT_STATUS_SUCCESS;
    // a set-up that failed part way may have left buffers behind
    hipFree(e->d_opsT); e->d_opsT = nullptr;
    for (auto& w : e->d_gws) { hipFree(w); w = nullptr; }
    size_t n_opsT = 0;
    for (EnsInst& x : e->in) { x.opsT_off = n_opsT; n_opsT += (size_t)x.amps * x.d * x.d; }
    if (hipMalloc(&e->d_opsT, n_opsT * sizeof(float2)) != hipSuccess) {
        hipGetLastError(); e->d_opsT = nullptr;
        return PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
    }
    for (const EnsInst& x : e->in)
        if (pa::launch_transpose<float>(e->d_ops + x.ops_off + (size_t)x.d * x.d, (int)x.amps, (int)x.d,
                                        e->d_opsT + x.opsT_off, e->stream) != hipSuccess)
            return PARAMENT_STATUS_CUBLAS_FAILED;
    // longest first onto the least loaded branch
    const int nb = kEnsSide;
    std::vector<int> order(e->in.size());
    std::vector<double> cost(e->in.size());
    for (size_t i = 0; i < e->in.size(); ++i) {
        const EnsInst& x = e->in[i];
        order[i] = (int)i;
        cost[i] = (double)x.d * x.d * x.d * x.N * (x.M + (x.d <= 32 ? 2.0 : 3.0 * x.M));
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    e->gbranch.assign(e->in.size(), 0);
    double load[kEnsSide] = {};
    size_t need[kEnsSide] = {};
    for (int i : order) {
        int k = 0;
        for (int q = 1; q < nb; ++q)
            if (load[q] < load[k]) k = q;
        e->gbranch[i] = k;
        load[k] += cost[i];
        need[k] = std::max(need[k], ens_grad_layout(e->in[i]).total);
    }
    for (int k = 0; k < nb; ++k) {
        if (!need[k]) continue;
        if (hipMalloc(&e->d_gws[k], need[k] * sizeof(float2)) != hipSuccess) {
            hipGetLastError(); e->d_gws[k] = nullptr;
            return PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
        }
    }
    e->grad_ready = true;
    return PARAMENT_STATUS_SUCCESS;
}

// Instance i's gradient on stream st in workspace ws; controls from e->d_c.
hipError_t ens_grad_instance(Ensemble* e, int i, float2* ws, const float2* tgt, float* fid, float* grad, hipStream_t st) {
    const EnsInst& x = e->in[i];
    const EnsGradLayout L = ens_grad_layout(x);
    pa::ChebArgs<float> ca;
    ca.ops = e->d_ops + x.ops_off; ca.ncoef = (int)x.amps; ca.d = (int)x.d; ca.N = x.N; ca.M = x.M;
    ca.s = (float)(2.0 / x.hnorm);
    ca.opsT = e->d_opsT + x.opsT_off;
    ca.ctl.c = e->d_c + x.c_off; ca.ctl.pts = (int)x.pts; ca.ctl.amps_in = (int)x.amps; ca.ctl.amps_set = (int)x.amps;
    ca.ctl.mode = e->quad; ca.ctl.magnus = 0; ca.ctl.tau = (float)x.tau;
    for (int k = 0; k <= x.M; ++k) { ca.are[k] = (float)x.are[k]; ca.aim[k] = (float)x.aim[k]; }
    ca.are[0] = (float)(x.are[0] - 1.0);
    if (power_form_ok(x.M, x.tau * x.hnorm)) {
        double mr[pa::kMaxCheb], mi[pa::kMaxCheb];
        power_coefficients(x.M, x.tau * x.hnorm, x.are, x.aim, mr, mi);
        ca.ps = 1;
        for (int k = 0; k <= x.M; ++k) { ca.pre[k] = (float)mr[k]; ca.pim[k] = (float)mi[k]; }
    }
    const float2* t = tgt + x.out_off;
    float* f = fid ? fid + i : nullptr;
    float* g = grad + x.c_off;
    if (L.fused)
        return grad_fused_enqueue<float>(ca, L.p, FusedGradBufs<float>{ws + L.off[0], ws + L.off[1], ws + L.off[2], ws + L.off[3],
                                                                      ws + L.off[4]},
                                         t, f, g, st, nullptr);
    StepGradBufs<float> sb;
    sb.E = ws + L.off[0]; sb.work = sb.E + (size_t)x.N * x.d * x.d; sb.V = ws + L.off[1]; sb.W = ws + L.off[2];
    sb.dg = ws + L.off[3]; sb.gsum = ws + L.off[4];
    sb.keep = L.sp.keep ? sb.work + (size_t)L.sp.nwork * L.sp.chunk * x.d * x.d : nullptr;
    return grad_stepwise_enqueue<float>(ca, L.sp, sb, t, f, g, st);
}

// All instances: branches forked from and joined to e->stream (graph mode), or one after the
// other on e->stream (loop mode).
hipError_t ens_grad_enqueue(Ensemble* e, const float2* tgt, float* fid, float* grad) {
    if (e->exec_mode == 1) {
        hipError_t r = hipSuccess;
        for (size_t i = 0; r == hipSuccess && i < e->in.size(); ++i)
            r = ens_grad_instance(e, (int)i, e->d_gws[e->gbranch[i]], tgt, fid, grad, e->stream);
        return r;
    }
    hipError_t r = hipEventRecord(e->fork_ev, e->stream);
    bool used[kEnsSide] = {};
    for (size_t i = 0; r == hipSuccess && i < e->in.size(); ++i) {
        const int k = e->gbranch[i];
        if (!used[k]) { r = hipStreamWaitEvent(e->side[k], e->fork_ev, 0); used[k] = true; }
        if (r == hipSuccess) r = ens_grad_instance(e, (int)i, e->d_gws[k], tgt, fid, grad, e->side[k]);
    }
    for (int k = 0; r == hipSuccess && k < kEnsSide; ++k) {
        if (!used[k]) continue;
        r = hipEventRecord(e->join_ev[k], e->side[k]);
        if (r == hipSuccess) r = hipStreamWaitEvent(e->stream, e->join_ev[k], 0);
    }
    return r;
}

// Gradients of all instances (controls already in e->d_c).  Graph mode captures the solve once
// per (targets, fidelity, gradient) pointer triple and replays it.
Parament_ErrorCode ens_grad_run(Ensemble* e, const float2* tgt, float* fid, float* grad) {
    if (!e->has_h) return PARAMENT_STATUS_NO_HAMILTONIAN;
    Parament_ErrorCode err = ens_grad_setup(e);
    if (err != PARAMENT_STATUS_SUCCESS) return err;
    if (e->exec_mode == 1)
        return ens_grad_enqueue(e, tgt, fid, grad) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
    if (!e->gexec_grad || e->gkey[0] != tgt || e->gkey[1] != fid || e->gkey[2] != grad) {
        if (e->gexec_grad) { hipGraphExecDestroy(e->gexec_grad); e->gexec_grad = nullptr; }
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return PARAMENT_STATUS_CUBLAS_FAILED;
        const hipError_t r = ens_grad_enqueue(e, tgt, fid, grad);
        const hipError_t r2 = hipStreamEndCapture(e->stream, &g);
        if (r != hipSuccess || r2 != hipSuccess || !g) { hipGetLastError(); if (g) hipGraphDestroy(g); return PARAMENT_STATUS_CUBLAS_FAILED; }
        const hipError_t r3 = hipGraphInstantiate(&e->gexec_grad, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (r3 != hipSuccess) { hipGetLastError(); e->gexec_grad = nullptr; return PARAMENT_STATUS_CUBLAS_FAILED; }
        e->gkey[0] = tgt; e->gkey[1] = fid; e->gkey[2] = grad;
    }
    return hipGraphLaunch(e->gexec_grad, e->stream) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
}

}  // namespace

// ======================================================================
// Exported C ABI (parament.h).  Complex64 contexts are Context<float>, complex128
// contexts Context<double> (reference parament.cpp:886-954).
// ======================================================================
#define CTX32(h) reinterpret_cast<Context<float>*>(h)
#define CTX64(h) reinterpret_cast<Context<double>*>(h)

extern "C" {

Parament_ErrorCode Parament_create(Parament_Context_f32** handle_p) {
    return create<float>(reinterpret_cast<Context<float>**>(handle_p));
}
Parament_ErrorCode Parament_destroy(Parament_Context_f32* handle) { return destroy<float>(CTX32(handle)); }
Parament_ErrorCode Parament_setHamiltonian(Parament_Context_f32* handle, cuComplex* H0, cuComplex* H1, unsigned int dim,
                                           unsigned int amps, bool use_magnus, Parament_QuadratureSpec quadrature_mode) {
    return set_hamiltonian<float>(CTX32(handle), H0, H1, dim, amps, use_magnus, (int)quadrature_mode);
}
Parament_ErrorCode Parament_equiprop(Parament_Context_f32* handle, cuComplex* carr, double dt, unsigned int pts,
                                     unsigned int amps, cuComplex* out) {
    return equiprop<float>(CTX32(handle), carr, dt, pts, amps, out);
}
int Parament_selectIterationCycles_fp32(double H_norm, double dt) { return select_fp32(H_norm, dt); }
Parament_ErrorCode Parament_setIterationCyclesManually(Parament_Context_f32* handle, unsigned int cycles) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    if (cycles < 3 || cycles % 2 == 0 || cycles >= (unsigned)pa::kMaxCheb) return CTX32(handle)->last_error = PARAMENT_STATUS_INVALID_VALUE;
    CTX32(handle)->mmax = (int)cycles;
    CTX32(handle)->manual = true;
    return CTX32(handle)->last_error = PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_automaticIterationCycles(Parament_Context_f32* handle) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    CTX32(handle)->mmax = 11;
    CTX32(handle)->manual = false;
    return CTX32(handle)->last_error = PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_peekAtLastError(Parament_Context_f32* handle) {
    return handle ? CTX32(handle)->last_error : PARAMENT_STATUS_INVALID_VALUE;
}

const char* Parament_errorMessage(Parament_ErrorCode errorCode) {
    switch (errorCode) {
        case PARAMENT_STATUS_SUCCESS: return "Success";
        case PARAMENT_STATUS_HOST_ALLOC_FAILED: return "Memory allocation on the host failed.";
        case PARAMENT_STATUS_DEVICE_ALLOC_FAILED: return "Memory allocation on the device failed.";
        case PARAMENT_STATUS_CUBLAS_INIT_FAILED: return "Failed to initialize the cuBLAS library.";
        case PARAMENT_STATUS_INVALID_VALUE: return "Invalid value.";
        case PARAMENT_STATUS_CUBLAS_FAILED: return "Failed to execute cuBLAS function.";
        case PARAMENT_STATUS_SELECT_SMALLER_DT: return "Timestep too large";
        case PARAMENT_STATUS_INVALID_QUADRATURE_SELECTION: return "Invalid quadrature selection.";
        case PARAMENT_STATUS_NO_HAMILTONIAN: return "No hamiltonian set";
        case PARAMENT_STATUS_WORKSPACE_TOO_SMALL: return "The problem needs more device memory than the fixed workspace holds.";
        default: return "Unknown error code";
    }
}

Parament_ErrorCode Parament_create_fp64(Parament_Context_f64** handle_p) {
    return create<double>(reinterpret_cast<Context<double>**>(handle_p));
}
Parament_ErrorCode Parament_destroy_fp64(Parament_Context_f64* handle) { return destroy<double>(CTX64(handle)); }
Parament_ErrorCode Parament_setHamiltonian_fp64(Parament_Context_f64* handle, cuDoubleComplex* H0, cuDoubleComplex* H1,
                                                unsigned int dim, unsigned int amps, bool use_magnus,
                                                Parament_QuadratureSpec quadrature_mode) {
    return set_hamiltonian<double>(CTX64(handle), H0, H1, dim, amps, use_magnus, (int)quadrature_mode);
}
Parament_ErrorCode Parament_equiprop_fp64(Parament_Context_f64* handle, cuDoubleComplex* carr, double dt,
                                          unsigned int pts, unsigned int amps, cuDoubleComplex* out) {
    return equiprop<double>(CTX64(handle), carr, dt, pts, amps, out);
}
Parament_ErrorCode Parament_setIterationCyclesManually_fp64(Parament_Context_f64* handle, unsigned int cycles) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    if (cycles < 3 || cycles % 2 == 0 || cycles >= (unsigned)pa::kMaxCheb) return CTX64(handle)->last_error = PARAMENT_STATUS_INVALID_VALUE;
    CTX64(handle)->mmax = (int)cycles;
    CTX64(handle)->manual = true;
    return CTX64(handle)->last_error = PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_automaticIterationCycles_fp64(Parament_Context_f64* handle) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    CTX64(handle)->mmax = 11;
    CTX64(handle)->manual = false;
    return CTX64(handle)->last_error = PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_peekAtLastError_fp64(Parament_Context_f64* handle) {
    return handle ? CTX64(handle)->last_error : PARAMENT_STATUS_INVALID_VALUE;
}
int Parament_selectIterationCycles_fp64(double H_norm, double dt) { return select_fp64(H_norm, dt); }

double OneNorm(cuComplex* mat, unsigned int dim) { return mat ? one_norm(mat, dim) : 0.0; }
double OneNorm_fp64(cuDoubleComplex* mat, unsigned int dim) { return mat ? one_norm(mat, dim) : 0.0; }

// ---- additions ----
Parament_ErrorCode Parament_setPrecisionMode(Parament_Context_f32* handle, Parament_PrecisionMode mode) {
    return handle ? set_precision(CTX32(handle), (int)mode) : PARAMENT_STATUS_INVALID_VALUE;
}
Parament_ErrorCode Parament_setPrecisionMode_fp64(Parament_Context_f64* handle, Parament_PrecisionMode mode) {
    return handle ? set_precision(CTX64(handle), (int)mode) : PARAMENT_STATUS_INVALID_VALUE;
}
int Parament_getPrecisionMode(Parament_Context_f32* handle) { return handle ? CTX32(handle)->precision : -1; }
int Parament_getPrecisionMode_fp64(Parament_Context_f64* handle) { return handle ? CTX64(handle)->precision : -1; }
int Parament_getIterationCycles(Parament_Context_f32* handle) { return handle ? CTX32(handle)->last_m : -1; }
int Parament_getIterationCycles_fp64(Parament_Context_f64* handle) { return handle ? CTX64(handle)->last_m : -1; }

Parament_ErrorCode Parament_equipropDevice(Parament_Context_f32* handle, const cuComplex* carr_dev, double dt,
                                           unsigned int pts, unsigned int amps, cuComplex* out_dev) {
    return equiprop_device<float>(CTX32(handle), carr_dev, dt, pts, amps, out_dev);
}
Parament_ErrorCode Parament_equipropDevice_fp64(Parament_Context_f64* handle, const cuDoubleComplex* carr_dev, double dt,
                                                unsigned int pts, unsigned int amps, cuDoubleComplex* out_dev) {
    return equiprop_device<double>(CTX64(handle), carr_dev, dt, pts, amps, out_dev);
}
Parament_ErrorCode Parament_equipropDeviceNoAlloc_fp64(Parament_Context_f64* handle, const cuDoubleComplex* carr_dev,
                                                       double dt, unsigned int pts, unsigned int amps,
                                                       cuDoubleComplex* out_dev) {
    return equiprop_device_no_alloc<double>(CTX64(handle), reinterpret_cast<const double2*>(carr_dev), dt, pts, amps,
                                            reinterpret_cast<double2*>(out_dev));
}
Parament_ErrorCode Parament_synchronize(Parament_Context_f32* handle) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    if (!CTX32(handle)->bound) return PARAMENT_STATUS_SUCCESS;
    return hipStreamSynchronize(CTX32(handle)->stream) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
}
Parament_ErrorCode Parament_synchronize_fp64(Parament_Context_f64* handle) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    if (!CTX64(handle)->bound) return PARAMENT_STATUS_SUCCESS;
    return hipStreamSynchronize(CTX64(handle)->stream) == hipSuccess ? PARAMENT_STATUS_SUCCESS : PARAMENT_STATUS_CUBLAS_FAILED;
}
Parament_ErrorCode Parament_setKernelPath(Parament_Context_f32* handle, int path) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    if (path != 0 && path != 1) return CTX32(handle)->last_error = PARAMENT_STATUS_INVALID_VALUE;
    CTX32(handle)->path = path;
    return CTX32(handle)->last_error = PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_setKernelPath_fp64(Parament_Context_f64* handle, int path) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    if (path != 0 && path != 1) return CTX64(handle)->last_error = PARAMENT_STATUS_INVALID_VALUE;
    CTX64(handle)->path = path;
    return CTX64(handle)->last_error = PARAMENT_STATUS_SUCCESS;
}

Parament_ErrorCode Parament_setStream(Parament_Context_f32* handle, void* hip_stream) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    auto* h = CTX32(handle);
    h->stream = (hipStream_t)hip_stream;
    h->user_stream = true;
    return h->last_error = PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_setStream_fp64(Parament_Context_f64* handle, void* hip_stream) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    auto* h = CTX64(handle);
    h->stream = (hipStream_t)hip_stream;
    h->user_stream = true;
    return h->last_error = PARAMENT_STATUS_SUCCESS;
}
void* Parament_getStream(Parament_Context_f32* handle) { return handle ? (void*)CTX32(handle)->stream : nullptr; }
void* Parament_getStream_fp64(Parament_Context_f64* handle) { return handle ? (void*)CTX64(handle)->stream : nullptr; }

Parament_ErrorCode Parament_setTiming(Parament_Context_f32* handle, int enable) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    CTX32(handle)->timing = enable != 0;
    return PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_setTiming_fp64(Parament_Context_f64* handle, int enable) {
    if (!handle) return PARAMENT_STATUS_INVALID_VALUE;
    CTX64(handle)->timing = enable != 0;
    return PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_getKernelTimes(Parament_Context_f32* handle, double* slice_ms, double* tree_ms) {
    return handle ? kernel_times(CTX32(handle), slice_ms, tree_ms) : PARAMENT_STATUS_INVALID_VALUE;
}
Parament_ErrorCode Parament_getKernelTimes_fp64(Parament_Context_f64* handle, double* slice_ms, double* tree_ms) {
    return handle ? kernel_times(CTX64(handle), slice_ms, tree_ms) : PARAMENT_STATUS_INVALID_VALUE;
}

Parament_ErrorCode Parament_gateFidelityGradient(Parament_Context_f32* handle, const cuComplex* carr, double dt,
                                               unsigned int pts, unsigned int amps, const cuComplex* target,
                                               float* fidelity, float* grad) {
    return gradient_host<float>(CTX32(handle), carr, dt, pts, amps, target, fidelity, grad);
}
Parament_ErrorCode Parament_gateFidelityGradient_fp64(Parament_Context_f64* handle, const cuDoubleComplex* carr, double dt,
                                                    unsigned int pts, unsigned int amps, const cuDoubleComplex* target,
                                                    double* fidelity, double* grad) {
    return gradient_host<double>(CTX64(handle), carr, dt, pts, amps, target, fidelity, grad);
}
Parament_ErrorCode Parament_gateFidelityGradientDevice(Parament_Context_f32* handle, const cuComplex* carr_dev, double dt,
                                                     unsigned int pts, unsigned int amps, const cuComplex* target_dev,
                                                     float* fidelity_dev, float* grad_dev) {
    return gradient_device<float>(CTX32(handle), carr_dev, dt, pts, amps, target_dev, fidelity_dev, grad_dev);
}
Parament_ErrorCode Parament_gateFidelityGradientDevice_fp64(Parament_Context_f64* handle, const cuDoubleComplex* carr_dev,
                                                          double dt, unsigned int pts, unsigned int amps,
                                                          const cuDoubleComplex* target_dev, double* fidelity_dev,
                                                          double* grad_dev) {
    return gradient_device<double>(CTX64(handle), carr_dev, dt, pts, amps, target_dev, fidelity_dev, grad_dev);
}
Parament_ErrorCode Parament_setGradientCheckpointing(Parament_Context_f32* handle, unsigned int segment_slices,
                                                     unsigned int chunk_slices) {
    return set_checkpointing(CTX32(handle), segment_slices, chunk_slices);
}
Parament_ErrorCode Parament_setGradientCheckpointing_fp64(Parament_Context_f64* handle, unsigned int segment_slices,
                                                          unsigned int chunk_slices) {
    return set_checkpointing(CTX64(handle), segment_slices, chunk_slices);
}
size_t Parament_getGradientWorkspace(Parament_Context_f32* handle) { return handle ? CTX32(handle)->grad_workspace : 0; }
size_t Parament_getGradientWorkspace_fp64(Parament_Context_f64* handle) { return handle ? CTX64(handle)->grad_workspace : 0; }

Parament_ErrorCode Parament_ensembleCreate(Parament_Ensemble** ensemble_p, unsigned int count, const unsigned int* dims,
                                           const unsigned int* amps, const unsigned int* pts, const double* dt,
                                           Parament_QuadratureSpec quadrature_mode) {
    return ens_create(reinterpret_cast<Ensemble**>(ensemble_p), count, dims, amps, pts, dt, (int)quadrature_mode);
}
Parament_ErrorCode Parament_ensembleDestroy(Parament_Ensemble* ensemble) { return ens_destroy(reinterpret_cast<Ensemble*>(ensemble)); }
Parament_ErrorCode Parament_ensembleSetHamiltonians(Parament_Ensemble* ensemble, const cuComplex* H) {
    if (!ensemble) return PARAMENT_STATUS_INVALID_VALUE;
    auto* e = reinterpret_cast<Ensemble*>(ensemble);
    return e->last_error = ens_set_hamiltonians(e, H);
}
Parament_ErrorCode Parament_ensembleSetExecution(Parament_Ensemble* ensemble, int mode) {
    if (!ensemble) return PARAMENT_STATUS_INVALID_VALUE;
    auto* e = reinterpret_cast<Ensemble*>(ensemble);
    if (mode != 0 && mode != 1) return e->last_error = PARAMENT_STATUS_INVALID_VALUE;
    e->exec_mode = mode;
    return e->last_error = PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_ensembleTune(Parament_Ensemble* ensemble) {
    if (!ensemble) return PARAMENT_STATUS_INVALID_VALUE;
    auto* e = reinterpret_cast<Ensemble*>(ensemble);
    return e->last_error = ens_tune(e);
}
Parament_ErrorCode Parament_ensembleGetTuning(Parament_Ensemble* ensemble, int* config) {
    if (!ensemble || !config) return PARAMENT_STATUS_INVALID_VALUE;
    auto* e = reinterpret_cast<Ensemble*>(ensemble);
    // classes 16, 32 and 64 (the 48 class is tuned like the others but not reported)
    const int reported[3] = {0, 1, 3};
    for (int q = 0; q < 3; ++q) {
        const EnsCfg& cf = e->cfg[reported[q]];
        config[3 * q] = cf.T; config[3 * q + 1] = cf.logb; config[3 * q + 2] = cf.grid_mult;
    }
    config[9] = e->branches;
    return PARAMENT_STATUS_SUCCESS;
}
Parament_ErrorCode Parament_ensembleEquiprop(Parament_Ensemble* ensemble, const cuComplex* carr, cuComplex* out) {
    if (!ensemble) return PARAMENT_STATUS_INVALID_VALUE;
    auto* e = reinterpret_cast<Ensemble*>(ensemble);
    if (!carr || !out) return e->last_error = PARAMENT_STATUS_INVALID_VALUE;
    if (hipSetDevice(e->device) != hipSuccess) return e->last_error = PARAMENT_FAIL;
    if (e->n_c && hipMemcpyAsync(e->d_c, carr, e->n_c * sizeof(float2), hipMemcpyHostToDevice, e->stream) != hipSuccess)
        return e->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    Parament_ErrorCode err = ens_run(e);
    if (err == PARAMENT_STATUS_SUCCESS &&
        (hipMemcpyAsync(out, e->d_out, e->n_out * sizeof(float2), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
         hipStreamSynchronize(e->stream) != hipSuccess))
        err = PARAMENT_STATUS_CUBLAS_FAILED;
    if (err != PARAMENT_STATUS_SUCCESS) hipStreamSynchronize(e->stream);
    return e->last_error = err;
}
Parament_ErrorCode Parament_ensembleEquipropDevice(Parament_Ensemble* ensemble, const cuComplex* carr_dev, cuComplex* out_dev) {
    if (!ensemble) return PARAMENT_STATUS_INVALID_VALUE;
    auto* e = reinterpret_cast<Ensemble*>(ensemble);
    if (!carr_dev || !out_dev) return e->last_error = PARAMENT_STATUS_INVALID_VALUE;
    if (hipSetDevice(e->device) != hipSuccess) return e->last_error = PARAMENT_FAIL;
    if (e->n_c && hipMemcpyAsync(e->d_c, carr_dev, e->n_c * sizeof(float2), hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
        return e->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    Parament_ErrorCode err = ens_run(e);
    if (err == PARAMENT_STATUS_SUCCESS &&
        hipMemcpyAsync(out_dev, e->d_out, e->n_out * sizeof(float2), hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
        err = PARAMENT_STATUS_CUBLAS_FAILED;
    return e->last_error = err;
}
Parament_ErrorCode Parament_ensembleSynchronize(Parament_Ensemble* ensemble) {
    if (!ensemble) return PARAMENT_STATUS_INVALID_VALUE;
    return hipStreamSynchronize(reinterpret_cast<Ensemble*>(ensemble)->stream) == hipSuccess ? PARAMENT_STATUS_SUCCESS
                                                                                            : PARAMENT_STATUS_CUBLAS_FAILED;
}
Parament_ErrorCode Parament_ensemblePeekAtLastError(Parament_Ensemble* ensemble) {
    return ensemble ? reinterpret_cast<Ensemble*>(ensemble)->last_error : PARAMENT_STATUS_INVALID_VALUE;
}
Parament_ErrorCode Parament_ensembleGateFidelityGradient(Parament_Ensemble* ensemble, const cuComplex* carr,
                                                         const cuComplex* targets, float* fidelity, float* grad) {
    if (!ensemble) return PARAMENT_STATUS_INVALID_VALUE;
    auto* e = reinterpret_cast<Ensemble*>(ensemble);
    if (!carr || !targets || !grad) return e->last_error = PARAMENT_STATUS_INVALID_VALUE;
    if (hipSetDevice(e->device) != hipSuccess) return e->last_error = PARAMENT_FAIL;
    if ((!e->d_tgt && hipMalloc(&e->d_tgt, e->n_out * sizeof(float2)) != hipSuccess) ||
        (!e->d_fid && hipMalloc(&e->d_fid, e->in.size() * sizeof(float)) != hipSuccess) ||
        (!e->d_grad && hipMalloc(&e->d_grad, std::max<size_t>(e->n_c, 1) * sizeof(float)) != hipSuccess)) {
        hipGetLastError();
        return e->last_error = PARAMENT_STATUS_DEVICE_ALLOC_FAILED;
    }
    if ((e->n_c && hipMemcpyAsync(e->d_c, carr, e->n_c * sizeof(float2), hipMemcpyHostToDevice, e->stream) != hipSuccess) ||
        hipMemcpyAsync(e->d_tgt, targets, e->n_out * sizeof(float2), hipMemcpyHostToDevice, e->stream) != hipSuccess)
        return e->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    Parament_ErrorCode err = ens_grad_run(e, e->d_tgt, e->d_fid, e->d_grad);
    if (err == PARAMENT_STATUS_SUCCESS &&
        (hipMemcpyAsync(grad, e->d_grad, e->n_c * sizeof(float), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
         (fidelity && hipMemcpyAsync(fidelity, e->d_fid, e->in.size() * sizeof(float), hipMemcpyDeviceToHost, e->stream) !=
                          hipSuccess) ||
         hipStreamSynchronize(e->stream) != hipSuccess))
        err = PARAMENT_STATUS_CUBLAS_FAILED;
    if (err != PARAMENT_STATUS_SUCCESS) hipStreamSynchronize(e->stream);
    return e->last_error = err;
}
Parament_ErrorCode Parament_ensembleGateFidelityGradientDevice(Parament_Ensemble* ensemble, const cuComplex* carr_dev,
                                                               const cuComplex* targets_dev, float* fidelity_dev,
                                                               float* grad_dev) {
    if (!ensemble) return PARAMENT_STATUS_INVALID_VALUE;
    auto* e = reinterpret_cast<Ensemble*>(ensemble);
    if (!carr_dev || !targets_dev || !grad_dev) return e->last_error = PARAMENT_STATUS_INVALID_VALUE;
    if (hipSetDevice(e->device) != hipSuccess) return e->last_error = PARAMENT_FAIL;
    if (e->n_c && hipMemcpyAsync(e->d_c, carr_dev, e->n_c * sizeof(float2), hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
        return e->last_error = PARAMENT_STATUS_CUBLAS_FAILED;
    return e->last_error = ens_grad_run(e, reinterpret_cast<const float2*>(targets_dev), fidelity_dev, grad_dev);
}

void device_info(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) { hipGetLastError(); n = 0; }
    printf("Number of devices: %d\n", n);
    for (int i = 0; i < n; ++i) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, i) != hipSuccess) continue;
        printf("Device %d: %s (%s)\n", i, p.name, p.gcnArchName);
        printf("  Compute units: %d, clock %d kHz, wavefront %d\n", p.multiProcessorCount, p.clockRate, p.warpSize);
        printf("  Global memory: %.1f GiB, LDS per block: %zu KiB, L2: %d KiB\n",
               p.totalGlobalMem / 1073741824.0, p.sharedMemPerBlock / 1024, p.l2CacheSize / 1024);
    }
    fflush(stdout);
}

}  // extern "C"
