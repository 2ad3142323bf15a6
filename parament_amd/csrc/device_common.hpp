// Device-side building blocks shared by the equiprop kernels (gfx950 / CDNA4).
//
//  * complex helpers on float2 / double2 (the API's interleaved complex layout),
//  * slice_coefficient(): the per-slice control coefficient c'_kj of every
//    quadrature mode, evaluated on the fly from the raw samples
//    (reference control_expansion.cu:27-170, fused here into the consumer),
//  * WaveCgemm: one wavefront computes a (16*TM) x (16*TN) tile of a complex
//    product C = A * B whose operands sit in LDS as separate real / imaginary
//    planes, using the 16x16x4 FP32 or FP64 matrix-core instruction (4M
//    decomposition: Cr = ArBr - AiBi, Ci = ArBi + AiBr, FP32/FP64 accumulate).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pa {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <class R> struct CplxOf;
template <> struct CplxOf<float> { typedef float2 type; };
template <> struct CplxOf<double> { typedef double2 type; };
template <class R> using cplx = typename CplxOf<R>::type;

template <class R> __host__ __device__ inline cplx<R> mk(R x, R y) { cplx<R> c; c.x = x; c.y = y; return c; }
template <class C> __host__ __device__ inline C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }
template <class C> __host__ __device__ inline C csub(C a, C b) { a.x -= b.x; a.y -= b.y; return a; }
template <class C> __host__ __device__ inline C cmul(C a, C b) {
    C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}

// Quadrature modes (values of Parament_QuadratureSpec).
enum : int { QUAD_NONE = 0x00000000, QUAD_MIDPOINT = 0x01000000, QUAD_SIMPSON = 0x02000000 };

// Control description handed to every kernel that assembles slice Hamiltonians.
template <class R> struct Controls {
    const cplx<R>* c;   // raw samples, c[k * pts + j], k < amps_in
    int pts;            // samples per control
    int amps_in;        // controls given to this equiprop call
    int amps_set;       // control Hamiltonians set (rows of H beyond amps_in are zero-amplitude)
    int mode;           // QUAD_*
    int magnus;         // 1: Simpson means + commutator coefficients
    R tau;              // slice length (2 dt for Simpson / Magnus)
};

// Number of operator coefficients per slice (controls, then Magnus terms).
__host__ __device__ inline int n_coeffs(int amps_set, int magnus) {
    return magnus ? 2 * amps_set + amps_set * (amps_set - 1) / 2 : amps_set;
}

template <class R> __device__ inline cplx<R> sample(const Controls<R>& q, int k, int j) {
    return (k < q.amps_in) ? q.c[(size_t)k * q.pts + j] : mk<R>(0, 0);
}

// Coefficient of operator t (0-based, excluding H0) for slice j.
//   none:     c[k, j]
//   midpoint: 0.5 * (c[k, j] + c[k, j+1])                  (control_expansion.cu:111-113)
//   simpson:  ((c[k,2j] + 4 c[k,2j+1]) + c[k,2j+2]) * (1/6)  (control_expansion.cu:147-149;
//             the reference's complex divide by (6,0) reduces to this multiply)
//   magnus:   rows [0, A): Simpson means; [A, 2A): i tau/12 (c[k,2j+2] - c[k,2j]);
//             [2A, ...): i tau/12 (c_a[2j] c_b[2j+2] - c_a[2j+2] c_b[2j]) for a < b in
//             lexicographic pair order (control_expansion.cu:38-59).
template <class R> __device__ inline cplx<R> slice_coefficient(const Controls<R>& q, int t, int j) {
    const int A = q.amps_set;
    if (!q.magnus) {
        if (q.mode == QUAD_NONE) return sample(q, t, j);
        if (q.mode == QUAD_MIDPOINT) {
            cplx<R> s = cadd(sample(q, t, j), sample(q, t, j + 1));
            return mk<R>(R(0.5) * s.x, R(0.5) * s.y);
        }
    }
    if (t < A) {
        cplx<R> c0 = sample(q, t, 2 * j), c1 = sample(q, t, 2 * j + 1), c2 = sample(q, t, 2 * j + 2);
        cplx<R> s = cadd(cadd(c0, mk<R>(R(4) * c1.x, R(4) * c1.y)), c2);
        const R sixth = R(1) / R(6);
        return mk<R>(s.x * sixth, s.y * sixth);
    }
    const cplx<R> fac = mk<R>(0, q.tau / R(12));
    if (t < 2 * A) {
        const int k = t - A;
        return cmul(csub(sample(q, k, 2 * j + 2), sample(q, k, 2 * j)), fac);
    }
    // pair index -> (a, b), a < b, lexicographic
    int p = t - 2 * A, a = 0;
    while (p >= A - 1 - a) { p -= A - 1 - a; ++a; }
    const int b = a + 1 + p;
    cplx<R> v = csub(cmul(sample(q, a, 2 * j), sample(q, b, 2 * j + 2)),
                     cmul(sample(q, a, 2 * j + 2), sample(q, b, 2 * j)));
    return cmul(v, fac);
}

// ---------------------------------------------------------------------------------------
// Matrix-core primitive: 16x16x4, A[i][k] with i = lane%16, k = lane/16; B[k][j] with
// k = lane/16, j = lane%16; accumulator layout: see acc_row().
template <class R> struct Mma;
template <> struct Mma<float> {
    typedef f32x4 acc;
    static __device__ __forceinline__ acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct Mma<double> {
    typedef f64x4 acc;
    static __device__ __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// A complex wave tile: TM x TN blocks of 16x16, real and imaginary accumulators.
template <class R, int TM, int TN> struct CTile {
    typename Mma<R>::acc re[TM][TN];
    typename Mma<R>::acc im[TM][TN];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int m = 0; m < TM; ++m)
#pragma unroll
            for (int n = 0; n < TN; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) { re[m][n][v] = R(0); im[m][n][v] = R(0); }
    }
};

// C(tile at row0, col0) = A * B over K = KD (ZERO), or C += A * B (!ZERO), operands in LDS
// planes with row strides SA, SB.  kd (a multiple of 4, at most KD) shortens the inner
// dimension at run time, for a last chunk that is zero past kd.
template <class R, int TM, int TN, int KD, int SA, int SB, bool ZERO = true>
__device__ __forceinline__ void wave_cgemm(const R* __restrict__ Ar, const R* __restrict__ Ai,
                                           const R* __restrict__ Br, const R* __restrict__ Bi,
                                           int row0, int col0, CTile<R, TM, TN>& c, int kd = KD) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
    if (ZERO) c.zero();
#pragma unroll 4
    for (int k0 = 0; k0 < kd; k0 += 4) {
        R ar[TM], ai[TM], an[TM], br[TN], bi[TN];
#pragma unroll
        for (int m = 0; m < TM; ++m) {
            const int idx = (row0 + 16 * m + r) * SA + k0 + q;
            ar[m] = Ar[idx]; ai[m] = Ai[idx]; an[m] = -ai[m];
        }
#pragma unroll
        for (int n = 0; n < TN; ++n) {
            const int idx = (k0 + q) * SB + col0 + 16 * n + r;
            br[n] = Br[idx]; bi[n] = Bi[idx];
        }
#pragma unroll
        for (int m = 0; m < TM; ++m)
#pragma unroll
            for (int n = 0; n < TN; ++n) c.re[m][n] = Mma<R>::op(ar[m], br[n], c.re[m][n]);
#pragma unroll
        for (int m = 0; m < TM; ++m)
#pragma unroll
            for (int n = 0; n < TN; ++n) c.im[m][n] = Mma<R>::op(ar[m], bi[n], c.im[m][n]);
#pragma unroll
        for (int m = 0; m < TM; ++m)
#pragma unroll
            for (int n = 0; n < TN; ++n) c.re[m][n] = Mma<R>::op(an[m], bi[n], c.re[m][n]);
#pragma unroll
        for (int m = 0; m < TM; ++m)
#pragma unroll
            for (int n = 0; n < TN; ++n) c.im[m][n] = Mma<R>::op(ai[m], br[n], c.im[m][n]);
    }
}

// Row / column of accumulator element v of block (m, n) for this lane.  The FP32
// instruction gives each lane four consecutive rows, C[4*(lane/16) + v][lane%16]; the
// FP64 instruction interleaves them, C[4*v + lane/16][lane%16].
template <class R>
__device__ __forceinline__ int acc_row(int row0, int m, int v) {
    const int q = (threadIdx.x & 63) >> 4;
    return row0 + 16 * m + (sizeof(R) == 8 ? 4 * v + q : 4 * q + v);
}
__device__ __forceinline__ int acc_col(int col0, int n) { return col0 + 16 * n + (threadIdx.x & 15); }

// LDS row padding (elements) of the A-operand and B-operand planes; chosen so the
// 16x16x4 operand reads of a half-wave fall into distinct banks.
template <class R> struct Pad;
template <> struct Pad<float> { static constexpr int A = 2, B = 16; };
template <> struct Pad<double> { static constexpr int A = 1, B = 8; };

// Work split of one D x D product (D = 16, 32, 48, 64): (D/16)^2 waves, each owning one 16x16
// output tile, row-major over the tile grid (D = 16: a single wave).
template <int D> struct Shape {
    static_assert(D % 16 == 0 && D >= 16 && D <= 64, "16x16 tiles");
    static constexpr int W = D / 16;
    static constexpr int G = W * W;
    static constexpr int T = 1;
    static __device__ __forceinline__ int row0(int w) { return 16 * (w / W); }
    static __device__ __forceinline__ int col0(int w) { return 16 * (w % W); }
};

}  // namespace pa
