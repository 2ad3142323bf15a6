// Mixed-precision Chebyshev slice kernel (gfx950 bf16 matrix cores + FP64 first-order terms).
//
// The slice propagator is split as (parament.cpp:569-652 recurrence, rearranged exactly):
//     U_j - I = (a_0 - 1) I + a_1 s X_j + [ s X_j b_1' - 2 b_2 ],   b_1' = s X_j b_2 - b_3
// where b_k are the Clenshaw states of the same degree-M Chebyshev series (b_1 = b_1' + a_1 I).
// The first two terms carry the O(tau |H|) part of the propagator; they are formed in FP64
// from the FP64 slice Hamiltonian, because any rounding there recurs identically in every
// slice and accumulates linearly over the product.  The bracket is O((tau |H|)^2) and is
// evaluated by the recurrence in FP32 arithmetic whose GEMMs run on the bf16 matrix cores.
//
// Exact matrix-core accumulation: each GEMM operand is scaled by a power of two and split in
// fixed point, x 2^(8-e) = s0 + s1 + s2 + r with s0 an integer (|s0| <= 256), s1 = r1 2^-8,
// s2 = r2 2^-16 (|r1|, |r2| <= 128 integers), all exact in bf16 and |r| <= 2^-17.  Products
// s0 t0 are integers below 2^16 and their sum over K = 64 (real and imaginary parts of the 4M
// form combined) stays below 2^24: that accumulator is exact on the matrix cores regardless of
// how the instruction rounds.  The five products with a non-leading slice (s0 t1, s1 t0, s0 t2,
// s1 t1, s2 t0) share a second accumulator whose sum needs more than 24 bits only below
// 2^-31 of the leading one, so whatever rounding the instruction applies stays there.  The two
// accumulators are recombined with IEEE arithmetic: the remaining roundings are the operand
// splits and the FP32 epilogues, all round-to-nearest.
// The slice result E_j is stored as complex64 for the FP64 product tree.
#include "kernels.hpp"

namespace pa {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

bool mixed_supported(int d) { return d >= 1 && d <= 64; }

template <int D> struct MShape {
    static constexpr int W = D / 32;         // wave tiles per side
    static constexpr int G = W * W;          // waves per slice
    static constexpr int S = D + 8;          // plane row stride (bf16 elements), 16-byte rows
    static constexpr int PL = D * S;         // plane size
};

// Fixed-point split of a = x 2^(8-e) (|x| < 2^e, so |a| < 256) into three bf16 numbers with
// a = s0 + s1 + s2 + O(2^-17): s0 = r0, s1 = r1 2^-8, s2 = r2 2^-16 with integers
// |r0| <= 256, |r1|, |r2| <= 128 (each exactly representable in bf16).
__device__ __forceinline__ void split_fixed(double x, double scale, __bf16& s0, __bf16& s1, __bf16& s2) {
    const double a = x * scale;                 // |a| <= 256
    const double r0 = rint(a);
    const double b = (a - r0) * 256.0;          // exact, |b| <= 128
    const double r1 = rint(b);
    const double r2 = rint((b - r1) * 256.0);   // |r2| <= 128
    s0 = (__bf16)(float)r0; s1 = (__bf16)(float)(r1 * (1.0 / 256.0)); s2 = (__bf16)(float)(r2 * (1.0 / 65536.0));
}
__device__ __forceinline__ void split_fixed(float x, float scale, __bf16& s0, __bf16& s1, __bf16& s2) {
    const float a = x * scale;
    const float r0 = rintf(a);
    const float b = (a - r0) * 256.f;
    const float r1 = rintf(b);
    const float r2 = rintf((b - r1) * 256.f);
    s0 = (__bf16)r0; s1 = (__bf16)(r1 * (1.f / 256.f)); s2 = (__bf16)(r2 * (1.f / 65536.f));
}

// Same split with the last slice rounded stochastically: r2 = floor(b + u), u uniform in
// [0, 1), so the representation error has zero mean.  Used for the slice Hamiltonian, whose
// constant entries would otherwise carry the same rounding error into every slice.
__device__ __forceinline__ void split_fixed_sr(double x, double scale, float u, __bf16& s0, __bf16& s1, __bf16& s2) {
    const double a = x * scale;
    const double r0 = rint(a);
    const double b = (a - r0) * 256.0;
    const double r1 = rint(b);
    const double r2 = floor((b - r1) * 256.0 + (double)u);
    s0 = (__bf16)(float)r0; s1 = (__bf16)(float)(r1 * (1.0 / 256.0)); s2 = (__bf16)(float)(r2 * (1.0 / 65536.0));
}

// Reproducible 32 random bits from (slice, element pair): an integer hash.  Each byte gives
// one dither value (k + 1/2) / 256, k uniform in 0..255 (mean 1/2: unbiased rounding).
__device__ __forceinline__ unsigned hash_bits(unsigned j, unsigned e) {
    unsigned h = j * 0x9E3779B1u ^ (e * 0x85EBCA77u + 0xC2B2AE3Du);
    h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
    return h;
}

// Exponent e with m < 2^e for the largest modulus m of a matrix (0 for a zero matrix).
__device__ __forceinline__ int scale_exp(double m) {
    int e = 0;
    if (m > 0) frexp(m, &e);   // m = f 2^e, f in [0.5, 1)
    return e;                   // |x| <= m < 2^e
}

// Accumulator element v of a 32x32x16 tile: row 8*(v/4) + 4*(lane/32) + v%4, column lane%32.
__device__ __forceinline__ int mrow(int row0, int v) { return row0 + 8 * (v >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (v & 3); }
__device__ __forceinline__ int mcol(int col0) { return col0 + (threadIdx.x & 31); }

struct MTile {
    f32x16 re, im;
};

// Accumulators of one 32x32 complex tile: [0] = leading-slice products (exact), [1] = the
// products involving a non-leading slice (carry their own 2^-8 / 2^-16 scales).
struct LTile {
    f32x16 re[2], im[2];
};

__device__ __forceinline__ bf16x8 negate(bf16x8 v) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

// C = X * B on the wave's 32x32 tile with exact level accumulation.  X slice planes are
// row-major (A operand), B slice planes transposed (Bt[col][k]); plane p = 2 * slice + {0: re, 1: im}.
template <int D>
__device__ __forceinline__ void wave_cgemm_exact(const __bf16* __restrict__ Xp, const __bf16* __restrict__ Bp,
                                                 int row0, int col0, LTile& c) {
    constexpr int S = MShape<D>::S, PL = MShape<D>::PL;
    const int lane = threadIdx.x & 63;
    const int r = lane & 31, kh = lane >> 5;
#pragma unroll
    for (int L = 0; L < 2; ++L)
#pragma unroll
        for (int v = 0; v < 16; ++v) { c.re[L][v] = 0.f; c.im[L][v] = 0.f; }
#pragma unroll
    for (int ks = 0; ks < D / 16; ++ks) {
        const int ka = (row0 + r) * S + 16 * ks + 8 * kh;
        const int kb = (col0 + r) * S + 16 * ks + 8 * kh;
        bf16x8 Ar[3], Ai[3], An[3], Br[3], Bi[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            Ar[p] = *reinterpret_cast<const bf16x8*>(Xp + (2 * p) * PL + ka);
            Ai[p] = *reinterpret_cast<const bf16x8*>(Xp + (2 * p + 1) * PL + ka);
            An[p] = negate(Ai[p]);
            Br[p] = *reinterpret_cast<const bf16x8*>(Bp + (2 * p) * PL + kb);
            Bi[p] = *reinterpret_cast<const bf16x8*>(Bp + (2 * p + 1) * PL + kb);
        }
        // slice pairs (p, q) with p + q = L <= 2
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            const int p = (t == 0) ? 0 : (t == 1) ? 0 : (t == 2) ? 1 : (t == 3) ? 0 : (t == 4) ? 1 : 2;
            const int q = (t == 0) ? 0 : (t == 1) ? 1 : (t == 2) ? 0 : (t == 3) ? 2 : (t == 4) ? 1 : 0;
            const int L = (p + q) > 0 ? 1 : 0;
            c.re[L] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ar[p], Br[q], c.re[L], 0, 0, 0);
            c.im[L] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ar[p], Bi[q], c.im[L], 0, 0, 0);
            c.re[L] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(An[p], Bi[q], c.re[L], 0, 0, 0);
            c.im[L] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ai[p], Br[q], c.im[L], 0, 0, 0);
        }
    }
}

template <int D>
__global__ __launch_bounds__(64 * MShape<D>::G) void cheb_mixed_kernel(MixedArgs a) {
    constexpr int G = MShape<D>::G, S = MShape<D>::S, PL = MShape<D>::PL, W = MShape<D>::W;
    __shared__ __attribute__((aligned(16))) __bf16 Xp[6 * PL];
    __shared__ __attribute__((aligned(16))) __bf16 Bp[6 * PL];
    __shared__ double2 coef[kMaxCoeffs];
    __shared__ float red[G];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, kh = lane >> 5;
    const int row0 = 32 * (wave / W), col0 = 32 * (wave % W);
    const int d = a.d, dd = d * d, M = a.M;
    const float sf = (float)a.s;  // exact: the host picks Hnorm so that 2/Hnorm is a float

    auto x64 = [&](int i, int k) -> double2 {
        double2 x = a.ops[i * d + k];
        for (int t = 0; t < a.ncoef; ++t) x = cadd(x, cmul(coef[t], a.ops[(size_t)(t + 1) * dd + i * d + k]));
        return x;
    };
    // b (accumulator layout, FP32) -> transposed slice planes Bt[col][row], scaled by the
    // work-group max published in red[] before the caller's barrier; returns the exponent.
    auto wave_max = [&](const MTile& b) {
        float m = 0.f;
#pragma unroll
        for (int v = 0; v < 16; ++v) m = fmaxf(m, fmaxf(fabsf(b.re[v]), fabsf(b.im[v])));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if (lane == 0) red[wave] = m;
    };
    auto write_b = [&](const MTile& b) -> int {
        float m = red[0];
        for (int w = 1; w < G; ++w) m = fmaxf(m, red[w]);
        const int eb = scale_exp((double)m);
        const float sc = ldexpf(1.f, 8 - eb);   // maps |b| < 2^eb into |.| < 256
        const int col = mcol(col0);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int rbase = row0 + 8 * g + 4 * kh;
            bf16x4 h[2], m1[2], l[2];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                __bf16 t0, t1, t2;
                split_fixed(b.re[4 * g + q], sc, t0, t1, t2);
                h[0][q] = t0; m1[0][q] = t1; l[0][q] = t2;
                split_fixed(b.im[4 * g + q], sc, t0, t1, t2);
                h[1][q] = t0; m1[1][q] = t1; l[1][q] = t2;
            }
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                *reinterpret_cast<bf16x4*>(Bp + (0 + c) * PL + col * S + rbase) = h[c];
                *reinterpret_cast<bf16x4*>(Bp + (2 + c) * PL + col * S + rbase) = m1[c];
                *reinterpret_cast<bf16x4*>(Bp + (4 + c) * PL + col * S + rbase) = l[c];
            }
        }
        return eb;
    };
    for (int j = blockIdx.x; j < a.N; j += gridDim.x) {
        for (int t = tid; t < a.ncoef; t += 64 * G) coef[t] = slice_coefficient(a.ctl, t, j);
        __syncthreads();
        // X_j in FP64, split in fixed point against the bound |X| <= |H0| + sum_t |c'_t| |O_t|
        // (per-operator largest moduli from the host), last slice rounded stochastically.
        double bound = a.opmax[0];
        for (int t = 0; t < a.ncoef; ++t) bound += hypot(coef[t].x, coef[t].y) * a.opmax[t + 1];
        const int ex = scale_exp(bound);
        const double xsc = ldexp(1.0, 8 - ex);
        typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
        for (int e2 = tid; e2 < D * D / 2; e2 += 64 * G) {
            const int i = (2 * e2) / D, k = (2 * e2) % D;   // elements (i, k), (i, k + 1)
            double2 x0 = mk<double>(0, 0), x1 = mk<double>(0, 0);
            if (i < d && k < d) x0 = x64(i, k);
            if (i < d && k + 1 < d) x1 = x64(i, k + 1);
            const unsigned hsh = hash_bits(j, e2);
            const float u[4] = {(float)(hsh & 0xFFu) * (1.f / 256.f) + (1.f / 512.f),
                                (float)((hsh >> 8) & 0xFFu) * (1.f / 256.f) + (1.f / 512.f),
                                (float)((hsh >> 16) & 0xFFu) * (1.f / 256.f) + (1.f / 512.f),
                                (float)(hsh >> 24) * (1.f / 256.f) + (1.f / 512.f)};
            bf16x2 p[6];
            __bf16 s0, s1, s2;
            split_fixed_sr(x0.x, xsc, u[0], s0, s1, s2); p[0][0] = s0; p[2][0] = s1; p[4][0] = s2;
            split_fixed_sr(x0.y, xsc, u[1], s0, s1, s2); p[1][0] = s0; p[3][0] = s1; p[5][0] = s2;
            split_fixed_sr(x1.x, xsc, u[2], s0, s1, s2); p[0][1] = s0; p[2][1] = s1; p[4][1] = s2;
            split_fixed_sr(x1.y, xsc, u[3], s0, s1, s2); p[1][1] = s0; p[3][1] = s1; p[5][1] = s2;
#pragma unroll
            for (int q = 0; q < 6; ++q) *reinterpret_cast<bf16x2*>(Xp + q * PL + i * S + k) = p[q];
        }
        __syncthreads();

        MTile b1, b2;
        LTile acc;
        {
            const float aMr = (float)a.are[M], aMi = (float)a.aim[M];
            const float aM1r = (float)a.are[M - 1], aM1i = (float)a.aim[M - 1];
            const float xinv = ldexpf(1.f, ex - 8);
            const int col = mcol(col0);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = mrow(row0, v);
                const bool dg = row == col;
                const int o = row * S + col;
                const float xr = (((float)Xp[4 * PL + o] + (float)Xp[2 * PL + o]) + (float)Xp[o]) * xinv;
                const float xi = (((float)Xp[5 * PL + o] + (float)Xp[3 * PL + o]) + (float)Xp[PL + o]) * xinv;
                b2.re[v] = dg ? aMr : 0.f;
                b2.im[v] = dg ? aMi : 0.f;
                // s a_M X + a_{M-1} I
                b1.re[v] = sf * (aMr * xr - aMi * xi) + (dg ? aM1r : 0.f);
                b1.im[v] = sf * (aMr * xi + aMi * xr) + (dg ? aM1i : 0.f);
            }
        }
        wave_max(b1);
        __syncthreads();
        int eb = write_b(b1);
        __syncthreads();

        // b_k = s X b_{k+1} - b_{k+2} + a_k I for k = M-2 .. 2, then b_1' (no a_1 I).
        // For k = 2, 3 the diagonal update adds a_k in FP64: a rounded a_k would put the same
        // error into every slice (|a_2|, |a_3| ~ (tau |H|)^2 / 8, (tau |H|)^3 / 48); for k >= 4 the
        // FP32 rounding of a_k is below 2^-24 |a_4| and negligible over any product length.
        for (int k = M - 2; k >= 1; --k) {
            wave_cgemm_exact<D>(Xp, Bp, row0, col0, acc);
            const float sc = sf * ldexpf(1.f, ex + eb - 16);   // s times the operand scales
            const int col = mcol(col0);
            if (k == 2 || k == 3) {
                const double akr = a.are[k], aki = a.aim[k];
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const float cr = sc * (acc.re[0][v] + acc.re[1][v]);
                    const float ci = sc * (acc.im[0][v] + acc.im[1][v]);
                    float nr = cr - b2.re[v], ni = ci - b2.im[v];
                    if (mrow(row0, v) == col) {
                        nr = (float)((double)cr - (double)b2.re[v] + akr);
                        ni = (float)((double)ci - (double)b2.im[v] + aki);
                    }
                    b2.re[v] = b1.re[v]; b2.im[v] = b1.im[v];
                    b1.re[v] = nr; b1.im[v] = ni;
                }
            } else {
                const float akr = (k >= 2) ? (float)a.are[k] : 0.f, aki = (k >= 2) ? (float)a.aim[k] : 0.f;
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const bool dg = mrow(row0, v) == col;
                    const float nr = sc * (acc.re[0][v] + acc.re[1][v]) - b2.re[v] + (dg ? akr : 0.f);
                    const float ni = sc * (acc.im[0][v] + acc.im[1][v]) - b2.im[v] + (dg ? aki : 0.f);
                    b2.re[v] = b1.re[v]; b2.im[v] = b1.im[v];
                    b1.re[v] = nr; b1.im[v] = ni;
                }
            }
            wave_max(b1);
            __syncthreads();
            eb = write_b(b1);
            __syncthreads();
        }
        // R = s X b_1' - 2 b_2 ;  E = (a_0 - 1) I + a_1 s X + R  (FP64)
        wave_cgemm_exact<D>(Xp, Bp, row0, col0, acc);
        {
            const double a0m1 = a.are[0];
            const double2 a1s = mk<double>(a.are[1] * a.s, a.aim[1] * a.s);
            const double f = ldexp(1.0, ex + eb - 16);
            float2* out = a.out + (size_t)j * dd;
            const int col = mcol(col0);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = mrow(row0, v);
                if (row < d && col < d) {
                    const double cr = ((double)acc.re[0][v] + (double)acc.re[1][v]) * f;
                    const double ci = ((double)acc.im[0][v] + (double)acc.im[1][v]) * f;
                    const double rr = a.s * cr - 2.0 * (double)b2.re[v];
                    const double ri = a.s * ci - 2.0 * (double)b2.im[v];
                    const double2 t = cmul(a1s, x64(row, col));
                    out[row * d + col] = make_float2((float)(t.x + rr + (row == col ? a0m1 : 0.0)), (float)(t.y + ri));
                }
            }
        }
        __syncthreads();
    }
}

hipError_t launch_cheb_mixed(const MixedArgs& a, hipStream_t st) {
    if (a.d <= 0 || a.d > 64 || a.N <= 0 || a.M < 3 || a.M >= kMaxCheb || a.ncoef > kMaxCoeffs)
        return hipErrorInvalidValue;
    if (a.d <= 32) hipLaunchKernelGGL((cheb_mixed_kernel<32>), dim3(a.N), dim3(64 * MShape<32>::G), 0, st, a);
    else hipLaunchKernelGGL((cheb_mixed_kernel<64>), dim3(a.N), dim3(64 * MShape<64>::G), 0, st, a);
    return hipGetLastError();
}

}  // namespace pa
