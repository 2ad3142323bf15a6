// Mixed-precision Chebyshev slice kernel (gfx950 bf16 matrix cores + FP64 first-order terms).
//
// The slice propagator is split as (parament.cpp:569-652 recurrence, rearranged exactly):
//     U_j - I = (a_0 - 1) I + a_1 s X_j + [ s X_j b_1' - 2 b_2 ],   b_1' = s X_j b_2 - b_3
// where b_k are the Clenshaw states of the same degree-M Chebyshev series (b_1 = b_1' + a_1 I).
// The first two terms carry the O(tau |H|) part of the propagator; they are formed in FP64
// from the FP64 slice Hamiltonian, because any rounding there recurs identically in every
// slice and accumulates linearly over the product.  The bracket is O((tau |H|)^2) and is
// evaluated by the recurrence in FP32 arithmetic whose GEMMs run on the bf16 matrix cores
// with three-way operand splits: x = x0 + x1 + x2 (bf16 each, ~24 significant bits), and
// each real product keeps the six terms x_a y_b with a + b <= 2, accumulated in FP32.
// Complex products use the 4M form with four accumulators (ArBr, AiBi, ArBi, AiBr).
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

// x ~= h + m + l with bf16 terms; the residuals are exact in the wider type.
__device__ __forceinline__ void split3(double x, __bf16& h, __bf16& m, __bf16& l) {
    h = (__bf16)(float)x;
    const double r1 = x - (double)(float)h;
    m = (__bf16)(float)r1;
    const double r2 = r1 - (double)(float)m;
    l = (__bf16)(float)r2;
}
__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
    h = (__bf16)x;
    const float r1 = x - (float)h;
    m = (__bf16)r1;
    const float r2 = r1 - (float)m;
    l = (__bf16)r2;
}

// Accumulator element v of a 32x32x16 tile: row 8*(v/4) + 4*(lane/32) + v%4, column lane%32.
__device__ __forceinline__ int mrow(int row0, int v) { return row0 + 8 * (v >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (v & 3); }
__device__ __forceinline__ int mcol(int col0) { return col0 + (threadIdx.x & 31); }

struct MTile {
    f32x16 re, im;
};

// C = X * B on the wave's 32x32 tile: X planes row-major (A operand), B planes transposed
// (Bt[col][k]); plane p = 2 * term + {0: real, 1: imag}.
template <int D>
__device__ __forceinline__ void wave_cgemm_bf16x3(const __bf16* __restrict__ Xp, const __bf16* __restrict__ Bp,
                                                  int row0, int col0, MTile& c) {
    constexpr int S = MShape<D>::S, PL = MShape<D>::PL;
    const int lane = threadIdx.x & 63;
    const int r = lane & 31, kh = lane >> 5;
    f32x16 rr, ii, ri, ir;
#pragma unroll
    for (int v = 0; v < 16; ++v) { rr[v] = 0.f; ii[v] = 0.f; ri[v] = 0.f; ir[v] = 0.f; }
#pragma unroll
    for (int ks = 0; ks < D / 16; ++ks) {
        const int ka = (row0 + r) * S + 16 * ks + 8 * kh;
        const int kb = (col0 + r) * S + 16 * ks + 8 * kh;
        bf16x8 A[6], B[6];
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            A[p] = *reinterpret_cast<const bf16x8*>(Xp + p * PL + ka);
            B[p] = *reinterpret_cast<const bf16x8*>(Bp + p * PL + kb);
        }
        // term pairs (a, b) with a + b <= 2
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            const int ta = (t == 0) ? 0 : (t == 1) ? 0 : (t == 2) ? 1 : (t == 3) ? 0 : (t == 4) ? 1 : 2;
            const int tb = (t == 0) ? 0 : (t == 1) ? 1 : (t == 2) ? 0 : (t == 3) ? 2 : (t == 4) ? 1 : 0;
            rr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[2 * ta], B[2 * tb], rr, 0, 0, 0);
            ri = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[2 * ta], B[2 * tb + 1], ri, 0, 0, 0);
            ii = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[2 * ta + 1], B[2 * tb + 1], ii, 0, 0, 0);
            ir = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[2 * ta + 1], B[2 * tb], ir, 0, 0, 0);
        }
    }
    c.re = rr - ii;
    c.im = ri + ir;
}

template <int D>
__global__ __launch_bounds__(64 * MShape<D>::G) void cheb_mixed_kernel(MixedArgs a) {
    constexpr int G = MShape<D>::G, S = MShape<D>::S, PL = MShape<D>::PL, W = MShape<D>::W;
    __shared__ __attribute__((aligned(16))) __bf16 Xp[6 * PL];
    __shared__ __attribute__((aligned(16))) __bf16 Bp[6 * PL];
    __shared__ double2 coef[kMaxCoeffs];
    const int tid = threadIdx.x, wave = tid >> 6, kh = (tid & 63) >> 5;
    const int row0 = 32 * (wave / W), col0 = 32 * (wave % W);
    const int d = a.d, dd = d * d, M = a.M;
    const float sf = (float)a.s;  // exact: the host picks Hnorm so that 2/Hnorm is a float

    auto x64 = [&](int i, int k) -> double2 {
        double2 x = a.ops[i * d + k];
        for (int t = 0; t < a.ncoef; ++t) x = cadd(x, cmul(coef[t], a.ops[(size_t)(t + 1) * dd + i * d + k]));
        return x;
    };
    // b (accumulator layout, FP32) -> transposed split planes Bt[col][row]
    auto write_b = [&](const MTile& b) {
        const int col = mcol(col0);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int rbase = row0 + 8 * g + 4 * kh;
            bf16x4 h[2], m[2], l[2];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                __bf16 th, tm, tl;
                split3(b.re[4 * g + q], th, tm, tl);
                h[0][q] = th; m[0][q] = tm; l[0][q] = tl;
                split3(b.im[4 * g + q], th, tm, tl);
                h[1][q] = th; m[1][q] = tm; l[1][q] = tl;
            }
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                *reinterpret_cast<bf16x4*>(Bp + (0 + c) * PL + col * S + rbase) = h[c];
                *reinterpret_cast<bf16x4*>(Bp + (2 + c) * PL + col * S + rbase) = m[c];
                *reinterpret_cast<bf16x4*>(Bp + (4 + c) * PL + col * S + rbase) = l[c];
            }
        }
    };

    for (int j = blockIdx.x; j < a.N; j += gridDim.x) {
        for (int t = tid; t < a.ncoef; t += 64 * G) coef[t] = slice_coefficient(a.ctl, t, j);
        __syncthreads();
        for (int e = tid; e < D * D; e += 64 * G) {
            const int i = e / D, k = e % D;
            double2 x = mk<double>(0, 0);
            if (i < d && k < d) x = x64(i, k);
            __bf16 h, m, l;
            split3(x.x, h, m, l);
            Xp[0 * PL + i * S + k] = h; Xp[2 * PL + i * S + k] = m; Xp[4 * PL + i * S + k] = l;
            split3(x.y, h, m, l);
            Xp[1 * PL + i * S + k] = h; Xp[3 * PL + i * S + k] = m; Xp[5 * PL + i * S + k] = l;
        }
        __syncthreads();

        MTile b1, b2, c;
        {
            const float aMr = (float)a.are[M], aMi = (float)a.aim[M];
            const float aM1r = (float)a.are[M - 1], aM1i = (float)a.aim[M - 1];
            const int col = mcol(col0);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = mrow(row0, v);
                const bool dg = row == col;
                const int o = row * S + col;
                const float xr = (float)Xp[o] + (float)Xp[2 * PL + o] + (float)Xp[4 * PL + o];
                const float xi = (float)Xp[PL + o] + (float)Xp[3 * PL + o] + (float)Xp[5 * PL + o];
                b2.re[v] = dg ? aMr : 0.f;
                b2.im[v] = dg ? aMi : 0.f;
                // s a_M X + a_{M-1} I
                b1.re[v] = sf * (aMr * xr - aMi * xi) + (dg ? aM1r : 0.f);
                b1.im[v] = sf * (aMr * xi + aMi * xr) + (dg ? aM1i : 0.f);
            }
        }
        write_b(b1);
        __syncthreads();

        // b_k = s X b_{k+1} - b_{k+2} + a_k I for k = M-2 .. 2, then b_1' (no a_1 I)
        for (int k = M - 2; k >= 1; --k) {
            wave_cgemm_bf16x3<D>(Xp, Bp, row0, col0, c);
            const float akr = (k >= 2) ? (float)a.are[k] : 0.f, aki = (k >= 2) ? (float)a.aim[k] : 0.f;
            const int col = mcol(col0);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const bool dg = mrow(row0, v) == col;
                const float nr = sf * c.re[v] - b2.re[v] + (dg ? akr : 0.f);
                const float ni = sf * c.im[v] - b2.im[v] + (dg ? aki : 0.f);
                b2.re[v] = b1.re[v]; b2.im[v] = b1.im[v];
                b1.re[v] = nr; b1.im[v] = ni;
            }
            __syncthreads();
            write_b(b1);
            __syncthreads();
        }
        // R = s X b_1' - 2 b_2 ;  E = (a_0 - 1) I + a_1 s X + R  (FP64)
        wave_cgemm_bf16x3<D>(Xp, Bp, row0, col0, c);
        {
            const double a0m1 = a.are[0];
            const double2 a1s = mk<double>(a.are[1] * a.s, a.aim[1] * a.s);
            float2* out = a.out + (size_t)j * dd;
            const int col = mcol(col0);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = mrow(row0, v);
                if (row < d && col < d) {
                    const double rr = (double)(sf * c.re[v] - 2.f * b2.re[v]);
                    const double ri = (double)(sf * c.im[v] - 2.f * b2.im[v]);
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
