// Mixed-precision Chebyshev slice kernel (gfx950 bf16 matrix cores + FP64 first-order terms),
// with the lowest levels of the pairwise product tree fused in.
//
// The slice propagator is split as (parament.cpp:569-652 recurrence, rearranged exactly):
//     U_j - I = (a_0 - 1) I + a_1 s X_j + [ s X_j b_1' - 2 b_2 ],   b_1' = s X_j b_2 - b_3
// where b_k are the Clenshaw states of the same degree-M Chebyshev series (b_1 = b_1' + a_1 I).
// The first two terms carry the O(tau |H|) part of the propagator; they are formed in FP64
// from the FP64 slice Hamiltonian, because any rounding there recurs identically in every
// slice and accumulates linearly over the product.  For the same reason the diagonal
// coefficients a_2, a_3 are added in FP64.  The bracket is O((tau |H|)^2) and is evaluated by
// the recurrence in FP32 arithmetic whose GEMMs run on the bf16 matrix cores.
//
// Split products: every FP32 operand x is written as three bf16 numbers, s0 = bf16(x),
// s1 = bf16(x - s0), s2 = bf16(x - s0 - s1) (round to nearest; both differences are exact in
// FP32), so x = s0 + s1 + s2 within 2^-27 |x|.  A product is the sum of the six slice products
// s_p t_q with p + q <= 2 (the three dropped ones are below 2^-26 of the leading one), each
// exact in FP32 and accumulated by the matrix cores in FP32: the leading products in one
// accumulator, the five smaller ones in a second, added at the end.  This is an FP32-accurate
// complex GEMM at six bf16 products per real product.
//
// Tree levels: work-group w computes the slices [w 2^F, (w+1) 2^F) one after the other and
// multiplies them on chip in the reference's pairing (later slice on the left), on deviations
// from identity: (I + A)(I + B) = I + (A + B + A B), the product A B being a split product.  A
// slice beyond N is a zero node, and combining with it returns the other node unchanged, as
// the reference's odd-carry rule does.  The node E = U - I of level F is stored as complex64
// for the FP64 levels above.
//
// Work split: one work-group per block of slices, one wave per 16x32 output tile (8 waves at
// dim 64, two per SIMD inside the one-work-group-per-CU LDS budget; 2 waves at dim 32), so that
// matrix-core, vector and LDS work of different waves overlap; the two 16x16 blocks of a wave
// share each A-operand fragment, which keeps the LDS reads per matrix-core operation low.
#include "kernels.hpp"

namespace pa {

typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

bool mixed_supported(int d) { return d >= 1 && d <= 64; }

// Wave tiles of 16 x (16 TN): TN column blocks share each A fragment (register blocking).
template <int D> struct MShape {
    static constexpr int TN = 2;             // 16-column blocks per wave
    static constexpr int WR = D / 16;        // wave rows
    static constexpr int WC = D / (16 * TN); // wave columns
    static constexpr int G = WR * WC;        // waves per slice
    // Plane layout.  Dim 64: unpadded 128-byte rows whose 16-byte chunks are permuted by
    // chunk ^ (row & 7), so that the 16-byte fragment reads of 16 rows and the 8-byte
    // transposed tile writes both spread over the banks (measured: 1.6x the fragment-read rate
    // of 144-byte padded rows).  Dim 32: rows padded by 16 bytes.
    static constexpr bool SWZ = (D == 64);
    static constexpr int S = SWZ ? D : D + 8;  // row stride (bf16 elements)
    static constexpr int PL = D * S;           // plane size
    // element offset of (row, col); col and col + 1 .. col + 7 within one aligned group of 8
    // stay contiguous
    static __device__ __forceinline__ int off(int row, int col) {
        if constexpr (SWZ) return row * D + ((((col >> 3) ^ row) & 7) << 3) + (col & 7);
        else return row * S + col;
    }
};

// Three-way bf16 split of two FP32 values (one packed conversion per slice pair).
__device__ __forceinline__ void split3(float x0, float x1, bf16x2& s0, bf16x2& s1, bf16x2& s2) {
    s0[0] = (__bf16)x0; s0[1] = (__bf16)x1;
    const float r0 = x0 - (float)s0[0], r1 = x1 - (float)s0[1];
    s1[0] = (__bf16)r0; s1[1] = (__bf16)r1;
    const float q0 = r0 - (float)s1[0], q1 = r1 - (float)s1[1];
    s2[0] = (__bf16)q0; s2[1] = (__bf16)q1;
}

// Accumulator element v of a 16x16x32 tile: row 4*(lane/16) + v, column lane%16.
__device__ __forceinline__ int mrow(int row0, int v) { return row0 + 4 * ((threadIdx.x & 63) >> 4) + v; }
__device__ __forceinline__ int mcol(int col0, int n) { return col0 + 16 * n + (threadIdx.x & 15); }

template <int TN> struct MTile {
    f32x4v re[TN], im[TN];
};

// Accumulators of the wave's complex tiles: [L][n], L = 0 leading-slice products, L = 1 the
// products involving a smaller slice.
template <int TN> struct LTile {
    f32x4v re[2][TN], im[2][TN];
};

__device__ __forceinline__ bf16x8 negate(bf16x8 v) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

// C = A * B on the wave's 16 x 16TN tile.  A slice planes are row-major, B slice planes
// transposed (Bt[col][k]); plane 2 * slice + {0: re, 1: im}.  Operand fragments of the
// 16x16x32 instruction: lane l holds A[l%16][8(l/16) .. +7] and Bt[l%16][8(l/16) .. +7].
// LV is the largest p + q of the slice pairs summed: 2 (six products, FP32 accurate), 1 (three
// products, relative error ~2^-16) or 0 (the leading product only, ~2^-8).
template <int D, int LV = 2>
__device__ __forceinline__ void wave_cgemm_split(const __bf16* __restrict__ Ap, const __bf16* __restrict__ Bp,
                                                 int row0, int col0, LTile<MShape<D>::TN>& c) {
    constexpr int PL = MShape<D>::PL, TN = MShape<D>::TN;
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int L = 0; L < 2; ++L)
#pragma unroll
        for (int n = 0; n < TN; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) { c.re[L][n][v] = 0.f; c.im[L][n][v] = 0.f; }
#pragma unroll 1
    for (int ks = 0; ks < D / 32; ++ks) {
        const int ka = MShape<D>::off(row0 + r, 32 * ks + 8 * kq);
        // B slice q (both column blocks) against the A slices p <= LV - q
#pragma unroll
        for (int q = 0; q <= LV; ++q) {
            bf16x8 Br[TN], Bi[TN];
#pragma unroll
            for (int n = 0; n < TN; ++n) {
                const int kb = MShape<D>::off(col0 + 16 * n + r, 32 * ks + 8 * kq);
                Br[n] = *reinterpret_cast<const bf16x8*>(Bp + (2 * q) * PL + kb);
                Bi[n] = *reinterpret_cast<const bf16x8*>(Bp + (2 * q + 1) * PL + kb);
            }
#pragma unroll
            for (int p = 0; p + q <= LV; ++p) {
                const bf16x8 Ar = *reinterpret_cast<const bf16x8*>(Ap + (2 * p) * PL + ka);
                const bf16x8 Ai = *reinterpret_cast<const bf16x8*>(Ap + (2 * p + 1) * PL + ka);
                const bf16x8 An = negate(Ai);
                const int L = (p + q) > 0 ? 1 : 0;
#pragma unroll
                for (int n = 0; n < TN; ++n) {
                    c.re[L][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ar, Br[n], c.re[L][n], 0, 0, 0);
                    c.im[L][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ar, Bi[n], c.im[L][n], 0, 0, 0);
                    c.re[L][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(An, Bi[n], c.re[L][n], 0, 0, 0);
                    c.im[L][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ai, Br[n], c.im[L][n], 0, 0, 0);
                }
            }
        }
    }
}

// Leading NL bf16 slices of two FP32 values (split3 for NL = 3).
template <int NL>
__device__ __forceinline__ void splitn(float x0, float x1, bf16x2 (&s)[3]) {
    if constexpr (NL == 3) {
        split3(x0, x1, s[0], s[1], s[2]);
    } else {
        s[0][0] = (__bf16)x0; s[0][1] = (__bf16)x1;
        if constexpr (NL == 2) {
            s[1][0] = (__bf16)(x0 - (float)s[0][0]);
            s[1][1] = (__bf16)(x1 - (float)s[0][1]);
        }
    }
}

// Tile (accumulator layout) -> split planes.  TRANSPOSED: the natural write of the layout,
// P[col][row] = b(row, col), four consecutive rows per lane in one 8-byte store (this plane is
// the B operand for the product (.) b, or the A operand for b^T (.)).  Otherwise row-major
// P[row][col] = b(row, col), 2-byte stores.  Only the leading NL slices are written: a plane
// read only by products of split level NL - 1 needs no more.
template <int D, bool TRANSPOSED, int NL = 3>
__device__ __forceinline__ void write_split(const MTile<MShape<D>::TN>& b, __bf16* __restrict__ P, int row0, int col0) {
    constexpr int PL = MShape<D>::PL, TN = MShape<D>::TN;
    const int rbase = mrow(row0, 0);
#pragma unroll
    for (int n = 0; n < TN; ++n) {
        const int col = mcol(col0, n);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const f32x4v& x = c ? b.im[n] : b.re[n];
            bf16x2 lo[3], hi[3];
            splitn<NL>(x[0], x[1], lo);
            splitn<NL>(x[2], x[3], hi);
#pragma unroll
            for (int t = 0; t < NL; ++t) {
                if (TRANSPOSED) {
                    const bf16x4 v = {lo[t][0], lo[t][1], hi[t][0], hi[t][1]};
                    *reinterpret_cast<bf16x4*>(P + (2 * t + c) * PL + MShape<D>::off(col, rbase)) = v;
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        P[(2 * t + c) * PL + MShape<D>::off(rbase + q, col)] = (q < 2 ? lo[t] : hi[t])[q & 1];
                }
            }
        }
    }
}

// A slice coefficient in two halves, so that its control samples can be loaded long before
// the value is needed: load the samples, then form exactly what slice_coefficient() returns
// (non-Magnus modes).
struct RawCoef {
    double2 c0, c1, c2;
};
__device__ __forceinline__ RawCoef coef_load(const Controls<double>& q, int t, int j) {
    RawCoef r;
    if (q.mode == QUAD_NONE) { r.c0 = sample(q, t, j); r.c1 = r.c2 = mk<double>(0, 0); }
    else if (q.mode == QUAD_MIDPOINT) { r.c0 = sample(q, t, j); r.c1 = sample(q, t, j + 1); r.c2 = mk<double>(0, 0); }
    else { r.c0 = sample(q, t, 2 * j); r.c1 = sample(q, t, 2 * j + 1); r.c2 = sample(q, t, 2 * j + 2); }
    return r;
}
__device__ __forceinline__ double2 coef_finish(const Controls<double>& q, const RawCoef& r) {
    if (q.mode == QUAD_NONE) return r.c0;
    if (q.mode == QUAD_MIDPOINT) {
        const double2 s = cadd(r.c0, r.c1);
        return mk<double>(0.5 * s.x, 0.5 * s.y);
    }
    const double2 s = cadd(cadd(r.c0, mk<double>(4.0 * r.c1.x, 4.0 * r.c1.y)), r.c2);
    const double sixth = 1.0 / 6.0;
    return mk<double>(s.x * sixth, s.y * sixth);
}

template <int D, int F, bool PS>
__global__ __launch_bounds__(64 * MShape<D>::G) void cheb_mixed_kernel(MixedArgs a) {
    constexpr int G = MShape<D>::G, PL = MShape<D>::PL, WC = MShape<D>::WC;
    constexpr int TN = MShape<D>::TN;
    constexpr int NB = 1 << F;
    typedef MTile<TN> Tile;
    __shared__ __attribute__((aligned(16))) __bf16 Xp[6 * PL];
    __shared__ __attribute__((aligned(16))) __bf16 Bp[6 * PL];
    // X_j - (s0 + s1 + s2), the FP64 residual of the split; also the FP32 stage of the tree products
    __shared__ __attribute__((aligned(16))) float2 Rx[D * (D + 2)];
    __shared__ double2 coef[kMaxCoeffs];
    constexpr int SR = D + 2;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int row0 = 16 * (wave / WC), col0 = 16 * TN * (wave % WC);
    const int d = a.d, dd = d * d, M = a.M;
    const float sf = (float)a.s;  // exact: the host picks Hnorm so that 2/Hnorm is a float
    const int n_nodes = (a.N + NB - 1) >> F;

    // Coefficients of the next slice are prefetched during the current one when every
    // coefficient has its own thread and the mode has no Magnus terms.
    const bool prefetch = !a.ctl.magnus && a.ncoef <= 64 * G;
    bool coef_ready = false;   // coef[] already holds the coefficients of the slice about to run

    // E_j = U_j - I of slice j into e (accumulator layout, FP32); jn is the slice this
    // work-group runs next (or -1).
    auto slice = [&](int j, int jn, Tile& e) __attribute__((always_inline)) {
        if (!coef_ready)
            for (int t = tid; t < a.ncoef; t += 64 * G) coef[t] = slice_coefficient(a.ctl, t, j);
        __syncthreads();
        if constexpr (PS) {
            // Y_j = tau (H0 + sum_t c'_t O_t) in FP64, as 2 x 2 blocks: wave region rg (16 x 16
            // elements), lane (bi, bk) = (l / 8, l % 8) forms rows i, i + 1 and columns k, k + 1, so
            // that both Y (Xp) and Z = Y^T (Bp) take two 4-byte stores per row pair, and the residual
            // of the FP32 sum ((s2 + s1) + s0) that the power form keeps in registers goes to Rx.
            constexpr int RT = D / 16;
#pragma unroll 1
            for (int it = 0; it < RT * RT / G; ++it) {
                const int rg = wave + G * it, lane = tid & 63;
                const int i = 16 * (rg / RT) + 2 * (lane >> 3), k = 16 * (rg % RT) + 2 * (lane & 7);
                bool v[4];
                double2 x[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int ii = i + (e >> 1), kk = k + (e & 1);
                    v[e] = ii < d && kk < d;
                    x[e] = v[e] ? a.ops[ii * d + kk] : mk<double>(0, 0);
                }
                for (int t0 = 0; t0 < a.ncoef; t0 += 2) {
                    double2 o[2][4];
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const size_t at = (size_t)(t0 + q + 1) * dd + (i + (e >> 1)) * d + k + (e & 1);
                            o[q][e] = (v[e] && t0 + q < a.ncoef) ? a.ops[at] : mk<double>(0, 0);
                        }
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        if (t0 + q < a.ncoef)
#pragma unroll
                            for (int e = 0; e < 4; ++e) x[e] = cadd(x[e], cmul(coef[t0 + q], o[q][e]));
                }
                const double tau = a.ctl.tau;
#pragma unroll
                for (int e = 0; e < 4; ++e) { x[e].x *= tau; x[e].y *= tau; }
                // slices of the row pairs: r[c][row][limb] (c: re / im), each bf16x2 = columns k, k + 1
                bf16x2 r[2][2][3];
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const double2 &xa = x[2 * h], &xb = x[2 * h + 1];
                        split3((float)(c ? xa.y : xa.x), (float)(c ? xb.y : xb.x), r[c][h][0], r[c][h][1], r[c][h][2]);
                    }
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int t = 0; t < 3; ++t) {
                        __bf16* P = Xp + (2 * t + c) * PL;
                        *reinterpret_cast<bf16x2*>(P + MShape<D>::off(i, k)) = r[c][0][t];
                        *reinterpret_cast<bf16x2*>(P + MShape<D>::off(i + 1, k)) = r[c][1][t];
                        __bf16* Q = Bp + (2 * t + c) * PL;
                        const bf16x2 z0 = {r[c][0][t][0], r[c][1][t][0]}, z1 = {r[c][0][t][1], r[c][1][t][1]};
                        *reinterpret_cast<bf16x2*>(Q + MShape<D>::off(k, i)) = z0;
                        *reinterpret_cast<bf16x2*>(Q + MShape<D>::off(k + 1, i)) = z1;
                    }
                auto res = [&](double xv, int c, int h, int col) -> float {
                    return (float)(xv - (double)(((float)r[c][h][2][col] + (float)r[c][h][1][col]) + (float)r[c][h][0][col]));
                };
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const double2 &xa = x[2 * h], &xb = x[2 * h + 1];
                    *reinterpret_cast<float4*>(Rx + (i + h) * SR + k) =
                        make_float4(res(xa.x, 0, h, 0), res(xa.y, 1, h, 0), res(xb.x, 0, h, 1), res(xb.y, 1, h, 1));
                }
            }
        } else {
            // X_j = H0 + sum_t c'_t O_t in FP64: each thread issues the operator loads of an element
            // pair together before the arithmetic, so their latencies overlap.
#pragma unroll 1
            for (int it = 0; it < D * D / 2 / (64 * G); ++it) {
                const int e2 = tid + it * 64 * G;
                const int i = (2 * e2) / D, k = (2 * e2) % D;   // elements (i, k), (i, k + 1)
                const bool v0 = i < d && k < d, v1 = i < d && k + 1 < d;
                double2 x0 = mk<double>(0, 0), x1 = mk<double>(0, 0);
                if (v0) x0 = a.ops[i * d + k];
                if (v1) x1 = a.ops[i * d + k + 1];
                for (int t0 = 0; t0 < a.ncoef; t0 += 4) {
                    double2 o0[4], o1[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const size_t base = (size_t)(t0 + q + 1) * dd + i * d + k;
                        o0[q] = (v0 && t0 + q < a.ncoef) ? a.ops[base] : mk<double>(0, 0);
                        o1[q] = (v1 && t0 + q < a.ncoef) ? a.ops[base + 1] : mk<double>(0, 0);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (t0 + q < a.ncoef) {
                            x0 = cadd(x0, cmul(coef[t0 + q], o0[q]));
                            x1 = cadd(x1, cmul(coef[t0 + q], o1[q]));
                        }
                }
                bf16x2 p[6];
                split3((float)x0.x, (float)x1.x, p[0], p[2], p[4]);
                split3((float)x0.y, (float)x1.y, p[1], p[3], p[5]);
#pragma unroll
                for (int q = 0; q < 6; ++q) *reinterpret_cast<bf16x2*>(Xp + q * PL + MShape<D>::off(i, k)) = p[q];
                // residuals for the FP64 first-order term (the sum of the three slices is exact in FP64)
                auto res = [&](double x, int h, int c) -> float {
                    return (float)(x - (((double)(float)p[4 + c][h] + (double)(float)p[2 + c][h]) + (double)(float)p[c][h]));
                };
                Rx[i * SR + k] = make_float2(res(x0.x, 0, 0), res(x0.y, 0, 1));
                Rx[i * SR + k + 1] = make_float2(res(x1.x, 1, 0), res(x1.y, 1, 1));
            }
        }
        __syncthreads();
        // coef[] is free from here on: start loading the next slice's control samples
        const bool pre = prefetch && jn >= 0 && jn < a.N;
        RawCoef raw;
        if (pre && tid < a.ncoef) raw = coef_load(a.ctl, tid, jn);

        if constexpr (PS) {
            // Power form of the same polynomial, Y = tau X, coefficients m_k = are/aim[k]
            // (m_0 - 1 in are[0]):  E = (m_0 - 1) I + m_1 Y + R,  R = Y^2 Q_1,
            // Q_J = m_{2J} I + m_{2J+1} Y,  Q_j = m_{2j} I + m_{2j+1} Y + Y^2 Q_{j+1}  (J = (M-1)/2):
            // (M + 1) / 2 products.  Q_1 (the largest terms) is formed in FP64.  The product
            // Y^2 Q_{j+1} enters E with weight ~ |Y|^(2j+2) / (2j+2)!, so the inner ones may drop
            // slice pairs (levels chosen by the host); Y Y and Y^2 Q_1 are always FP32 accurate.
            //
            // The matrix cores form P_A P_B^T of two row-major planes, and the natural (8-byte)
            // write of a result tile C is the plane C^T.  With Xp = Y and Bp = Z = Y^T, the
            // product Z Z (Bp, Xp) therefore writes naturally as the plane Y^2, and each
            // Y^2 Q(Y) takes the natural plane of Q(Y): every plane write is a natural one.
            const int J = (M - 1) / 2;
            Tile ye, q;
            LTile<TN> acc;
#pragma unroll
            for (int n = 0; n < TN; ++n) {
                const int col = mcol(col0, n);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int o = MShape<D>::off(mrow(row0, v), col);
                    ye.re[n][v] = ((float)Xp[4 * PL + o] + (float)Xp[2 * PL + o]) + (float)Xp[o];
                    ye.im[n][v] = ((float)Xp[5 * PL + o] + (float)Xp[3 * PL + o]) + (float)Xp[PL + o];
                }
            }
            wave_cgemm_split<D>(Bp, Xp, row0, col0, acc);   // Z Z = (Y^2)^T
#pragma unroll
            for (int n = 0; n < TN; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    q.re[n][v] = acc.re[0][n][v] + acc.re[1][n][v];
                    q.im[n][v] = acc.im[0][n][v] + acc.im[1][n][v];
                }
            __syncthreads();
            write_split<D, true>(q, Xp, row0, col0);   // A := Y^2 for the rest of the slice
            // Q_j = m_{2j} I + m_{2j+1} Y (+ acc): FP32, or FP64 for j = 1
            auto bracket = [&](int jj, bool with_acc) {
                const double mr = a.are[2 * jj], mi = a.aim[2 * jj];
                const double nr = a.are[2 * jj + 1], ni = a.aim[2 * jj + 1];
#pragma unroll
                for (int n = 0; n < TN; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const bool dg = mrow(row0, v) == mcol(col0, n);
                        const float cr = with_acc ? acc.re[0][n][v] + acc.re[1][n][v] : 0.f;
                        const float ci = with_acc ? acc.im[0][n][v] + acc.im[1][n][v] : 0.f;
                        if (jj == 1) {
                            const double yr = ye.re[n][v], yi = ye.im[n][v];
                            q.re[n][v] = (float)((nr * yr - ni * yi) + (double)cr + (dg ? mr : 0.0));
                            q.im[n][v] = (float)((nr * yi + ni * yr) + (double)ci + (dg ? mi : 0.0));
                        } else {
                            const float yr = ye.re[n][v], yi = ye.im[n][v];
                            q.re[n][v] = ((float)nr * yr - (float)ni * yi) + cr + (dg ? (float)mr : 0.f);
                            q.im[n][v] = ((float)nr * yi + (float)ni * yr) + ci + (dg ? (float)mi : 0.f);
                        }
                    }
            };
            // level of the product Y^2 Q_{jj+1} (jj >= 1); Y^2 Q_1 (jj = 0) is full precision
            auto level = [&](int jj) { return jj >= 1 ? (a.ps_levels >> (2 * jj)) & 3 : 2; };
            // Q_jj as the B plane, with the slices that the product reading it (level(jj - 1)) uses
            auto write_q = [&](int jj) {
                const int nl = level(jj - 1) + 1;
                if (nl == 1) write_split<D, true, 1>(q, Bp, row0, col0);
                else if (nl == 2) write_split<D, true, 2>(q, Bp, row0, col0);
                else write_split<D, true, 3>(q, Bp, row0, col0);
            };
            bracket(J, false);
            write_q(J);
            __syncthreads();
            for (int jj = J - 1; jj >= 1; --jj) {
                // Y^2 Q_{jj+1}, at the precision level the host chose for its weight in E
                const int lv = level(jj);
                if (lv == 0) wave_cgemm_split<D, 0>(Xp, Bp, row0, col0, acc);
                else if (lv == 1) wave_cgemm_split<D, 1>(Xp, Bp, row0, col0, acc);
                else wave_cgemm_split<D, 2>(Xp, Bp, row0, col0, acc);
                bracket(jj, true);
                __syncthreads();
                write_q(jj);
                __syncthreads();
            }
            if (pre && tid < a.ncoef) coef[tid] = coef_finish(a.ctl, raw);
            coef_ready = pre;
            wave_cgemm_split<D>(Xp, Bp, row0, col0, acc);       // R = Y^2 Q_1
            const double m0m1 = a.are[0];
            const double2 m1 = mk<double>(a.are[1], a.aim[1]);
#pragma unroll
            for (int n = 0; n < TN; ++n) {
                const int col = mcol(col0, n);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = mrow(row0, v);
                    double er = 0.0, ei = 0.0;
                    if (row < d && col < d) {
                        const float2 rx = Rx[row * SR + col];
                        const double2 y = mk<double>((double)ye.re[n][v] + (double)rx.x, (double)ye.im[n][v] + (double)rx.y);
                        const double2 t = cmul(m1, y);
                        er = t.x + ((double)acc.re[0][n][v] + (double)acc.re[1][n][v]) + (row == col ? m0m1 : 0.0);
                        ei = t.y + ((double)acc.im[0][n][v] + (double)acc.im[1][n][v]);
                    }
                    e.re[n][v] = (float)er;
                    e.im[n][v] = (float)ei;
                }
            }
            __syncthreads();
            return;
        }
        Tile b1, b2;
        LTile<TN> acc;
        {
            const float aMr = (float)a.are[M], aMi = (float)a.aim[M];
            const float aM1r = (float)a.are[M - 1], aM1i = (float)a.aim[M - 1];
#pragma unroll
            for (int n = 0; n < TN; ++n) {
                const int col = mcol(col0, n);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = mrow(row0, v);
                    const bool dg = row == col;
                    const int o = MShape<D>::off(row, col);
                    const float xr = ((float)Xp[4 * PL + o] + (float)Xp[2 * PL + o]) + (float)Xp[o];
                    const float xi = ((float)Xp[5 * PL + o] + (float)Xp[3 * PL + o]) + (float)Xp[PL + o];
                    b2.re[n][v] = dg ? aMr : 0.f;
                    b2.im[n][v] = dg ? aMi : 0.f;
                    // s a_M X + a_{M-1} I
                    b1.re[n][v] = sf * (aMr * xr - aMi * xi) + (dg ? aM1r : 0.f);
                    b1.im[n][v] = sf * (aMr * xi + aMi * xr) + (dg ? aM1i : 0.f);
                }
            }
        }
        write_split<D, true>(b1, Bp, row0, col0);
        __syncthreads();

        // b_k = s X b_{k+1} - b_{k+2} + a_k I for k = M-2 .. 2, then b_1' (no a_1 I).  One step:
        // bc = b_{k+1} (staged in Bp), bp = b_{k+2} -> b_k, staged for the next step.
        auto step = [&](const Tile& bc, Tile& bp, int k) {
            wave_cgemm_split<D>(Xp, Bp, row0, col0, acc);
            if (k == 2 || k == 3) {
                const double akr = a.are[k], aki = a.aim[k];
#pragma unroll
                for (int n = 0; n < TN; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float cr = sf * (acc.re[0][n][v] + acc.re[1][n][v]);
                        const float ci = sf * (acc.im[0][n][v] + acc.im[1][n][v]);
                        float nr = cr - bp.re[n][v], ni = ci - bp.im[n][v];
                        if (mrow(row0, v) == mcol(col0, n)) {
                            nr = (float)((double)cr - (double)bp.re[n][v] + akr);
                            ni = (float)((double)ci - (double)bp.im[n][v] + aki);
                        }
                        bp.re[n][v] = nr; bp.im[n][v] = ni;
                    }
            } else {
                const float akr = (k >= 2) ? (float)a.are[k] : 0.f, aki = (k >= 2) ? (float)a.aim[k] : 0.f;
#pragma unroll
                for (int n = 0; n < TN; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const bool dg = mrow(row0, v) == mcol(col0, n);
                        bp.re[n][v] = sf * (acc.re[0][n][v] + acc.re[1][n][v]) - bp.re[n][v] + (dg ? akr : 0.f);
                        bp.im[n][v] = sf * (acc.im[0][n][v] + acc.im[1][n][v]) - bp.im[n][v] + (dg ? aki : 0.f);
                    }
            }
            __syncthreads();
            write_split<D, true>(bp, Bp, row0, col0);
            __syncthreads();
        };
        // M is odd: the M - 2 steps are pairs k, k - 1 (k = M-2, M-4, .., 3) and a last k = 1.
        for (int k = M - 2; k >= 2; k -= 2) {
            step(b1, b2, k);        // b2 <- b_k
            step(b2, b1, k - 1);    // b1 <- b_{k-1}
        }
        step(b1, b2, 1);            // b2 <- b_1', b1 = b_2
        if (pre && tid < a.ncoef) coef[tid] = coef_finish(a.ctl, raw);
        coef_ready = pre;
        // R = s X b_1' - 2 b_2 ;  E = (a_0 - 1) I + a_1 s X + R  (FP64)
        wave_cgemm_split<D>(Xp, Bp, row0, col0, acc);
        const double a0m1 = a.are[0];
        const double2 a1s = mk<double>(a.are[1] * a.s, a.aim[1] * a.s);
#pragma unroll
        for (int n = 0; n < TN; ++n) {
            const int col = mcol(col0, n);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = mrow(row0, v);
                double er = 0.0, ei = 0.0;
                if (row < d && col < d) {
                    const double cr = (double)acc.re[0][n][v] + (double)acc.re[1][n][v];
                    const double ci = (double)acc.im[0][n][v] + (double)acc.im[1][n][v];
                    // X_j in FP64 from its split and residual
                    const int o = MShape<D>::off(row, col);
                    const float2 rx = Rx[row * SR + col];
                    const double xr = (((double)(float)Xp[4 * PL + o] + (double)(float)Xp[2 * PL + o]) +
                                       (double)(float)Xp[o]) + (double)rx.x;
                    const double xi = (((double)(float)Xp[5 * PL + o] + (double)(float)Xp[3 * PL + o]) +
                                       (double)(float)Xp[PL + o]) + (double)rx.y;
                    const double2 t = cmul(a1s, mk<double>(xr, xi));
                    er = t.x + (a.s * cr - 2.0 * (double)b1.re[n][v]) + (row == col ? a0m1 : 0.0);
                    ei = t.y + (a.s * ci - 2.0 * (double)b1.im[n][v]);
                }
                e.re[n][v] = (float)er;
                e.im[n][v] = (float)ei;
            }
        }
        __syncthreads();
    };

    // A := A + B + A B for the later node A and the earlier node B (split product, in place).
    auto combine = [&](Tile& A, const Tile& B) __attribute__((always_inline)) {
        write_split<D, false>(A, Xp, row0, col0);
        write_split<D, true>(B, Bp, row0, col0);
        __syncthreads();
        LTile<TN> acc;
        wave_cgemm_split<D>(Xp, Bp, row0, col0, acc);
#pragma unroll
        for (int n = 0; n < TN; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                A.re[n][v] = (A.re[n][v] + B.re[n][v]) + (acc.re[0][n][v] + acc.re[1][n][v]);
                A.im[n][v] = (A.im[n][v] + B.im[n][v]) + (acc.im[0][n][v] + acc.im[1][n][v]);
            }
        __syncthreads();
    };

    for (int w = blockIdx.x; w < n_nodes; w += gridDim.x) {
        // next slice this work-group runs after slice j: j + 1 inside the block, else the first
        // slice of its next block (or none)
        auto next_of = [&](int j) -> int {
            if (j + 1 < a.N && j + 1 < (w + 1) * NB) return j + 1;
            const int nb = (w + (int)gridDim.x) * NB;
            return nb < a.N ? nb : -1;
        };
        // The block's slices run one after another (a loop: the slice code is emitted once).  After
        // slice p the carry chain over the levels l (unrolled: every pending node stays in fixed
        // registers) combines with the pending earlier sibling pend[l] while bit l of p is set and
        // otherwise parks the node in pend[l].  Slices past N are zero nodes, which a combine
        // returns unchanged (the reference's odd carry).
        Tile pend[F > 0 ? F : 1];
#pragma unroll 1
        for (int p = 0; p < NB; ++p) {
            const int j = w * NB + p;
            Tile e;
            if (j < a.N) {   // uniform
                slice(j, next_of(j), e);
            } else {
#pragma unroll
                for (int n = 0; n < TN; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) { e.re[n][v] = 0.f; e.im[n][v] = 0.f; }
            }
            bool carry = true;
#pragma unroll
            for (int l = 0; l < F; ++l) {
                if (carry) {
                    if ((p >> l) & 1) {
                        combine(e, pend[l]);
                    } else {
                        pend[l] = e;
                        carry = false;
                    }
                }
            }
            if (carry) {   // p = 2^F - 1: e is the block's node
                float2* out = a.out + (size_t)w * dd;
#pragma unroll
                for (int n = 0; n < TN; ++n) {
                    const int col = mcol(col0, n);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = mrow(row0, v);
                        if (row < d && col < d) out[row * d + col] = make_float2(e.re[n][v], e.im[n][v]);
                    }
                }
            }
        }
    }
}

int mixed_fused_levels(int N) { return N >= 4096 ? 3 : (N >= 1024 ? 2 : (N >= 256 ? 1 : 0)); }

hipError_t launch_cheb_mixed(const MixedArgs& a, hipStream_t st) {
    if (a.d <= 0 || a.d > 64 || a.N <= 0 || a.M < 3 || a.M >= kMaxCheb || a.ncoef > kMaxCoeffs)
        return hipErrorInvalidValue;
    const int F = mixed_fused_levels(a.N);
    const int grid = (a.N + (1 << F) - 1) >> F;
    auto go = [&](auto kern, int g) { hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * g), 0, st, a); };
#define PA_MIXED_GO(DD, PSV)                                                        \
    switch (F) {                                                                    \
        case 3: go(cheb_mixed_kernel<DD, 3, PSV>, MShape<DD>::G); break;            \
        case 2: go(cheb_mixed_kernel<DD, 2, PSV>, MShape<DD>::G); break;            \
        case 1: go(cheb_mixed_kernel<DD, 1, PSV>, MShape<DD>::G); break;            \
        default: go(cheb_mixed_kernel<DD, 0, PSV>, MShape<DD>::G); break;           \
    }
    if (a.d <= 32) {
        if (a.ps) { PA_MIXED_GO(32, true) } else { PA_MIXED_GO(32, false) }
    } else {
        if (a.ps) { PA_MIXED_GO(64, true) } else { PA_MIXED_GO(64, false) }
    }
#undef PA_MIXED_GO
    return hipGetLastError();
}

}  // namespace pa
