// Mixed-precision slice kernel, power form, on scaled fp16 two-way split products (gfx950 fp16
// matrix cores + FP64 first-order terms), with the lowest levels of the pairwise product tree
// fused in.  Same algorithm, wave layout and tree fold as cheb_mixed_kernel (mixed.hip); what
// differs is the representation of the FP32 operands of the products:
//
//     x = (h0 + h1) / S,   h0 = fp16(S x),  h1 = fp16(S x - h0),   S = 2^14
//
// fp16 keeps 11 significant bits, so the two parts carry 22 bits (the difference S x - h0 is
// exact in FP32), and a product is the sum h0 h0' + h0 h1' + h1 h0' (+ h1 h1' when FULL = 4) of
// fp16 products accumulated in FP32: three (four) matrix-core products per FP32-accurate product
// instead of the six of the bf16 three-way split, and two operand planes instead of three.  The
// scale keeps the operands in fp16's range: an entry is at most the spectral norm of its matrix,
// |Y|, |Y^2| <= ||tau X||_2^2 and |Q_j| < 1 while ||tau X||_2 <= 1 (for controls of modulus <= 1,
// ||tau X||_2 <= tau Hnorm <= 1 here), and the tree nodes |E| = |U - I| <= 2 for unitary slices,
// so S x stays below 2^15 < 65504 (overflow needs an entry of 4).  An entry whose low part falls
// below fp16's normal range loses at most 2^-25 / S = 2^-39 in absolute terms.  The 22-bit
// products' rounding grows with ||tau X||_2 (controls well above 1 in modulus): see the precision
// modes in parament.h.  The accumulated products carry S^2, removed exactly by the
// power-of-two factor 2^-28.  The Horner brackets Q_j, the slice result E and the fold's tree
// nodes are held multiplied by S (exact), so that their planes are written without a multiply;
// the work-group's node is divided by S when it is stored.
#include "kernels.hpp"

namespace pa {

namespace {

typedef float f32x4h __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

constexpr float kScale = 16384.0f;                       // S = 2^14
constexpr float kInvScale = 1.0f / 16384.0f;
constexpr float kInvScale2 = 1.0f / (16384.0f * 16384.0f);  // products carry S^2

// Wave tiles of 16 x 32 (two 16x16 blocks sharing each A fragment), as in mixed.hip.
template <int D> struct HShape {
    static constexpr int TN = 2;
    static constexpr int WR = D / 16;
    static constexpr int WC = D / 32;
    static constexpr int G = WR * WC;
    static constexpr bool SWZ = (D == 64);
    static constexpr int S = SWZ ? D : D + 8;
    static constexpr int PL = D * S;
    static __device__ __forceinline__ int off(int row, int col) {
        if constexpr (SWZ) return row * D + ((((col >> 3) ^ row) & 7) << 3) + (col & 7);
        else return row * S + col;
    }
};

// The two fp16 parts of the scaled values f x0, f x1 (round to nearest even; f = S for plain
// values, S / S^2 for raw product accumulators, 1 for values held scaled by S).
__device__ __forceinline__ void split2(float x0, float x1, f16x2& h0, f16x2& h1, float f = kScale) {
    const float s0 = x0 * f, s1 = x1 * f;
    h0[0] = (_Float16)s0;
    h0[1] = (_Float16)s1;
    h1[0] = (_Float16)(s0 - (float)h0[0]);
    h1[1] = (_Float16)(s1 - (float)h0[1]);
}
// The same for values already scaled by S.  one holds 1.0 where the compiler cannot see it, so
// that s - h0 still becomes one mixed-precision FMA (s * one - h0, rounded to fp16) instead of a
// conversion, a subtraction and a conversion.
__device__ __forceinline__ void split2s(float s0, float s1, float one, f16x2& h0, f16x2& h1) {
    h0[0] = (_Float16)s0;
    h0[1] = (_Float16)s1;
    h1[0] = (_Float16)__builtin_fmaf(s0, one, -(float)h0[0]);
    h1[1] = (_Float16)__builtin_fmaf(s1, one, -(float)h0[1]);
}
// The same from FP64 values already scaled by S (s = S x); also returns what the parts leave out,
// x - (h0 + h1) / S.  The scaled value is rounded to FP32 first (y, with its rounding residual e)
// and y is split with the hardware FP32 -> fp16 conversions; the empty asm statement keeps the
// compiler from merging the two roundings into a double -> half conversion, which has no
// instruction and is emulated.
__device__ __forceinline__ void split2d(double s0, double s1, f16x2& h0, f16x2& h1, float& r0, float& r1) {
    float y0 = (float)s0, y1 = (float)s1;
    const float e0 = (float)(s0 - (double)y0), e1 = (float)(s1 - (double)y1);
    asm volatile("" : "+v"(y0), "+v"(y1));
    split2(y0, y1, h0, h1, 1.0f);
    // (y - h0) and then (y - h0) - h1 are exact in FP32
    r0 = (((y0 - (float)h0[0]) - (float)h1[0]) + e0) * kInvScale;
    r1 = (((y1 - (float)h0[1]) - (float)h1[1]) + e1) * kInvScale;
}

__device__ __forceinline__ int hrow(int row0, int v) { return row0 + 4 * ((threadIdx.x & 63) >> 4) + v; }
__device__ __forceinline__ int hcol(int col0, int n) { return col0 + 16 * n + (threadIdx.x & 15); }

template <int TN> struct HTile {
    f32x4h re[TN], im[TN];
};

__device__ __forceinline__ f16x8 hneg(f16x8 v) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(f16x8, u);
}

// c = S^2 (A B) on the wave's 16 x 32 tile (consumers remove S^2, or S for a result held scaled
// by S): A planes row-major, B planes transposed
// (Bt[col][k]); plane 2 * part + {0: re, 1: im}.  NP = 1: h0 h0' only (relative error ~2^-11);
// 3: with the two cross terms (~2^-21); 4: all four products.  NEGP: A's planes 4 + part hold
// its negated imaginary parts (otherwise each fragment is negated here).
template <int D, int NP, bool NEGP = false>
__device__ __forceinline__ void wave_cgemm_h2(const _Float16* __restrict__ Ap, const _Float16* __restrict__ Bp, int row0,
                                              int col0, HTile<HShape<D>::TN>& c) {
    constexpr int PL = HShape<D>::PL, TN = HShape<D>::TN;
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int n = 0; n < TN; ++n)
#pragma unroll
        for (int v = 0; v < 4; ++v) { c.re[n][v] = 0.f; c.im[n][v] = 0.f; }
#pragma unroll 1
    for (int ks = 0; ks < D / 32; ++ks) {
        const int ka = HShape<D>::off(row0 + r, 32 * ks + 8 * kq);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (q == 1 && NP == 1) break;
            f16x8 Br[TN], Bi[TN];
#pragma unroll
            for (int n = 0; n < TN; ++n) {
                const int kb = HShape<D>::off(col0 + 16 * n + r, 32 * ks + 8 * kq);
                Br[n] = *reinterpret_cast<const f16x8*>(Bp + (2 * q) * PL + kb);
                Bi[n] = *reinterpret_cast<const f16x8*>(Bp + (2 * q + 1) * PL + kb);
            }
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const bool use = (p == 0 && q == 0) || (NP >= 3 && p + q == 1) || (NP == 4 && p == 1 && q == 1);
                if (!use) continue;
                const f16x8 Ar = *reinterpret_cast<const f16x8*>(Ap + (2 * p) * PL + ka);
                const f16x8 Ai = *reinterpret_cast<const f16x8*>(Ap + (2 * p + 1) * PL + ka);
                const f16x8 An = NEGP ? *reinterpret_cast<const f16x8*>(Ap + (4 + p) * PL + ka) : hneg(Ai);
#pragma unroll
                for (int n = 0; n < TN; ++n) {
                    c.re[n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ar, Br[n], c.re[n], 0, 0, 0);
                    c.im[n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ar, Bi[n], c.im[n], 0, 0, 0);
                    c.re[n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(An, Bi[n], c.re[n], 0, 0, 0);
                    c.im[n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ai, Br[n], c.im[n], 0, 0, 0);
                }
            }
        }
    }
}

// Tile (accumulator layout) -> fp16 planes of its NL leading parts.  TRANSPOSED: the natural
// write, P[col][row] = b(row, col), four consecutive rows per lane in one 8-byte store;
// otherwise row-major P[row][col], 2-byte stores.  The tile is multiplied by f before the split,
// unless f is an opaque 1.0 (a tile held scaled by S): then split2s.  NEGIM (natural write):
// planes 4 + part also receive the negated imaginary parts.
template <int D, bool TRANSPOSED, int NL = 2, bool SCALED = false, bool NEGIM = false>
__device__ __forceinline__ void write_h2(const HTile<HShape<D>::TN>& b, _Float16* __restrict__ P, int row0, int col0,
                                         float f = kScale) {
    constexpr int PL = HShape<D>::PL, TN = HShape<D>::TN;
    const int rbase = hrow(row0, 0);
#pragma unroll
    for (int n = 0; n < TN; ++n) {
        const int col = hcol(col0, n);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const f32x4h& x = c ? b.im[n] : b.re[n];
            f16x2 lo0, lo1, hi0, hi1;
            if (SCALED) {
                split2s(x[0], x[1], f, lo0, lo1);
                split2s(x[2], x[3], f, hi0, hi1);
            } else {
                split2(x[0], x[1], lo0, lo1, f);
                split2(x[2], x[3], hi0, hi1, f);
            }
#pragma unroll
            for (int t = 0; t < NL; ++t) {
                const f16x2& a = t ? lo1 : lo0;
                const f16x2& z = t ? hi1 : hi0;
                if (TRANSPOSED) {
                    const f16x4 v = {a[0], a[1], z[0], z[1]};
                    *reinterpret_cast<f16x4*>(P + (2 * t + c) * PL + HShape<D>::off(col, rbase)) = v;
                    if (NEGIM && c == 1) {
                        typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
                        const u32x2 u = __builtin_bit_cast(u32x2, v) ^ 0x80008000u;
                        *reinterpret_cast<u32x2*>(P + (4 + t) * PL + HShape<D>::off(col, rbase)) = u;
                    }
                } else {
                    P[(2 * t + c) * PL + HShape<D>::off(rbase, col)] = a[0];
                    P[(2 * t + c) * PL + HShape<D>::off(rbase + 1, col)] = a[1];
                    P[(2 * t + c) * PL + HShape<D>::off(rbase + 2, col)] = z[0];
                    P[(2 * t + c) * PL + HShape<D>::off(rbase + 3, col)] = z[1];
                }
            }
        }
    }
}

struct RawCoefH {
    double2 c0, c1, c2;
};
__device__ __forceinline__ RawCoefH coef_load_h(const Controls<double>& q, int t, int j) {
    RawCoefH r;
    if (q.mode == QUAD_NONE) { r.c0 = sample(q, t, j); r.c1 = r.c2 = mk<double>(0, 0); }
    else if (q.mode == QUAD_MIDPOINT) { r.c0 = sample(q, t, j); r.c1 = sample(q, t, j + 1); r.c2 = mk<double>(0, 0); }
    else { r.c0 = sample(q, t, 2 * j); r.c1 = sample(q, t, 2 * j + 1); r.c2 = sample(q, t, 2 * j + 2); }
    return r;
}
__device__ __forceinline__ double2 coef_finish_h(const Controls<double>& q, const RawCoefH& r) {
    if (q.mode == QUAD_NONE) return r.c0;
    if (q.mode == QUAD_MIDPOINT) {
        const double2 s = cadd(r.c0, r.c1);
        return mk<double>(0.5 * s.x, 0.5 * s.y);
    }
    const double2 s = cadd(cadd(r.c0, mk<double>(4.0 * r.c1.x, 4.0 * r.c1.y)), r.c2);
    const double sixth = 1.0 / 6.0;
    return mk<double>(s.x * sixth, s.y * sixth);
}

}  // namespace

// Power form only (tau Hnorm <= 1): E = (m_0 - 1) I + m_1 Y + Y^2 Q_1 with Q_J = m_{2J} I +
// m_{2J+1} Y and Q_j = m_{2j} I + m_{2j+1} Y + Y^2 Q_{j+1} (J = (M - 1) / 2), Y = tau X; Q_1 and
// the first-order part in FP64, the inner products Y^2 Q_{j+1} at the part count the host chose
// (a.ps_levels: 0 = the leading products only, otherwise FULL).  EXACT: d == D (no padding).
//
// Tree fold: the work-group's 2^F slices run one after another (a loop, so the slice code is
// emitted once); after slice p, the carry chain over the levels l (unrolled, so every pending
// node sits in fixed registers) combines with the pending earlier sibling pend[l] while bit l of
// p is set, and otherwise parks the node in pend[l].  Slices past N are zero nodes, which the
// combine returns unchanged bit for bit (the reference's odd carry).
template <int D, int F, int FULL, bool EXACT>
__global__ __launch_bounds__(64 * HShape<D>::G) void cheb_h2_kernel(MixedArgs a) {
    constexpr int G = HShape<D>::G, PL = HShape<D>::PL, WC = HShape<D>::WC;
    constexpr int TN = HShape<D>::TN;
    constexpr int NB = 1 << F;
    typedef HTile<TN> Tile;
    // A operands: planes 0-3 as in wave_cgemm_h2; planes 4, 5: Y^2's negated imaginary parts
    __shared__ __attribute__((aligned(16))) _Float16 Xp[6 * PL];
    // B operands, two sets: consecutive Horner steps read one and write the other, so a step needs
    // one barrier (the one after its write)
    __shared__ __attribute__((aligned(16))) _Float16 Bpp[2][4 * PL];
    _Float16* const Bp = Bpp[0];
    // Y_j - (h0 + h1) / S: the FP64 residual of the split, for the first-order term
    __shared__ __attribute__((aligned(16))) float2 Rx[D * (D + 2)];
    __shared__ double2 coef[kMaxCoeffs];
    constexpr int SR = D + 2;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int row0 = 16 * (wave / WC), col0 = 16 * TN * (wave % WC);
    const int d = EXACT ? D : a.d, dd = d * d, M = a.M;
    const int n_nodes = (a.N + NB - 1) >> F;
    const bool prefetch = !a.ctl.magnus && a.ncoef <= 64 * G;
    bool coef_ready = false;
    // diagonal positions of the wave's tile and the in-range mask (fixed for the launch)
    bool dg[TN][4], in[TN][4];
#pragma unroll
    for (int n = 0; n < TN; ++n)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dg[n][v] = hrow(row0, v) == hcol(col0, n);
            in[n][v] = EXACT || (hrow(row0, v) < d && hcol(col0, n) < d);
        }

    float one = 1.0f;   // opaque to the compiler (split2s)
    asm volatile("" : "+v"(one));
    // the assembly forms S Y = (tau S) H0 + sum_t (tau S c'_t) O_t: coef[] holds tau S c'_t
    const double ts = a.ctl.tau * (double)kScale;
    auto scaled = [&](double2 c) { return mk<double>(c.x * ts, c.y * ts); };
    auto slice = [&](int j, int jn, Tile& e) __attribute__((always_inline)) {
        if (!coef_ready)
            for (int t = tid; t < a.ncoef; t += 64 * G) coef[t] = scaled(slice_coefficient(a.ctl, t, j));
        __syncthreads();
        // Y_j = tau (H0 + sum_t c'_t O_t) in FP64 as 2 x 2 blocks (lane (bi, bk) = (l / 8, l % 8) of
        // a 16 x 16 region): both Y (Xp) and Z = Y^T (Bp) take 4-byte stores per row pair
        {
            constexpr int RT = D / 16;
#pragma unroll 1
            for (int it = 0; it < RT * RT / G; ++it) {
                const int rg = wave + G * it, lane = tid & 63;
                const int i = 16 * (rg / RT) + 2 * (lane >> 3), k = 16 * (rg % RT) + 2 * (lane & 7);
                bool v[4];
                double2 x[4];
                if (a.ncoef <= 3) {
                    // every operator element of the iteration requested before any arithmetic (one
                    // load latency): the loads are unconditional, an operator slot past ncoef reads
                    // H0's element again and is weighted by zero, and a padded element reads the
                    // operators' element 0 and is set to zero afterwards
                    double2 h[4], o[3][4];
#pragma unroll
                    for (int e4 = 0; e4 < 4; ++e4) {
                        const int ii = i + (e4 >> 1), kk = k + (e4 & 1);
                        v[e4] = EXACT || (ii < d && kk < d);
                        const int at = v[e4] ? ii * d + kk : 0;
                        h[e4] = a.ops[at];
#pragma unroll
                        for (int q = 0; q < 3; ++q) o[q][e4] = a.ops[(q < a.ncoef ? (size_t)(q + 1) * dd : 0) + at];
                    }
#pragma unroll
                    for (int e4 = 0; e4 < 4; ++e4) x[e4] = mk<double>(h[e4].x * ts, h[e4].y * ts);
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const double2 cc = (q < a.ncoef) ? coef[q] : mk<double>(0, 0);
#pragma unroll
                        for (int e4 = 0; e4 < 4; ++e4) {
                            x[e4].x = __builtin_fma(cc.x, o[q][e4].x, x[e4].x);
                            x[e4].x = __builtin_fma(-cc.y, o[q][e4].y, x[e4].x);
                            x[e4].y = __builtin_fma(cc.x, o[q][e4].y, x[e4].y);
                            x[e4].y = __builtin_fma(cc.y, o[q][e4].x, x[e4].y);
                        }
                    }
#pragma unroll
                    for (int e4 = 0; e4 < 4; ++e4)
                        if (!v[e4]) x[e4] = mk<double>(0, 0);
                } else {
#pragma unroll
                    for (int e4 = 0; e4 < 4; ++e4) {
                        const int ii = i + (e4 >> 1), kk = k + (e4 & 1);
                        v[e4] = EXACT || (ii < d && kk < d);
                        x[e4] = v[e4] ? a.ops[ii * d + kk] : mk<double>(0, 0);
                        x[e4].x *= ts;
                        x[e4].y *= ts;
                    }
                }
                for (int t0 = 0; a.ncoef > 3 && t0 < a.ncoef; t0 += 2) {
                    double2 o[2][4];
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int e4 = 0; e4 < 4; ++e4) {
                            const size_t at = (size_t)(t0 + q + 1) * dd + (i + (e4 >> 1)) * d + k + (e4 & 1);
                            o[q][e4] = (v[e4] && t0 + q < a.ncoef) ? a.ops[at] : mk<double>(0, 0);
                        }
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        if (t0 + q < a.ncoef) {
                            const double2 cc = coef[t0 + q];
#pragma unroll
                            for (int e4 = 0; e4 < 4; ++e4) {   // x += cc o, four FMAs
                                x[e4].x = __builtin_fma(cc.x, o[q][e4].x, x[e4].x);
                                x[e4].x = __builtin_fma(-cc.y, o[q][e4].y, x[e4].x);
                                x[e4].y = __builtin_fma(cc.x, o[q][e4].y, x[e4].y);
                                x[e4].y = __builtin_fma(cc.y, o[q][e4].x, x[e4].y);
                            }
                        }
                }
                // parts of the row pairs: p[c][h][part] (c: re / im, h: row i + h), columns k, k + 1
                f16x2 p[2][2][2];
                float rr[2][2][2];
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const double2 &xa = x[2 * h], &xb = x[2 * h + 1];
                        split2d(c ? xa.y : xa.x, c ? xb.y : xb.x, p[c][h][0], p[c][h][1], rr[c][h][0], rr[c][h][1]);
                    }
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        _Float16* P = Xp + (2 * t + c) * PL;
                        *reinterpret_cast<f16x2*>(P + HShape<D>::off(i, k)) = p[c][0][t];
                        *reinterpret_cast<f16x2*>(P + HShape<D>::off(i + 1, k)) = p[c][1][t];
                        _Float16* Q = Bp + (2 * t + c) * PL;
                        const f16x2 z0 = {p[c][0][t][0], p[c][1][t][0]}, z1 = {p[c][0][t][1], p[c][1][t][1]};
                        *reinterpret_cast<f16x2*>(Q + HShape<D>::off(k, i)) = z0;
                        *reinterpret_cast<f16x2*>(Q + HShape<D>::off(k + 1, i)) = z1;
                    }
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    *reinterpret_cast<float4*>(Rx + (i + h) * SR + k) =
                        make_float4(rr[0][h][0], rr[1][h][0], rr[0][h][1], rr[1][h][1]);
            }
        }
        __syncthreads();
        const bool pre = prefetch && jn >= 0 && jn < a.N;
        RawCoefH raw;
        if (pre && tid < a.ncoef) raw = coef_load_h(a.ctl, tid, jn);

        const int J = (M - 1) / 2;
        Tile ye, q, acc;
#pragma unroll
        for (int n = 0; n < TN; ++n) {
            const int col = hcol(col0, n);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int o = HShape<D>::off(hrow(row0, v), col);
                ye.re[n][v] = ((float)Xp[o] + (float)Xp[2 * PL + o]) * kInvScale;
                ye.im[n][v] = ((float)Xp[PL + o] + (float)Xp[3 * PL + o]) * kInvScale;
            }
        }
        wave_cgemm_h2<D, FULL>(Bp, Xp, row0, col0, acc);   // S^2 Z Z = S^2 (Y^2)^T
        __syncthreads();
        write_h2<D, true, 2, false, true>(acc, Xp, row0, col0, kScale * kInvScale2);   // A := Y^2 for the rest of the slice
        // S Q_jj = S (m_{2jj} I + m_{2jj+1} Y) (+ the product, raw S^2 Y^2 Q_{jj+1}): FP32, except
        // Q_1 (the largest terms, whose constant diagonal m_2 recurs in every slice), formed in FP64
        // and rounded once.  Q is held scaled by S (exact), so its planes are written without a
        // multiply.
        auto bracket = [&](int jj, bool with_acc) {
            const double mr = a.are[2 * jj] * kScale, mi = a.aim[2 * jj] * kScale;
            const double nr = a.are[2 * jj + 1] * kScale, ni = a.aim[2 * jj + 1] * kScale;
#pragma unroll
            for (int n = 0; n < TN; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float cr = with_acc ? acc.re[n][v] : 0.f;
                    const float ci = with_acc ? acc.im[n][v] : 0.f;
                    if (jj == 1) {
                        const double yr = ye.re[n][v], yi = ye.im[n][v];
                        q.re[n][v] = (float)((nr * yr - ni * yi) + (double)(cr * kInvScale) + (dg[n][v] ? mr : 0.0));
                        q.im[n][v] = (float)((nr * yi + ni * yr) + (double)(ci * kInvScale) + (dg[n][v] ? mi : 0.0));
                    } else {
                        const float yr = ye.re[n][v], yi = ye.im[n][v];
                        q.re[n][v] = __builtin_fmaf(cr, kInvScale, ((float)nr * yr - (float)ni * yi) + (dg[n][v] ? (float)mr : 0.f));
                        q.im[n][v] = __builtin_fmaf(ci, kInvScale, ((float)nr * yi + (float)ni * yr) + (dg[n][v] ? (float)mi : 0.f));
                    }
                }
        };
        // whether the product Y^2 Q_{jj+1} (jj >= 1) runs on the leading parts only
        auto lead_only = [&](int jj) { return jj >= 1 && ((a.ps_levels >> (2 * jj)) & 3) == 0; };
        auto write_q = [&](int jj, _Float16* dstp) {
            if (lead_only(jj - 1)) write_h2<D, true, 1, true>(q, dstp, row0, col0, one);
            else write_h2<D, true, 2, true>(q, dstp, row0, col0, one);
        };
        int cur = 1;   // B set holding the current Q (set 0 held Z, read by the product above)
        bracket(J, false);
        write_q(J, Bpp[cur]);
        __syncthreads();
        for (int jj = J - 1; jj >= 1; --jj) {
            if (lead_only(jj)) wave_cgemm_h2<D, 1, true>(Xp, Bpp[cur], row0, col0, acc);
            else wave_cgemm_h2<D, FULL, true>(Xp, Bpp[cur], row0, col0, acc);
            bracket(jj, true);
            cur ^= 1;
            write_q(jj, Bpp[cur]);   // every wave finished reading this set before the last barrier
            __syncthreads();
        }
        if (pre && tid < a.ncoef) coef[tid] = scaled(coef_finish_h(a.ctl, raw));
        coef_ready = pre;
        wave_cgemm_h2<D, FULL, true>(Xp, Bpp[cur], row0, col0, acc);   // S^2 R = S^2 Y^2 Q_1
        // E = (m_0 - 1) I + m_1 Y + Y^2 Q_1 in FP64, held scaled by S like every tree node of the
        // fold (exact), so that the fold's plane writes need no multiply
        const double m0m1 = a.are[0] * kScale;
        const double2 m1 = mk<double>(a.are[1] * kScale, a.aim[1] * kScale);
#pragma unroll
        for (int n = 0; n < TN; ++n) {
            const int col = hcol(col0, n);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = hrow(row0, v);
                double er = 0.0, ei = 0.0;
                if (in[n][v]) {
                    const float2 rx = Rx[row * SR + col];
                    const double2 y = mk<double>((double)ye.re[n][v] + (double)rx.x, (double)ye.im[n][v] + (double)rx.y);
                    const double2 t = cmul(m1, y);
                    er = t.x + (double)(acc.re[n][v] * kInvScale) + (dg[n][v] ? m0m1 : 0.0);
                    ei = t.y + (double)(acc.im[n][v] * kInvScale);
                }
                e.re[n][v] = (float)er;
                e.im[n][v] = (float)ei;
            }
        }
        __syncthreads();
    };

    // A := A + B + A B for the later node A and the earlier node B (in place; both held scaled by S)
    auto combine = [&](Tile& A, const Tile& B) __attribute__((always_inline)) {
        write_h2<D, false, 2, true>(A, Xp, row0, col0, one);
        write_h2<D, true, 2, true>(B, Bp, row0, col0, one);
        __syncthreads();
        Tile acc;
        wave_cgemm_h2<D, FULL>(Xp, Bp, row0, col0, acc);
#pragma unroll
        for (int n = 0; n < TN; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                A.re[n][v] = __builtin_fmaf(acc.re[n][v], kInvScale, A.re[n][v] + B.re[n][v]);
                A.im[n][v] = __builtin_fmaf(acc.im[n][v], kInvScale, A.im[n][v] + B.im[n][v]);
            }
        __syncthreads();
    };

    for (int w = blockIdx.x; w < n_nodes; w += gridDim.x) {
        auto next_of = [&](int j) -> int {
            if (j + 1 < a.N && j + 1 < (w + 1) * NB) return j + 1;
            const int nb = (w + (int)gridDim.x) * NB;
            return nb < a.N ? nb : -1;
        };
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
                    const int col = hcol(col0, n);
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (in[n][v])
                            out[hrow(row0, v) * d + col] = make_float2(e.re[n][v] * kInvScale, e.im[n][v] * kInvScale);
                }
            }
        }
    }
}

// fp16 two-way split form of the mixed slice kernel (power form only, tau Hnorm' <= 1); full = 3 or
// 4 products per FP32-accurate product.
hipError_t launch_cheb_h2(const MixedArgs& a, int full, hipStream_t st) {
    if (a.d <= 0 || a.d > 64 || a.N <= 0 || a.M < 3 || a.M >= kMaxCheb || a.ncoef > kMaxCoeffs || !a.ps ||
        (full != 3 && full != 4))
        return hipErrorInvalidValue;
    const int F = mixed_fused_levels(a.N);
    const int grid = (a.N + (1 << F) - 1) >> F;
    auto go = [&](auto kern, int g) { hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * g), 0, st, a); };
#define PA_H2_GO(DD, FL, EX)                                                         \
    switch (F) {                                                                     \
        case 3: go(cheb_h2_kernel<DD, 3, FL, EX>, HShape<DD>::G); break;             \
        case 2: go(cheb_h2_kernel<DD, 2, FL, EX>, HShape<DD>::G); break;             \
        case 1: go(cheb_h2_kernel<DD, 1, FL, EX>, HShape<DD>::G); break;             \
        default: go(cheb_h2_kernel<DD, 0, FL, EX>, HShape<DD>::G); break;            \
    }
    if (a.d <= 32) {
        if (a.d == 32) { if (full == 3) { PA_H2_GO(32, 3, true) } else { PA_H2_GO(32, 4, true) } }
        else { if (full == 3) { PA_H2_GO(32, 3, false) } else { PA_H2_GO(32, 4, false) } }
    } else {
        if (a.d == 64) { if (full == 3) { PA_H2_GO(64, 3, true) } else { PA_H2_GO(64, 4, true) } }
        else { if (full == 3) { PA_H2_GO(64, 3, false) } else { PA_H2_GO(64, 4, false) } }
    }
#undef PA_H2_GO
    return hipGetLastError();
}

}  // namespace pa
