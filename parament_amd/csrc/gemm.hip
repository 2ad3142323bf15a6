// Batched complex GEMM on the gfx950 matrix cores for the step-wise (any-dim) path:
//     C_b = alpha A_b B_b + beta Cin_b + beta2 Cin2_b + diag I  (+ A_b + B_b when add_ab)
// over a batch of d x d row-major complex matrices, with the arithmetic of the products set by
// the mode: FP64 (16x16x4 f64 instruction), FP32 (16x16x4 f32), bf16 split products
// (16x16x32 bf16; every FP32 operand is split into 2 or 3 bf16 slices, s0 = bf16(x),
// s1 = bf16(x - s0), s2 = bf16(x - s0 - s1), and the slice products with p + q < slices are
// accumulated in FP32), or scaled fp16 two-way split products (16x16x32 f16: x S = h0 + h1 + r
// with S = 2^14, h0 = fp16(x S), h1 = fp16(x S - h0), 22 bits in all; the three products
// h0 h0', h0 h1', h1 h0' accumulated in FP32 and the sum scaled by 2^-28; operand entries must
// stay below 4 in modulus, which holds for the power form's operands and the tree's deviations
// from identity).  The epilogue is evaluated in FP64 and rounded once to the output type,
// so a Clenshaw step's a_k I is added exactly.  Used for the Chebyshev steps
// (parament.cpp:569-652) and the product-tree levels (parament.cpp:657-718) at dims where a
// slice does not fit on chip.
//
// Tiling: a work-group computes a 64x64 output tile with 4 waves of 32x32 (2x2 blocks of
// 16x16), or a 32x32 tile with 4 waves of 16x16 (less padding for dims just above a multiple
// of 64); K advances in chunks of 32 staged in LDS (zero-padded past d).  Work items (batch b,
// tile t) are assigned so that every XCD receives a contiguous range of them: the tiles of one
// slice then share that XCD's L2 for their A rows and B columns.
#include "kernels.hpp"

#include <type_traits>

namespace pa {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8g __attribute__((ext_vector_type(8)));
typedef float f32x4g __attribute__((ext_vector_type(4)));
constexpr float kF16Scale = 16384.0f;                        // S of the fp16 split
constexpr double kF16Unscale = 1.0 / (16384.0 * 16384.0);    // the products carry S^2

constexpr int KT = 32;   // K chunk

// Work item of this block: XCD x (= blockIdx.x mod 8) takes items [x * per, (x + 1) * per).
__device__ __forceinline__ long work_item(long total) {
    const long nb = gridDim.x, b = blockIdx.x;
    const long per = nb / 8, rem = nb % 8;     // blocks per XCD (the first `rem` XCDs get one more)
    const long x = b % 8, i = b / 8;
    const long start = x * per + (x < rem ? x : rem);
    const long w = start + i;
    return w < total ? w : -1;
}

// Output tile of a launch: the caller's choice (32 or 64), else the one with the smaller
// padded product volume, 32-tiles counted 1.5x for their lower matrix-core efficiency.
int gemm_tile(const GemmArgs& g) {
    if (g.tile == 32 || g.tile == 64) return g.tile;
    auto vol = [&](int bt) { const double p = (double)((g.d + bt - 1) / bt) * bt; return p * p * (double)((g.d + KT - 1) / KT) * KT; };
    return 1.5 * vol(32) < vol(64) ? 32 : 64;
}

template <class T> __device__ __forceinline__ double2 ld(const T* p, long i);
template <> __device__ __forceinline__ double2 ld<double2>(const double2* p, long i) { return p[i]; }
template <> __device__ __forceinline__ double2 ld<float2>(const float2* p, long i) {
    const float2 v = p[i];
    return mk<double>(v.x, v.y);
}
__device__ __forceinline__ void st(double2* p, long i, double2 v) { p[i] = v; }
__device__ __forceinline__ void st(float2* p, long i, double2 v) { p[i] = make_float2((float)v.x, (float)v.y); }

template <class TA, class TB, class TC, class TO>
struct EpiIO {
    const TA* A; const TB* B; const TC* Cin; const TC* Cin2; TO* C;
};

// Output element o of product b from the accumulated x: alpha x + beta Cin + beta2 Cin2
// (+ A + B) + diag, in FP64, rounded once.  Each element reads its own Cin / Cin2 positions
// before it is written, so C may be one of them.  A, B point at product b's operands.
template <class TA, class TB, class TC, class TO>
__device__ __forceinline__ void epi_store(const GemmArgs& g, const EpiIO<TA, TB, TC, TO>& io, const TA* A,
                                          const TB* B, long b, long dd, long o, bool on_diag, double xr, double xi) {
    double2 y = mk<double>(g.alpha_re * xr - g.alpha_im * xi, g.alpha_re * xi + g.alpha_im * xr);
    if (io.Cin) {
        const double2 c = ld(io.Cin, b * g.sc * dd + o);
        y.x += g.beta_re * c.x - g.beta_im * c.y;
        y.y += g.beta_re * c.y + g.beta_im * c.x;
    }
    if (io.Cin2) {
        const double2 c = ld(io.Cin2, b * g.sc2 * dd + o);
        y.x += g.beta2_re * c.x - g.beta2_im * c.y;
        y.y += g.beta2_re * c.y + g.beta2_im * c.x;
    }
    if (g.add_ab) {
        const double2 a = ld(A, o), bb = ld(B, o);
        y.x += a.x + bb.x; y.y += a.y + bb.y;
    }
    if (on_diag) { y.x += g.diag_re; y.y += g.diag_im; }
    st(io.C, b * g.so * dd + o, y);
}

// FP64 / FP32 matrix-core tiles: A chunk [64][32] (stride SA), B chunk [32][64] (stride SB).
// The next K chunk is fetched into registers while the matrix cores work on the current one.
// EDGE: d is not a multiple of BT, so loads and stores past d are masked (otherwise none is).
template <class R, int BT, bool EDGE, bool DUAL, class TA, class TB, class TC, class TO>
__global__ __launch_bounds__(256) void zgemm_fp_kernel(GemmArgs g, EpiIO<TA, TB, TC, TO> io) {
    constexpr int TW = BT / 32;   // 16x16 blocks per wave side (4 waves, 2 x 2)
    constexpr int SA = KT + Pad<R>::A, SB = BT + Pad<R>::B;
    constexpr int NE = BT * KT / 256;   // chunk elements per thread, of A and of B
    __shared__ R Ar[BT * SA], Ai[BT * SA], Br[KT * SB], Bi[KT * SB];
    const int d = g.d, tiles = (d + BT - 1) / BT;
    const long total = (long)g.batch * tiles * tiles;
    const long w = work_item(total);
    if (w < 0) return;
    const long b = w / (tiles * tiles);
    const int t = (int)(w % (tiles * tiles)), ti = t / tiles, tj = t % tiles;
    const long dd = (long)d * d;
    const TA* A = io.A + b * g.sa * dd;
    const TB* B = io.B + b * g.sb * dd;
    // DUAL: a second product g.A2 g.B2 accumulated by a second pass of the K loop (its own
    // instantiation); KA, KB: the pair the K loop reads (A and B stay the epilogue's)
    const TA* KA = A;
    const TB* KB = B;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int wr = (BT / 2) * (wave >> 1), wc = (BT / 2) * (wave & 1);
    TA ra[NE];
    TB rb[NE];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = tid + 256 * u;
            const int gi = ti * BT + e / KT, gk = k0 + e % KT;
            ra[u].x = 0; ra[u].y = 0;
            if (!EDGE || (gi < d && gk < d)) ra[u] = KA[(long)gi * d + gk];
        }
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = tid + 256 * u;
            const int gk = k0 + e / BT, gj = tj * BT + e % BT;
            rb[u].x = 0; rb[u].y = 0;
            if (!EDGE || (gk < d && gj < d)) rb[u] = KB[(long)gk * d + gj];
        }
    };
    CTile<R, TW, TW> acc;
    acc.zero();
#pragma unroll 1
    for (int pass = 0; pass < (DUAL ? 2 : 1); ++pass) {
    if (DUAL && pass) {
        KA = static_cast<const TA*>(g.A2) + b * g.sa2 * dd;
        KB = static_cast<const TB*>(g.B2) + b * g.sb2 * dd;
    }
    fetch(0);
    for (int k0 = 0; k0 < d; k0 += KT) {
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = tid + 256 * u;
            Ar[(e / KT) * SA + e % KT] = (R)ra[u].x; Ai[(e / KT) * SA + e % KT] = (R)ra[u].y;
            Br[(e / BT) * SB + e % BT] = (R)rb[u].x; Bi[(e / BT) * SB + e % BT] = (R)rb[u].y;
        }
        __syncthreads();
        if (k0 + KT < d) fetch(k0 + KT);
        // a last chunk runs only the k-steps that reach d (the rest of it is zero)
        wave_cgemm<R, TW, TW, KT, SA, SB, false>(Ar, Ai, Br, Bi, wr, wc, acc, EDGE ? min(KT, (d - k0 + 3) & ~3) : KT);
        __syncthreads();
    }
    }   // pass
    // epilogue in FP64
#pragma unroll
    for (int m = 0; m < TW; ++m)
#pragma unroll
        for (int n = 0; n < TW; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = ti * BT + acc_row<R>(wr, m, v), gj = tj * BT + acc_col(wc, n);
                if (EDGE && (gi >= d || gj >= d)) continue;
                epi_store(g, io, A, B, b, dd, (long)gi * d + gj, gi == gj, acc.re[m][n][v], acc.im[m][n][v]);
            }
}

// Whole-product FP32 kernel for dims 33-128 (complex64 operands): work-group b forms the full C_b with NW = ceil(d/16)
// waves, wave w the 16-row band [16 w, 16 w + 16) over all 16 NW columns, so the padding is to a
// multiple of 16 and every operand element is fetched once per product.  K in chunks of 32,
// the next one fetched into registers while the matrix cores work on the current one.
template <int NW, class TA, class TB, class TC, class TO>
__global__ __launch_bounds__(64 * NW) void zgemm_row_kernel(GemmArgs g, EpiIO<TA, TB, TC, TO> io) {
    typedef float R;
    constexpr int DP = 16 * NW, NT = 64 * NW;
    constexpr int SA = KT + Pad<R>::A, SB = DP + Pad<R>::B;
    constexpr int NE = DP * KT / NT;   // 8 elements of A and of B per thread and chunk
    __shared__ R Ar[DP * SA], Ai[DP * SA], Br[KT * SB], Bi[KT * SB];
    const int d = g.d;
    const long b = work_item(g.batch);
    if (b < 0) return;
    const long dd = (long)d * d;
    const TA* A = io.A + b * g.sa * dd;
    const TB* B = io.B + b * g.sb * dd;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int wr = 16 * wave;
    TA ra[NE];
    TB rb[NE];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = tid + NT * u;
            const int gi = e / KT, gk = k0 + e % KT;
            ra[u].x = 0; ra[u].y = 0;
            if (gi < d && gk < d) ra[u] = A[(long)gi * d + gk];
        }
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = tid + NT * u;
            const int gk = k0 + e / DP, gj = e % DP;
            rb[u].x = 0; rb[u].y = 0;
            if (gk < d && gj < d) rb[u] = B[(long)gk * d + gj];
        }
    };
    CTile<R, 1, NW> acc;
    acc.zero();
    fetch(0);
    for (int k0 = 0; k0 < d; k0 += KT) {
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = tid + NT * u;
            Ar[(e / KT) * SA + e % KT] = (R)ra[u].x; Ai[(e / KT) * SA + e % KT] = (R)ra[u].y;
            Br[(e / DP) * SB + e % DP] = (R)rb[u].x; Bi[(e / DP) * SB + e % DP] = (R)rb[u].y;
        }
        __syncthreads();
        if (k0 + KT < d) fetch(k0 + KT);
        wave_cgemm<R, 1, NW, KT, SA, SB, false>(Ar, Ai, Br, Bi, wr, 0, acc, min(KT, (d - k0 + 3) & ~3));
        __syncthreads();
    }
#pragma unroll
    for (int n = 0; n < NW; ++n)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int gi = acc_row<R>(wr, 0, v), gj = acc_col(0, n);
            if (gi >= d || gj >= d) continue;
            epi_store(g, io, A, B, b, dd, (long)gi * d + gj, gi == gj, acc.re[0][n][v], acc.im[0][n][v]);
        }
}

// Split products: A chunk planes [slice][re/im][64][32 + 8] row-major, B chunk planes
// transposed [slice][re/im][64 cols][32 + 8].  Operand fragments of the 16x16x32 instruction:
// lane l holds A[l%16][8(l/16) .. +7] and Bt[l%16][8(l/16) .. +7]; accumulator element v is
// C[4(l/16) + v][l%16].  F16: the scaled fp16 two-way split (NS = 2), else bf16 slices.  EDGE:
// d is not a multiple of BT, so loads and stores past d are masked (otherwise none is).
template <int NS, int BT, bool F16, bool EDGE, bool DUAL, class TA, class TB, class TC, class TO>
__global__ __launch_bounds__(256) void zgemm_split_kernel(GemmArgs g, EpiIO<TA, TB, TC, TO> io) {
    typedef std::conditional_t<F16, _Float16, __bf16> H;
    typedef std::conditional_t<F16, f16x8g, bf16x8> H8;
    constexpr int TW = BT / 32;   // 16x16 blocks per wave side (4 waves, 2 x 2)
    constexpr int S = KT + 8, PL = BT * S;
    constexpr float pre = F16 ? kF16Scale : 1.0f;
    __shared__ __attribute__((aligned(16))) H Ap[2 * NS * PL];
    __shared__ __attribute__((aligned(16))) H Bp[2 * NS * PL];
    const int d = g.d, tiles = (d + BT - 1) / BT;
    const long total = (long)g.batch * tiles * tiles;
    const long w = work_item(total);
    if (w < 0) return;
    const long b = w / (tiles * tiles);
    const int t = (int)(w % (tiles * tiles)), ti = t / tiles, tj = t % tiles;
    const long dd = (long)d * d;
    const TA* A = io.A + b * g.sa * dd;
    const TB* B = io.B + b * g.sb * dd;
    // DUAL: a second product g.A2 g.B2 accumulated by a second pass of the K loop (its own
    // instantiation, so that the single-product kernels keep their registers); KA, KB: the pair
    // the K loop reads (A and B stay the epilogue's)
    const TA* KA = A;
    const TB* KB = B;
    constexpr int npass = DUAL ? 2 : 1;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wr = (BT / 2) * (wave >> 1), wc = (BT / 2) * (wave & 1);
    const int r16 = lane & 15, kq = lane >> 4;
    // acc[L][m][n]: L = 0 leading products, 1 the rest (bf16: the rest carries up to 16 bits below
    // the leading products; fp16: one accumulator, whose FP32 sum keeps the split's 22 bits)
    constexpr int NL = F16 ? 1 : 2;
    f32x4g accr[NL][TW][TW], acci[NL][TW][TW];
#pragma unroll
    for (int L = 0; L < NL; ++L)
#pragma unroll
        for (int m = 0; m < TW; ++m)
#pragma unroll
            for (int n = 0; n < TW; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) { accr[L][m][n][v] = 0.f; acci[L][m][n][v] = 0.f; }
    // two elements adjacent in k split together: each level's pair goes to LDS in one 4-byte store
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    typedef _Float16 f16x2g __attribute__((ext_vector_type(2)));
    typedef std::conditional_t<F16, f16x2g, bf16x2> H2;
    auto split_store2 = [&](H* P, int o, float x0, float x1, int c) {
        float r0 = x0 * pre, r1 = x1 * pre;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            H2 h;
            h[0] = (H)r0; h[1] = (H)r1;
            *reinterpret_cast<H2*>(P + (2 * s + c) * PL + o) = h;
            r0 -= (float)h[0]; r1 -= (float)h[1];
        }
    };
    // the next K chunk is fetched into registers while the matrix cores work on the current one;
    // thread element pairs: A (r, k), (r, k+1) and B (k, c), (k+1, c), k even.  Which thread takes
    // which pair: with 64-wide tiles, the 16-lane groups g of a wave take A rows 4 apart and B pair
    // columns rotated by g, so that each 4-byte LDS store of a wave lands in 64 different banks (row
    // stride S / 2 = 20 words); the loads stay coalesced in runs of 16 lanes.
    constexpr int NP = BT * KT / 512;
    const int g16 = lane >> 4;
    auto a_row = [&](int u) { return BT == 64 ? 16 * u + 4 * g16 + wave : (tid + 256 * u) / (KT / 2); };
    auto a_kp = [&](int u) { return BT == 64 ? (lane & 15) : (tid + 256 * u) % (KT / 2); };
    auto b_col = [&](int u) { return BT == 64 ? lane : (tid + 256 * u) % BT; };
    auto b_kp = [&](int u) { return BT == 64 ? 4 * u + ((wave + g16) & 3) : (tid + 256 * u) / BT; };
    TA ra[NP][2];
    TB rb[NP][2];
    // without edges, each thread's element pointers are formed once and advanced by one K chunk
    // per fetch (KT along A's rows, KT rows of B)
    const TA* pa[NP];
    const TB* pb[NP];
    auto init_ptrs = [&]() {
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            pa[u] = KA + (long)(ti * BT + a_row(u)) * d + 2 * a_kp(u);
            pb[u] = KB + (long)(2 * b_kp(u)) * d + tj * BT + b_col(u);
        }
    };
    init_ptrs();
    auto fetch = [&](int k0) {
        if constexpr (!EDGE) {
#pragma unroll
            for (int u = 0; u < NP; ++u) {
                ra[u][0] = pa[u][0]; ra[u][1] = pa[u][1];
                rb[u][0] = pb[u][0]; rb[u][1] = pb[u][d];
                pa[u] += KT;
                pb[u] += (long)KT * d;
            }
            return;
        }
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int gi = ti * BT + a_row(u), gk = k0 + 2 * a_kp(u);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                ra[u][h].x = 0; ra[u][h].y = 0;
                if (gi < d && gk + h < d) ra[u][h] = KA[(long)gi * d + gk + h];
            }
        }
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int gk = k0 + 2 * b_kp(u), gj = tj * BT + b_col(u);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                rb[u][h].x = 0; rb[u][h].y = 0;
                if (gk + h < d && gj < d) rb[u][h] = KB[(long)(gk + h) * d + gj];
            }
        }
    };
#pragma unroll 1
    for (int pass = 0; pass < npass; ++pass) {
    if (DUAL && pass) {
        KA = static_cast<const TA*>(g.A2) + b * g.sa2 * dd;
        KB = static_cast<const TB*>(g.B2) + b * g.sb2 * dd;
        init_ptrs();
    }
    fetch(0);
    for (int k0 = 0; k0 < d; k0 += KT) {
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int oa = a_row(u) * S + 2 * a_kp(u), ob = b_col(u) * S + 2 * b_kp(u);
            split_store2(Ap, oa, (float)ra[u][0].x, (float)ra[u][1].x, 0);
            split_store2(Ap, oa, (float)ra[u][0].y, (float)ra[u][1].y, 1);
            split_store2(Bp, ob, (float)rb[u][0].x, (float)rb[u][1].x, 0);
            split_store2(Bp, ob, (float)rb[u][0].y, (float)rb[u][1].y, 1);
        }
        __syncthreads();
        if (k0 + KT < d) fetch(k0 + KT);
        H8 br[TW][NS], bi[TW][NS];
#pragma unroll
        for (int n = 0; n < TW; ++n)
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                const int o = (wc + 16 * n + r16) * S + 8 * kq;
                br[n][q] = *reinterpret_cast<const H8*>(Bp + (2 * q) * PL + o);
                bi[n][q] = *reinterpret_cast<const H8*>(Bp + (2 * q + 1) * PL + o);
            }
        auto mfma = [](H8 a, H8 b, f32x4g c) {
            if constexpr (F16) return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
            else return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
        };
#pragma unroll
        for (int m = 0; m < TW; ++m)
#pragma unroll
            for (int p = 0; p < NS; ++p) {
                const int o = (wr + 16 * m + r16) * S + 8 * kq;
                const H8 ar = *reinterpret_cast<const H8*>(Ap + (2 * p) * PL + o);
                const H8 ai = *reinterpret_cast<const H8*>(Ap + (2 * p + 1) * PL + o);
                typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
                u32x4 u = __builtin_bit_cast(u32x4, ai);
                u ^= 0x80008000u;   // the sign bit of every 16-bit element
                const H8 an = __builtin_bit_cast(H8, u);
#pragma unroll
                for (int n = 0; n < TW; ++n)
#pragma unroll
                    for (int q = 0; q + p < NS; ++q) {
                        const int L = (NL > 1 && p + q > 0) ? 1 : 0;
                        accr[L][m][n] = mfma(ar, br[n][q], accr[L][m][n]);
                        acci[L][m][n] = mfma(ar, bi[n][q], acci[L][m][n]);
                        accr[L][m][n] = mfma(an, bi[n][q], accr[L][m][n]);
                        acci[L][m][n] = mfma(ai, br[n][q], acci[L][m][n]);
                    }
            }
        __syncthreads();
    }
    }   // pass
    if constexpr (F16 && sizeof(TC) == 8 && sizeof(TO) == 8) {
        // the fp16 split serves the power form (complex64 in and out): its epilogue in FP32, the
        // 2^-28 of the products folded into alpha (each term rounded once more than in FP64,
        // below the output's own rounding)
        const float a_re = (float)(g.alpha_re * kF16Unscale), a_im = (float)(g.alpha_im * kF16Unscale);
        const float b_re = (float)g.beta_re, b_im = (float)g.beta_im;
        const float c_re = (float)g.beta2_re, c_im = (float)g.beta2_im;
        const float g_re = (float)g.diag_re, g_im = (float)g.diag_im;
        const long ib = b * g.sc * dd, ib2 = b * g.sc2 * dd, ob = b * g.so * dd;
#pragma unroll
        for (int m = 0; m < TW; ++m)
#pragma unroll
            for (int n = 0; n < TW; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int gi = ti * BT + wr + 16 * m + 4 * kq + v, gj = tj * BT + wc + 16 * n + r16;
                    if (EDGE && (gi >= d || gj >= d)) continue;
                    const long o = (long)gi * d + gj;
                    const float xr = accr[0][m][n][v], xi = acci[0][m][n][v];
                    float yr = a_re * xr - a_im * xi, yi = a_re * xi + a_im * xr;
                    if (io.Cin) {
                        const TC cv = io.Cin[ib + o];
                        yr = __builtin_fmaf(b_re, cv.x, __builtin_fmaf(-b_im, cv.y, yr));
                        yi = __builtin_fmaf(b_re, cv.y, __builtin_fmaf(b_im, cv.x, yi));
                    }
                    if (io.Cin2) {
                        const TC cv = io.Cin2[ib2 + o];
                        yr = __builtin_fmaf(c_re, cv.x, __builtin_fmaf(-c_im, cv.y, yr));
                        yi = __builtin_fmaf(c_re, cv.y, __builtin_fmaf(c_im, cv.x, yi));
                    }
                    if (g.add_ab) {
                        const TA av = A[o];
                        const TB bv = B[o];
                        yr += av.x + bv.x; yi += av.y + bv.y;
                    }
                    if (gi == gj) { yr += g_re; yi += g_im; }
                    TO out;
                    out.x = yr; out.y = yi;
                    io.C[ob + o] = out;
                }
        return;
    }
#pragma unroll
    for (int m = 0; m < TW; ++m)
#pragma unroll
        for (int n = 0; n < TW; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = ti * BT + wr + 16 * m + 4 * kq + v, gj = tj * BT + wc + 16 * n + r16;
                if (EDGE && (gi >= d || gj >= d)) continue;
                constexpr double post = F16 ? kF16Unscale : 1.0;
                epi_store(g, io, A, B, b, dd, (long)gi * d + gj, gi == gj,
                          ((double)accr[0][m][n][v] + (NL > 1 ? (double)accr[NL - 1][m][n][v] : 0.0)) * post,
                          ((double)acci[0][m][n][v] + (NL > 1 ? (double)acci[NL - 1][m][n][v] : 0.0)) * post);
            }
}

template <int BT, class TA, class TB, class TC, class TO>
hipError_t launch_tiled(const GemmArgs& g, int mode, const EpiIO<TA, TB, TC, TO>& io, hipStream_t st) {
    const int tiles = (g.d + BT - 1) / BT;
    const long total = (long)g.batch * tiles * tiles;
    if (total <= 0) return hipSuccess;
    if (total > 0x7fffffffL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)total), block(256);
    switch (mode) {
#define PA_FP2(R, DU)                                                                                                   \
    if (g.d % BT) hipLaunchKernelGGL((zgemm_fp_kernel<R, BT, true, DU, TA, TB, TC, TO>), grid, block, 0, st, g, io);   \
    else hipLaunchKernelGGL((zgemm_fp_kernel<R, BT, false, DU, TA, TB, TC, TO>), grid, block, 0, st, g, io);
#define PA_FP(R) if (g.A2) { PA_FP2(R, true) } else { PA_FP2(R, false) }
        case GEMM_F64: PA_FP(double) break;
        case GEMM_F32: PA_FP(float) break;
#undef PA_FP
#undef PA_FP2
#define PA_SPLIT(NS, F16, DU)                                                                                          \
    if (g.d % BT) hipLaunchKernelGGL((zgemm_split_kernel<NS, BT, F16, true, DU, TA, TB, TC, TO>), grid, block, 0, st, g, io); \
    else hipLaunchKernelGGL((zgemm_split_kernel<NS, BT, F16, false, DU, TA, TB, TC, TO>), grid, block, 0, st, g, io);
        case GEMM_BF16X2: PA_SPLIT(2, false, false) break;
        case GEMM_BF16X3: PA_SPLIT(3, false, false) break;
        case GEMM_F16X2:
            if (g.A2) { PA_SPLIT(2, true, true) }
            else { PA_SPLIT(2, true, false) }
            break;
#undef PA_SPLIT
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class TA, class TB, class TC, class TO>
hipError_t launch_row(const GemmArgs& g, const EpiIO<TA, TB, TC, TO>& io, hipStream_t st) {
    const int nw = (g.d + 15) / 16;
    if (g.batch <= 0) return hipSuccess;
    const dim3 grid((unsigned)g.batch);
    switch (nw) {
        case 3: hipLaunchKernelGGL((zgemm_row_kernel<3, TA, TB, TC, TO>), grid, dim3(64 * 3), 0, st, g, io); break;
        case 4: hipLaunchKernelGGL((zgemm_row_kernel<4, TA, TB, TC, TO>), grid, dim3(64 * 4), 0, st, g, io); break;
        case 5: hipLaunchKernelGGL((zgemm_row_kernel<5, TA, TB, TC, TO>), grid, dim3(64 * 5), 0, st, g, io); break;
        case 6: hipLaunchKernelGGL((zgemm_row_kernel<6, TA, TB, TC, TO>), grid, dim3(64 * 6), 0, st, g, io); break;
        case 7: hipLaunchKernelGGL((zgemm_row_kernel<7, TA, TB, TC, TO>), grid, dim3(64 * 7), 0, st, g, io); break;
        case 8: hipLaunchKernelGGL((zgemm_row_kernel<8, TA, TB, TC, TO>), grid, dim3(64 * 8), 0, st, g, io); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class TA, class TB, class TC, class TO>
hipError_t launch_typed(const GemmArgs& g, int mode, hipStream_t st) {
    EpiIO<TA, TB, TC, TO> io;
    io.A = static_cast<const TA*>(g.A); io.B = static_cast<const TB*>(g.B);
    io.Cin = static_cast<const TC*>(g.Cin); io.Cin2 = static_cast<const TC*>(g.Cin2); io.C = static_cast<TO*>(g.C);
    // complex64 FP32 products at dims 33-128: the whole-product kernel when asked for (tile 16),
    // and by default where it measured faster than the 32/64 tiles (dims 33-48: 27-28%,
    // 97-112: 5-6%; at 49-96 and 113-128 it was level or slower)
    if constexpr (sizeof(TA) == 8 && sizeof(TB) == 8 && sizeof(TO) == 8) {
        const bool row_default = (g.d > 32 && g.d <= 48) || (g.d > 96 && g.d <= 112);
        if (mode == GEMM_F32 && !g.A2 && g.d > 32 && g.d <= 128 && (g.tile == 16 || (g.tile == 0 && row_default)))
            return launch_row(g, io, st);
    }
    return gemm_tile(g) == 32 ? launch_tiled<32>(g, mode, io, st) : launch_tiled<64>(g, mode, io, st);
}

}  // namespace

// Operand / output element types: complex128 when the flag is set, complex64 otherwise; Cin
// and Cin2 have the output's type.  FP64 products take complex64 operands (the transition from
// low-precision tree levels) and complex128 ones; the FP32 and split modes run on complex64.
hipError_t launch_zgemm(const GemmArgs& g, int mode, hipStream_t st) {
    if (g.d <= 0 || g.batch < 0) return hipErrorInvalidValue;
    if ((g.A2 || g.B2) && (!g.A2 || !g.B2 || mode == GEMM_BF16X2 || mode == GEMM_BF16X3 || g.add_ab))
        return hipErrorInvalidValue;
    if (mode == GEMM_F64) {
        if (!g.out_c128) return hipErrorInvalidValue;
        if (g.a_c128 && g.b_c128) return launch_typed<double2, double2, double2, double2>(g, mode, st);
        if (!g.a_c128 && !g.b_c128) return launch_typed<float2, float2, double2, double2>(g, mode, st);
        return hipErrorInvalidValue;
    }
    if (g.a_c128 || g.b_c128 || g.out_c128) return hipErrorInvalidValue;
    return launch_typed<float2, float2, float2, float2>(g, mode, st);
}

}  // namespace pa
