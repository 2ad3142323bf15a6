// Device bodies of the fused on-chip kernels, shared by the single-problem kernels
// (kernels.hip) and the multi-instance ensemble kernels (ensemble.hip).  The wave split SH
// provides G (waves), T (16x16 blocks per wave side) and row0/col0 (tile origin of a wave).
#pragma once
#include "kernels.hpp"

namespace pa {

// Square wave split of a D x D product: W x W waves (W = D / (16 T)), each a (16 T)^2 tile.
template <int D, int T_> struct TileShape {
    static constexpr int T = T_;
    static constexpr int W = D / (16 * T_);
    static constexpr int G = W * W;
    static __device__ __forceinline__ int row0(int w) { return 16 * T_ * (w / W); }
    static __device__ __forceinline__ int col0(int w) { return 16 * T_ * (w % W); }
};

// One slice of the fused kernels with the wave split SH (G waves, T x T blocks of 16x16 each),
// from its operator coefficients already in coef[]: returns E_j = U_j - I in the accumulator
// layout.  ops (global or LDS) holds H0 and the operators; are / aim hold the Chebyshev
// coefficients with are[0] = a_0 - 1.  LDS: planes Xr, Xi (stride D + Pad::A), Br, Bi (stride
// D + Pad::B).  Ends with a barrier: the planes and coef[] are free afterwards.
//
// NS > 1 (d <= D / NS): NS slices at once as the diagonal blocks of one D x D tile, block q at
// rows and columns [q D/NS, (q+1) D/NS) with coefficients coef[64 q ..].  Every step of the
// recurrence maps block-diagonal matrices to block-diagonal matrices, and a block's products
// only add exact zeros from the others, so each block holds what the slice alone would.
template <class R, int D, class SH, int NS = 1>
__device__ __forceinline__ CTile<R, SH::T, SH::T> cheb_slice_tile(const cplx<R>* __restrict__ ops, int ncoef, int d,
                                                                  int M, R s, const R* are, const R* aim, R* Xr,
                                                                  R* Xi, R* Br, R* Bi, const cplx<R>* coef) {
    constexpr int G = SH::G, T = SH::T;
    constexpr int SA = D + Pad<R>::A, SB = D + Pad<R>::B;
    constexpr int HB = D / NS;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int row0 = SH::row0(wave), col0 = SH::col0(wave);
    const int dd = d * d;
    for (int e = tid; e < D * D; e += 64 * G) {
        const int i = e / D, k = e % D;
        const int q = i / HB, ii = i % HB, kk = k % HB;
        R xr = R(0), xi = R(0);
        if (k / HB == q && ii < d && kk < d) {
            const cplx<R>* cq = coef + 64 * q;
            cplx<R> x = ops[ii * d + kk];
            for (int t = 0; t < ncoef; ++t)
                x = cadd(x, cmul(cq[t], ops[(size_t)(t + 1) * dd + ii * d + kk]));
            xr = x.x; xi = x.y;
        }
        Xr[i * SA + k] = xr; Xi[i * SA + k] = xi;
    }
    __syncthreads();

    CTile<R, T, T> b1, b2, acc;
    {
        const cplx<R> aM = mk<R>(are[M], aim[M]);
        const cplx<R> aM1 = mk<R>(are[M - 1], aim[M - 1]);
        const cplx<R> saM = mk<R>(s * aM.x, s * aM.y);
#pragma unroll
        for (int m = 0; m < T; ++m)
#pragma unroll
            for (int n = 0; n < T; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                    const bool dg = (row == col);
                    b2.re[m][n][v] = dg ? aM.x : R(0);
                    b2.im[m][n][v] = dg ? aM.y : R(0);
                    const cplx<R> t = cmul(saM, mk<R>(Xr[row * SA + col], Xi[row * SA + col]));
                    b1.re[m][n][v] = t.x + (dg ? aM1.x : R(0));
                    b1.im[m][n][v] = t.y + (dg ? aM1.y : R(0));
                    Br[row * SB + col] = b1.re[m][n][v];
                    Bi[row * SB + col] = b1.im[m][n][v];
                }
    }
    __syncthreads();

    for (int k = M - 2; k >= 0; --k) {
        wave_cgemm<R, T, T, D, SA, SB>(Xr, Xi, Br, Bi, row0, col0, acc);
        const R beta = (k == 0) ? R(-2) : R(-1);
        // k == 0: are[0] holds a_0 - 1 (rounded from FP64), so the result is E_j = U_j - I
        const R akr = are[k], aki = aim[k];
#pragma unroll
        for (int m = 0; m < T; ++m)
#pragma unroll
            for (int n = 0; n < T; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const bool dg = acc_row<R>(row0, m, v) == acc_col(col0, n);
                    const R nr = s * acc.re[m][n][v] + beta * b2.re[m][n][v] + (dg ? akr : R(0));
                    const R ni = s * acc.im[m][n][v] + beta * b2.im[m][n][v] + (dg ? aki : R(0));
                    b2.re[m][n][v] = b1.re[m][n][v];
                    b2.im[m][n][v] = b1.im[m][n][v];
                    b1.re[m][n][v] = nr;
                    b1.im[m][n][v] = ni;
                }
        if (k > 0) {
            __syncthreads();
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                        Br[row * SB + col] = b1.re[m][n][v];
                        Bi[row * SB + col] = b1.im[m][n][v];
                    }
            __syncthreads();
        }
    }
    __syncthreads();
    return b1;
}

// The same slice in the power form, for tau Hnorm <= 1: p(X) = sum_k m_k Y^k with Y = tau X by
// Horner's rule in S = Y^2,
//   Q_J = m_{2J} I + m_{2J+1} Y,  Q_j = m_{2j} I + m_{2j+1} Y + S Q_{j+1},  p(X) = Q_0  (J = (M-1)/2),
// (M+1)/2 products instead of the recurrence's M - 1.  pre[0] holds m_0 - 1, so Q_0 = E_j.  X goes
// into both operand layouts; after S = Y Y the A planes hold S and the B planes Q, and Y is taken
// from registers at the accumulator positions.  Ends with a barrier.
template <class R, int D, class SH>
__device__ __forceinline__ CTile<R, SH::T, SH::T> power_slice_tile(const cplx<R>* __restrict__ ops, int ncoef, int d,
                                                                   int M, R tau, const R* pre, const R* pim, R* Xr,
                                                                   R* Xi, R* Br, R* Bi, const cplx<R>* coef) {
    constexpr int G = SH::G, T = SH::T;
    constexpr int SA = D + Pad<R>::A, SB = D + Pad<R>::B;
    typedef CTile<R, T, T> Tile;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int row0 = SH::row0(wave), col0 = SH::col0(wave);
    const int dd = d * d;
    for (int e = tid; e < D * D; e += 64 * G) {
        const int i = e / D, k = e % D;
        R xr = R(0), xi = R(0);
        if (i < d && k < d) {
            cplx<R> x = ops[i * d + k];
            for (int t = 0; t < ncoef; ++t) x = cadd(x, cmul(coef[t], ops[(size_t)(t + 1) * dd + i * d + k]));
            xr = x.x; xi = x.y;
        }
        Xr[i * SA + k] = xr; Xi[i * SA + k] = xi;
        Br[i * SB + k] = xr; Bi[i * SB + k] = xi;
    }
    __syncthreads();
    Tile xt, acc, q;
#pragma unroll
    for (int m = 0; m < T; ++m)
#pragma unroll
        for (int n = 0; n < T; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                xt.re[m][n][v] = Xr[row * SA + col]; xt.im[m][n][v] = Xi[row * SA + col];
            }
    wave_cgemm<R, T, T, D, SA, SB>(Xr, Xi, Br, Bi, row0, col0, acc);   // X X
    // m_{2jj} I + m_{2jj+1} tau X (+ the product)
    auto horner = [&](int jj, bool with_product) {
        const cplx<R> mA = mk<R>(pre[2 * jj], pim[2 * jj]);
        const cplx<R> mB = mk<R>(tau * pre[2 * jj + 1], tau * pim[2 * jj + 1]);
#pragma unroll
        for (int m = 0; m < T; ++m)
#pragma unroll
            for (int n = 0; n < T; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const bool dg = acc_row<R>(row0, m, v) == acc_col(col0, n);
                    const cplx<R> y = cmul(mB, mk<R>(xt.re[m][n][v], xt.im[m][n][v]));
                    q.re[m][n][v] = y.x + (dg ? mA.x : R(0)) + (with_product ? acc.re[m][n][v] : R(0));
                    q.im[m][n][v] = y.y + (dg ? mA.y : R(0)) + (with_product ? acc.im[m][n][v] : R(0));
                }
    };
    const int J = (M - 1) / 2;
    const R tau2 = tau * tau;
    horner(J, false);
    __syncthreads();   // every wave has read the X planes
#pragma unroll
    for (int m = 0; m < T; ++m)
#pragma unroll
        for (int n = 0; n < T; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                Xr[row * SA + col] = tau2 * acc.re[m][n][v]; Xi[row * SA + col] = tau2 * acc.im[m][n][v];   // S
                Br[row * SB + col] = q.re[m][n][v]; Bi[row * SB + col] = q.im[m][n][v];                     // Q_J
            }
    __syncthreads();
    for (int jj = J - 1; jj >= 0; --jj) {
        wave_cgemm<R, T, T, D, SA, SB>(Xr, Xi, Br, Bi, row0, col0, acc);   // S Q_{jj+1}
        horner(jj, true);
        if (jj > 0) {
            __syncthreads();
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                        Br[row * SB + col] = q.re[m][n][v]; Bi[row * SB + col] = q.im[m][n][v];
                    }
            __syncthreads();
        }
    }
    __syncthreads();
    return q;
}

// One slice of the fused kernel (see cheb_fused_kernel): E_j = U_j - I of slice j into out
// (d x d); LDS as for cheb_slice_tile, plus coef[ncoef].
template <class R, int D, class SH>
__device__ __forceinline__ void cheb_slice(const cplx<R>* __restrict__ ops, int ncoef, int d, const Controls<R>& ctl,
                                           int j, int M, R s, const R* are, const R* aim, R* Xr, R* Xi, R* Br,
                                           R* Bi, cplx<R>* coef, cplx<R>* __restrict__ out) {
    constexpr int G = SH::G, T = SH::T;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int row0 = SH::row0(wave), col0 = SH::col0(wave);
    for (int t = tid; t < ncoef; t += 64 * G) coef[t] = slice_coefficient(ctl, t, j);
    __syncthreads();
    const CTile<R, T, T> e = cheb_slice_tile<R, D, SH>(ops, ncoef, d, M, s, are, aim, Xr, Xi, Br, Bi, coef);
#pragma unroll
    for (int m = 0; m < T; ++m)
#pragma unroll
        for (int n = 0; n < T; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                if (row < d && col < d) out[row * d + col] = mk<R>(e.re[m][n][v], e.im[m][n][v]);
            }
}

// Block blk of a tree level with the wave split SH: reduces nodes [blk B, blk B + B) of src (n_in
// nodes) and writes one node to dst (+ I when final_level), and, given store, every node it
// completes below that (the levels a gradient sweep reads back).  LDS planes Ar, Ai (stride
// D + Pad::A), Br, Bi (stride D + Pad::B).
template <class R, int D, int LOGB, class SrcT, class SH>
__device__ __forceinline__ void tree_block(const SrcT* __restrict__ src, int n_in, int d, cplx<R>* __restrict__ dst,
                                           int final_level, long blk, R* Ar, R* Ai, R* Br, R* Bi,
                                           const TreeStore<R>* store = nullptr) {
    constexpr int G = SH::G, T = SH::T;
    constexpr int SA = D + Pad<R>::A, SB = D + Pad<R>::B;
    constexpr int B = 1 << LOGB;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int row0 = SH::row0(wave), col0 = SH::col0(wave);
    const int dd = d * d;
    const long base = blk * B;

    auto stage = [&](long j, R* Pr, R* Pi, int S) {
        const bool live = j < n_in;
        const SrcT* p = src + (size_t)(live ? j : 0) * dd;
        for (int e = tid; e < D * D; e += 64 * G) {
            const int i = e / D, k = e % D;
            R xr = R(0), xi = R(0);
            if (live && i < d && k < d) { xr = R(p[i * d + k].x); xi = R(p[i * d + k].y); }
            Pr[i * S + k] = xr; Pi[i * S + k] = xi;
        }
    };
    auto tile_to_lds = [&](const CTile<R, T, T>& t, R* Pr, R* Pi, int S) {
#pragma unroll
        for (int m = 0; m < T; ++m)
#pragma unroll
            for (int n = 0; n < T; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                    Pr[row * S + col] = t.re[m][n][v];
                    Pi[row * S + col] = t.im[m][n][v];
                }
    };
    // out = A + B + A B from the staged planes (A later, B earlier)
    auto combine_staged = [&](CTile<R, T, T>& out) {
        __syncthreads();
        wave_cgemm<R, T, T, D, SA, SB>(Ar, Ai, Br, Bi, row0, col0, out);
#pragma unroll
        for (int m = 0; m < T; ++m)
#pragma unroll
            for (int n = 0; n < T; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                    out.re[m][n][v] += Ar[row * SA + col] + Br[row * SB + col];
                    out.im[m][n][v] += Ai[row * SA + col] + Bi[row * SB + col];
                }
        __syncthreads();
    };

    // node `idx` of level k (1 .. LOGB - 1) to the caller's array, when it exists at that level
    auto keep = [&](const CTile<R, T, T>& t, int k, long idx) {
        if (!store || idx >= store->n[k]) return;
        cplx<R>* o = store->lev[k] + (size_t)idx * dd;
#pragma unroll
        for (int m = 0; m < T; ++m)
#pragma unroll
            for (int n = 0; n < T; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                    if (row < d && col < d) o[row * d + col] = mk<R>(t.re[m][n][v], t.im[m][n][v]);
                }
    };
    CTile<R, T, T> st[LOGB > 1 ? LOGB : 1];  // st[l]: completed level-l node waiting for its sibling
    CTile<R, T, T> cur;
#pragma unroll
    for (int p = 0; p < B / 2; ++p) {
        stage(base + 2 * p + 1, Ar, Ai, SA);
        stage(base + 2 * p, Br, Bi, SB);
        combine_staged(cur);  // level-1 node p
        if (LOGB > 1) keep(cur, 1, blk * (B / 2) + p);
#pragma unroll
        for (int l = 1; l < LOGB; ++l) {
            if (((p >> (l - 1)) & 1) == 0) { st[l] = cur; break; }
            tile_to_lds(cur, Ar, Ai, SA);
            tile_to_lds(st[l], Br, Bi, SB);
            combine_staged(cur);   // level-(l+1) node p >> l
            if (l + 1 < LOGB) keep(cur, l + 1, blk * (B >> (l + 1)) + (p >> l));
        }
    }
    cplx<R>* out = dst;
    const R diag = final_level ? R(1) : R(0);
#pragma unroll
    for (int m = 0; m < T; ++m)
#pragma unroll
        for (int n = 0; n < T; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                if (row < d && col < d)
                    out[row * d + col] = mk<R>(cur.re[m][n][v] + (row == col ? diag : R(0)), cur.im[m][n][v]);
            }
    __syncthreads();
}

}  // namespace pa
