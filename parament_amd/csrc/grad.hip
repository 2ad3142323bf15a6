// Adjoint (GRAPE) gradient of the gate fidelity F = |Tr(U_t^H U)|^2 / d^2 with respect to the
// control samples, for U = p(X_{N-1}) ... p(X_0) as equiprop computes it (p = the degree-M
// Chebyshev approximant of exp(-i tau .), parament.cpp:569-652).  No reference counterpart.
//
// With g = Tr(U_t^H U), P_j = U_{j-1}..U_0 and S_j = U_{N-1}..U_{j+1}:
//     dg / dc'_tj = Tr(U_t^H S_j Dp(X_j)[O_t] P_j) = Tr(W_j Dp(X_j)[O_t]),  W_j = P_j U_t^H S_j,
// and, because Dp(X)[E] = sum_m alpha_m sum_r X^r E X^(m-1-r) (p in monomial form) and the trace
// is cyclic, Tr(W Dp(X)[O]) = Tr(Dp(X)[W] O).  So one directional derivative per slice,
//     Q_j = Dp(X_j)[W_j]   (forward-mode differentiation of the Clenshaw recurrence:
//                           db_k = s (W b_{k+1} + X db_{k+1}) - db_{k+2}),
// gives every control's gradient through the traces Tr(Q_j O_t).
//
// The costates W come from a log-depth scan over the pairwise product tree (node (l, i) =
// product of slices [i 2^l, (i+1) 2^l), later slices on the left, an odd last node carried up
// alone).  For a node n covering [a, b), W(n) = P_a U_t^H S_{b-1}; the root has W = U_t^H, and a
// parent with children c0 (earlier) and c1 (later) gives
//     W(c0) = W(parent) V(c1),   W(c1) = V(c0) W(parent),
// an only child inheriting W(parent).  The up-sweep keeps the tree's levels, the down-sweep
// runs from the root to level 1, and the slice kernel forms each leaf's W_j from its parent's
// W and its sibling's E = U - I.  The quadrature's chain rule then maps dF/dc'_tj to the raw
// samples.
#include "kernels.hpp"
#include "fused_device.hpp"

namespace pa {

namespace {

template <class R, int D> struct GL {
    static constexpr int G = Shape<D>::G, T = Shape<D>::T;
    static constexpr int SA = D + Pad<R>::A, SB = D + Pad<R>::B;
};

// Stage a d x d row-major matrix into zero-padded LDS planes, adding diag to its diagonal.
template <class R, int D>
__device__ __forceinline__ void stage(const cplx<R>* __restrict__ src, int d, R* Pr, R* Pi, int S, R diag = R(0)) {
    for (int e = threadIdx.x; e < D * D; e += 64 * GL<R, D>::G) {
        const int i = e / D, k = e % D;
        R xr = R(0), xi = R(0);
        if (i < d && k < d) {
            const cplx<R> v = src[i * d + k];
            xr = v.x + (i == k ? diag : R(0));
            xi = v.y;
        }
        Pr[i * S + k] = xr; Pi[i * S + k] = xi;
    }
}

// The same for the conjugate transpose of a d x d row-major matrix (no diagonal term).
template <class R, int D>
__device__ __forceinline__ void stage_ct(const cplx<R>* __restrict__ src, int d, R* Pr, R* Pi, int S) {
    for (int e = threadIdx.x; e < D * D; e += 64 * GL<R, D>::G) {
        const int i = e / D, k = e % D;
        R xr = R(0), xi = R(0);
        if (i < d && k < d) {
            const cplx<R> v = src[k * d + i];
            xr = v.x;
            xi = -v.y;
        }
        Pr[i * S + k] = xr; Pi[i * S + k] = xi;
    }
}

template <class R, int D>
__device__ __forceinline__ void tile_store(const CTile<R, GL<R, D>::T, GL<R, D>::T>& w, int d, cplx<R>* dst) {
    constexpr int T = GL<R, D>::T;
    const int wave = threadIdx.x >> 6;
    const int row0 = Shape<D>::row0(wave), col0 = Shape<D>::col0(wave);
#pragma unroll
    for (int m = 0; m < T; ++m)
#pragma unroll
        for (int n = 0; n < T; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                if (row < d && col < d) dst[row * d + col] = mk<R>(w.re[m][n][v], w.im[m][n][v]);
            }
}

template <class R, int D>
__device__ __forceinline__ void tile_to_lds(const CTile<R, GL<R, D>::T, GL<R, D>::T>& t, R* Pr, R* Pi, int S) {
    constexpr int T = GL<R, D>::T;
    const int wave = threadIdx.x >> 6;
    const int row0 = Shape<D>::row0(wave), col0 = Shape<D>::col0(wave);
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
}

// One derivative step's products in a single pass over K, sharing the operand loads:
//     cb = X b        (skipped when with_b is false),
//     cd = W b + X db
// X, W are A operands (stride SA), b, db are B operands (stride SB); 4M complex products.
template <class R, int T, int KD, int SA, int SB>
__device__ __forceinline__ void wave_step_products(const R* __restrict__ Xr, const R* __restrict__ Xi,
                                                   const R* __restrict__ Wr, const R* __restrict__ Wi,
                                                   const R* __restrict__ Br, const R* __restrict__ Bi,
                                                   const R* __restrict__ Dr, const R* __restrict__ Di, int row0,
                                                   int col0, bool with_b, CTile<R, T, T>& cb, CTile<R, T, T>& cd) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
    cb.zero();
    cd.zero();
#pragma unroll 2
    for (int k0 = 0; k0 < KD; k0 += 4) {
        R xr[T], xi[T], xn[T], wr[T], wi[T], wn[T], br[T], bi[T], dr[T], di[T];
#pragma unroll
        for (int m = 0; m < T; ++m) {
            const int idx = (row0 + 16 * m + r) * SA + k0 + q;
            xr[m] = Xr[idx]; xi[m] = Xi[idx]; xn[m] = -xi[m];
            wr[m] = Wr[idx]; wi[m] = Wi[idx]; wn[m] = -wi[m];
        }
#pragma unroll
        for (int n = 0; n < T; ++n) {
            const int idx = (k0 + q) * SB + col0 + 16 * n + r;
            br[n] = Br[idx]; bi[n] = Bi[idx];
            dr[n] = Dr[idx]; di[n] = Di[idx];
        }
#pragma unroll
        for (int m = 0; m < T; ++m)
#pragma unroll
            for (int n = 0; n < T; ++n) {
                if (with_b) {
                    cb.re[m][n] = Mma<R>::op(xr[m], br[n], cb.re[m][n]);
                    cb.im[m][n] = Mma<R>::op(xr[m], bi[n], cb.im[m][n]);
                    cb.re[m][n] = Mma<R>::op(xn[m], bi[n], cb.re[m][n]);
                    cb.im[m][n] = Mma<R>::op(xi[m], br[n], cb.im[m][n]);
                }
                cd.re[m][n] = Mma<R>::op(wr[m], br[n], cd.re[m][n]);
                cd.im[m][n] = Mma<R>::op(wr[m], bi[n], cd.im[m][n]);
                cd.re[m][n] = Mma<R>::op(wn[m], bi[n], cd.re[m][n]);
                cd.im[m][n] = Mma<R>::op(wi[m], br[n], cd.im[m][n]);
                cd.re[m][n] = Mma<R>::op(xr[m], dr[n], cd.re[m][n]);
                cd.im[m][n] = Mma<R>::op(xr[m], di[n], cd.im[m][n]);
                cd.re[m][n] = Mma<R>::op(xn[m], di[n], cd.re[m][n]);
                cd.im[m][n] = Mma<R>::op(xi[m], dr[n], cd.im[m][n]);
            }
    }
}

}  // namespace

// Up-sweep, LOGB levels per launch: block w reduces input nodes [w 2^LOGB, (w+1) 2^LOGB) with
// the forward tree's block body (pairing, odd carry and deviation-from-identity form are the
// same), keeping every node of the intermediate levels; the last level goes to st.lev[LOGB].
// All levels hold E = V - I.
template <class R, int D, int LOGB>
__global__ __launch_bounds__(64 * Shape<D>::G) void up_block_kernel(const cplx<R>* __restrict__ in, int n_in, int d,
                                                                       TreeStore<R> st) {
    constexpr int SA = D + Pad<R>::A, SB = D + Pad<R>::B;
    __shared__ R Ar[D * SA], Ai[D * SA], Br[D * SB], Bi[D * SB];
    tree_block<R, D, LOGB, cplx<R>, Shape<D>>(in, n_in, d, st.lev[LOGB] + (size_t)blockIdx.x * d * d, 0, blockIdx.x,
                                              Ar, Ai, Br, Bi, &st);
}

// Down-sweep level: child c of parent c/2 gets W(c) = W(parent) V(c+1) if it is the earlier
// child (c even), W(c) = V(c-1) W(parent) if it is the later one, and W(parent) if it has no
// sibling (V = I + E from the stored deviations).  One work-group per child.
template <class R, int D>
__global__ __launch_bounds__(64 * Shape<D>::G) void down_level_kernel(const cplx<R>* __restrict__ V, int n_child, int d,
                                                                         const cplx<R>* __restrict__ Wp,
                                                                         cplx<R>* __restrict__ Wc) {
    constexpr int G = GL<R, D>::G, T = GL<R, D>::T, SA = GL<R, D>::SA, SB = GL<R, D>::SB;
    __shared__ R Ar[D * SA], Ai[D * SA], Br[D * SB], Bi[D * SB];
    const int c = blockIdx.x, sib = c ^ 1, dd = d * d, wave = threadIdx.x >> 6;
    const cplx<R>* P = Wp + (size_t)(c >> 1) * dd;
    cplx<R>* dst = Wc + (size_t)c * dd;
    if (sib >= n_child) {
        for (int e = threadIdx.x; e < dd; e += 64 * G) dst[e] = P[e];
        return;
    }
    const cplx<R>* S = V + (size_t)sib * dd;   // E = V - I of the sibling
    if ((c & 1) == 0) { stage<R, D>(P, d, Ar, Ai, SA); stage<R, D>(S, d, Br, Bi, SB, R(1)); }
    else { stage<R, D>(S, d, Ar, Ai, SA, R(1)); stage<R, D>(P, d, Br, Bi, SB); }
    __syncthreads();
    CTile<R, T, T> w;
    wave_cgemm<R, T, T, D, SA, SB>(Ar, Ai, Br, Bi, Shape<D>::row0(wave), Shape<D>::col0(wave), w);
    tile_store<R, D>(w, d, dst);
}

// The root's children and g in one launch: child c (block c < 2) gets W(c) = t^H V(c + 1) or
// V(c - 1) t^H (the root's costate is t^H itself, staged from t), and block 2 forms
// g = Tr(t^H (I + e)) from the root's deviation e (FP64 partial sums, fixed reduction order) into
// gout[0].  Replaces the root launch and the first down-sweep level.
template <class R, int D>
__global__ __launch_bounds__(64 * Shape<D>::G) void down_root_kernel(const cplx<R>* __restrict__ V,
                                                                        const cplx<R>* __restrict__ t,
                                                                        const cplx<R>* __restrict__ u, int d,
                                                                        cplx<R>* __restrict__ Wc,
                                                                        cplx<R>* __restrict__ gout) {
    constexpr int G = GL<R, D>::G, T = GL<R, D>::T, SA = GL<R, D>::SA, SB = GL<R, D>::SB;
    __shared__ R Ar[D * SA], Ai[D * SA], Br[D * SB], Bi[D * SB];
    __shared__ double sr[64 * G], si[64 * G];
    const int c = blockIdx.x, dd = d * d, wave = threadIdx.x >> 6;
    if (c == 2) {
        double pr = 0.0, pi = 0.0;
        for (int e = threadIdx.x; e < dd; e += 64 * G) {
            const cplx<R> a = t[e];
            const double bx = (double)u[e].x + (e / d == e % d ? 1.0 : 0.0), by = (double)u[e].y;
            pr += (double)a.x * bx + (double)a.y * by;   // conj(a) b
            pi += (double)a.x * by - (double)a.y * bx;
        }
        sr[threadIdx.x] = pr; si[threadIdx.x] = pi;
        __syncthreads();
        for (int o = 32 * G; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) { sr[threadIdx.x] += sr[threadIdx.x + o]; si[threadIdx.x] += si[threadIdx.x + o]; }
            __syncthreads();
        }
        if (threadIdx.x == 0) gout[0] = mk<R>((R)sr[0], (R)si[0]);
        return;
    }
    const cplx<R>* S = V + (size_t)(c ^ 1) * dd;   // E = V - I of the sibling
    if (c == 0) { stage_ct<R, D>(t, d, Ar, Ai, SA); stage<R, D>(S, d, Br, Bi, SB, R(1)); }
    else { stage<R, D>(S, d, Ar, Ai, SA, R(1)); stage_ct<R, D>(t, d, Br, Bi, SB); }
    __syncthreads();
    CTile<R, T, T> w;
    wave_cgemm<R, T, T, D, SA, SB>(Ar, Ai, Br, Bi, Shape<D>::row0(wave), Shape<D>::col0(wave), w);
    tile_store<R, D>(w, d, Wc + (size_t)c * dd);
}

// Per slice j = j0 + jl (jl < a.N; j0 even): X_j, the costate W_j from the level-1 costate of
// its parent and the E = U - I of its sibling (E holds this launch's leaves, W1 their
// parents), then Q_j = Dp(X_j)[W_j] and dg[t * dg_stride + j] = Tr(Q_j O_t).
template <class R, int D, bool PS>
__global__ __launch_bounds__(64 * Shape<D>::G, 2) void grad_slice_kernel(ChebArgs<R> a, const cplx<R>* __restrict__ E,
                                                                         const cplx<R>* __restrict__ W1,
                                                                         cplx<R>* __restrict__ dg, int dg_stride,
                                                                         GradDirect<R> gd) {
    constexpr int G = GL<R, D>::G, T = GL<R, D>::T, SA = GL<R, D>::SA, SB = GL<R, D>::SB;
    __shared__ R Xr[D * SA], Xi[D * SA], Wr[D * SA], Wi[D * SA], Br[D * SB], Bi[D * SB], Dr[D * SB], Di[D * SB];
    __shared__ cplx<R> coef[kGradCoeffs];
    __shared__ cplx<R> red[kGradCoeffs * G];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int row0 = Shape<D>::row0(wave), col0 = Shape<D>::col0(wave);
    const int d = a.d, dd = d * d, M = a.M;
    const R s = a.s;
    typedef CTile<R, T, T> Tile;

    for (int jl = blockIdx.x; jl < a.N; jl += gridDim.x) {
        const int j = a.j0 + jl;
        for (int t = tid; t < a.ncoef; t += 64 * G) coef[t] = slice_coefficient(a.ctl, t, j);
        __syncthreads();
        for (int e = tid; e < D * D; e += 64 * G) {
            const int i = e / D, k = e % D;
            cplx<R> x = mk<R>(0, 0);
            if (i < d && k < d) {
                x = a.ops[i * d + k];
                for (int t = 0; t < a.ncoef; ++t) x = cadd(x, cmul(coef[t], a.ops[(size_t)(t + 1) * dd + i * d + k]));
            }
            Xr[i * SA + k] = x.x; Xi[i * SA + k] = x.y;
        }
        // W_j = W(parent) (I + E_sib) for the earlier child, (I + E_sib) W(parent) for the later
        const int sib = jl ^ 1;
        const cplx<R>* Wp = W1 + (size_t)(jl >> 1) * dd;
        if (sib < a.N) {
            const bool earlier = (jl & 1) == 0;
            const cplx<R>* Es = E + (size_t)sib * dd;
            if (earlier) { stage<R, D>(Wp, d, Wr, Wi, SA); stage<R, D>(Es, d, Br, Bi, SB); }
            else { stage<R, D>(Es, d, Wr, Wi, SA); stage<R, D>(Wp, d, Br, Bi, SB); }
            __syncthreads();
            Tile acc;
            wave_cgemm<R, T, T, D, SA, SB>(Wr, Wi, Br, Bi, row0, col0, acc);
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                        acc.re[m][n][v] += earlier ? Wr[row * SA + col] : Br[row * SB + col];
                        acc.im[m][n][v] += earlier ? Wi[row * SA + col] : Bi[row * SB + col];
                    }
            __syncthreads();
            tile_to_lds<R, D>(acc, Wr, Wi, SA);
        } else {
            stage<R, D>(Wp, d, Wr, Wi, SA);
        }
        __syncthreads();

        Tile b1, b2, db1, db2;
        if constexpr (PS) {
            // Forward-mode derivative of the power form P(Y) = sum_k m_k Y^k, Y = tau X, by Horner's
            // rule in S = Y^2 with direction V = tau W:  dS = Y V + V Y,
            //   Q_J = m_{2J} I + m_{2J+1} Y,  dQ_J = m_{2J+1} V,
            //   Q_j = m_{2j} I + m_{2j+1} Y + S Q_{j+1},  dQ_j = m_{2j+1} V + dS Q_{j+1} + S dQ_{j+1},
            //   Dp(X)[W] = m_1 V + dS Q_1 + S dQ_1.
            // Each step's three products are one pass of wave_step_products (A planes S, dS; B planes
            // Q, dQ); X X and W X + X W come from the same pass on (X, W; X, W).
            const R tau = a.ctl.tau, tau2 = tau * tau;
            const int J = (M - 1) / 2;
            Tile xt, wt;   // Y / tau and V / tau at the accumulator positions
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                        xt.re[m][n][v] = Xr[row * SA + col]; xt.im[m][n][v] = Xi[row * SA + col];
                        wt.re[m][n][v] = Wr[row * SA + col]; wt.im[m][n][v] = Wi[row * SA + col];
                    }
            for (int e = tid; e < D * D; e += 64 * G) {
                const int i = e / D, k = e % D;
                Br[i * SB + k] = Xr[i * SA + k]; Bi[i * SB + k] = Xi[i * SA + k];
                Dr[i * SB + k] = Wr[i * SA + k]; Di[i * SB + k] = Wi[i * SA + k];
            }
            __syncthreads();
            Tile cb, cd;
            wave_step_products<R, T, D, SA, SB>(Xr, Xi, Wr, Wi, Br, Bi, Dr, Di, row0, col0, true, cb, cd);
            // Q (b1) and dQ (db1) of the current Horner level from m_{2j}, m_{2j+1} and the products
            auto horner = [&](int jj, bool with_products) {
                const cplx<R> mA = mk<R>(a.pre[2 * jj], a.pim[2 * jj]);
                const cplx<R> mB = mk<R>(tau * a.pre[2 * jj + 1], tau * a.pim[2 * jj + 1]);
#pragma unroll
                for (int m = 0; m < T; ++m)
#pragma unroll
                    for (int n = 0; n < T; ++n)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const bool dg_ = acc_row<R>(row0, m, v) == acc_col(col0, n);
                            const cplx<R> y = cmul(mB, mk<R>(xt.re[m][n][v], xt.im[m][n][v]));
                            const cplx<R> w = cmul(mB, mk<R>(wt.re[m][n][v], wt.im[m][n][v]));
                            b1.re[m][n][v] = y.x + (dg_ ? mA.x : R(0)) + (with_products ? cb.re[m][n][v] : R(0));
                            b1.im[m][n][v] = y.y + (dg_ ? mA.y : R(0)) + (with_products ? cb.im[m][n][v] : R(0));
                            db1.re[m][n][v] = w.x + (with_products ? cd.re[m][n][v] : R(0));
                            db1.im[m][n][v] = w.y + (with_products ? cd.im[m][n][v] : R(0));
                        }
            };
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        b2.re[m][n][v] = tau2 * cb.re[m][n][v]; b2.im[m][n][v] = tau2 * cb.im[m][n][v];     // S
                        db2.re[m][n][v] = tau2 * cd.re[m][n][v]; db2.im[m][n][v] = tau2 * cd.im[m][n][v];   // dS
                    }
            horner(J, false);
            __syncthreads();
            tile_to_lds<R, D>(b2, Xr, Xi, SA);
            tile_to_lds<R, D>(db2, Wr, Wi, SA);
            tile_to_lds<R, D>(b1, Br, Bi, SB);
            tile_to_lds<R, D>(db1, Dr, Di, SB);
            __syncthreads();
            for (int jj = J - 1; jj >= 1; --jj) {
                wave_step_products<R, T, D, SA, SB>(Xr, Xi, Wr, Wi, Br, Bi, Dr, Di, row0, col0, true, cb, cd);
                horner(jj, true);
                __syncthreads();
                tile_to_lds<R, D>(b1, Br, Bi, SB);
                tile_to_lds<R, D>(db1, Dr, Di, SB);
                __syncthreads();
            }
            wave_step_products<R, T, D, SA, SB>(Xr, Xi, Wr, Wi, Br, Bi, Dr, Di, row0, col0, false, cb, cd);
            {
                const cplx<R> m1 = mk<R>(tau * a.pre[1], tau * a.pim[1]);
#pragma unroll
                for (int m = 0; m < T; ++m)
#pragma unroll
                    for (int n = 0; n < T; ++n)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const cplx<R> w = cmul(m1, mk<R>(wt.re[m][n][v], wt.im[m][n][v]));
                            db1.re[m][n][v] = w.x + cd.re[m][n][v];
                            db1.im[m][n][v] = w.y + cd.im[m][n][v];
                        }
            }
        } else {
        // b_M = a_M I, db_M = 0; b_{M-1} = s a_M X + a_{M-1} I, db_{M-1} = s a_M W
        {
            const cplx<R> saM = mk<R>(s * a.are[M], s * a.aim[M]);
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                        const bool dg_ = row == col;
                        b2.re[m][n][v] = dg_ ? a.are[M] : R(0); b2.im[m][n][v] = dg_ ? a.aim[M] : R(0);
                        db2.re[m][n][v] = R(0); db2.im[m][n][v] = R(0);
                        const cplx<R> xv = cmul(saM, mk<R>(Xr[row * SA + col], Xi[row * SA + col]));
                        b1.re[m][n][v] = xv.x + (dg_ ? a.are[M - 1] : R(0));
                        b1.im[m][n][v] = xv.y + (dg_ ? a.aim[M - 1] : R(0));
                        const cplx<R> wv = cmul(saM, mk<R>(Wr[row * SA + col], Wi[row * SA + col]));
                        db1.re[m][n][v] = wv.x; db1.im[m][n][v] = wv.y;
                    }
        }
        tile_to_lds<R, D>(b1, Br, Bi, SB);
        tile_to_lds<R, D>(db1, Dr, Di, SB);
        __syncthreads();
        Tile cb, cd;
        // One step k: bc = b_{k+1}, dbc = db_{k+1} (both staged in LDS); bp = b_{k+2} -> b_k,
        // dbp = db_{k+2} -> db_k:  b_k = s X b_{k+1} + beta b_{k+2} + a_k I (not needed at k = 0),
        // db_k = s (W b_{k+1} + X db_{k+1}) + beta db_{k+2}.
        auto step = [&](Tile& bp, Tile& dbp, int k) {
            const R beta = (k == 0) ? R(-2) : R(-1);
            wave_step_products<R, T, D, SA, SB>(Xr, Xi, Wr, Wi, Br, Bi, Dr, Di, row0, col0, k > 0, cb, cd);
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const bool dg_ = acc_row<R>(row0, m, v) == acc_col(col0, n);
                        bp.re[m][n][v] = s * cb.re[m][n][v] + beta * bp.re[m][n][v] + (dg_ ? a.are[k] : R(0));
                        bp.im[m][n][v] = s * cb.im[m][n][v] + beta * bp.im[m][n][v] + (dg_ ? a.aim[k] : R(0));
                        dbp.re[m][n][v] = s * cd.re[m][n][v] + beta * dbp.re[m][n][v];
                        dbp.im[m][n][v] = s * cd.im[m][n][v] + beta * dbp.im[m][n][v];
                    }
            if (k > 0) {
                __syncthreads();
                tile_to_lds<R, D>(bp, Br, Bi, SB);
                tile_to_lds<R, D>(dbp, Dr, Di, SB);
                __syncthreads();
            }
        };
        // M is odd, so the M - 1 steps k = M-2 .. 0 come in pairs and the roles alternate.
        for (int k = M - 2; k >= 1; k -= 2) {
            step(b2, db2, k);       // b2 <- b_k, db2 <- db_k
            step(b1, db1, k - 1);   // b1 <- b_{k-1}, db1 <- db_{k-1}
        }
        }
        // db1 = Q_j; dg[t][j] = sum_{r,c} Q[r][c] O_t[c][r] = sum Q[r][c] O_t^T[r][c]: the transposed
        // operators give coalesced loads, two controls' loads are issued together, and each
        // wave's partial sums are reduced across lanes, then across waves through LDS
        for (int t0 = 0; t0 < a.ncoef; t0 += 2) {
            cplx<R> o[2][4];
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = acc_row<R>(row0, 0, v), col = acc_col(col0, 0);
                    o[q][v] = (t0 + q < a.ncoef && row < d && col < d)
                                  ? a.opsT[(size_t)(t0 + q) * dd + row * d + col] : mk<R>(0, 0);
                }
            R pr[2], pi[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                pr[q] = R(0); pi[q] = R(0);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    pr[q] += db1.re[0][0][v] * o[q][v].x - db1.im[0][0][v] * o[q][v].y;
                    pi[q] += db1.re[0][0][v] * o[q][v].y + db1.im[0][0][v] * o[q][v].x;
                }
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
                for (int q = 0; q < 2; ++q) { pr[q] += __shfl_xor(pr[q], off, 64); pi[q] += __shfl_xor(pi[q], off, 64); }
            if (lane == 0)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    if (t0 + q >= a.ncoef) break;
                    if (G > 1) {
                        red[(t0 + q) * G + wave] = mk<R>(pr[q], pi[q]);
                    } else if (gd.grad) {   // no quadrature: dF/dc_tj = 2 Re(conj(g) dg_tj) / d^2
                        if (t0 + q < gd.amps) {
                            const cplx<R> g = gd.g[0];
                            gd.grad[(size_t)(t0 + q) * gd.pts + j] = R(2) * (g.x * pr[q] + g.y * pi[q]) * (R(1) / ((R)d * (R)d));
                        }
                    } else {
                        dg[(size_t)(t0 + q) * dg_stride + j] = mk<R>(pr[q], pi[q]);
                    }
                }
        }
        if (G == 1 && gd.grad && gd.fid && j == 0 && tid == 0) {
            const cplx<R> g = gd.g[0];
            gd.fid[0] = (g.x * g.x + g.y * g.y) * (R(1) / ((R)d * (R)d));
        }
        if (G > 1) {
            __syncthreads();
            for (int t = tid; t < a.ncoef; t += 64 * G) {
                cplx<R> sum = red[t * G];
                for (int w = 1; w < G; ++w) sum = cadd(sum, red[t * G + w]);
                if (gd.grad) {
                    // no quadrature: sample j is slice j, dF/dc_tj = 2 Re(conj(g) dg_tj) / d^2
                    if (t < gd.amps) {
                        const cplx<R> g = gd.g[0];
                        gd.grad[(size_t)t * gd.pts + j] = R(2) * (g.x * sum.x + g.y * sum.y) * (R(1) / ((R)d * (R)d));
                    }
                } else {
                    dg[(size_t)t * dg_stride + j] = sum;
                }
            }
            if (gd.grad && gd.fid && j == 0 && tid == 0) {
                const cplx<R> g = gd.g[0];
                gd.fid[0] = (g.x * g.x + g.y * g.y) * (R(1) / ((R)d * (R)d));
            }
        }
        __syncthreads();
    }
}

// Root of the gradient sweep, one launch: blocks 0 .. nb-1 write the root costate W = t^H (times
// wscale[0].y when wscale is given), block nb forms g = Tr(t^H (I + e)) from the root's deviation
// e (FP64 partial sums, fixed reduction order) into gout[0].
template <class R>
__global__ __launch_bounds__(256) void grad_root_kernel(const cplx<R>* __restrict__ t, const cplx<R>* __restrict__ u,
                                                        int d, int nb, cplx<R>* __restrict__ wroot,
                                                        cplx<R>* __restrict__ gout, const cplx<R>* __restrict__ wscale) {
    __shared__ double sr[256], si[256];
    if ((int)blockIdx.x < nb) {
        const int e = blockIdx.x * 256 + threadIdx.x;
        if (e >= d * d) return;
        const int i = e / d, k = e % d;
        const cplx<R> v = t[k * d + i];
        const R f = wscale ? wscale[0].y : R(1);
        wroot[e] = mk<R>(f * v.x, -f * v.y);
        return;
    }
    double pr = 0.0, pi = 0.0;
    for (int e = threadIdx.x; e < d * d; e += 256) {
        const cplx<R> a = t[e];
        const double bx = (double)u[e].x + (e / d == e % d ? 1.0 : 0.0), by = (double)u[e].y;
        pr += (double)a.x * bx + (double)a.y * by;   // conj(a) b
        pi += (double)a.x * by - (double)a.y * bx;
    }
    sr[threadIdx.x] = pr; si[threadIdx.x] = pi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sr[threadIdx.x] += sr[threadIdx.x + o]; si[threadIdx.x] += si[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) gout[0] = mk<R>((R)sr[0], (R)si[0]);
}

// out[0] = (c, 1 / c) with c = ||t||_F (1 for t = 0), one work-group: every entry of a costate
// formed from t^H / c by products of (near-)unitary factors is below 1 in modulus, since it is
// at most ||t||_2 / c.  FP64 partial sums, fixed reduction order.
template <class R>
__global__ __launch_bounds__(256) void target_scale_kernel(const cplx<R>* __restrict__ t, int d,
                                                           cplx<R>* __restrict__ out) {
    __shared__ double sq[256];
    double a = 0.0;
    for (long e = threadIdx.x; e < (long)d * d; e += 256) {
        const cplx<R> v = t[e];
        a += (double)v.x * v.x + (double)v.y * v.y;
    }
    sq[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sq[threadIdx.x] += sq[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double c = sq[0] > 0.0 ? sqrt(sq[0]) : 1.0;
        out[0] = mk<R>((R)c, (R)(1.0 / c));
    }
}

// grad[k * pts + p] = sum_j w_jp 2 Re(conj(g) dg[k][j]) / d^2 (quadrature chain rule), F = |g|^2 / d^2.
template <class R>
__global__ void grad_finish_kernel(const cplx<R>* __restrict__ dg, const cplx<R>* __restrict__ gptr, int N, int pts,
                                   int amps, int mode, int d, R* __restrict__ grad, R* __restrict__ fid) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    const cplx<R> g = gptr[0];
    const R inv = R(1) / ((R)d * (R)d);
    if (idx == 0 && fid) fid[0] = (g.x * g.x + g.y * g.y) * inv;
    if (idx >= (long)amps * pts) return;
    const int k = (int)(idx / pts), p = (int)(idx % pts);
    auto term = [&](int j, R w) -> R {
        if (j < 0 || j >= N) return R(0);
        const cplx<R> v = dg[(size_t)k * N + j];
        return w * R(2) * (g.x * v.x + g.y * v.y) * inv;   // 2 Re(conj(g) v) / d^2
    };
    R r = R(0);
    if (mode == QUAD_NONE) r = term(p, R(1));
    else if (mode == QUAD_MIDPOINT) r = term(p - 1, R(0.5)) + term(p, R(0.5));
    else {
        if (p & 1) r = term((p - 1) / 2, R(4) / R(6));
        else r = term(p / 2, R(1) / R(6)) + term(p / 2 - 1, R(1) / R(6));
    }
    grad[idx] = r;
}

// Traces of the step-wise gradient: dg[t * dg_stride + j0 + j] = Tr(Q_j O_t) for the n slices
// of Q (one work-group per slice), from the transposed operators, so that both operands are
// read along rows, times scale[0].x when scale is given.  Eight operators per pass over Q_j.
template <class R>
__global__ __launch_bounds__(256) void grad_trace_kernel(const cplx<R>* __restrict__ Q, int d,
                                                         const cplx<R>* __restrict__ opsT, int ncoef,
                                                         cplx<R>* __restrict__ dg, int dg_stride, int j0,
                                                         const cplx<R>* __restrict__ scale) {
    constexpr int KB = 8;
    __shared__ R sr[KB][4], si[KB][4];
    const long dd = (long)d * d;
    const int j = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const cplx<R>* q = Q + (size_t)j * dd;
    for (int t0 = 0; t0 < ncoef; t0 += KB) {
        R pr[KB], pi[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) { pr[u] = R(0); pi[u] = R(0); }
        for (long e = threadIdx.x; e < dd; e += 256) {
            const cplx<R> x = q[e];
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                if (t0 + u >= ncoef) break;
                const cplx<R> o = opsT[(size_t)(t0 + u) * dd + e];
                pr[u] += x.x * o.x - x.y * o.y;
                pi[u] += x.x * o.y + x.y * o.x;
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
            for (int u = 0; u < KB; ++u) { pr[u] += __shfl_xor(pr[u], off, 64); pi[u] += __shfl_xor(pi[u], off, 64); }
        if (lane == 0)
#pragma unroll
            for (int u = 0; u < KB; ++u) { sr[u][wave] = pr[u]; si[u][wave] = pi[u]; }
        __syncthreads();
        if (threadIdx.x < KB && t0 + (int)threadIdx.x < ncoef) {
            const int u = threadIdx.x;
            const R f = scale ? scale[0].x : R(1);
            dg[(size_t)(t0 + u) * dg_stride + j0 + j] =
                mk<R>(f * ((sr[u][0] + sr[u][1]) + (sr[u][2] + sr[u][3])), f * ((si[u][0] + si[u][1]) + (si[u][2] + si[u][3])));
        }
        __syncthreads();
    }
}

template <class R>
hipError_t launch_grad_trace(const cplx<R>* Q, int n, int d, const cplx<R>* opsT, int ncoef, cplx<R>* dg, int dg_stride,
                             int j0, hipStream_t st, const cplx<R>* scale) {
    if (n <= 0 || d <= 0 || ncoef <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((grad_trace_kernel<R>), dim3(n), dim3(256), 0, st, Q, d, opsT, ncoef, dg, dg_stride, j0, scale);
    return hipGetLastError();
}

// Magnus chain rule (control_expansion.cu:27-99 differentiated): with G_t(j) = dg[t][j] the
// derivative of g with respect to coefficient t of slice j, g is holomorphic in the
// coefficients, so dg/dc_kp = sum_(t,j) G_t(j) dcoef_t(j)/dc_kp, and dF/dc_kp = 2 Re(conj(g) dg/dc_kp) / d^2.
// Sample p of control k enters slice j = p/2 (p = 2j, even), j = (p-1)/2 (odd) and j = p/2 - 1
// (p = 2j + 2, even):
//   Simpson mean t = k:        1/6, 4/6, 1/6;
//   gamma_k, t = A + k:        -i tau/12 at p = 2j, +i tau/12 at p = 2j + 2;
//   delta_ab, t = 2A + pair:   i tau/12 (c_a[2j] c_b[2j+2] - c_a[2j+2] c_b[2j]) differentiated:
//                              for k = a: +c_b[2j+2] at p = 2j, -c_b[2j] at p = 2j + 2;
//                              for k = b: -c_a[2j+2] at p = 2j, +c_a[2j] at p = 2j + 2 (times i tau/12).
template <class R>
__global__ void grad_finish_magnus_kernel(const cplx<R>* __restrict__ dg, const cplx<R>* __restrict__ gptr,
                                          Controls<R> q, int N, int d, R* __restrict__ grad, R* __restrict__ fid) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    const cplx<R> g = gptr[0];
    const R inv = R(1) / ((R)d * (R)d);
    if (idx == 0 && fid) fid[0] = (g.x * g.x + g.y * g.y) * inv;
    if (idx >= (long)q.amps_in * q.pts) return;
    const int k = (int)(idx / q.pts), p = (int)(idx % q.pts), A = q.amps_set;
    const R f = q.tau / R(12);
    // accumulate dg/dc_kp; (re, im) of the factor multiplying G
    cplx<R> acc = mk<R>(0, 0);
    auto add = [&](int t, int j, cplx<R> w) {
        if (j < 0 || j >= N) return;
        acc = cadd(acc, cmul(w, dg[(size_t)t * N + j]));
    };
    auto smp = [&](int b, int i) -> R { return (b < q.amps_in && i >= 0 && i < q.pts) ? q.c[(size_t)b * q.pts + i].x : R(0); };
    if (p & 1) {
        add(k, (p - 1) / 2, mk<R>(R(4) / R(6), 0));
    } else {
        const int j0 = p / 2, j1 = p / 2 - 1;   // p = 2 j0, p = 2 j1 + 2
        add(k, j0, mk<R>(R(1) / R(6), 0));
        add(k, j1, mk<R>(R(1) / R(6), 0));
        add(A + k, j0, mk<R>(0, -f));
        add(A + k, j1, mk<R>(0, f));
        int t = 2 * A;
        for (int a = 0; a < A; ++a)
            for (int b = a + 1; b < A; ++b, ++t) {
                if (k == a) {
                    add(t, j0, mk<R>(0, f * smp(b, 2 * j0 + 2)));
                    add(t, j1, mk<R>(0, -f * smp(b, 2 * j1)));
                } else if (k == b) {
                    add(t, j0, mk<R>(0, -f * smp(a, 2 * j0 + 2)));
                    add(t, j1, mk<R>(0, f * smp(a, 2 * j1)));
                }
            }
    }
    grad[idx] = R(2) * (g.x * acc.x + g.y * acc.y) * inv;   // 2 Re(conj(g) acc) / d^2
}

template <class R>
hipError_t launch_grad_finish_magnus(const cplx<R>* dg, const cplx<R>* g, const Controls<R>& q, int N, int d, R* grad,
                                     R* fid, hipStream_t st) {
    const long total = (long)q.amps_in * q.pts;
    const unsigned grid = (unsigned)(total > 0 ? (total + 255) / 256 : 1);
    hipLaunchKernelGGL((grad_finish_magnus_kernel<R>), dim3(grid), dim3(256), 0, st, dg, g, q, N, d, grad, fid);
    return hipGetLastError();
}

bool grad_supported(int d) { return d >= 1 && d <= 32; }

template <class R>
hipError_t launch_up_block(const cplx<R>* in, int n_in, int d, int logb, const TreeStore<R>& st, hipStream_t s) {
    if (d <= 0 || d > 32 || n_in <= 0 || logb < 1 || logb > 3) return hipErrorInvalidValue;
    const int grid = (n_in + (1 << logb) - 1) >> logb;
#define PA_UP_BLOCK(DD)                                                                                        \
    if (logb == 1) hipLaunchKernelGGL((up_block_kernel<R, DD, 1>), dim3(grid), dim3(64 * Shape<DD>::G), 0, s, in, n_in, d, st); \
    else if (logb == 2) hipLaunchKernelGGL((up_block_kernel<R, DD, 2>), dim3(grid), dim3(64 * Shape<DD>::G), 0, s, in, n_in, d, st); \
    else hipLaunchKernelGGL((up_block_kernel<R, DD, 3>), dim3(grid), dim3(64 * Shape<DD>::G), 0, s, in, n_in, d, st);
    if (d <= 16) { PA_UP_BLOCK(16) } else { PA_UP_BLOCK(32) }
#undef PA_UP_BLOCK
    return hipGetLastError();
}

template <class R>
hipError_t launch_down_level(const cplx<R>* V, int n_child, int d, const cplx<R>* Wp, cplx<R>* Wc, hipStream_t st) {
    if (d <= 0 || d > 32 || n_child <= 0) return hipErrorInvalidValue;
    if (d <= 16) hipLaunchKernelGGL((down_level_kernel<R, 16>), dim3(n_child), dim3(64 * Shape<16>::G), 0, st, V, n_child, d, Wp, Wc);
    else hipLaunchKernelGGL((down_level_kernel<R, 32>), dim3(n_child), dim3(64 * Shape<32>::G), 0, st, V, n_child, d, Wp, Wc);
    return hipGetLastError();
}

// Level top - 1 (two children) of the down-sweep together with g (down_root_kernel).
template <class R>
hipError_t launch_down_root(const cplx<R>* V, const cplx<R>* t, const cplx<R>* u, int d, cplx<R>* Wc, cplx<R>* g,
                            hipStream_t st) {
    if (d <= 0 || d > 32) return hipErrorInvalidValue;
    if (d <= 16) hipLaunchKernelGGL((down_root_kernel<R, 16>), dim3(3), dim3(64 * Shape<16>::G), 0, st, V, t, u, d, Wc, g);
    else hipLaunchKernelGGL((down_root_kernel<R, 32>), dim3(3), dim3(64 * Shape<32>::G), 0, st, V, t, u, d, Wc, g);
    return hipGetLastError();
}

template <class R>
hipError_t launch_grad_slices(const ChebArgs<R>& a, const cplx<R>* E, const cplx<R>* W1, cplx<R>* dg, int dg_stride,
                              hipStream_t st, GradDirect<R> gd) {
    if (a.d <= 0 || a.d > 32 || a.N <= 0 || (a.j0 & 1) || a.M < 3 || a.M >= kMaxCheb || a.ncoef > kGradCoeffs)
        return hipErrorInvalidValue;
    // the direct gradient needs one sample per slice and the traces of plain controls
    if (gd.grad && (a.ctl.mode != QUAD_NONE || a.ctl.magnus || !gd.g || gd.amps > a.ncoef)) return hipErrorInvalidValue;
    const dim3 grid(a.N);
    if (a.d <= 16) {
        if (a.ps) hipLaunchKernelGGL((grad_slice_kernel<R, 16, true>), grid, dim3(64 * Shape<16>::G), 0, st, a, E, W1, dg, dg_stride, gd);
        else hipLaunchKernelGGL((grad_slice_kernel<R, 16, false>), grid, dim3(64 * Shape<16>::G), 0, st, a, E, W1, dg, dg_stride, gd);
    } else {
        if (a.ps) hipLaunchKernelGGL((grad_slice_kernel<R, 32, true>), grid, dim3(64 * Shape<32>::G), 0, st, a, E, W1, dg, dg_stride, gd);
        else hipLaunchKernelGGL((grad_slice_kernel<R, 32, false>), grid, dim3(64 * Shape<32>::G), 0, st, a, E, W1, dg, dg_stride, gd);
    }
    return hipGetLastError();
}

template <class R>
__global__ __launch_bounds__(256) void transpose_kernel(const cplx<R>* __restrict__ in, int count, int d,
                                                        cplx<R>* __restrict__ out) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x, dd = (long)d * d;
    if (e >= count * dd) return;
    const long t = e / dd, r = (e % dd) / d, c = e % d;
    out[e] = in[t * dd + c * d + r];
}

template <class R>
hipError_t launch_transpose(const cplx<R>* in, int count, int d, cplx<R>* out, hipStream_t st) {
    const long total = (long)count * d * d;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL((transpose_kernel<R>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, in, count, d, out);
    return hipGetLastError();
}

template <class R>
hipError_t launch_grad_root(const cplx<R>* t, const cplx<R>* u, int d, cplx<R>* wroot, cplx<R>* g, hipStream_t st,
                            const cplx<R>* wscale) {
    const int nb = (d * d + 255) / 256;
    hipLaunchKernelGGL((grad_root_kernel<R>), dim3(nb + 1), dim3(256), 0, st, t, u, d, nb, wroot, g, wscale);
    return hipGetLastError();
}

template <class R>
hipError_t launch_target_scale(const cplx<R>* t, int d, cplx<R>* out, hipStream_t st) {
    if (d <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((target_scale_kernel<R>), dim3(1), dim3(256), 0, st, t, d, out);
    return hipGetLastError();
}

template <class R>
hipError_t launch_grad_finish(const cplx<R>* dg, const cplx<R>* g, int N, int pts, int amps, int mode, int d, R* grad,
                              R* fid, hipStream_t st) {
    const long total = (long)amps * pts;
    const unsigned grid = (unsigned)(total > 0 ? (total + 255) / 256 : 1);
    hipLaunchKernelGGL((grad_finish_kernel<R>), dim3(grid), dim3(256), 0, st, dg, g, N, pts, amps, mode, d, grad, fid);
    return hipGetLastError();
}

#define PA_GRAD_INSTANTIATE(R)                                                                                       \
    template hipError_t launch_up_block<R>(const cplx<R>*, int, int, int, const TreeStore<R>&, hipStream_t);         \
    template hipError_t launch_down_level<R>(const cplx<R>*, int, int, const cplx<R>*, cplx<R>*, hipStream_t);      \
    template hipError_t launch_down_root<R>(const cplx<R>*, const cplx<R>*, const cplx<R>*, int, cplx<R>*, cplx<R>*, \
                                            hipStream_t);                                                            \
    template hipError_t launch_grad_slices<R>(const ChebArgs<R>&, const cplx<R>*, const cplx<R>*, cplx<R>*, int,     \
                                              hipStream_t, GradDirect<R>);                                           \
    template hipError_t launch_grad_root<R>(const cplx<R>*, const cplx<R>*, int, cplx<R>*, cplx<R>*, hipStream_t,   \
                                            const cplx<R>*);                                                         \
    template hipError_t launch_target_scale<R>(const cplx<R>*, int, cplx<R>*, hipStream_t);                          \
    template hipError_t launch_transpose<R>(const cplx<R>*, int, int, cplx<R>*, hipStream_t);                        \
    template hipError_t launch_grad_trace<R>(const cplx<R>*, int, int, const cplx<R>*, int, cplx<R>*, int, int,     \
                                             hipStream_t, const cplx<R>*);                                           \
    template hipError_t launch_grad_finish<R>(const cplx<R>*, const cplx<R>*, int, int, int, int, int, R*, R*,       \
                                              hipStream_t);                                                          \
    template hipError_t launch_grad_finish_magnus<R>(const cplx<R>*, const cplx<R>*, const Controls<R>&, int, int,  \
                                                     R*, R*, hipStream_t);
PA_GRAD_INSTANTIATE(float)
PA_GRAD_INSTANTIATE(double)
#undef PA_GRAD_INSTANTIATE

}  // namespace pa
