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
// gives every control's gradient through the traces Tr(Q_j O_t).  The prefix products P_j and
// the costates L_j = U_t^H S_j come from a log-depth scan: an up-sweep keeps every level of
// the pairwise product tree, a down-sweep pushes P and L from the root (P = I, L = U_t^H) to
// the leaves: for a parent with children c0 (earlier) and c1 (later),
//     P(c0) = P, P(c1) = V(c0) P,   L(c1) = L, L(c0) = L V(c1).
// Then W_j = P_j L_j.  The quadrature's chain rule maps dF/dc'_tj to the raw samples.
#include "kernels.hpp"

namespace pa {

namespace {

template <int D> struct GS {
    static constexpr int G = Shape<D>::G, T = Shape<D>::T;
    static constexpr int SA = D + Pad<double>::A, SB = D + Pad<double>::B;
};

// Split of the derivative kernel: it holds four state tiles (b, b', db, db'), so every dim
// uses quadrants -- D = 32: four waves of 16x16; D = 16: one wave.
template <int D> struct GradShape {
    static constexpr int G = (D == 16) ? 1 : 4;
    static constexpr int T = (D == 16) ? 1 : D / 32;
    static __device__ __forceinline__ int row0(int w) { return D == 16 ? 0 : (D / 2) * (w >> 1); }
    static __device__ __forceinline__ int col0(int w) { return D == 16 ? 0 : (D / 2) * (w & 1); }
};

// Stage a d x d complex128 matrix (row-major, or conjugate-transposed) into LDS planes, zero-padded.
template <int D, int NT = 64 * GS<D>::G>
__device__ __forceinline__ void stage(const double2* __restrict__ src, int d, double* Pr, double* Pi, int S,
                                      bool conj_t = false) {
    for (int e = threadIdx.x; e < D * D; e += NT) {
        const int i = e / D, k = e % D;
        double xr = 0.0, xi = 0.0;
        if (i < d && k < d) {
            if (conj_t) { const double2 v = src[k * d + i]; xr = v.x; xi = -v.y; }
            else { const double2 v = src[i * d + k]; xr = v.x; xi = v.y; }
        }
        Pr[i * S + k] = xr; Pi[i * S + k] = xi;
    }
}

template <int D>
__device__ __forceinline__ void tile_store(const CTile<double, GS<D>::T, GS<D>::T>& w, int d, double2* dst, double diag = 0.0) {
    constexpr int T = GS<D>::T;
    const int wave = threadIdx.x >> 6;
    const int row0 = Shape<D>::row0(wave), col0 = Shape<D>::col0(wave);
#pragma unroll
    for (int m = 0; m < T; ++m)
#pragma unroll
        for (int n = 0; n < T; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = acc_row<double>(row0, m, v), col = acc_col(col0, n);
                if (row < d && col < d)
                    dst[row * d + col] = mk<double>(w.re[m][n][v] + (row == col ? diag : 0.0), w.im[m][n][v]);
            }
}

template <int D>
__device__ __forceinline__ void copy_mat(const double2* __restrict__ src, double2* __restrict__ dst, int d) {
    for (int e = threadIdx.x; e < d * d; e += 64 * GS<D>::G) dst[e] = src[e];
}

}  // namespace

// E_j -> U_j = I + E_j in place, for all N leaves.
__global__ void add_identity_kernel(double2* m, int d, long total) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= total) return;
    const int e = (int)(idx % ((long)d * d));
    if (e / d == e % d) m[idx].x += 1.0;
}

// Up-sweep level: out[i] = in[2i+1] in[2i], or a copy of in[2i] for an odd last node.
template <int D>
__global__ __launch_bounds__(64 * Shape<D>::G) void pair_level_kernel(const double2* __restrict__ in, int n_in, int d,
                                                                         double2* __restrict__ out) {
    constexpr int T = GS<D>::T, SA = GS<D>::SA, SB = GS<D>::SB;
    __shared__ double Ar[D * SA], Ai[D * SA], Br[D * SB], Bi[D * SB];
    const int i = blockIdx.x, dd = d * d, wave = threadIdx.x >> 6;
    if (2 * i + 1 >= n_in) { copy_mat<D>(in + (size_t)(2 * i) * dd, out + (size_t)i * dd, d); return; }
    stage<D>(in + (size_t)(2 * i + 1) * dd, d, Ar, Ai, SA);
    stage<D>(in + (size_t)(2 * i) * dd, d, Br, Bi, SB);
    __syncthreads();
    CTile<double, T, T> w;
    wave_cgemm<double, T, T, D, SA, SB>(Ar, Ai, Br, Bi, Shape<D>::row0(wave), Shape<D>::col0(wave), w);
    tile_store<D>(w, d, out + (size_t)i * dd);
}

// Down-sweep level: parent i (values Pp, Lp) -> children 2i, 2i+1 of the level below (values Vc).
template <int D>
__global__ __launch_bounds__(64 * Shape<D>::G) void down_level_kernel(const double2* __restrict__ Vc, int n_child, int d,
                                                                         const double2* __restrict__ Pp,
                                                                         const double2* __restrict__ Lp,
                                                                         double2* __restrict__ Pc, double2* __restrict__ Lc) {
    constexpr int T = GS<D>::T, SA = GS<D>::SA, SB = GS<D>::SB;
    __shared__ double Ar[D * SA], Ai[D * SA], Br[D * SB], Bi[D * SB];
    const int i = blockIdx.x, dd = d * d, wave = threadIdx.x >> 6;
    const int c0 = 2 * i, c1 = 2 * i + 1;
    const double2* P = Pp + (size_t)i * dd;
    const double2* L = Lp + (size_t)i * dd;
    copy_mat<D>(P, Pc + (size_t)c0 * dd, d);
    if (c1 >= n_child) { copy_mat<D>(L, Lc + (size_t)c0 * dd, d); return; }
    copy_mat<D>(L, Lc + (size_t)c1 * dd, d);
    CTile<double, T, T> w;
    const int row0 = Shape<D>::row0(wave), col0 = Shape<D>::col0(wave);
    // P(c1) = V(c0) P
    stage<D>(Vc + (size_t)c0 * dd, d, Ar, Ai, SA);
    stage<D>(P, d, Br, Bi, SB);
    __syncthreads();
    wave_cgemm<double, T, T, D, SA, SB>(Ar, Ai, Br, Bi, row0, col0, w);
    tile_store<D>(w, d, Pc + (size_t)c1 * dd);
    __syncthreads();
    // L(c0) = L V(c1)
    stage<D>(L, d, Ar, Ai, SA);
    stage<D>(Vc + (size_t)c1 * dd, d, Br, Bi, SB);
    __syncthreads();
    wave_cgemm<double, T, T, D, SA, SB>(Ar, Ai, Br, Bi, row0, col0, w);
    tile_store<D>(w, d, Lc + (size_t)c0 * dd);
}

// L at the root: U_t^H.
__global__ void conj_transpose_kernel(const double2* __restrict__ t, int d, double2* __restrict__ out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= d * d) return;
    const int i = e / d, k = e % d;
    const double2 v = t[k * d + i];
    out[e] = mk<double>(v.x, -v.y);
}

// Per slice: X_j, W_j = P_j L_j, then Q_j = Dp(X_j)[W_j] and dg[t][j] = Tr(Q_j O_t).
template <int D>
__global__ __launch_bounds__(64 * GradShape<D>::G) void grad_slice_kernel(ChebArgs<double> a, const double2* __restrict__ Pl,
                                                                         const double2* __restrict__ Ll,
                                                                         double2* __restrict__ dg) {
    constexpr int G = GradShape<D>::G, T = GradShape<D>::T, SA = GS<D>::SA, SB = GS<D>::SB;
    __shared__ double Xr[D * SA], Xi[D * SA], Wr[D * SA], Wi[D * SA], Br[D * SB], Bi[D * SB];
    __shared__ double2 coef[kMaxCoeffs];
    __shared__ double2 red[64][G];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int row0 = GradShape<D>::row0(wave), col0 = GradShape<D>::col0(wave);
    const int d = a.d, dd = d * d, M = a.M;
    const double s = a.s;
    typedef CTile<double, T, T> Tile;

    auto to_lds = [&](const Tile& t, double* Pr, double* Pi, int S) {
#pragma unroll
        for (int m = 0; m < T; ++m)
#pragma unroll
            for (int n = 0; n < T; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = acc_row<double>(row0, m, v), col = acc_col(col0, n);
                    Pr[row * S + col] = t.re[m][n][v];
                    Pi[row * S + col] = t.im[m][n][v];
                }
    };

    for (int j = blockIdx.x; j < a.N; j += gridDim.x) {
        for (int t = tid; t < a.ncoef; t += 64 * G) coef[t] = slice_coefficient(a.ctl, t, j);
        __syncthreads();
        for (int e = tid; e < D * D; e += 64 * G) {
            const int i = e / D, k = e % D;
            double2 x = mk<double>(0, 0);
            if (i < d && k < d) {
                x = a.ops[i * d + k];
                for (int t = 0; t < a.ncoef; ++t) x = cadd(x, cmul(coef[t], a.ops[(size_t)(t + 1) * dd + i * d + k]));
            }
            Xr[i * SA + k] = x.x; Xi[i * SA + k] = x.y;
        }
        // W_j = P_j L_j
        stage<D, 64 * G>(Pl + (size_t)j * dd, d, Wr, Wi, SA);
        stage<D, 64 * G>(Ll + (size_t)j * dd, d, Br, Bi, SB);
        __syncthreads();
        {
            Tile acc;
            wave_cgemm<double, T, T, D, SA, SB>(Wr, Wi, Br, Bi, row0, col0, acc);
            __syncthreads();
            to_lds(acc, Wr, Wi, SA);
        }

        // b_M = a_M I, db_M = 0; b_{M-1} = s a_M X + a_{M-1} I, db_{M-1} = s a_M W
        Tile b1, b2, db1, db2;
        {
            const double2 saM = mk<double>(s * a.are[M], s * a.aim[M]);
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = acc_row<double>(row0, m, v), col = acc_col(col0, n);
                        const bool dg_ = row == col;
                        b2.re[m][n][v] = dg_ ? a.are[M] : 0.0; b2.im[m][n][v] = dg_ ? a.aim[M] : 0.0;
                        db2.re[m][n][v] = 0.0; db2.im[m][n][v] = 0.0;
                        const double2 xv = cmul(saM, mk<double>(Xr[row * SA + col], Xi[row * SA + col]));
                        b1.re[m][n][v] = xv.x + (dg_ ? a.are[M - 1] : 0.0);
                        b1.im[m][n][v] = xv.y + (dg_ ? a.aim[M - 1] : 0.0);
                        const double2 wv = cmul(saM, mk<double>(Wr[row * SA + col], Wi[row * SA + col]));
                        db1.re[m][n][v] = wv.x; db1.im[m][n][v] = wv.y;
                    }
        }
        __syncthreads();
        to_lds(b1, Br, Bi, SB);
        __syncthreads();
        Tile acc2;
        // One step k: bc = b_{k+1} (staged in B), bp = b_{k+2} -> b_k; dbc = db_{k+1}, dbp = db_{k+2} -> db_k.
        auto step = [&](Tile& bc, Tile& bp, Tile& dbc, Tile& dbp, int k) {
            const double beta = (k == 0) ? -2.0 : -1.0;
            if (k > 0) {   // b_k = s X b_{k+1} + beta b_{k+2} + a_k I (not needed at k = 0)
                wave_cgemm<double, T, T, D, SA, SB>(Xr, Xi, Br, Bi, row0, col0, acc2);
#pragma unroll
                for (int m = 0; m < T; ++m)
#pragma unroll
                    for (int n = 0; n < T; ++n)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const bool dg_ = acc_row<double>(row0, m, v) == acc_col(col0, n);
                            bp.re[m][n][v] = s * acc2.re[m][n][v] + beta * bp.re[m][n][v] + (dg_ ? a.are[k] : 0.0);
                            bp.im[m][n][v] = s * acc2.im[m][n][v] + beta * bp.im[m][n][v] + (dg_ ? a.aim[k] : 0.0);
                        }
            }
            // db_k = s W b_{k+1} + s X db_{k+1} + beta db_{k+2}
            wave_cgemm<double, T, T, D, SA, SB>(Wr, Wi, Br, Bi, row0, col0, acc2);
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        dbp.re[m][n][v] = s * acc2.re[m][n][v] + beta * dbp.re[m][n][v];
                        dbp.im[m][n][v] = s * acc2.im[m][n][v] + beta * dbp.im[m][n][v];
                    }
            __syncthreads();
            to_lds(dbc, Br, Bi, SB);
            __syncthreads();
            wave_cgemm<double, T, T, D, SA, SB>(Xr, Xi, Br, Bi, row0, col0, acc2);
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        dbp.re[m][n][v] += s * acc2.re[m][n][v];
                        dbp.im[m][n][v] += s * acc2.im[m][n][v];
                    }
            if (k > 0) {
                __syncthreads();
                to_lds(bp, Br, Bi, SB);
                __syncthreads();
            }
        };
        // M is odd, so the M - 1 steps k = M-2 .. 0 come in pairs and the roles alternate.
        for (int k = M - 2; k >= 1; k -= 2) {
            step(b1, b2, db1, db2, k);       // b2 <- b_k, db2 <- db_k
            step(b2, b1, db2, db1, k - 1);   // b1 <- b_{k-1}, db1 <- db_{k-1}
        }
        // db1 = Q_j; dg[t][j] = sum_{r,c} Q[r][c] O_t[c][r]
        for (int t = 0; t < a.ncoef; ++t) {
            const double2* O = a.ops + (size_t)(t + 1) * dd;
            double pr = 0.0, pi = 0.0;
#pragma unroll
            for (int m = 0; m < T; ++m)
#pragma unroll
                for (int n = 0; n < T; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = acc_row<double>(row0, m, v), col = acc_col(col0, n);
                        if (row < d && col < d) {
                            const double2 o = O[col * d + row];
                            pr += db1.re[m][n][v] * o.x - db1.im[m][n][v] * o.y;
                            pi += db1.re[m][n][v] * o.y + db1.im[m][n][v] * o.x;
                        }
                    }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) { pr += __shfl_xor(pr, off, 64); pi += __shfl_xor(pi, off, 64); }
            if (G == 1) {
                if (lane == 0) dg[(size_t)t * a.N + j] = mk<double>(pr, pi);
            } else {
                if (lane == 0) red[t & 63][wave] = mk<double>(pr, pi);
                __syncthreads();
                if (tid == 0) {
                    double2 sum = red[t & 63][0];
                    for (int w = 1; w < G; ++w) sum = cadd(sum, red[t & 63][w]);
                    dg[(size_t)t * a.N + j] = sum;
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
}

// g = Tr(U_t^H U) -> out[0]
__global__ void target_trace_kernel(const double2* __restrict__ t, const double2* __restrict__ u, int d,
                                    double2* __restrict__ out) {
    __shared__ double sr[256], si[256];
    double pr = 0.0, pi = 0.0;
    for (int e = threadIdx.x; e < d * d; e += 256) {
        const double2 a = t[e], b = u[e];
        pr += a.x * b.x + a.y * b.y;   // conj(a) b
        pi += a.x * b.y - a.y * b.x;
    }
    sr[threadIdx.x] = pr; si[threadIdx.x] = pi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sr[threadIdx.x] += sr[threadIdx.x + o]; si[threadIdx.x] += si[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = mk<double>(sr[0], si[0]);
}

// grad[k * pts + p] = sum_j w_jp 2 Re(conj(g) dg[k][j]) / d^2 (quadrature chain rule), F = |g|^2 / d^2.
__global__ void grad_finish_kernel(const double2* __restrict__ dg, const double2* __restrict__ gptr, int N, int pts,
                                   int amps, int mode, int d, double* __restrict__ grad, double* __restrict__ fid) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    const double2 g = gptr[0];
    const double inv = 1.0 / ((double)d * d);
    if (idx == 0 && fid) fid[0] = (g.x * g.x + g.y * g.y) * inv;
    if (idx >= (long)amps * pts) return;
    const int k = (int)(idx / pts), p = (int)(idx % pts);
    auto term = [&](int j, double w) -> double {
        if (j < 0 || j >= N) return 0.0;
        const double2 v = dg[(size_t)k * N + j];
        return w * 2.0 * (g.x * v.x + g.y * v.y) * inv;   // 2 Re(conj(g) v) / d^2
    };
    double r = 0.0;
    if (mode == QUAD_NONE) r = term(p, 1.0);
    else if (mode == QUAD_MIDPOINT) r = term(p - 1, 0.5) + term(p, 0.5);
    else {
        if (p & 1) r = term((p - 1) / 2, 4.0 / 6.0);
        else r = term(p / 2, 1.0 / 6.0) + term(p / 2 - 1, 1.0 / 6.0);
    }
    grad[idx] = r;
}

bool grad_supported(int d) { return d >= 1 && d <= 32; }

hipError_t launch_add_identity(double2* m, int d, int n, hipStream_t st) {
    const long total = (long)n * d * d;
    hipLaunchKernelGGL(add_identity_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, m, d, total);
    return hipGetLastError();
}

hipError_t launch_pair_level(const double2* in, int n_in, int d, double2* out, hipStream_t st) {
    const int grid = (n_in + 1) / 2;
    if (d <= 16) hipLaunchKernelGGL((pair_level_kernel<16>), dim3(grid), dim3(64 * Shape<16>::G), 0, st, in, n_in, d, out);
    else if (d <= 32) hipLaunchKernelGGL((pair_level_kernel<32>), dim3(grid), dim3(64 * Shape<32>::G), 0, st, in, n_in, d, out);
    else hipLaunchKernelGGL((pair_level_kernel<64>), dim3(grid), dim3(64 * Shape<64>::G), 0, st, in, n_in, d, out);
    return hipGetLastError();
}

hipError_t launch_down_level(const double2* Vc, int n_child, int d, const double2* Pp, const double2* Lp, double2* Pc,
                             double2* Lc, hipStream_t st) {
    const int grid = (n_child + 1) / 2;
    if (d <= 16) hipLaunchKernelGGL((down_level_kernel<16>), dim3(grid), dim3(64 * Shape<16>::G), 0, st, Vc, n_child, d, Pp, Lp, Pc, Lc);
    else if (d <= 32) hipLaunchKernelGGL((down_level_kernel<32>), dim3(grid), dim3(64 * Shape<32>::G), 0, st, Vc, n_child, d, Pp, Lp, Pc, Lc);
    else hipLaunchKernelGGL((down_level_kernel<64>), dim3(grid), dim3(64 * Shape<64>::G), 0, st, Vc, n_child, d, Pp, Lp, Pc, Lc);
    return hipGetLastError();
}

hipError_t launch_conj_transpose(const double2* t, int d, double2* out, hipStream_t st) {
    hipLaunchKernelGGL(conj_transpose_kernel, dim3((d * d + 255) / 256), dim3(256), 0, st, t, d, out);
    return hipGetLastError();
}

hipError_t launch_grad_slices(const ChebArgs<double>& a, const double2* P, const double2* L, double2* dg, hipStream_t st) {
    if (a.d <= 0 || a.d > 32 || a.N <= 0 || a.M < 3 || a.M >= kMaxCheb || a.ncoef > 64) return hipErrorInvalidValue;
    if (a.d <= 16) hipLaunchKernelGGL((grad_slice_kernel<16>), dim3(a.N), dim3(64 * GradShape<16>::G), 0, st, a, P, L, dg);
    else hipLaunchKernelGGL((grad_slice_kernel<32>), dim3(a.N), dim3(64 * GradShape<32>::G), 0, st, a, P, L, dg);
    return hipGetLastError();
}

hipError_t launch_target_trace(const double2* t, const double2* u, int d, double2* out, hipStream_t st) {
    hipLaunchKernelGGL(target_trace_kernel, dim3(1), dim3(256), 0, st, t, u, d, out);
    return hipGetLastError();
}

hipError_t launch_grad_finish(const double2* dg, const double2* g, int N, int pts, int amps, int mode, int d, double* grad,
                              double* fid, hipStream_t st) {
    const long total = (long)amps * pts;
    hipLaunchKernelGGL(grad_finish_kernel, dim3((unsigned)((total + 255) / 256 > 0 ? (total + 255) / 256 : 1)), dim3(256),
                       0, st, dg, g, N, pts, amps, mode, d, grad, fid);
    return hipGetLastError();
}

}  // namespace pa
