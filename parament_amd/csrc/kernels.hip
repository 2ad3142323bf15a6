// Equiprop device kernels for gfx950 (MI355X): fused Chebyshev slice kernel,
// single-launch pairwise product tree, and the generic step-wise kernels.
#include "kernels.hpp"
#include "fused_device.hpp"

namespace pa {

// ---------------------------------------------------------------------------------------
// Fused slice kernel: for each slice j, assemble X_j = H0 + sum_t c'_tj O_t in LDS, run the
// Clenshaw recurrence of parament.cpp:569-652 with all states on chip, and write the slice
// propagator as its deviation from identity, E_j = U_j - I (the last step adds
// (a_0 - 1) I instead of a_0 I), which keeps the rounding of the near-identity slices
// relative to |E_j| instead of |U_j|.
//   b_M = a_M I, b_{M-1} = s a_M X + a_{M-1} I      (elementwise: b_M is diagonal)
//   b_k = s X b_{k+1} - b_{k+2} + a_k I, k = M-2..1  (one complex GEMM each)
//   U   = s X b_1 - 2 b_2 + a_0 I
// X stays in LDS (A operand), b_{k+1} is staged in LDS as the B operand, and b_{k+1},
// b_{k+2} are kept in registers in the accumulator layout for the epilogue.
template <class R, int D>
__global__ __launch_bounds__(64 * Shape<D>::G) void cheb_fused_kernel(ChebArgs<R> a) {
    constexpr int SA = D + Pad<R>::A, SB = D + Pad<R>::B;
    __shared__ R Xr[D * SA], Xi[D * SA], Br[D * SB], Bi[D * SB];
    __shared__ cplx<R> coef[kMaxCoeffs];
    for (int j = blockIdx.x; j < a.N; j += gridDim.x)
        cheb_slice<R, D, Shape<D>>(a.ops, a.ncoef, a.d, a.ctl, a.j0 + j, a.M, a.s, a.are, a.aim, Xr, Xi, Br, Bi, coef,
                                   a.out + (size_t)j * a.d * a.d);
}

// The fused slice kernel in the power form (power_slice_tile): for callers that set a.ps with
// the power coefficients pre/pim (pre[0] = m_0 - 1), when tau Hnorm <= 1.
template <class R, int D>
__global__ __launch_bounds__(64 * Shape<D>::G) void power_fused_kernel(ChebArgs<R> a) {
    constexpr int SA = D + Pad<R>::A, SB = D + Pad<R>::B;
    __shared__ R Xr[D * SA], Xi[D * SA], Br[D * SB], Bi[D * SB];
    __shared__ cplx<R> coef[kMaxCoeffs];
    typedef Shape<D> SH;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int row0 = SH::row0(wave), col0 = SH::col0(wave);
    for (int jl = blockIdx.x; jl < a.N; jl += gridDim.x) {
        for (int t = tid; t < a.ncoef; t += 64 * SH::G) coef[t] = slice_coefficient(a.ctl, t, a.j0 + jl);
        __syncthreads();
        const CTile<R, SH::T, SH::T> e =
            power_slice_tile<R, D, SH>(a.ops, a.ncoef, a.d, a.M, a.ctl.tau, a.pre, a.pim, Xr, Xi, Br, Bi, coef);
        cplx<R>* out = a.out + (size_t)jl * a.d * a.d;
#pragma unroll
        for (int m = 0; m < SH::T; ++m)
#pragma unroll
            for (int n = 0; n < SH::T; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = acc_row<R>(row0, m, v), col = acc_col(col0, n);
                    if (row < a.d && col < a.d) out[row * a.d + col] = mk<R>(e.re[m][n][v], e.im[m][n][v]);
                }
    }
}

template <class R>
hipError_t launch_cheb_fused(const ChebArgs<R>& a, hipStream_t st) {
    if (a.d <= 0 || a.d > 64 || a.N <= 0 || a.M < 3 || a.M >= kMaxCheb || a.ncoef > kMaxCoeffs)
        return hipErrorInvalidValue;
    const int grid = a.N;
    // dims <= 16: one wave per slice; a persistent grid of 32768 work-groups (each looping over
    // slices) measured ~10% faster than one work-group per slice at a million slices
    if (a.ps) {
        if (a.d <= 16) hipLaunchKernelGGL((power_fused_kernel<R, 16>), dim3(grid < 32768 ? grid : 32768), dim3(64 * Shape<16>::G), 0, st, a);
        else if (a.d <= 32) hipLaunchKernelGGL((power_fused_kernel<R, 32>), dim3(grid), dim3(64 * Shape<32>::G), 0, st, a);
        else if (a.d <= 48) hipLaunchKernelGGL((power_fused_kernel<R, 48>), dim3(grid), dim3(64 * Shape<48>::G), 0, st, a);
        else hipLaunchKernelGGL((power_fused_kernel<R, 64>), dim3(grid), dim3(64 * Shape<64>::G), 0, st, a);
        return hipGetLastError();
    }
    if (a.d <= 16) hipLaunchKernelGGL((cheb_fused_kernel<R, 16>), dim3(grid < 32768 ? grid : 32768), dim3(64 * Shape<16>::G), 0, st, a);
    else if (a.d <= 32) hipLaunchKernelGGL((cheb_fused_kernel<R, 32>), dim3(grid), dim3(64 * Shape<32>::G), 0, st, a);
    else if (a.d <= 48) hipLaunchKernelGGL((cheb_fused_kernel<R, 48>), dim3(grid), dim3(64 * Shape<48>::G), 0, st, a);
    else hipLaunchKernelGGL((cheb_fused_kernel<R, 64>), dim3(grid), dim3(64 * Shape<64>::G), 0, st, a);
    return hipGetLastError();
}
template hipError_t launch_cheb_fused<float>(const ChebArgs<float>&, hipStream_t);
template hipError_t launch_cheb_fused<double>(const ChebArgs<double>&, hipStream_t);

// ---------------------------------------------------------------------------------------
// Small dims (d <= 16): one wave per work-group on a persistent grid.  The operators are the
// same for every slice, so they stay in LDS for the whole launch; each lane t < ncoef loads the
// next slice's coefficient t while the current slice runs; and the F lowest tree levels are
// folded on chip in the reference's pairing (binary counter, later node on the left, on
// deviations from identity, a slice beyond N acting as the zero node of the odd carry).  Node w
// of level F covers slices [w 2^F, (w+1) 2^F) and goes to a.out as E = U - I.  For d <= 16 / NS
// a tile carries NS independent slices as diagonal blocks (cheb_slice_tile): group w folds
// nodes w, w + groups, .. side by side, so the 16 x 16 matrix-core products that pad a
// d <= 8 slice do the work of NS slices.
template <class R, int F, int NS>
__global__ __launch_bounds__(64) void small_fold_kernel(ChebArgs<R> a, int n_nodes) {
    constexpr int D = 16, NB = 1 << F, HB = D / NS;
    constexpr int SA = D + Pad<R>::A, SB = D + Pad<R>::B;
    typedef Shape<16> SH;
    typedef CTile<R, 1, 1> Tile;
    __shared__ R Xr[D * SA], Xi[D * SA], Br[D * SB], Bi[D * SB];
    __shared__ cplx<R> coef[64 * NS];
    extern __shared__ __align__(16) unsigned char small_ops_lds[];
    cplx<R>* ops = reinterpret_cast<cplx<R>*>(small_ops_lds);
    const int tid = threadIdx.x, d = a.d, dd = d * d;
    for (int e = tid; e < (1 + a.ncoef) * dd; e += 64) ops[e] = a.ops[e];
    // tile group w holds nodes w + q groups (q < NS) in its diagonal blocks
    const int groups = (n_nodes + NS - 1) / NS;

    // out = A + B + A B for the later node A and the earlier node B
    auto combine = [&](const Tile& A, const Tile& B, Tile& out) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = acc_row<R>(0, 0, v), col = acc_col(0, 0);
            Xr[row * SA + col] = A.re[0][0][v]; Xi[row * SA + col] = A.im[0][0][v];
            Br[row * SB + col] = B.re[0][0][v]; Bi[row * SB + col] = B.im[0][0][v];
        }
        __syncthreads();
        Tile acc;
        wave_cgemm<R, 1, 1, D, SA, SB>(Xr, Xi, Br, Bi, 0, 0, acc);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            out.re[0][0][v] = acc.re[0][0][v] + (A.re[0][0][v] + B.re[0][0][v]);
            out.im[0][0][v] = acc.im[0][0][v] + (A.im[0][0][v] + B.im[0][0][v]);
        }
        __syncthreads();
    };
    // slice p of block q of group w (may lie beyond N: then a zero node)
    auto slice_of = [&](int w, int q, int p) { return (w + q * groups) * NB + p; };
    // lane t: coefficient t of each block's slice about to run (zero beyond N)
    cplx<R> next[NS];
    auto prefetch = [&](int w, int p) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            next[q] = mk<R>(0, 0);
            const int j = (w < groups) ? slice_of(w, q, p) : a.N;
            if (tid < a.ncoef && j < a.N) next[q] = slice_coefficient(a.ctl, tid, a.j0 + j);
        }
    };
    prefetch(blockIdx.x, 0);
    for (int w = blockIdx.x; w < groups; w += gridDim.x) {
        Tile stk[F > 0 ? F : 1];
        Tile node;
        // steps below slice N of the block each accumulator element lies in
        int lim[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) lim[v] = a.N - (w + (acc_row<R>(0, 0, v) / HB) * groups) * NB;
        // one slice per tile: steps up to slice N - 1; several: every step of every group
        const int jcap = NS > 1 ? groups * NB : a.N;
#pragma unroll
        for (int p = 0; p < NB; ++p) {
            const int j = slice_of(w, 0, p);
            // one slice per tile: slices beyond N are skipped (uniform), the ragged block below;
            // several: every block runs all NB steps, slices beyond N as zero nodes, which the
            // binary counter combines exactly as the odd-carry rule
            if (j < jcap) {
                if (tid < a.ncoef)
#pragma unroll
                    for (int q = 0; q < NS; ++q) coef[64 * q + tid] = next[q];
                if (p + 1 < NB && (NS > 1 || j + 1 < a.N)) prefetch(w, p + 1);
                else prefetch(w + (int)gridDim.x, 0);
                __syncthreads();
                Tile e = cheb_slice_tile<R, D, SH, NS>(ops, a.ncoef, d, a.M, a.s, a.are, a.aim, Xr, Xi, Br, Bi, coef);
                if (NS > 1) {
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (p >= lim[v]) { e.re[0][0][v] = R(0); e.im[0][0][v] = R(0); }
                }
                int merged = 0;
#pragma unroll
                for (int l = 0; l < F; ++l) {
                    const int mask = (2 << l) - 1;
                    if ((p & mask) == mask) { Tile m; combine(e, stk[l], m); e = m; merged = l + 1; }
                }
                if (merged == F) {
                    node = e;
                } else {
#pragma unroll
                    for (int l = 0; l < F; ++l)
                        if (l == merged) stk[l] = e;
                }
            }
        }
        if (NS == 1) {
            const int live = min(a.N - w * NB, NB);
            if (live < NB) {   // ragged block: the pending nodes fold from the lowest level up
                bool have = false;
#pragma unroll
                for (int l = 0; l < F; ++l) {
                    if ((live >> l) & 1) {
                        if (!have) { node = stk[l]; have = true; }
                        else { Tile m; combine(node, stk[l], m); node = m; }
                    }
                }
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = acc_row<R>(0, 0, v), col = acc_col(0, 0);
            const int q = row / HB, i = row % HB, k = col % HB, nd = w + q * groups;
            if (col / HB == q && i < d && k < d && nd < n_nodes)
                a.out[(size_t)nd * dd + i * d + k] = mk<R>(node.re[0][0][v], node.im[0][0][v]);
        }
    }
}

int small_fold_levels(int N) {
    // fold while at least 2048 nodes remain (measured at dim 8: 4096 slices run fastest with one
    // folded level, 27 us per call against 28 with none and 31 with two)
    int f = 0;
    while (f < 5 && (N >> (f + 1)) >= 2048) ++f;
    return f;
}

bool small_fold_supported(int d, int ncoef, size_t elem) {
    return d >= 1 && d <= 16 && ncoef <= 64 && (size_t)(1 + ncoef) * d * d * elem <= 48 * 1024;
}

template <class R>
hipError_t launch_small_fold(const ChebArgs<R>& a, hipStream_t st) {
    if (!small_fold_supported(a.d, a.ncoef, sizeof(cplx<R>)) || a.N <= 0 || a.M < 3 || a.M >= kMaxCheb)
        return hipErrorInvalidValue;
    const int F = small_fold_levels(a.N);
    const int n_nodes = (a.N + (1 << F) - 1) >> F;
    // slices per 16 x 16 tile: 4 for d <= 4, 2 for d <= 8
    const int NS = a.d <= 4 ? 4 : (a.d <= 8 ? 2 : 1);
    const int groups = (n_nodes + NS - 1) / NS;
    // up to 16,384 one-wave work-groups (16 per SIMD): at a million dim-8 slices, 4,096 of them
    // (four groups each) left 2.2 waves resident per SIMD and took 0.78 ms against 0.69 ms
    const int grid = groups < 16384 ? groups : 16384;
    const size_t lds = (size_t)(1 + a.ncoef) * a.d * a.d * sizeof(cplx<R>);
#define PA_SMALL(FF)                                                                                              \
    if (NS == 4) hipLaunchKernelGGL((small_fold_kernel<R, FF, 4>), dim3(grid), dim3(64), lds, st, a, n_nodes);    \
    else if (NS == 2) hipLaunchKernelGGL((small_fold_kernel<R, FF, 2>), dim3(grid), dim3(64), lds, st, a, n_nodes); \
    else hipLaunchKernelGGL((small_fold_kernel<R, FF, 1>), dim3(grid), dim3(64), lds, st, a, n_nodes);
    switch (F) {
        case 5: PA_SMALL(5) break;
        case 4: PA_SMALL(4) break;
        case 3: PA_SMALL(3) break;
        case 2: PA_SMALL(2) break;
        case 1: PA_SMALL(1) break;
        default: PA_SMALL(0) break;
    }
#undef PA_SMALL
    return hipGetLastError();
}
template hipError_t launch_small_fold<float>(const ChebArgs<float>&, hipStream_t);
template hipError_t launch_small_fold<double>(const ChebArgs<double>&, hipStream_t);

// ---------------------------------------------------------------------------------------
// Pairwise product tree (parament.cpp:657-718 pairing, later slice on the left) on
// deviations from identity: (I + A)(I + B) = I + (A + B + A B), every node is stored as
// E = node - I and the root is written as I + E.
//
// One launch per 2^LOGB-fold reduction: work-group w reduces nodes [w B, w B + B) of the
// current level entirely on chip and writes node w of the level LOGB above.  Inside the
// block the pairing is the reference's (binary-counter order, completed nodes kept in
// registers).  A ragged last block is padded with E = 0 nodes: combining with a zero node
// returns the other node bit for bit (0 * X = 0, X + 0 + 0 = X), which reproduces the
// reference's odd-carry rule exactly.  Node (l, i) therefore covers slices
// [i 2^l, min((i+1) 2^l, N)) as in the reference.  No work-group depends on another
// inside a launch; kernel boundaries order the levels.
template <class R, int D, int LOGB, class SrcT>
__global__ __launch_bounds__(64 * Shape<D>::G) void tree_block_kernel(const SrcT* __restrict__ src, int n_in, int d,
                                                                         cplx<R>* __restrict__ dst, int final_level) {
    constexpr int SA = D + Pad<R>::A, SB = D + Pad<R>::B;
    __shared__ R Ar[D * SA], Ai[D * SA], Br[D * SB], Bi[D * SB];
    tree_block<R, D, LOGB, SrcT, Shape<D>>(src, n_in, d, dst + (size_t)blockIdx.x * d * d, final_level, blockIdx.x, Ar,
                                           Ai, Br, Bi);
}


// Top levels of the tree, where few products remain and a work-group per block of nodes would
// leave most CUs idle: one pairwise level per launch, each product spread over (D/16)^2
// work-groups of four waves.  Work-group (p, t) forms the 16 x 16 output tile t of node p of the
// next level, E = A + B + A B (A = node 2p + 1, the later, on the left; B = node 2p), its waves
// each taking a quarter of the inner dimension straight from memory and adding their partial
// tiles through LDS.  A last unpaired node is carried up unchanged (the odd-carry rule).
template <class R, int D, class SrcT>
__global__ __launch_bounds__(256) void tree_split_kernel(const SrcT* __restrict__ src, int n_in, int d,
                                                         cplx<R>* __restrict__ dst, int final_level) {
    constexpr int TT = D / 16, KW = D / 4;
    __shared__ cplx<R> part[4][256];
    const int prod = blockIdx.x / (TT * TT), t = blockIdx.x % (TT * TT);
    const int r0 = 16 * (t / TT), c0 = 16 * (t % TT);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r = lane & 15, q = lane >> 4;
    const int dd = d * d;
    const bool pair = 2L * prod + 1 < n_in;
    const SrcT* A = src + (size_t)(2L * prod + (pair ? 1 : 0)) * dd;
    const SrcT* B = src + (size_t)(2L * prod) * dd;
    typename Mma<R>::acc cr, ci;
#pragma unroll
    for (int v = 0; v < 4; ++v) { cr[v] = R(0); ci[v] = R(0); }
    if (pair) {
#pragma unroll
        for (int k0 = wave * KW; k0 < wave * KW + KW; k0 += 4) {
            const int k = k0 + q, ra = r0 + r, cb = c0 + r;
            R ar = R(0), ai = R(0), br = R(0), bi = R(0);
            if (ra < d && k < d) { ar = R(A[ra * d + k].x); ai = R(A[ra * d + k].y); }
            if (k < d && cb < d) { br = R(B[k * d + cb].x); bi = R(B[k * d + cb].y); }
            cr = Mma<R>::op(ar, br, cr);
            ci = Mma<R>::op(ar, bi, ci);
            cr = Mma<R>::op(-ai, bi, cr);
            ci = Mma<R>::op(ai, br, ci);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) part[wave][(acc_row<R>(0, 0, v)) * 16 + r] = mk<R>(cr[v], ci[v]);
    __syncthreads();
    const int i = r0 + (tid >> 4), j = c0 + (tid & 15);
    if (i < d && j < d) {
        cplx<R> e = cadd(cadd(part[0][tid], part[1][tid]), cadd(part[2][tid], part[3][tid]));
        if (pair) e = cadd(e, mk<R>(R(A[i * d + j].x), R(A[i * d + j].y)));
        e = cadd(e, mk<R>(R(B[i * d + j].x), R(B[i * d + j].y)));
        if (final_level && i == j) e.x += R(1);
        dst[(size_t)prod * dd + i * d + j] = e;
    }
}

// Dims <= 16: block reductions of up to 64 nodes per work-group.  The block's nodes sit in LDS and
// are reduced level by level, each product's d x d outputs spread over the threads (vector units:
// an 8 x 8 complex product is 512 multiply-adds, too small for a matrix-core tile), with the
// reference's pairing (later node on the left) and zero nodes past the end, which return their
// partner bit for bit (the odd-carry rule).  Every thread reads its operands before the barrier
// that precedes the in-place write of the level's results.  Two launches reduce up to 4096 nodes.
constexpr int kSmallTreeMaxPer = 8;   // outputs per thread and level (B * D * D <= 16 * 256)

// nodes per work-group for nodes padded to D x D
__host__ __device__ inline int small_tree_block(int D) {
    int b = 64;
    while (b > 1 && b * D * D > 2 * kSmallTreeMaxPer * 256) b >>= 1;
    return b;
}

// Nodes are held padded to D x D with zeros (D = 4, 8, 16): a zero row or column of E = U - I
// stays zero through A + B + A B, so the d x d block is what the unpadded products give.
template <class R, int D, class SrcT>
__global__ __launch_bounds__(256) void small_tree_kernel(const SrcT* __restrict__ src, int n_in, int d, int B,
                                                         cplx<R>* __restrict__ dst, int final_level) {
    extern __shared__ __attribute__((aligned(16))) unsigned char small_tree_lds[];
    cplx<R>* nd = reinterpret_cast<cplx<R>*>(small_tree_lds);
    constexpr int DD = D * D;
    const int dd = d * d;
    const int first = blockIdx.x * B;
    const int live = min(B, n_in - first);
    // all of the block's loads in flight before the first LDS store (B * D * D <= 16 * 256)
    SrcT ld[2 * kSmallTreeMaxPer];
#pragma unroll
    for (int t = 0; t < 2 * kSmallTreeMaxPer; ++t) {
        const int e = threadIdx.x + 256 * t;
        const int node = e / DD, ik = e % DD, i = ik / D, k = ik % D;
        ld[t].x = 0;
        ld[t].y = 0;
        if (e < B * DD && node < live && i < d && k < d) ld[t] = src[(size_t)(first + node) * dd + i * d + k];
    }
#pragma unroll
    for (int t = 0; t < 2 * kSmallTreeMaxPer; ++t) {
        const int e = threadIdx.x + 256 * t;
        if (e < B * DD) nd[e] = mk<R>(R(ld[t].x), R(ld[t].y));
    }
    __syncthreads();
    for (int cnt = B; cnt > 1; cnt >>= 1) {
        const int total = (cnt >> 1) * DD;
        cplx<R> res[kSmallTreeMaxPer];
        // with nodes padded to D x D, total is a multiple of 64 unless it is below 64, so the
        // guard never splits a wave that has work past the first
#pragma unroll
        for (int t = 0; t < kSmallTreeMaxPer; ++t) {
            const int e = threadIdx.x + 256 * t;
            if (e < total) {
                const int p = e / DD, ik = e % DD, i = ik / D, k = ik % D;
                const cplx<R>* A = nd + (size_t)(2 * p + 1) * DD;   // later node
                const cplx<R>* Bm = nd + (size_t)(2 * p) * DD;      // earlier node
                cplx<R> acc = cadd(A[ik], Bm[ik]);
#pragma unroll
                for (int m = 0; m < D; ++m) acc = cadd(acc, cmul(A[i * D + m], Bm[m * D + k]));
                res[t] = acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kSmallTreeMaxPer; ++t) {
            const int e = threadIdx.x + 256 * t;
            if (e < total) nd[e] = res[t];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < dd; e += 256) {
        const int i = e / d, k = e % d;
        cplx<R> v = nd[i * D + k];
        if (final_level && i == k) v.x += R(1);
        dst[(size_t)blockIdx.x * dd + e] = v;
    }
}

template <class R, class SrcT>
static void launch_small_tree(int D, const SrcT* src, int n, int d, int b, cplx<R>* out, bool fin, hipStream_t st) {
    const int n_out = (n + b - 1) / b;
    const size_t lds = (size_t)b * D * D * sizeof(cplx<R>);
    if (D == 4)
        hipLaunchKernelGGL((small_tree_kernel<R, 4, SrcT>), dim3(n_out), dim3(256), lds, st, src, n, d, b, out, fin ? 1 : 0);
    else if (D == 8)
        hipLaunchKernelGGL((small_tree_kernel<R, 8, SrcT>), dim3(n_out), dim3(256), lds, st, src, n, d, b, out, fin ? 1 : 0);
    else
        hipLaunchKernelGGL((small_tree_kernel<R, 16, SrcT>), dim3(n_out), dim3(256), lds, st, src, n, d, b, out, fin ? 1 : 0);
}

template <class R> constexpr int tree_logb() { return 3; }
// levels with at most this many nodes go through tree_split_kernel
constexpr int kTreeSplitNodes = 1024;

template <class R, int D, class SrcT>
static void launch_tree_level(const SrcT* src, int n, int d, cplx<R>* dst, bool final_level, hipStream_t st) {
    constexpr int LOGB = tree_logb<R>();
    const int grid = (n + (1 << LOGB) - 1) >> LOGB;
    hipLaunchKernelGGL((tree_block_kernel<R, D, LOGB, SrcT>), dim3(grid), dim3(64 * Shape<D>::G), 0, st, src, n, d,
                       dst, final_level ? 1 : 0);
}

template <class R, int D, class SrcT>
static void launch_tree_split(const SrcT* src, int n, int d, cplx<R>* dst, bool final_level, hipStream_t st) {
    const int grid = ((n + 1) / 2) * (D / 16) * (D / 16);
    hipLaunchKernelGGL((tree_split_kernel<R, D, SrcT>), dim3(grid), dim3(256), 0, st, src, n, d, dst,
                       final_level ? 1 : 0);
}

// nodes of the widest level a scratch buffer receives
template <class R>
int tree_scratch_nodes(int N) {
    constexpr int B = 1 << tree_logb<R>();
    return N <= kTreeSplitNodes ? (N + 1) / 2 : (N + B - 1) / B;
}
template int tree_scratch_nodes<float>(int);
template int tree_scratch_nodes<double>(int);

template <class R>
hipError_t launch_tree(const void* leaves, bool leaves_are_c64, int N, int d, cplx<R>* scratch0, cplx<R>* scratch1,
                       cplx<R>* root, hipStream_t st) {
    if (d <= 0 || d > 64 || N <= 0) return hipErrorInvalidValue;
    if (d <= 16) {
        // block reductions in LDS: ceil(n / B) nodes per launch, nodes padded to D x D
        const int D = d <= 4 ? 4 : (d <= 8 ? 8 : 16);
        const int B = small_tree_block(D);
        cplx<R>* bufs[2] = {scratch0, scratch1};
        int which = 0, n = N;
        bool first = true;
        while (true) {
            const int n_out = (n + B - 1) / B;
            const bool fin = n_out == 1;
            int b = B;   // a single block needs only the next power of two above its node count
            if (fin) while (b > 1 && (b >> 1) >= n) b >>= 1;
            cplx<R>* out = fin ? root : bufs[which];
            if (first && leaves_are_c64)
                launch_small_tree<R>(D, reinterpret_cast<const float2*>(leaves), n, d, b, out, fin, st);
            else
                launch_small_tree<R>(D, first ? reinterpret_cast<const cplx<R>*>(leaves) : bufs[which ^ 1], n, d, b, out,
                                     fin, st);
            if (fin) break;
            first = false;
            n = n_out;
            which ^= 1;
        }
        return hipGetLastError();
    }
    constexpr int B = 1 << tree_logb<R>();
    // a level of n nodes: B-fold blocks on one work-group each, or one pairwise level split over
    // work-groups once few nodes remain; returns the node count of the level it writes
    auto level = [&](const auto* src, int n, cplx<R>* root_or_dst, cplx<R>* dst) -> int {
        const bool split = n <= kTreeSplitNodes;
        const int n_out = split ? (n + 1) / 2 : (n + B - 1) / B;
        const bool fin = n_out == 1;
        cplx<R>* out = fin ? root_or_dst : dst;
        if (split) {
            if (d <= 16) launch_tree_split<R, 16>(src, n, d, out, fin, st);
            else if (d <= 32) launch_tree_split<R, 32>(src, n, d, out, fin, st);
            else if (d <= 48) launch_tree_split<R, 48>(src, n, d, out, fin, st);
            else launch_tree_split<R, 64>(src, n, d, out, fin, st);
        } else {
            if (d <= 16) launch_tree_level<R, 16>(src, n, d, out, fin, st);
            else if (d <= 32) launch_tree_level<R, 32>(src, n, d, out, fin, st);
            else if (d <= 48) launch_tree_level<R, 48>(src, n, d, out, fin, st);
            else launch_tree_level<R, 64>(src, n, d, out, fin, st);
        }
        return n_out;
    };
    cplx<R>* bufs[2] = {scratch0, scratch1};
    int which = 0;
    // first level reads the leaves
    int n = leaves_are_c64 ? level(reinterpret_cast<const float2*>(leaves), N, root, bufs[which])
                           : level(reinterpret_cast<const cplx<R>*>(leaves), N, root, bufs[which]);
    while (n > 1) {
        n = level(static_cast<const cplx<R>*>(bufs[which]), n, root, bufs[which ^ 1]);
        which ^= 1;
    }
    return hipGetLastError();
}
template hipError_t launch_tree<float>(const void*, bool, int, int, float2*, float2*, float2*, hipStream_t);
template hipError_t launch_tree<double>(const void*, bool, int, int, double2*, double2*, double2*, hipStream_t);

// ---------------------------------------------------------------------------------------
// Generic step-wise kernels (any dim): the reference's launch structure
// (parament.cpp:486-718) with the diagonal add and the scaling fused into the GEMM.
// Assembly X_j = H0 + sum_t c'_tj O_t, one thread per matrix element over kAsmSlices consecutive
// slices: the element's operator values are read once and used for every slice of the group
// (one read per slice had the operators streamed from the cache hierarchy once per slice,
// (1 + ncoef) x 16 bytes per element of every slice).  Per slice the sum keeps its order.
// Dynamic LDS: the group's coefficients, kAsmSlices x ncoef.
constexpr int kAsmSlices = 16;

template <class R>
__device__ __forceinline__ void asm_group_coefficients(const ChebArgs<R>& a, int j0, int ns, cplx<R>* coef,
                                                       bool local_index) {
    for (int idx = threadIdx.x; idx < ns * a.ncoef; idx += blockDim.x) {
        const int s = idx / a.ncoef, t = idx % a.ncoef;
        coef[idx] = slice_coefficient(a.ctl, t, (local_index ? 0 : a.j0) + j0 + s);
    }
    __syncthreads();
}

// x[s] = X_{j0+s} at element e (s < kAsmSlices; entries past the group's ns slices are unused)
template <class R>
__device__ __forceinline__ void asm_group_element(const ChebArgs<R>& a, const cplx<R>* coef, int e, int dd,
                                                  cplx<R> (&x)[kAsmSlices]) {
    const cplx<R> h = a.ops[e];
#pragma unroll
    for (int s = 0; s < kAsmSlices; ++s) x[s] = h;
    for (int t = 0; t < a.ncoef; ++t) {
        const cplx<R> o = a.ops[(size_t)(t + 1) * dd + e];
#pragma unroll
        for (int s = 0; s < kAsmSlices; ++s) x[s] = cadd(x[s], cmul(coef[s * a.ncoef + t], o));
    }
}

template <class R, class TO>
__global__ __launch_bounds__(256) void assemble_kernel(ChebArgs<R> a, TO* __restrict__ X, int blocks_per) {
    extern __shared__ __attribute__((aligned(16))) unsigned char asm_lds[];
    cplx<R>* coef = reinterpret_cast<cplx<R>*>(asm_lds);
    const int j0 = (blockIdx.x / blocks_per) * kAsmSlices, b = blockIdx.x % blocks_per;
    const int ns = min(kAsmSlices, a.N - j0);
    asm_group_coefficients<R>(a, j0, ns, coef, false);
    const int dd = a.d * a.d;
    const int e = b * 256 + threadIdx.x;
    if (e >= dd) return;
    cplx<R> x[kAsmSlices];
    asm_group_element<R>(a, coef, e, dd, x);
#pragma unroll
    for (int s = 0; s < kAsmSlices; ++s)
        if (s < ns) {
            TO o;
            o.x = x[s].x * a.asm_scale; o.y = x[s].y * a.asm_scale;
            X[(size_t)(j0 + s) * dd + e] = o;
        }
}

template <class R, class TO>
static hipError_t launch_assemble_to(const ChebArgs<R>& a, TO* X, hipStream_t st) {
    if (a.ncoef > kMaxCoeffs || a.N <= 0) return hipErrorInvalidValue;
    const int blocks_per = (a.d * a.d + 255) / 256;
    const unsigned groups = (unsigned)((a.N + kAsmSlices - 1) / kAsmSlices);
    const size_t lds = (size_t)kAsmSlices * a.ncoef * sizeof(cplx<R>);
    hipLaunchKernelGGL((assemble_kernel<R, TO>), dim3(groups * blocks_per), dim3(256), lds, st, a, X, blocks_per);
    return hipGetLastError();
}

template <class R>
hipError_t launch_assemble(const ChebArgs<R>& a, cplx<R>* X, hipStream_t st) {
    return launch_assemble_to<R, cplx<R>>(a, X, st);
}
template hipError_t launch_assemble<float>(const ChebArgs<float>&, float2*, hipStream_t);
template hipError_t launch_assemble<double>(const ChebArgs<double>&, double2*, hipStream_t);

hipError_t launch_assemble_c64(const ChebArgs<double>& a, float2* X, hipStream_t st) {
    return launch_assemble_to<double, float2>(a, X, st);
}

__global__ __launch_bounds__(256) void widen_kernel(const float2* __restrict__ in, long total, double2* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < total) out[i] = mk<double>(in[i].x, in[i].y);
}

hipError_t launch_widen(const float2* in, int d, int n, double2* out, hipStream_t st) {
    const long total = (long)n * d * d;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, in, total, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void narrow_kernel(const double2* __restrict__ in, long total, float2* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < total) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}

hipError_t launch_narrow(const double2* in, long n, float2* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(narrow_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, n, out);
    return hipGetLastError();
}

// E_j = (a_0 - 1) I + a_1 s X_j + R_j in FP64 (X_j recomputed from the operators, kAsmSlices
// slices per thread as in assemble_kernel), rounded once to complex64: the first-order part of
// the mixed-precision slice propagator.
__global__ __launch_bounds__(256) void first_order_kernel(ChebArgs<double> a, const float2* __restrict__ R,
                                                          float2* __restrict__ E, int blocks_per) {
    extern __shared__ __attribute__((aligned(16))) unsigned char fo_lds[];
    double2* coef = reinterpret_cast<double2*>(fo_lds);
    const int j0 = (blockIdx.x / blocks_per) * kAsmSlices, b = blockIdx.x % blocks_per;
    const int ns = min(kAsmSlices, a.N - j0);
    asm_group_coefficients<double>(a, j0, ns, coef, true);
    const int d = a.d, dd = d * d;
    const int e = b * 256 + threadIdx.x;
    if (e >= dd) return;
    double2 x[kAsmSlices];
    asm_group_element<double>(a, coef, e, dd, x);
    const double2 a1s = mk<double>(a.are[1] * a.s, a.aim[1] * a.s);
    const double dg = (e / d == e % d) ? a.are[0] : 0.0;
#pragma unroll
    for (int s = 0; s < kAsmSlices; ++s)
        if (s < ns) {
            const size_t o = (size_t)(j0 + s) * dd + e;
            const double2 f = cmul(a1s, x[s]);
            const float2 r = R[o];
            const double er = f.x + (double)r.x + dg;
            const double ei = f.y + (double)r.y;
            E[o] = make_float2((float)er, (float)ei);
        }
}

hipError_t launch_first_order(const ChebArgs<double>& a, const float2* R, float2* E, hipStream_t st) {
    if (a.ncoef > kMaxCoeffs || a.N <= 0) return hipErrorInvalidValue;
    const int blocks_per = (a.d * a.d + 255) / 256;
    const unsigned groups = (unsigned)((a.N + kAsmSlices - 1) / kAsmSlices);
    const size_t lds = (size_t)kAsmSlices * a.ncoef * sizeof(double2);
    hipLaunchKernelGGL(first_order_kernel, dim3(groups * blocks_per), dim3(256), lds, st, a, R, E, blocks_per);
    return hipGetLastError();
}

template <class R>
__global__ __launch_bounds__(256) void fill_kernel(const cplx<R>* __restrict__ X, cplx<R> alpha, cplx<R> diag, int d,
                                                   long total, cplx<R>* __restrict__ out) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= total) return;
    const int dd = d * d;
    const int e = (int)(idx % dd), i = e / d, k = e % d;
    cplx<R> v = mk<R>(0, 0);
    if (X) v = cmul(alpha, X[idx]);
    if (i == k) v = cadd(v, diag);
    out[idx] = v;
}

template <class R>
hipError_t launch_fill(const cplx<R>* X, cplx<R> alpha, cplx<R> diag, int d, int batch, cplx<R>* out,
                       hipStream_t st) {
    const long total = (long)batch * d * d;
    hipLaunchKernelGGL((fill_kernel<R>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, X, alpha, diag, d,
                       total, out);
    return hipGetLastError();
}
template hipError_t launch_fill<float>(const float2*, float2, float2, int, int, float2*, hipStream_t);
template hipError_t launch_fill<double>(const double2*, double2, double2, int, int, double2*, hipStream_t);

}  // namespace pa
