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

template <class R>
hipError_t launch_cheb_fused(const ChebArgs<R>& a, hipStream_t st) {
    if (a.d <= 0 || a.d > 64 || a.N <= 0 || a.M < 3 || a.M >= kMaxCheb || a.ncoef > kMaxCoeffs)
        return hipErrorInvalidValue;
    const int grid = a.N;
    if (a.d <= 16) hipLaunchKernelGGL((cheb_fused_kernel<R, 16>), dim3(grid), dim3(64 * Shape<16>::G), 0, st, a);
    else if (a.d <= 32) hipLaunchKernelGGL((cheb_fused_kernel<R, 32>), dim3(grid), dim3(64 * Shape<32>::G), 0, st, a);
    else hipLaunchKernelGGL((cheb_fused_kernel<R, 64>), dim3(grid), dim3(64 * Shape<64>::G), 0, st, a);
    return hipGetLastError();
}
template hipError_t launch_cheb_fused<float>(const ChebArgs<float>&, hipStream_t);
template hipError_t launch_cheb_fused<double>(const ChebArgs<double>&, hipStream_t);

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
            else launch_tree_split<R, 64>(src, n, d, out, fin, st);
        } else {
            if (d <= 16) launch_tree_level<R, 16>(src, n, d, out, fin, st);
            else if (d <= 32) launch_tree_level<R, 32>(src, n, d, out, fin, st);
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
template <class R, class TO>
__global__ __launch_bounds__(256) void assemble_kernel(ChebArgs<R> a, TO* __restrict__ X, int blocks_per) {
    __shared__ cplx<R> coef[kMaxCoeffs];
    const int j = blockIdx.x / blocks_per, b = blockIdx.x % blocks_per;
    for (int t = threadIdx.x; t < a.ncoef; t += blockDim.x) coef[t] = slice_coefficient(a.ctl, t, j);
    __syncthreads();
    const int dd = a.d * a.d;
    const int e = b * 256 + threadIdx.x;
    if (e >= dd) return;
    cplx<R> x = a.ops[e];
    for (int t = 0; t < a.ncoef; ++t) x = cadd(x, cmul(coef[t], a.ops[(size_t)(t + 1) * dd + e]));
    TO o;
    o.x = x.x; o.y = x.y;
    X[(size_t)j * dd + e] = o;
}

template <class R>
hipError_t launch_assemble(const ChebArgs<R>& a, cplx<R>* X, hipStream_t st) {
    if (a.ncoef > kMaxCoeffs) return hipErrorInvalidValue;
    const int blocks_per = (a.d * a.d + 255) / 256;
    hipLaunchKernelGGL((assemble_kernel<R, cplx<R>>), dim3((unsigned)a.N * blocks_per), dim3(256), 0, st, a, X, blocks_per);
    return hipGetLastError();
}
template hipError_t launch_assemble<float>(const ChebArgs<float>&, float2*, hipStream_t);
template hipError_t launch_assemble<double>(const ChebArgs<double>&, double2*, hipStream_t);

hipError_t launch_assemble_c64(const ChebArgs<double>& a, float2* X, hipStream_t st) {
    if (a.ncoef > kMaxCoeffs) return hipErrorInvalidValue;
    const int blocks_per = (a.d * a.d + 255) / 256;
    hipLaunchKernelGGL((assemble_kernel<double, float2>), dim3((unsigned)a.N * blocks_per), dim3(256), 0, st, a, X,
                       blocks_per);
    return hipGetLastError();
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

// E_j = (a_0 - 1) I + a_1 s X_j + R_j in FP64 (X_j recomputed from the operators), rounded once
// to complex64: the first-order part of the mixed-precision slice propagator.
__global__ __launch_bounds__(256) void first_order_kernel(ChebArgs<double> a, const float2* __restrict__ R,
                                                          float2* __restrict__ E, int blocks_per) {
    __shared__ double2 coef[kMaxCoeffs];
    const int j = blockIdx.x / blocks_per, b = blockIdx.x % blocks_per;
    for (int t = threadIdx.x; t < a.ncoef; t += blockDim.x) coef[t] = slice_coefficient(a.ctl, t, j);
    __syncthreads();
    const int d = a.d, dd = d * d;
    const int e = b * 256 + threadIdx.x;
    if (e >= dd) return;
    double2 x = a.ops[e];
    for (int t = 0; t < a.ncoef; ++t) x = cadd(x, cmul(coef[t], a.ops[(size_t)(t + 1) * dd + e]));
    const double2 a1s = mk<double>(a.are[1] * a.s, a.aim[1] * a.s);
    const double2 f = cmul(a1s, x);
    const float2 r = R[(size_t)j * dd + e];
    const double er = f.x + (double)r.x + ((e / d == e % d) ? a.are[0] : 0.0);
    const double ei = f.y + (double)r.y;
    E[(size_t)j * dd + e] = make_float2((float)er, (float)ei);
}

hipError_t launch_first_order(const ChebArgs<double>& a, const float2* R, float2* E, hipStream_t st) {
    if (a.ncoef > kMaxCoeffs) return hipErrorInvalidValue;
    const int blocks_per = (a.d * a.d + 255) / 256;
    hipLaunchKernelGGL(first_order_kernel, dim3((unsigned)a.N * blocks_per), dim3(256), 0, st, a, R, E, blocks_per);
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
