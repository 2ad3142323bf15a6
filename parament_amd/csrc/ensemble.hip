// Multi-instance kernels of the ensemble solver: many independent equiprop problems of
// different dims, controls and lengths, each dim class (16 / 32 / 48 / 64 after padding) served by
// one persistent kernel that walks a task table.  A slice task (instance i, slice j) runs the
// fused Chebyshev slice body (parament.cpp:569-652) with instance i's operators, controls and
// coefficients; a tree task (level descriptor l, block b) reduces 2^LOGB nodes of one
// instance's pairwise product tree (parament.cpp:657-718) exactly as the single-problem kernels
// do.  The wave split (T) and the tree block size (LOGB) are template parameters chosen per
// class at run time by the host's tuner.  complex64.
#include "fused_device.hpp"

namespace pa {

template <int D, int T>
__global__ __launch_bounds__((64 * TileShape<D, T>::G)) void ens_slice_kernel(const EnsInstance* __restrict__ inst,
                                                                             const int2* __restrict__ tasks, int ntasks) {
    typedef TileShape<D, T> SH;
    constexpr int SA = D + Pad<float>::A, SB = D + Pad<float>::B;
    __shared__ float Xr[D * SA], Xi[D * SA], Br[D * SB], Bi[D * SB];
    __shared__ float2 coef[kMaxCoeffs];
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const int2 tk = tasks[t];
        const EnsInstance& I = inst[tk.x];
        Controls<float> ctl;
        ctl.c = I.c; ctl.pts = I.pts; ctl.amps_in = I.amps; ctl.amps_set = I.amps;
        ctl.mode = I.mode; ctl.magnus = 0; ctl.tau = I.tau;
        float2* out = I.leaves + (size_t)tk.y * I.d * I.d;
        if (!I.ps) {
            cheb_slice<float, D, SH>(I.ops, I.ncoef, I.d, ctl, tk.y, I.M, I.s, I.are, I.aim, Xr, Xi, Br, Bi, coef, out);
            continue;
        }
        const int tid = threadIdx.x, wave = tid >> 6;
        const int row0 = SH::row0(wave), col0 = SH::col0(wave), d = I.d;
        for (int c = tid; c < I.ncoef; c += 64 * SH::G) coef[c] = slice_coefficient(ctl, c, tk.y);
        __syncthreads();
        const CTile<float, SH::T, SH::T> e =
            power_slice_tile<float, D, SH>(I.ops, I.ncoef, d, I.M, I.tau, I.pre, I.pim, Xr, Xi, Br, Bi, coef);
#pragma unroll
        for (int m = 0; m < SH::T; ++m)
#pragma unroll
            for (int n = 0; n < SH::T; ++n)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = acc_row<float>(row0, m, v), col = acc_col(col0, n);
                    if (row < d && col < d) out[row * d + col] = mk<float>(e.re[m][n][v], e.im[m][n][v]);
                }
    }
}

template <int D, int T, int LOGB>
__global__ __launch_bounds__((64 * TileShape<D, T>::G)) void ens_tree_kernel(const EnsLevel* __restrict__ lv,
                                                                            const int2* __restrict__ tasks, int ntasks) {
    typedef TileShape<D, T> SH;
    constexpr int SA = D + Pad<float>::A, SB = D + Pad<float>::B;
    __shared__ float Ar[D * SA], Ai[D * SA], Br[D * SB], Bi[D * SB];
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const int2 tk = tasks[t];
        const EnsLevel& L = lv[tk.x];
        tree_block<float, D, LOGB, float2, SH>(L.src, L.n_in, L.d, L.dst + (size_t)tk.y * L.d * L.d, L.final_level, tk.y,
                                               Ar, Ai, Br, Bi);
    }
}

namespace {
template <int D, int T>
hipError_t slice_launch(const EnsInstance* inst, const int2* tasks, int ntasks, int grid, hipStream_t st) {
    hipLaunchKernelGGL((ens_slice_kernel<D, T>), dim3(grid), dim3(64 * TileShape<D, T>::G), 0, st, inst, tasks, ntasks);
    return hipGetLastError();
}
template <int D, int T, int LOGB>
hipError_t tree_launch(const EnsLevel* lv, const int2* tasks, int ntasks, int grid, hipStream_t st) {
    hipLaunchKernelGGL((ens_tree_kernel<D, T, LOGB>), dim3(grid), dim3(64 * TileShape<D, T>::G), 0, st, lv, tasks, ntasks);
    return hipGetLastError();
}
}  // namespace

bool ens_shape_valid(int D, int T) {
    return (D == 16 && T == 1) || (D == 32 && (T == 1 || T == 2)) || (D == 48 && T == 1) ||
           (D == 64 && (T == 1 || T == 2));
}

hipError_t launch_ens_slices(int D, int T, const EnsInstance* inst, const int2* tasks, int ntasks, int grid, hipStream_t st) {
    if (ntasks <= 0) return hipSuccess;
    if (grid <= 0 || !ens_shape_valid(D, T)) return hipErrorInvalidValue;
    if (D == 16) return slice_launch<16, 1>(inst, tasks, ntasks, grid, st);
    if (D == 32) return T == 1 ? slice_launch<32, 1>(inst, tasks, ntasks, grid, st) : slice_launch<32, 2>(inst, tasks, ntasks, grid, st);
    if (D == 48) return slice_launch<48, 1>(inst, tasks, ntasks, grid, st);
    return T == 1 ? slice_launch<64, 1>(inst, tasks, ntasks, grid, st) : slice_launch<64, 2>(inst, tasks, ntasks, grid, st);
}

hipError_t launch_ens_tree(int D, int T, int logb, const EnsLevel* lv, const int2* tasks, int ntasks, int grid,
                           hipStream_t st) {
    if (ntasks <= 0) return hipSuccess;
    if (grid <= 0 || !ens_shape_valid(D, T) || (logb != 2 && logb != 3)) return hipErrorInvalidValue;
#define PA_ENS_TREE(DD, TT)                                                                         \
    return logb == 2 ? tree_launch<DD, TT, 2>(lv, tasks, ntasks, grid, st)                          \
                     : tree_launch<DD, TT, 3>(lv, tasks, ntasks, grid, st);
    if (D == 16) { PA_ENS_TREE(16, 1) }
    if (D == 32) {
        if (T == 1) { PA_ENS_TREE(32, 1) }
        PA_ENS_TREE(32, 2)
    }
    if (D == 48) { PA_ENS_TREE(48, 1) }
    if (T == 1) { PA_ENS_TREE(64, 1) }
    PA_ENS_TREE(64, 2)
#undef PA_ENS_TREE
}

}  // namespace pa
