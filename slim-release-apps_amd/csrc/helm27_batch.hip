// Work-list launches of the 27-point stencil: one launch applies the stencil
// to many (operator, vector) items, possibly on grids of different sizes
// (batched right-hand sides, several frequencies or models at once).
#include "helm27_body.hpp"

namespace helm3d {

namespace {

// One launch over a work list of stencil applications (possibly on different
// grids): item it owns blocks [first[it], first[it+1]); items with
// active[it] == 0 are skipped.
template <typename R, bool ADJ, bool DERIV, bool WNC, int TY, int PF, int EPI, int RY>
__global__ __launch_bounds__(TX* TY / RY) void helm27_batch_kernel(const StencilLaunch<R>* __restrict__ items,
                                                              const int* __restrict__ first, int nitems,
                                                              const int* __restrict__ active)
{
    const int bid = blockIdx.x;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= bid) lo = mid;
        else hi = mid - 1;
    }
    if (active && !active[lo]) return;
    helm27_body<R, ADJ, DERIV, WNC, TY, PF, EPI, RY>(items[lo], bid - first[lo], first[lo + 1] - first[lo]);
}


template <typename R, bool A, bool W, int EPI>
hipError_t launch_b(const StencilLaunch<R>* items, const int* first, int nitems, int blocks, const int* active,
                    int ty, int pf, int ry, hipStream_t s)
{
    const dim3 grid(blocks);
    if (ty == 8 && ry == 2) {
        if (pf >= 2)
            hipLaunchKernelGGL((helm27_batch_kernel<R, A, false, W, 8, 2, EPI, 2>), grid, dim3(TX * 4), 0, s, items,
                               first, nitems, active);
        else
            hipLaunchKernelGGL((helm27_batch_kernel<R, A, false, W, 8, 1, EPI, 2>), grid, dim3(TX * 4), 0, s, items,
                               first, nitems, active);
    } else if (ty == 8) {
        if (pf >= 2)
            hipLaunchKernelGGL((helm27_batch_kernel<R, A, false, W, 8, 2, EPI, 1>), grid, dim3(TX * 8), 0, s, items,
                               first, nitems, active);
        else
            hipLaunchKernelGGL((helm27_batch_kernel<R, A, false, W, 8, 1, EPI, 1>), grid, dim3(TX * 8), 0, s, items,
                               first, nitems, active);
    } else {
        if (pf >= 2)
            hipLaunchKernelGGL((helm27_batch_kernel<R, A, false, W, 4, 2, EPI, 1>), grid, dim3(TX * 4), 0, s, items,
                               first, nitems, active);
        else
            hipLaunchKernelGGL((helm27_batch_kernel<R, A, false, W, 4, 1, EPI, 1>), grid, dim3(TX * 4), 0, s, items,
                               first, nitems, active);
    }
    return hipGetLastError();
}

template <typename R, bool A, bool W>
hipError_t launch_v(int epi, const StencilLaunch<R>* items, const int* first, int nitems, int blocks,
                    const int* active, int ty, int pf, int ry, hipStream_t s)
{
    switch (epi) {
    case 0: return launch_b<R, A, W, 0>(items, first, nitems, blocks, active, ty, pf, ry, s);
    case 1: return launch_b<R, A, W, 1>(items, first, nitems, blocks, active, ty, pf, ry, s);
    case 3: return launch_b<R, A, W, 3>(items, first, nitems, blocks, active, ty, pf, ry, s);
    case 4: return launch_b<R, A, W, 4>(items, first, nitems, blocks, active, ty, pf, ry, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <typename R>
hipError_t launch_stencil_batch(int variant, int epi, const StencilLaunch<R>* items, const int* first, int nitems,
                                int blocks, const int* active, int ty, int pf, hipStream_t s, int ry)
{
    if (nitems <= 0 || blocks <= 0) return hipSuccess;
    if (variant & VAR_DERIV) return hipErrorInvalidValue;
    const bool adj = variant & VAR_ADJ, wnc = variant & VAR_WNC;
    if (!adj && !wnc) return launch_v<R, false, false>(epi, items, first, nitems, blocks, active, ty, pf, ry, s);
    if (!adj && wnc) return launch_v<R, false, true>(epi, items, first, nitems, blocks, active, ty, pf, ry, s);
    if (adj && !wnc) return launch_v<R, true, false>(epi, items, first, nitems, blocks, active, ty, pf, ry, s);
    return launch_v<R, true, true>(epi, items, first, nitems, blocks, active, ty, pf, ry, s);
}

template hipError_t launch_stencil_batch<double>(int, int, const StencilLaunch<double>*, const int*, int, int,
                                                 const int*, int, int, hipStream_t, int);
template hipError_t launch_stencil_batch<float>(int, int, const StencilLaunch<float>*, const int*, int, int,
                                                const int*, int, int, hipStream_t, int);

}  // namespace helm3d
