// 27-point Helmholtz stencil with PML on gfx950: forward, adjoint, derivative
// and complex-wavenumber forms in one z-marching kernel template.
//
// Reference semantics: do_Hmvp of SLIM-release-apps
// tools/algorithms/3DFreqModeling/Helm3d_27pt.h:827-875 with the coefficients
// of :680-714 (forward) and :718-822 (adjoint), the DERIV / WNCMPLX variants of
// :4-16, 202-206, and zero padding outside the grid (:263-317).
//
// Formulation.  Every coefficient is linear in the PML values of one point, so
// the row of H at p = (i,j,k) splits into
//   * per-column complex coefficient sets F (in-plane neighbours) and G
//     (neighbours in planes k+-1) that carry the x/y PML of the column and are
//     fixed for the whole z-march,
//   * a real 9-point in-plane stencil Qz that carries the z PML, weighted per
//     plane by xi_z^lo, xi_z^hi, xi_z (uniform across the block),
//   * the mass-lumping stencil (wn_coef, -WM2, -WM3, -WM4 by |d|_1) acting on
//     z = wn .* x (forward) or on x with conj(wn) at the centre (adjoint).
// The adjoint is the exact conjugate transpose: F/G take the neighbour's x/y
// PML values and the z weights come from the neighbouring planes.
//
// Tiling.  A block of 64 x TY / RY threads (TY = 4 or 8 rows, RY = 1 or 2 rows
// per thread) owns a 64 x TY column tile and marches over a z chunk.  Each plane (with a one-point x/y halo) is
// staged once in LDS (double-buffered, one barrier per plane); the global
// loads of the next PF planes (PF = 1..4) are in flight while a plane is
// computed, so HBM latency overlaps the march.  Block
// ids are remapped so consecutive tiles of the same chunk land on the same XCD
// and share halo lines in its L2.
#include "helm27_body.hpp"

namespace helm3d {

namespace {

template <typename R, bool ADJ, bool DERIV, bool WNC, int TY, int PF, int EPI, int RY>
__global__ __launch_bounds__(TX* TY / RY) void helm27_kernel(StencilLaunch<R> L)
{
    helm27_body<R, ADJ, DERIV, WNC, TY, PF, EPI, RY>(L, blockIdx.x, gridDim.x);
}

}  // namespace

template <typename R>
StencilGrid stencil_grid(int nx, int ny, int nz, int kchunk, int ty, int pf, int ry)
{
    StencilGrid g;
    g.ty = (ty == 8) ? 8 : 4;
    g.pf = std::max(1, std::min(pf, 4));
    g.ry = (ry == 2 && g.ty == 8) ? 2 : 1;
    g.tiles_x = (nx + TX - 1) / TX;
    g.tiles_y = (ny + g.ty - 1) / g.ty;
    if (kchunk <= 0) {
        // about two blocks per CU (256 CUs) with long z marches: measured best
        // on 240x240x140 (fewer, longer marches amortise the two pipeline
        // fill planes and keep halo planes in L2); chunks of >= 8 planes
        const long tiles = long(g.tiles_x) * g.tiles_y;
        int nch = int((480 + tiles - 1) / tiles);
        nch = std::max(1, std::min(nch, (nz + 7) / 8));
        kchunk = (nz + nch - 1) / nch;
    }
    g.kchunk = std::max(1, std::min(std::min(kchunk, nz), kMaxChunk));
    g.nchunks = (nz + g.kchunk - 1) / g.kchunk;
    g.blocks = g.tiles_x * g.tiles_y * g.nchunks;
    return g;
}

namespace {
template <typename R, bool A, bool D, bool W, int EPI>
hipError_t launch_shape(const StencilLaunch<R>& L, const StencilGrid& g, hipStream_t s)
{
    const dim3 grid(g.blocks);
    if (g.ty == 8 && g.ry == 2) {
        switch (g.pf) {
        case 1: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 1, EPI, 2>), grid, dim3(TX * 4), 0, s, L); break;
        case 2: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 2, EPI, 2>), grid, dim3(TX * 4), 0, s, L); break;
        case 3: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 3, EPI, 2>), grid, dim3(TX * 4), 0, s, L); break;
        default: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 4, EPI, 2>), grid, dim3(TX * 4), 0, s, L); break;
        }
    } else if (g.ty == 8) {
        switch (g.pf) {
        case 1: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 1, EPI, 1>), grid, dim3(TX * 8), 0, s, L); break;
        case 2: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 2, EPI, 1>), grid, dim3(TX * 8), 0, s, L); break;
        case 3: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 3, EPI, 1>), grid, dim3(TX * 8), 0, s, L); break;
        default: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 4, EPI, 1>), grid, dim3(TX * 8), 0, s, L); break;
        }
    } else {
        switch (g.pf) {
        case 1: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 4, 1, EPI, 1>), grid, dim3(TX * 4), 0, s, L); break;
        case 2: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 4, 2, EPI, 1>), grid, dim3(TX * 4), 0, s, L); break;
        case 3: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 4, 3, EPI, 1>), grid, dim3(TX * 4), 0, s, L); break;
        default: hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 4, 4, EPI, 1>), grid, dim3(TX * 4), 0, s, L); break;
        }
    }
    return hipGetLastError();
}

// Epilogue by operand presence: the addend form exists for the wavenumber
// operators (residuals), the gradient form for the FP64 adjoint derivative.
template <typename R, bool A, bool D, bool W>
hipError_t launch_epi(const StencilLaunch<R>& L, const StencilGrid& g, hipStream_t s)
{
    if constexpr (std::is_same<R, double>::value && A && D && !W) {
        if (L.gg) return launch_shape<R, A, D, W, 2>(L, g, s);
    }
    if constexpr (!D) {
        if (L.b) return launch_shape<R, A, D, W, 1>(L, g, s);
    }
    if (L.gg || L.b) return hipErrorInvalidValue;
    return launch_shape<R, A, D, W, 0>(L, g, s);
}
}  // namespace

template <typename R>
hipError_t launch_stencil(int variant, StencilLaunch<R> L, const StencilGrid& g, hipStream_t s)
{
    L.tiles_x = g.tiles_x;
    L.tiles_y = g.tiles_y;
    L.kchunk = g.kchunk;
    const bool adj = variant & VAR_ADJ, der = variant & VAR_DERIV, wnc = variant & VAR_WNC;
    if (!adj && !der && !wnc) return launch_epi<R, false, false, false>(L, g, s);
    if (!adj && !der && wnc) return launch_epi<R, false, false, true>(L, g, s);
    if (!adj && der && !wnc) return launch_epi<R, false, true, false>(L, g, s);
    if (!adj && der && wnc) return launch_epi<R, false, true, true>(L, g, s);
    if (adj && !der && !wnc) return launch_epi<R, true, false, false>(L, g, s);
    if (adj && !der && wnc) return launch_epi<R, true, false, true>(L, g, s);
    if (adj && der && !wnc) return launch_epi<R, true, true, false>(L, g, s);
    return launch_epi<R, true, true, true>(L, g, s);
}

template StencilGrid stencil_grid<double>(int, int, int, int, int, int, int);
template StencilGrid stencil_grid<float>(int, int, int, int, int, int, int);
template hipError_t launch_stencil<double>(int, StencilLaunch<double>, const StencilGrid&, hipStream_t);
template hipError_t launch_stencil<float>(int, StencilLaunch<float>, const StencilGrid&, hipStream_t);

}  // namespace helm3d
