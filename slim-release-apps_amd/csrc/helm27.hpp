// Launch interface of the stencil and Krylov kernels (internal to the library).
#pragma once

#include <algorithm>

#include "helm3d_common.hpp"

namespace helm3d {

// Operator variant bits; the six combinations built as MEX binaries in
// ibin/install_MEX:88-105 are {0, ADJ} x {real wn, WNC, DERIV}.
enum Variant : int { VAR_FORW = 0, VAR_ADJ = 1, VAR_DERIV = 2, VAR_WNC = 4 };

template <typename R>
struct StencilLaunch {
    OpDesc<R> op;
    const cpx<R>* x;     // input vector (device, interleaved complex)
    cpx<R>* y;           // output: y = alpha * (H x) + beta * b
    const cpx<R>* b;     // optional addend (nullptr: none)
    R alpha, beta;
    const double* anrm;  // optional (EPI 0/1): alpha is divided by sqrt(*anrm) (device scalar)
    double* norm_part;   // optional: per-block partial sums of |y|^2
    // optional FWI-gradient epilogue: instead of storing y, accumulate
    // gg[phys] += gscale * Re(conj(gu) * y) over the physical box
    // [p0, p0+pn) of each axis (comp2phys truncation of the PML).
    const cpx<R>* gu;
    double* gg;
    double gscale;
    int p0[3], pn[3];
    // write-only sink of kSinkSlots values for steps/lanes with no output
    cpx<R>* sink;
    int tiles_x, tiles_y, kchunk;
    // fused BiCGStab forms, scalars in sc:
    //   EPI 3: f0 = r, f1 = p, f2 = v, f3 = t of the previous iteration;
    //          y = new v, o2 = new p, o3 = new r, zz = z (updated in place),
    //          rh = rhat
    //   EPI 4: f0 = r, f1 = v; y = t, rh = rhat
    const cpx<R>* f0;
    const cpx<R>* f1;
    const cpx<R>* f2;
    const cpx<R>* f3;
    cpx<R>* o2;
    cpx<R>* o3;
    cpx<R>* zz;
    const cpx<R>* rh;
    double* sc;
    int rho_new, rho_old, first;
    const int* skip;  // optional (device): the launch does nothing while *skip != 0
};

// Launch shape: 64 x ty column tiles, z chunks of kchunk planes, pf planes of
// loads in flight, ry rows per thread (2 only with ty = 8: 256 threads per
// tile instead of 512).
constexpr int kSinkSlots = 65536;

struct StencilGrid {
    int tiles_x, tiles_y, kchunk, nchunks, blocks, ty, pf, ry;
};

// kchunk <= 0 picks a chunk giving ~480 blocks; ty in {4, 8}; pf in 1..4;
// ry in {1, 2} (2 needs ty = 8; otherwise 1 is used).
template <typename R>
StencilGrid stencil_grid(int nx, int ny, int nz, int kchunk, int ty = 4, int pf = 2, int ry = 1);

template <typename R>
hipError_t launch_stencil(int variant, StencilLaunch<R> L, const StencilGrid& g, hipStream_t s);

// One launch over nitems device-resident StencilLaunch items (operators
// without DERIV); item it owns blocks [first[it], first[it+1]) of `blocks`,
// launched with its own tiles_x/tiles_y/kchunk and the common shape (ty, pf,
// ry: two rows per thread with ty = 8).
// epi: 0 (y = alpha H x), 1 (+ beta b), 3, 4 (fused BiCGStab forms).
// active (nullable): items with active[it] == 0 do nothing.
template <typename R>
hipError_t launch_stencil_batch(int variant, int epi, const StencilLaunch<R>* items, const int* first, int nitems,
                                int blocks, const int* active, int ty, int pf, hipStream_t s, int ry = 1);

}  // namespace helm3d
