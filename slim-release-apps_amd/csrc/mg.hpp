// Multigrid grid transfers: launch interface (internal).
#pragma once

#include <algorithm>

#include "helm3d_common.hpp"

namespace helm3d {

// One axis of the trilinear transfer between a fine grid of nf points and a
// coarse grid of nc points (device arrays).
struct MgAxis {
    const int* I;      // [nf] coarse bracket of each fine index
    const double* wa;  // [nf] weight of I
    const double* wb;  // [nf] weight of I+1 (0 when I = nc-1)
    const int* rp;     // [nc+1] row pointers of the transposed lists
    const int* fi;     // fine indices using each coarse index
    const double* fw;  // their weights
    // the transposed lists padded to pw entries per coarse index (weight 0,
    // a valid fine index): [nc * pw]
    const int* pfi;
    const double* pfw;
    int pw;
};

struct MgAxes {
    int nf[3], nc[3];
    MgAxis ax[3];
};

// xf += P xc (R = double or float)
template <typename R>
hipError_t launch_prolong_add(const MgAxes& T, const cpx<R>* xc, cpx<R>* xf, hipStream_t s);
// yc = scale * P^T rf
template <typename R>
hipError_t launch_restrict(const MgAxes& T, double scale, const cpx<R>* rf, cpx<R>* yc, hipStream_t s);

}  // namespace helm3d
