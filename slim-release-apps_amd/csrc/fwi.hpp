// FWI set-up and gradient support kernels: launch interface (internal).
#pragma once

#include "helm3d_common.hpp"

namespace helm3d {

// Grid of the PML-extended model: physical sizes and lower PML widths.
struct ExtGrid {
    int np[3];  // physical points per axis
    int lo[3];  // PML points below the physical box per axis
    int nt[3];  // extended points per axis (np + lo + hi)
};

// wn = c * f(v[clamp(p)]) for every point p of the extended grid (edge
// replication), with f chosen by unit; dwn / ddwn (nullable) likewise.
hipError_t launch_param2wavenum(const ExtGrid& G, const double* v, int unit, double c, double* wn, double* dwn,
                                double* ddwn, hipStream_t s);

// g[phys] (+)= sum of gpml over the points that edge replication maps onto
// phys (the adjoint of the extension).
hipError_t launch_extension_adjoint(const ExtGrid& G, const double* gpml, double* g, bool accumulate,
                                    hipStream_t s);

// Row-compressed real-weight operator on complex128 vectors:
//   out[oidx ? oidx[r] : r] (= or +=) alpha * sum_{e in [rowptr[r], rowptr[r+1])} w[e] * in[col[e]]
hipError_t launch_csr_apply(int nrows, const long long* rowptr, const long long* col, const double* w,
                            const long long* oidx, double alpha, const cpx<double>* in, cpx<double>* out,
                            bool accumulate, hipStream_t s);

}  // namespace helm3d
