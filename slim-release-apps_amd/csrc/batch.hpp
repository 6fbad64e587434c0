// Work-list forms of the Krylov and BiCGStab vector kernels (internal).
//
// A launch covers every item of a batch: blockIdx.y is the item, blockIdx.x
// a block of that item's streaming range (items have their own lengths and
// block counts; surplus blocks return at once).  Vectors of an item live in
// its workspace as slots of `ld` elements; the kernels address them by slot
// index, uniform over the batch.  Items with active[it] == 0 are skipped.
#pragma once

#include "helm3d_common.hpp"

namespace helm3d {

template <typename R>
struct VecItem {
    cpx<R>* ws;            // workspace: slot s at ws + s*ld
    long ld;
    long n;
    int nblk;              // blocks of this item's streaming kernels
    int sblk;              // blocks of this item's stencil launches
    double* part;          // partial sums (columns of nblk, or of the stencil's block count)
    double* small;         // reduced values
    double* sc;            // BiCGStab scalar slots
    const cpx<double>* b;  // right-hand side (complex128)
    cpx<double>* x;        // solution (complex128)
};

constexpr int kBatchMaxItems = 1024;

// reduction target: small (values for the host) or sc (BiCGStab slots)
enum BatchTarget : int { BT_SMALL = 0, BT_SC = 1 };

template <typename R>
hipError_t launch_bdots(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int k, int vslot,
                        int wslot, hipStream_t s, bool nt = false);
// w -= sum_c a_c u_c with a_c = d_c / small[nrm + c] (d_c = u_c^H w in
// small[2c..2c+1], ||u_c||^2 in small[nrm + c]: the projection on the raw
// basis vectors); partial ||w||^2
template <typename R>
hipError_t launch_bupdate_nrm(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int k,
                              int vslot, int wslot, int nrm, hipStream_t s, bool nt = false);
// x += sum_c a_c v_c, a_c = coef[it*2*kMaxBasis + 2c .. +1]  (R = double: x is the item's solution)
// (nt on these three: streaming basis loads, for bases larger than the
// last-level cache)
hipError_t launch_bcombine(const VecItem<double>* items, const int* active, int nitems, int nblk_max, int k,
                           int vslot, const double* coef, hipStream_t s, bool nt = false);
// small/sc[off + q] = sum_b part[q*nb + b]; nb = sblk for partials written by
// a stencil launch, else nblk
template <typename R>
hipError_t launch_breduce(const VecItem<R>* items, const int* active, int nitems, int nq, int off, int target,
                          bool from_stencil, hipStream_t s);
template <typename R>
hipError_t launch_bnorm2(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int slot,
                         hipStream_t s);
// slot := b (complex128 items), x := 0
hipError_t launch_bstart(const VecItem<double>* items, const int* active, int nitems, int nblk_max, int slot,
                         hipStream_t s);
// ws32[dst] = complex64(ws64[src] / sqrt(small[nrm]))  (the normalised basis vector)
hipError_t launch_bdown_norm(const VecItem<double>* i64, const VecItem<float>* i32, const int* active, int nitems,
                             int nblk_max, int nrm, int src, int dst, hipStream_t s);
// ws64[dst] = complex128(ws32[src])
hipError_t launch_bup(const VecItem<float>* i32, const VecItem<double>* i64, const int* active, int nitems,
                      int nblk_max, int src, int dst, hipStream_t s);
// BiCGStab start from z = 0 with r = rhat = b (the b slot serves as both):
// p = v = t = z = 0 (slots given), partials of ||b||^2 in columns 0 and 2
// (column 1 zero): rho_0 and ||r_0||^2
template <typename R>
hipError_t launch_bbicg_init(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int bslot,
                             int p, int v, int t, int z, hipStream_t s);
// Last z update of the fused iteration: z += alpha p + omega (r - alpha v)
// with alpha and omega = (t, s) / ||t||^2 from the item's scalar slots
template <typename R>
hipError_t launch_bbicg_fin(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int p, int r,
                            int v, int z, hipStream_t s);

}  // namespace helm3d
