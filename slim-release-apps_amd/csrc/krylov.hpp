// Krylov vector kernels: launch interface (internal to the library).
#pragma once

#include "helm3d_common.hpp"

namespace helm3d {

constexpr int kMaxBasis = 32;      // largest restart length (basis slots per cycle)
constexpr int kMaxPartials = 2048; // blocks of the streaming reductions

// Up to kMaxBasis complex coefficients passed by value to a kernel.
struct CoefPack {
    double v[2 * kMaxBasis];
};

int krylov_blocks(long n);

// Blocks of the Gram-Schmidt passes over complex128 vectors long enough for
// the separate dots reduction (krylov_blocks(n) > kInBlockDots; shorter ones
// get krylov_blocks(n)).  The dots pass reduces 2K partial sums per block and
// streams fastest with about 32 points per thread (at least min(krylov_blocks,
// 512) blocks, at most kMaxPartials); the update pass, one partial per block,
// with about 2 (at most kMaxUpdateBlocks).  Measured on 8,064,000 points:
// dots 6.0 -> 6.4 TB/s at K = 20, update 5.2 -> 5.4 TB/s.
constexpr int kMaxUpdateBlocks = 16384;
static_assert(kMaxUpdateBlocks <= kMaxPartials * 2 * kMaxBasis, "update partials fit the dots partials buffer");
int dots_blocks(long n);
int update_blocks(long n);

// Device-side state of a result-free restarted GMRES (the multigrid
// smoothers and coarsest solves): the host enqueues every cycle, these values
// decide on the device which of them run.  skip, when set, turns the gated
// kernels of a cycle into no-ops.
struct GmresDevState {
    double normr0, beta, hlast;
    int stop, skip, kept;
    double coef[2 * kMaxBasis];  // combine coefficients of the last cycle
};

// skip (optional, device): the kernel does nothing while *skip != 0.
template <typename R>
hipError_t launch_dots(const cpx<R>* V, long ldv, int k, const cpx<R>* w, long n, double* part, int nblk,
                       hipStream_t s, const int* skip = nullptr);
// a_c = scl.v[c] * (d[2c] + i d[2c+1]); d is device memory.  With nrm (device
// memory, ||v_c||^2 per slot) the scale is (1 / sqrt(nrm[c]))^2 instead of
// scl.v[c], so the pass needs nothing from the host.
// dpart (optional, short columns): the dots partials (nblk per column, 2k
// columns); each block sums them itself instead of reading d, and block 0
// writes the 2k sums to hrow.
template <typename R>
hipError_t launch_update(const cpx<R>* V, long ldv, int k, const double* d, const CoefPack& scl, cpx<R>* w,
                         long n, double* part, int nblk, hipStream_t s, const double* nrm = nullptr,
                         const double* dpart = nullptr, double* hrow = nullptr, const int* skip = nullptr);
constexpr int kInBlockDots = 256;  // dots columns of at most this many partials are summed in the update pass
template <typename R>
hipError_t launch_combine(const cpx<R>* V, long ldv, int k, const CoefPack& a, cpx<R>* x, long n, int nblk,
                          hipStream_t s);
// x += sum_{c < min(k, *kept)} a_c v_c with a (2k doubles) and kept in
// device memory.
template <typename R>
hipError_t launch_combine_dev(const cpx<R>* V, long ldv, int k, const double* a, const int* kept, cpx<R>* x,
                              long n, int nblk, hipStream_t s, const int* skip = nullptr);
// Device-driven result-free GMRES (see GmresDevState): start of the solve,
// gate and restart test of a later cycle, least squares of a cycle of m
// steps whose rows (stride ldr) hold the dots and ||w_j||^2 (one thread, its
// work arrays in LDS: no private scratch).
hipError_t launch_gmres_start(const double* nrm, GmresDevState* st, hipStream_t s);
hipError_t launch_gmres_gate(GmresDevState* st, hipStream_t s);
hipError_t launch_gmres_restart(const double* nrm, double tol, GmresDevState* st, hipStream_t s);
hipError_t launch_gmres_lsq(const double* rows, int ldr, int m, double tol, GmresDevState* st, hipStream_t s);
hipError_t launch_interleave(const double* re, const double* im, cpx<double>* z, long n, hipStream_t s);
hipError_t launch_deinterleave(const cpx<double>* z, double* re, double* im, long n, hipStream_t s);
hipError_t launch_cast_real(const double* a, float* b, long n, hipStream_t s);
// out[q] = sum_b part[q*nblk + b] for q < nq (block order, deterministic)
// out[q] = sum of partial column q; out2 (optional) receives a copy.
hipError_t launch_reduce(const double* part, int nblk, int nq, double* out, hipStream_t s, double* out2 = nullptr,
                         const int* skip = nullptr);
template <typename RX, typename RY>
hipError_t launch_axpby(double ar, double ai, const cpx<RX>* x, double br, double bi, cpx<RY>* y, long n,
                        bool use_y, hipStream_t s);
template <typename R>
hipError_t launch_norm2(const cpx<R>* x, long n, double* part, int nblk, hipStream_t s);
// w = b, x = 0 and the partials of ||b||^2 (those of launch_norm2 on w) in one pass.
template <typename R>
hipError_t launch_start(const cpx<R>* b, cpx<R>* w, cpx<R>* x, long n, double* part, int nblk, hipStream_t s);

}  // namespace helm3d
