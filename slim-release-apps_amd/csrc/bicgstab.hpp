// BiCGStab kernels: launch interface and device scalar slots (internal).
#pragma once

#include "helm3d_common.hpp"

namespace helm3d {

// Device scalar slots (doubles).  rho and ||r||^2 alternate between two
// triples by iteration parity so a kernel never reads a slot the same
// launch rewrites.
constexpr int BS_RHO0 = 0;   // (rhat, r) re, im ; BS_RHO0 + 2: ||r||^2
constexpr int BS_RHO1 = 3;   // second triple
constexpr int BS_RV = 6;     // (rhat, v)
constexpr int BS_TS = 8;     // (t, s) re, im ; BS_TS + 2: ||t||^2
constexpr int BS_TT = 10;
constexpr int BS_RS = 11;    // (rhat, s) re, im (fused batched form)
constexpr int BS_RT = 13;    // (rhat, t) re, im (fused batched form)
constexpr int BS_ALPHA = 15;
constexpr int BS_OMEGA = 17;
constexpr int BS_SLOTS = 20;

// Up to two (a, b) vector pairs for the dot kernel.
template <typename R>
struct PairArgs {
    const cpx<R>* a[2];
    const cpx<R>* b[2];
};

template <typename R>
hipError_t launch_dots_pairs(const PairArgs<R>& P, int np, bool norm, long n, double* part, int nblk,
                             hipStream_t s);
template <typename R>
hipError_t launch_bicg_s(const cpx<R>* r, const cpx<R>* v, cpx<R>* s, long n, double* sc, int rho_slot,
                         hipStream_t st);
template <typename R>
hipError_t launch_bicg_xr(const cpx<R>* p, const cpx<R>* s, const cpx<R>* t, const cpx<R>* rhat, cpx<R>* z,
                          cpx<R>* r, long n, double* sc, double* part, int nblk, hipStream_t st);
template <typename R>
hipError_t launch_bicg_p(const cpx<R>* r, const cpx<R>* v, cpx<R>* p, long n, const double* sc, int rho_new,
                         int rho_old, hipStream_t st);

}  // namespace helm3d
