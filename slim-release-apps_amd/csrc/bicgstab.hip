// BiCGStab vector kernels (storage precision R, FP64 accumulation).
//
// The iteration of tools/solvers/Krylov/BICGSTAB.m:37-74 (van der Vorst's
// BiCGStab, identity preconditioner) with every scalar kept in device memory:
//   v = A p ; a = rho / (rhat, v) ; s = r - a v ; t = A s ;
//   w = (t, s) / (t, t) ; z += a p + w s ; r = s - w t ;
//   rho' = (rhat, r) ; beta = (rho' / rho)(a / w) ; p = r + beta (p - w v)
// The consumer kernels form a, w, beta themselves from the reduced dots, so a
// whole iteration is launched without the host reading any value.  beta uses
// the textbook ratio rho'/rho; BICGSTAB.m:59 divides (rnew, r0_bar) by itself,
// which makes beta = a/w -- that quirk is not reproduced.
//
// Scalar slots (doubles, device): see the BS_* offsets in bicgstab.hpp.
#include "bicgstab.hpp"
#include "kbody.hpp"
#include "krylov.hpp"

namespace helm3d {

namespace {

using kb::KB;
using kb::C;
using kb::cdiv;
using kb::cmulc;
using kb::ld;

// Partial sums of conj(a_q) . b_q for NP pairs; optionally ||a_last||^2.
template <typename R, int NP, bool NORM>
__global__ __launch_bounds__(KB) void dots_pairs_kernel(PairArgs<R> P, long n, double* __restrict__ part)
{
    kb::dots_pairs_body<R, NP, NORM>(P, n, part, blockIdx.x, gridDim.x);
}

// s = r - a v with a = rho / (rhat, v); block 0 records a.
template <typename R>
__global__ __launch_bounds__(KB) void bicg_s_kernel(const cpx<R>* __restrict__ r, const cpx<R>* __restrict__ v,
                                                    cpx<R>* __restrict__ s, long n, double* __restrict__ sc,
                                                    int rho_slot)
{
    const C a = cdiv(ld(sc, rho_slot), ld(sc, BS_RV));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc[BS_ALPHA] = a.re;
        sc[BS_ALPHA + 1] = a.im;
    }
    for (long t = long(blockIdx.x) * KB + threadIdx.x; t < n; t += long(gridDim.x) * KB) {
        const cpx<R> rv = r[t], vv = v[t];
        const double re = double(rv.re) - (a.re * double(vv.re) - a.im * double(vv.im));
        const double im = double(rv.im) - (a.re * double(vv.im) + a.im * double(vv.re));
        s[t] = cpx<R>{R(re), R(im)};
    }
}

// w = (t, s) / (t, t); z += a p + w s; r = s - w t; partials of (rhat, r) and
// ||r||^2.  Block 0 records w.
template <typename R>
__global__ __launch_bounds__(KB) void bicg_xr_kernel(const cpx<R>* __restrict__ p, const cpx<R>* __restrict__ s,
                                                     const cpx<R>* __restrict__ tt, const cpx<R>* __restrict__ rhat,
                                                     cpx<R>* __restrict__ z, cpx<R>* __restrict__ r, long n,
                                                     double* __restrict__ sc, double* __restrict__ part)
{
    kb::bicg_xr_body<R>(p, s, tt, rhat, z, r, n, sc, part, blockIdx.x, gridDim.x);
}

// p = r + beta (p - w v), beta = (rho_new / rho_old)(a / w).
template <typename R>
__global__ __launch_bounds__(KB) void bicg_p_kernel(const cpx<R>* __restrict__ r, const cpx<R>* __restrict__ v,
                                                    cpx<R>* __restrict__ p, long n, const double* __restrict__ sc,
                                                    int rho_new, int rho_old)
{
    const C a = ld(sc, BS_ALPHA), w = ld(sc, BS_OMEGA);
    const C b = cmulc(cdiv(ld(sc, rho_new), ld(sc, rho_old)), cdiv(a, w));
    for (long t = long(blockIdx.x) * KB + threadIdx.x; t < n; t += long(gridDim.x) * KB) {
        const cpx<R> rv = r[t], vv = v[t], pv = p[t];
        const double qr = double(pv.re) - (w.re * double(vv.re) - w.im * double(vv.im));
        const double qi = double(pv.im) - (w.re * double(vv.im) + w.im * double(vv.re));
        const double pr = double(rv.re) + b.re * qr - b.im * qi;
        const double pi = double(rv.im) + b.re * qi + b.im * qr;
        p[t] = cpx<R>{R(pr), R(pi)};
    }
}

}  // namespace

template <typename R>
hipError_t launch_dots_pairs(const PairArgs<R>& P, int np, bool norm, long n, double* part, int nblk,
                             hipStream_t s)
{
    if (np == 1 && !norm) hipLaunchKernelGGL((dots_pairs_kernel<R, 1, false>), dim3(nblk), dim3(KB), 0, s, P, n, part);
    else if (np == 1 && norm) hipLaunchKernelGGL((dots_pairs_kernel<R, 1, true>), dim3(nblk), dim3(KB), 0, s, P, n, part);
    else if (np == 2 && !norm) hipLaunchKernelGGL((dots_pairs_kernel<R, 2, false>), dim3(nblk), dim3(KB), 0, s, P, n, part);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename R>
hipError_t launch_bicg_s(const cpx<R>* r, const cpx<R>* v, cpx<R>* s, long n, double* sc, int rho_slot,
                         hipStream_t st)
{
    hipLaunchKernelGGL((bicg_s_kernel<R>), dim3(krylov_blocks(n)), dim3(KB), 0, st, r, v, s, n, sc, rho_slot);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_bicg_xr(const cpx<R>* p, const cpx<R>* s, const cpx<R>* t, const cpx<R>* rhat, cpx<R>* z,
                          cpx<R>* r, long n, double* sc, double* part, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL((bicg_xr_kernel<R>), dim3(nblk), dim3(KB), 0, st, p, s, t, rhat, z, r, n, sc, part);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_bicg_p(const cpx<R>* r, const cpx<R>* v, cpx<R>* p, long n, const double* sc, int rho_new,
                         int rho_old, hipStream_t st)
{
    hipLaunchKernelGGL((bicg_p_kernel<R>), dim3(krylov_blocks(n)), dim3(KB), 0, st, r, v, p, n, sc, rho_new, rho_old);
    return hipGetLastError();
}

#define HELM3D_BICG_INST(R)                                                                                      \
    template hipError_t launch_dots_pairs<R>(const PairArgs<R>&, int, bool, long, double*, int, hipStream_t);     \
    template hipError_t launch_bicg_s<R>(const cpx<R>*, const cpx<R>*, cpx<R>*, long, double*, int, hipStream_t); \
    template hipError_t launch_bicg_xr<R>(const cpx<R>*, const cpx<R>*, const cpx<R>*, const cpx<R>*, cpx<R>*,    \
                                          cpx<R>*, long, double*, double*, int, hipStream_t);                    \
    template hipError_t launch_bicg_p<R>(const cpx<R>*, const cpx<R>*, cpx<R>*, long, const double*, int, int,   \
                                         hipStream_t);
HELM3D_BICG_INST(double)
HELM3D_BICG_INST(float)

}  // namespace helm3d
