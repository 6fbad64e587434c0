// Work-list forms of the Krylov and BiCGStab vector kernels: one launch per
// kernel type covers every item of a batch (see batch.hpp).  The arithmetic
// is that of the single-vector kernels (kbody.hpp), so an item computes the
// same values whether it runs alone or in a batch of any composition.
#include <type_traits>

#include "batch.hpp"
#include "kbody.hpp"
#include "krylov.hpp"

namespace helm3d {

namespace {

using kb::KB;

// The item of this block, or nullptr when the block has no work.
template <typename R>
__device__ inline const VecItem<R>* item_of(const VecItem<R>* items, const int* active)
{
    const int it = blockIdx.y;
    if (active && !active[it]) return nullptr;
    const VecItem<R>* I = items + it;
    if (int(blockIdx.x) >= I->nblk) return nullptr;
    return I;
}

template <typename R>
__device__ inline cpx<R>* slot(const VecItem<R>* I, int s) { return I->ws + long(s) * I->ld; }

template <typename R, int K, bool NT>
__global__ __launch_bounds__(KB) void bdots_kernel(const VecItem<R>* __restrict__ items,
                                                   const int* __restrict__ active, int vslot, int wslot)
{
    const VecItem<R>* I = item_of(items, active);
    if (!I) return;
    kb::dots_body<R, K, NT>(slot(I, vslot), I->ld, slot(I, wslot), I->n, I->part, blockIdx.x, I->nblk);
}

template <typename R, int K, bool NT>
__global__ __launch_bounds__(KB) void bupdate_kernel(const VecItem<R>* __restrict__ items,
                                                     const int* __restrict__ active, int vslot, int wslot, int nrm)
{
    const VecItem<R>* I = item_of(items, active);
    if (!I) return;
    const double* d = I->small;
    double ar[K], ai[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const double u2 = d[nrm + c];
        const double inv = u2 > 0.0 ? 1.0 / u2 : 0.0;
        ar[c] = inv * d[2 * c];
        ai[c] = inv * d[2 * c + 1];
    }
    kb::update_body<R, K, NT>(slot(I, vslot), I->ld, ar, ai, slot(I, wslot), I->n, I->part, blockIdx.x, I->nblk);
}

template <int K, bool NT>
__global__ __launch_bounds__(KB) void bcombine_kernel(const VecItem<double>* __restrict__ items,
                                                      const int* __restrict__ active, int vslot,
                                                      const double* __restrict__ coef)
{
    const VecItem<double>* I = item_of(items, active);
    if (!I) return;
    const double* a = coef + long(blockIdx.y) * 2 * kMaxBasis;
    double pr[K], pi[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        pr[c] = a[2 * c];
        pi[c] = a[2 * c + 1];
    }
    kb::combine_body<double, K, NT>(slot(I, vslot), I->ld, pr, pi, I->x, I->n, blockIdx.x, I->nblk);
}

// grid (nq, nitems)
template <typename R>
__global__ __launch_bounds__(KB) void breduce_kernel(const VecItem<R>* __restrict__ items,
                                                     const int* __restrict__ active, int off, int target,
                                                     int from_stencil)
{
    const int it = blockIdx.y;
    if (active && !active[it]) return;
    const VecItem<R>* I = items + it;
    const double s = kb::reduce_body(I->part, from_stencil ? I->sblk : I->nblk, blockIdx.x);
    if (threadIdx.x == 0) (target == BT_SC ? I->sc : I->small)[off + blockIdx.x] = s;
}

template <typename R>
__global__ __launch_bounds__(KB) void bnorm2_kernel(const VecItem<R>* __restrict__ items,
                                                    const int* __restrict__ active, int s)
{
    const VecItem<R>* I = item_of(items, active);
    if (!I) return;
    kb::norm2_body<R>(slot(I, s), I->n, I->part, blockIdx.x, I->nblk);
}

__global__ __launch_bounds__(KB) void bstart_kernel(const VecItem<double>* __restrict__ items,
                                                    const int* __restrict__ active, int s)
{
    const VecItem<double>* I = item_of(items, active);
    if (!I) return;
    cpx<double>* r = slot(I, s);
    for (long t = long(blockIdx.x) * KB + threadIdx.x; t < I->n; t += long(I->nblk) * KB) {
        r[t] = I->b[t];
        I->x[t] = cpx<double>{0.0, 0.0};
    }
}

__global__ __launch_bounds__(KB) void bdown_kernel(const VecItem<double>* __restrict__ i64,
                                                   const VecItem<float>* __restrict__ i32,
                                                   const int* __restrict__ active, int nrm, int src, int dst)
{
    const VecItem<double>* I = item_of(i64, active);
    if (!I) return;
    const VecItem<float>* J = i32 + blockIdx.y;
    const double u2 = I->small[nrm];
    const double a = u2 > 0.0 ? 1.0 / sqrt(u2) : 0.0;
    const cpx<double>* x = slot(I, src);
    cpx<float>* y = slot(J, dst);
    for (long t = long(blockIdx.x) * KB + threadIdx.x; t < I->n; t += long(I->nblk) * KB) {
        const cpx<double> v = x[t];
        y[t] = cpx<float>{float(a * v.re), float(a * v.im)};
    }
}

__global__ __launch_bounds__(KB) void bup_kernel(const VecItem<float>* __restrict__ i32,
                                                 const VecItem<double>* __restrict__ i64,
                                                 const int* __restrict__ active, int src, int dst)
{
    const VecItem<float>* J = item_of(i32, active);
    if (!J) return;
    const VecItem<double>* I = i64 + blockIdx.y;
    const cpx<float>* x = slot(J, src);
    cpx<double>* y = slot(I, dst);
    for (long t = long(blockIdx.x) * KB + threadIdx.x; t < J->n; t += long(J->nblk) * KB) {
        const cpx<float> v = x[t];
        y[t] = cpx<double>{double(v.re), double(v.im)};
    }
}

template <typename R>
__global__ __launch_bounds__(KB) void bbicg_init_kernel(const VecItem<R>* __restrict__ items,
                                                        const int* __restrict__ active, int bs, int ps, int vs,
                                                        int ts, int zs)
{
    const VecItem<R>* I = item_of(items, active);
    if (!I) return;
    __shared__ double red[KB / 64];
    double s2[1] = {0.0};
    const cpx<R>* b = slot(I, bs);
    cpx<R>* p = slot(I, ps);
    cpx<R>* v = slot(I, vs);
    cpx<R>* tt = slot(I, ts);
    cpx<R>* z = slot(I, zs);
    long lo, hi;
    kb::range_plain(I->n, blockIdx.x, I->nblk, lo, hi);
    const cpx<R> zero{R(0), R(0)};
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        const cpx<R> bv = b[t];
        p[t] = zero;
        v[t] = zero;
        tt[t] = zero;
        z[t] = zero;
        s2[0] += double(bv.re) * double(bv.re) + double(bv.im) * double(bv.im);
    }
    kb::block_sum<1>(s2, red);
    if (threadIdx.x == 0) {
        const long G = I->nblk;
        I->part[blockIdx.x] = s2[0];
        I->part[G + blockIdx.x] = 0.0;
        I->part[2 * G + blockIdx.x] = s2[0];
    }
}

template <typename R>
__global__ __launch_bounds__(KB) void bbicg_fin_kernel(const VecItem<R>* __restrict__ items,
                                                       const int* __restrict__ active, int ps, int rs, int vs,
                                                       int zs)
{
    const VecItem<R>* I = item_of(items, active);
    if (!I) return;
    const double* sc = I->sc;
    const double ar = sc[BS_ALPHA], ai = sc[BS_ALPHA + 1];
    const double tn = sc[BS_TT];
    const double wr = tn > 0.0 ? sc[BS_TS] / tn : 0.0, wi = tn > 0.0 ? sc[BS_TS + 1] / tn : 0.0;
    const cpx<R>* p = slot(I, ps);
    const cpx<R>* r = slot(I, rs);
    const cpx<R>* v = slot(I, vs);
    cpx<R>* z = slot(I, zs);
    long lo, hi;
    kb::range_plain(I->n, blockIdx.x, I->nblk, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        const cpx<R> pv = p[t], rv = r[t], vv = v[t], zv = z[t];
        // s = r - alpha v
        const double sr = double(rv.re) - (ar * double(vv.re) - ai * double(vv.im));
        const double si = double(rv.im) - (ar * double(vv.im) + ai * double(vv.re));
        const double zr = double(zv.re) + ar * double(pv.re) - ai * double(pv.im) + wr * sr - wi * si;
        const double zi = double(zv.im) + ar * double(pv.im) + ai * double(pv.re) + wr * si + wi * sr;
        z[t] = cpx<R>{R(zr), R(zi)};
    }
}

template <int K, typename Fn>
hipError_t dispatch_k(int k, Fn&& fn)
{
    if constexpr (K >= kMaxBasis) {
        return hipErrorInvalidValue;
    } else {
        if (k == K) return fn(std::integral_constant<int, K>{});
        return dispatch_k<K + 1>(k, fn);
    }
}

}  // namespace

template <typename R>
hipError_t launch_bdots(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int k, int vslot,
                        int wslot, hipStream_t s, bool nt)
{
    if (k < 1) return hipErrorInvalidValue;
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        if (nt)
            hipLaunchKernelGGL((bdots_kernel<R, K, true>), dim3(nblk_max, nitems), dim3(KB), 0, s, items, active,
                               vslot, wslot);
        else
            hipLaunchKernelGGL((bdots_kernel<R, K, false>), dim3(nblk_max, nitems), dim3(KB), 0, s, items, active,
                               vslot, wslot);
        return hipGetLastError();
    });
}

template <typename R>
hipError_t launch_bupdate_nrm(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int k,
                              int vslot, int wslot, int nrm, hipStream_t s, bool nt)
{
    if (k < 1) return hipErrorInvalidValue;
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        if (nt)
            hipLaunchKernelGGL((bupdate_kernel<R, K, true>), dim3(nblk_max, nitems), dim3(KB), 0, s, items, active,
                               vslot, wslot, nrm);
        else
            hipLaunchKernelGGL((bupdate_kernel<R, K, false>), dim3(nblk_max, nitems), dim3(KB), 0, s, items, active,
                               vslot, wslot, nrm);
        return hipGetLastError();
    });
}

hipError_t launch_bcombine(const VecItem<double>* items, const int* active, int nitems, int nblk_max, int k,
                           int vslot, const double* coef, hipStream_t s, bool nt)
{
    if (k <= 0) return hipSuccess;
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        if (nt)
            hipLaunchKernelGGL((bcombine_kernel<K, true>), dim3(nblk_max, nitems), dim3(KB), 0, s, items, active,
                               vslot, coef);
        else
            hipLaunchKernelGGL((bcombine_kernel<K, false>), dim3(nblk_max, nitems), dim3(KB), 0, s, items, active,
                               vslot, coef);
        return hipGetLastError();
    });
}

template <typename R>
hipError_t launch_breduce(const VecItem<R>* items, const int* active, int nitems, int nq, int off, int target,
                          bool from_stencil, hipStream_t s)
{
    hipLaunchKernelGGL((breduce_kernel<R>), dim3(nq, nitems), dim3(KB), 0, s, items, active, off, target,
                       int(from_stencil));
    return hipGetLastError();
}

template <typename R>
hipError_t launch_bnorm2(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int slot_,
                         hipStream_t s)
{
    hipLaunchKernelGGL((bnorm2_kernel<R>), dim3(nblk_max, nitems), dim3(KB), 0, s, items, active, slot_);
    return hipGetLastError();
}

hipError_t launch_bstart(const VecItem<double>* items, const int* active, int nitems, int nblk_max, int slot_,
                         hipStream_t s)
{
    hipLaunchKernelGGL(bstart_kernel, dim3(nblk_max, nitems), dim3(KB), 0, s, items, active, slot_);
    return hipGetLastError();
}

hipError_t launch_bdown_norm(const VecItem<double>* i64, const VecItem<float>* i32, const int* active, int nitems,
                             int nblk_max, int nrm, int src, int dst, hipStream_t s)
{
    hipLaunchKernelGGL(bdown_kernel, dim3(nblk_max, nitems), dim3(KB), 0, s, i64, i32, active, nrm, src, dst);
    return hipGetLastError();
}

hipError_t launch_bup(const VecItem<float>* i32, const VecItem<double>* i64, const int* active, int nitems,
                      int nblk_max, int src, int dst, hipStream_t s)
{
    hipLaunchKernelGGL(bup_kernel, dim3(nblk_max, nitems), dim3(KB), 0, s, i32, i64, active, src, dst);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_bbicg_init(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int bslot,
                             int p, int v, int t, int z, hipStream_t s)
{
    hipLaunchKernelGGL((bbicg_init_kernel<R>), dim3(nblk_max, nitems), dim3(KB), 0, s, items, active, bslot, p, v,
                       t, z);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_bbicg_fin(const VecItem<R>* items, const int* active, int nitems, int nblk_max, int p, int r,
                            int v, int z, hipStream_t s)
{
    hipLaunchKernelGGL((bbicg_fin_kernel<R>), dim3(nblk_max, nitems), dim3(KB), 0, s, items, active, p, r, v, z);
    return hipGetLastError();
}

#define HELM3D_BATCH_INST(R)                                                                                     \
    template hipError_t launch_bdots<R>(const VecItem<R>*, const int*, int, int, int, int, int, hipStream_t,     \
                                        bool);                                                                   \
    template hipError_t launch_bupdate_nrm<R>(const VecItem<R>*, const int*, int, int, int, int, int, int,       \
                                              hipStream_t, bool);                                                \
    template hipError_t launch_breduce<R>(const VecItem<R>*, const int*, int, int, int, int, bool, hipStream_t); \
    template hipError_t launch_bnorm2<R>(const VecItem<R>*, const int*, int, int, int, hipStream_t);             \
    template hipError_t launch_bbicg_init<R>(const VecItem<R>*, const int*, int, int, int, int, int, int, int,   \
                                             hipStream_t);                                                       \
    template hipError_t launch_bbicg_fin<R>(const VecItem<R>*, const int*, int, int, int, int, int, int,         \
                                            hipStream_t);
HELM3D_BATCH_INST(double)
HELM3D_BATCH_INST(float)

}  // namespace helm3d
