// Fused Krylov vector kernels for the restarted GMRES of FGMRES.m on gfx950.
//
// The basis is kept unnormalised: slot c holds a raw vector v_c and the host
// holds its scale s_c = 1/||v_c||, so no pass is spent normalising
// (FGMRES.m:104-106 becomes a scalar).  Per Arnoldi step k the passes are
//   dots    d_c = v_c^H w, c < k            (reads w and k slots)
//   update  w <- w - sum_c a_c v_c, ||w||^2  (reads w and k slots, writes w)
// i.e. the classical Gram-Schmidt of FGMRES.m:100-103 in two HBM passes.
//
// Reductions are deterministic: every block reduces a fixed index range in a
// fixed order into a partial slot, and a single block sums the partials in
// block order; results do not depend on scheduling.  Partial sums are FP64 for
// both storage precisions.
#include "helm27.hpp"
#include "krylov.hpp"

namespace helm3d {

namespace {

constexpr int KB = 256;  // threads per block

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of NV values per thread; result valid in thread 0 slots.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* red /* [4][NV] */)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) red[wid * NV + q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = red[q];
            for (int w = 1; w < KB / 64; ++w) s += red[w * NV + q];
            v[q] = s;
        }
    }
}

// Contiguous range of the block: [lo, hi)
__device__ inline void block_range(long n, long& lo, long& hi)
{
    const long per = (n + gridDim.x - 1) / gridDim.x;
    lo = long(blockIdx.x) * per;
    hi = min(n, lo + per);
}

// d_c = sum conj(v_c) w for c < k (K compile-time upper bound); partials are
// laid out with the fixed stride 2*kMaxBasis per block.
template <typename R, int K>
__global__ __launch_bounds__(KB) void dots_kernel(const cpx<R>* __restrict__ V, long ldv, int k,
                                                  const cpx<R>* __restrict__ w, long n,
                                                  double* __restrict__ part)
{
    __shared__ double red[(KB / 64) * 2 * K];
    double acc[2 * K];
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) acc[q] = 0.0;
    long lo, hi;
    block_range(n, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        const cpx<R> wt = w[t];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            if (c < k) {
                const cpx<R> v = V[c * ldv + t];
                acc[2 * c] += double(v.re) * double(wt.re) + double(v.im) * double(wt.im);
                acc[2 * c + 1] += double(v.re) * double(wt.im) - double(v.im) * double(wt.re);
            }
        }
    }
    block_sum<2 * K>(acc, red);
    if (threadIdx.x == 0) {
        for (int q = 0; q < 2 * k; ++q) part[long(blockIdx.x) * 2 * kMaxBasis + q] = acc[q];
    }
}

// w <- w - sum_{c<k} a_c v_c with a_c = scl_c * d_c, d read from device memory
// (the reduced dot products), so the host need not see d before this pass;
// partial ||w||^2.
template <typename R, int K>
__global__ __launch_bounds__(KB) void update_kernel(const cpx<R>* __restrict__ V, long ldv, int k,
                                                    const double* __restrict__ d, CoefPack scl,
                                                    cpx<R>* __restrict__ w, long n, double* __restrict__ part)
{
    __shared__ double red[KB / 64];
    double s2[1] = {0.0};
    double a[2 * K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        a[2 * c] = c < k ? scl.v[c] * d[2 * c] : 0.0;
        a[2 * c + 1] = c < k ? scl.v[c] * d[2 * c + 1] : 0.0;
    }
    long lo, hi;
    block_range(n, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        cpx<R> wt = w[t];
        double re = wt.re, im = wt.im;
#pragma unroll
        for (int c = 0; c < K; ++c) {
            if (c < k) {
                const cpx<R> v = V[c * ldv + t];
                const double ar = a[2 * c], ai = a[2 * c + 1];
                re -= ar * double(v.re) - ai * double(v.im);
                im -= ar * double(v.im) + ai * double(v.re);
            }
        }
        wt.re = R(re);
        wt.im = R(im);
        w[t] = wt;
        s2[0] += double(wt.re) * double(wt.re) + double(wt.im) * double(wt.im);
    }
    block_sum<1>(s2, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s2[0];
}

// x <- x + sum_{c<k} a_c v_c (the solution update of FGMRES.m:115-119).
template <typename R, int K>
__global__ __launch_bounds__(KB) void combine_kernel(const cpx<R>* __restrict__ V, long ldv, int k,
                                                     CoefPack a, cpx<R>* __restrict__ x, long n)
{
    long lo, hi;
    block_range(n, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        cpx<R> xt = x[t];
        double re = xt.re, im = xt.im;
#pragma unroll
        for (int c = 0; c < K; ++c) {
            if (c < k) {
                const cpx<R> v = V[c * ldv + t];
                const double ar = a.v[2 * c], ai = a.v[2 * c + 1];
                re += ar * double(v.re) - ai * double(v.im);
                im += ar * double(v.im) + ai * double(v.re);
            }
        }
        xt.re = R(re);
        xt.im = R(im);
        x[t] = xt;
    }
}

// Sum part[b*stride + q] over b in block order for q < nq; out[q] = sum.
__global__ __launch_bounds__(KB) void reduce_kernel(const double* __restrict__ part, int nblk, int stride,
                                                    int nq, double* __restrict__ out)
{
    __shared__ double red[KB];
    for (int q = 0; q < nq; ++q) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += KB) s += part[long(b) * stride + q];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int o = KB / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = red[0];
        __syncthreads();
    }
}

// y = alpha x + beta y (complex alpha, beta); beta == 0 ignores y's contents.
template <typename RX, typename RY>
__global__ __launch_bounds__(KB) void axpby_kernel(double ar, double ai, const cpx<RX>* __restrict__ x,
                                                   double br, double bi, cpx<RY>* __restrict__ y,
                                                   long n, int use_y)
{
    long lo, hi;
    block_range(n, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        const cpx<RX> xv = x ? x[t] : cpx<RX>{RX(0), RX(0)};
        double re = ar * double(xv.re) - ai * double(xv.im);
        double im = ar * double(xv.im) + ai * double(xv.re);
        if (use_y) {
            const cpx<RY> yv = y[t];
            re += br * double(yv.re) - bi * double(yv.im);
            im += br * double(yv.im) + bi * double(yv.re);
        }
        y[t] = cpx<RY>{RY(re), RY(im)};
    }
}

// partial ||x||^2
template <typename R>
__global__ __launch_bounds__(KB) void norm2_kernel(const cpx<R>* __restrict__ x, long n,
                                                   double* __restrict__ part)
{
    __shared__ double red[KB / 64];
    double s2[1] = {0.0};
    long lo, hi;
    block_range(n, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        const cpx<R> v = x[t];
        s2[0] += double(v.re) * double(v.re) + double(v.im) * double(v.im);
    }
    block_sum<1>(s2, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s2[0];
}

// Split real/imaginary arrays (the MEX boundary layout) <-> interleaved complex.
__global__ __launch_bounds__(KB) void interleave_kernel(const double* __restrict__ re, const double* __restrict__ im,
                                                        cpx<double>* __restrict__ z, long n)
{
    for (long t = long(blockIdx.x) * KB + threadIdx.x; t < n; t += long(gridDim.x) * KB)
        z[t] = cpx<double>{re[t], im ? im[t] : 0.0};
}
__global__ __launch_bounds__(KB) void deinterleave_kernel(const cpx<double>* __restrict__ z, double* __restrict__ re,
                                                          double* __restrict__ im, long n)
{
    for (long t = long(blockIdx.x) * KB + threadIdx.x; t < n; t += long(gridDim.x) * KB) {
        const cpx<double> v = z[t];
        re[t] = v.re;
        im[t] = v.im;
    }
}
__global__ __launch_bounds__(KB) void cast_real_kernel(const double* __restrict__ a, float* __restrict__ b, long n)
{
    for (long t = long(blockIdx.x) * KB + threadIdx.x; t < n; t += long(gridDim.x) * KB) b[t] = float(a[t]);
}

template <typename R, int K>
hipError_t dots_k(const cpx<R>* V, long ldv, int k, const cpx<R>* w, long n, double* part, int nblk,
                  hipStream_t s)
{
    hipLaunchKernelGGL((dots_kernel<R, K>), dim3(nblk), dim3(KB), 0, s, V, ldv, k, w, n, part);
    return hipGetLastError();
}
template <typename R, int K>
hipError_t update_k(const cpx<R>* V, long ldv, int k, const double* d, const CoefPack& a, cpx<R>* w, long n,
                    double* part, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL((update_kernel<R, K>), dim3(nblk), dim3(KB), 0, s, V, ldv, k, d, a, w, n, part);
    return hipGetLastError();
}
template <typename R, int K>
hipError_t combine_k(const cpx<R>* V, long ldv, int k, const CoefPack& a, cpx<R>* x, long n, int nblk,
                     hipStream_t s)
{
    hipLaunchKernelGGL((combine_kernel<R, K>), dim3(nblk), dim3(KB), 0, s, V, ldv, k, a, x, n);
    return hipGetLastError();
}

}  // namespace

int krylov_blocks(long n)
{
    long b = (n + KB * 4 - 1) / (KB * 4);
    return int(std::max(1L, std::min(b, long(kMaxPartials))));
}

// Dispatch on k to a kernel whose register accumulators cover k (bucketed).
template <typename R>
hipError_t launch_dots(const cpx<R>* V, long ldv, int k, const cpx<R>* w, long n, double* part, int nblk,
                       hipStream_t s)
{
    if (k <= 4) return dots_k<R, 4>(V, ldv, k, w, n, part, nblk, s);
    if (k <= 8) return dots_k<R, 8>(V, ldv, k, w, n, part, nblk, s);
    if (k <= 16) return dots_k<R, 16>(V, ldv, k, w, n, part, nblk, s);
    if (k <= 24) return dots_k<R, 24>(V, ldv, k, w, n, part, nblk, s);
    if (k <= kMaxBasis) return dots_k<R, kMaxBasis>(V, ldv, k, w, n, part, nblk, s);
    return hipErrorInvalidValue;
}
template <typename R>
hipError_t launch_update(const cpx<R>* V, long ldv, int k, const double* d, const CoefPack& scl, cpx<R>* w,
                         long n, double* part, int nblk, hipStream_t s)
{
    if (k <= 4) return update_k<R, 4>(V, ldv, k, d, scl, w, n, part, nblk, s);
    if (k <= 8) return update_k<R, 8>(V, ldv, k, d, scl, w, n, part, nblk, s);
    if (k <= 16) return update_k<R, 16>(V, ldv, k, d, scl, w, n, part, nblk, s);
    if (k <= 24) return update_k<R, 24>(V, ldv, k, d, scl, w, n, part, nblk, s);
    if (k <= kMaxBasis) return update_k<R, kMaxBasis>(V, ldv, k, d, scl, w, n, part, nblk, s);
    return hipErrorInvalidValue;
}
template <typename R>
hipError_t launch_combine(const cpx<R>* V, long ldv, int k, const CoefPack& a, cpx<R>* x, long n, int nblk,
                          hipStream_t s)
{
    if (k <= 0) return hipSuccess;
    if (k <= 4) return combine_k<R, 4>(V, ldv, k, a, x, n, nblk, s);
    if (k <= 8) return combine_k<R, 8>(V, ldv, k, a, x, n, nblk, s);
    if (k <= 16) return combine_k<R, 16>(V, ldv, k, a, x, n, nblk, s);
    if (k <= 24) return combine_k<R, 24>(V, ldv, k, a, x, n, nblk, s);
    if (k <= kMaxBasis) return combine_k<R, kMaxBasis>(V, ldv, k, a, x, n, nblk, s);
    return hipErrorInvalidValue;
}

hipError_t launch_interleave(const double* re, const double* im, cpx<double>* z, long n, hipStream_t s)
{
    hipLaunchKernelGGL(interleave_kernel, dim3(krylov_blocks(n)), dim3(KB), 0, s, re, im, z, n);
    return hipGetLastError();
}
hipError_t launch_deinterleave(const cpx<double>* z, double* re, double* im, long n, hipStream_t s)
{
    hipLaunchKernelGGL(deinterleave_kernel, dim3(krylov_blocks(n)), dim3(KB), 0, s, z, re, im, n);
    return hipGetLastError();
}
hipError_t launch_cast_real(const double* a, float* b, long n, hipStream_t s)
{
    hipLaunchKernelGGL(cast_real_kernel, dim3(krylov_blocks(n)), dim3(KB), 0, s, a, b, n);
    return hipGetLastError();
}

hipError_t launch_reduce(const double* part, int nblk, int stride, int nq, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(KB), 0, s, part, nblk, stride, nq, out);
    return hipGetLastError();
}

template <typename RX, typename RY>
hipError_t launch_axpby(double ar, double ai, const cpx<RX>* x, double br, double bi, cpx<RY>* y, long n,
                        bool use_y, hipStream_t s)
{
    hipLaunchKernelGGL((axpby_kernel<RX, RY>), dim3(krylov_blocks(n)), dim3(KB), 0, s, ar, ai, x, br, bi, y, n,
                       int(use_y));
    return hipGetLastError();
}

template <typename R>
hipError_t launch_norm2(const cpx<R>* x, long n, double* part, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL((norm2_kernel<R>), dim3(nblk), dim3(KB), 0, s, x, n, part);
    return hipGetLastError();
}

#define HELM3D_KRYLOV_INST(R)                                                                               \
    template hipError_t launch_dots<R>(const cpx<R>*, long, int, const cpx<R>*, long, double*, int,        \
                                       hipStream_t);                                                        \
    template hipError_t launch_update<R>(const cpx<R>*, long, int, const double*, const CoefPack&, cpx<R>*,  \
                                         long, double*, int, hipStream_t);                                  \
    template hipError_t launch_combine<R>(const cpx<R>*, long, int, const CoefPack&, cpx<R>*, long, int,    \
                                          hipStream_t);                                                     \
    template hipError_t launch_norm2<R>(const cpx<R>*, long, double*, int, hipStream_t);
HELM3D_KRYLOV_INST(double)
HELM3D_KRYLOV_INST(float)
template hipError_t launch_axpby<double, double>(double, double, const cpx<double>*, double, double,
                                                 cpx<double>*, long, bool, hipStream_t);
template hipError_t launch_axpby<float, float>(double, double, const cpx<float>*, double, double, cpx<float>*,
                                               long, bool, hipStream_t);
template hipError_t launch_axpby<double, float>(double, double, const cpx<double>*, double, double,
                                                cpx<float>*, long, bool, hipStream_t);
template hipError_t launch_axpby<float, double>(double, double, const cpx<float>*, double, double,
                                                cpx<double>*, long, bool, hipStream_t);

}  // namespace helm3d
