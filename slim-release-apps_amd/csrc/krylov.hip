// Fused Krylov vector kernels for the restarted GMRES of FGMRES.m on gfx950.
//
// The basis is kept unnormalised: slot c holds a raw vector v_c and the host
// holds its scale s_c = 1/||v_c||, so no pass is spent normalising
// (FGMRES.m:104-106 becomes a scalar).  Per Arnoldi step k the passes are
//   dots    d_c = v_c^H w, c < k            (reads w and k slots)
//   update  w <- w - sum_c a_c v_c, ||w||^2  (reads w and k slots, writes w)
// i.e. the classical Gram-Schmidt of FGMRES.m:100-103 in two HBM passes.
// Each kernel is instantiated for the exact basis count k (1..31) so all k+1
// loads of a point are issued back to back, and every thread streams two
// points per iteration.
//
// Reductions are deterministic: every block reduces a fixed index range in a
// fixed order into a partial slot (column-major: part[q*nblk + block]), and a
// second kernel with one block per column sums the partials in block order.
// Partial sums are FP64 for both storage precisions.
#include <type_traits>

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

// Block-wide sums of NV values per thread; results valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* red /* [KB/64][NV] */)
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

// Contiguous range [lo, hi) of this block (even length, for paired points).
__device__ inline void block_range(long n, long& lo, long& hi)
{
    const long per = ((n + gridDim.x - 1) / gridDim.x + 1) & ~1L;
    lo = long(blockIdx.x) * per;
    hi = min(n, lo + per);
}

template <typename R>
__device__ inline void acc_dot(double& re, double& im, cpx<R> v, cpx<R> w)
{
    // += conj(v) * w
    re = fma(double(v.re), double(w.re), re);
    re = fma(double(v.im), double(w.im), re);
    im = fma(double(v.re), double(w.im), im);
    im = fma(-double(v.im), double(w.re), im);
}

// d_c = sum conj(v_c) w, c < K.
template <typename R, int K>
__global__ __launch_bounds__(KB) void dots_kernel(const cpx<R>* __restrict__ V, long ldv,
                                                  const cpx<R>* __restrict__ w, long n,
                                                  double* __restrict__ part)
{
    __shared__ double red[(KB / 64) * 2 * K];
    double acc[2 * K];
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) acc[q] = 0.0;
    long lo, hi;
    block_range(n, lo, hi);
    long t = lo + 2 * threadIdx.x;
    for (; t + 1 < hi; t += 2 * KB) {
        const cpx<R> w0 = w[t], w1 = w[t + 1];
        cpx<R> v0[K], v1[K];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            v0[c] = V[c * ldv + t];
            v1[c] = V[c * ldv + t + 1];
        }
#pragma unroll
        for (int c = 0; c < K; ++c) {
            acc_dot(acc[2 * c], acc[2 * c + 1], v0[c], w0);
            acc_dot(acc[2 * c], acc[2 * c + 1], v1[c], w1);
        }
    }
    if (t < hi) {
        const cpx<R> w0 = w[t];
#pragma unroll
        for (int c = 0; c < K; ++c) acc_dot(acc[2 * c], acc[2 * c + 1], V[c * ldv + t], w0);
    }
    block_sum<2 * K>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 2 * K; ++q) part[long(q) * gridDim.x + blockIdx.x] = acc[q];
    }
}

// w <- w - sum_{c<K} a_c v_c with a_c = scl_c * d_c (d: reduced dots in device
// memory, so the host need not see them before this pass); partial ||w||^2.
template <typename R, int K>
__global__ __launch_bounds__(KB) void update_kernel(const cpx<R>* __restrict__ V, long ldv,
                                                    const double* __restrict__ d, CoefPack scl,
                                                    cpx<R>* __restrict__ w, long n, double* __restrict__ part)
{
    __shared__ double red[KB / 64];
    double s2[1] = {0.0};
    double ar[K], ai[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        ar[c] = scl.v[c] * d[2 * c];
        ai[c] = scl.v[c] * d[2 * c + 1];
    }
    auto upd = [&](cpx<R> wt, const cpx<R>(&v)[K]) {
        double re = wt.re, im = wt.im;
#pragma unroll
        for (int c = 0; c < K; ++c) {
            re = fma(-ar[c], double(v[c].re), re);
            re = fma(ai[c], double(v[c].im), re);
            im = fma(-ar[c], double(v[c].im), im);
            im = fma(-ai[c], double(v[c].re), im);
        }
        const cpx<R> o{R(re), R(im)};
        s2[0] += double(o.re) * double(o.re) + double(o.im) * double(o.im);
        return o;
    };
    long lo, hi;
    block_range(n, lo, hi);
    long t = lo + 2 * threadIdx.x;
    for (; t + 1 < hi; t += 2 * KB) {
        const cpx<R> w0 = w[t], w1 = w[t + 1];
        cpx<R> v0[K], v1[K];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            v0[c] = V[c * ldv + t];
            v1[c] = V[c * ldv + t + 1];
        }
        w[t] = upd(w0, v0);
        w[t + 1] = upd(w1, v1);
    }
    if (t < hi) {
        cpx<R> v0[K];
#pragma unroll
        for (int c = 0; c < K; ++c) v0[c] = V[c * ldv + t];
        w[t] = upd(w[t], v0);
    }
    block_sum<1>(s2, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s2[0];
}

// x <- x + sum_{c<K} a_c v_c (the solution update of FGMRES.m:115-119).
template <typename R, int K>
__global__ __launch_bounds__(KB) void combine_kernel(const cpx<R>* __restrict__ V, long ldv, CoefPack a,
                                                     cpx<R>* __restrict__ x, long n)
{
    long lo, hi;
    block_range(n, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        cpx<R> v[K];
#pragma unroll
        for (int c = 0; c < K; ++c) v[c] = V[c * ldv + t];
        const cpx<R> xt = x[t];
        double re = xt.re, im = xt.im;
#pragma unroll
        for (int c = 0; c < K; ++c) {
            const double pr = a.v[2 * c], pi = a.v[2 * c + 1];
            re = fma(pr, double(v[c].re), re);
            re = fma(-pi, double(v[c].im), re);
            im = fma(pr, double(v[c].im), im);
            im = fma(pi, double(v[c].re), im);
        }
        x[t] = cpx<R>{R(re), R(im)};
    }
}

// out[q] = sum_b part[q*nblk + b] in block order; one block per column q.
__global__ __launch_bounds__(KB) void reduce_kernel(const double* __restrict__ part, int nblk,
                                                    double* __restrict__ out)
{
    __shared__ double red[KB];
    const int q = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += KB) s += part[long(q) * nblk + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = KB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
}

// y = alpha x + beta y (complex alpha, beta); use_y == 0 ignores y's contents.
template <typename RX, typename RY>
__global__ __launch_bounds__(KB) void axpby_kernel(double ar, double ai, const cpx<RX>* __restrict__ x,
                                                   double br, double bi, cpx<RY>* __restrict__ y,
                                                   long n, int use_y)
{
    for (long t = long(blockIdx.x) * KB + threadIdx.x; t < n; t += long(gridDim.x) * KB) {
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

// Runtime k -> fn(integral_constant<int, k>) for k in [K, kMaxBasis).
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

int krylov_blocks(long n)
{
    long b = (n + KB * 8 - 1) / (KB * 8);
    return int(std::max(1L, std::min(b, long(kMaxPartials))));
}

template <typename R>
hipError_t launch_dots(const cpx<R>* V, long ldv, int k, const cpx<R>* w, long n, double* part, int nblk,
                       hipStream_t s)
{
    if (k < 1) return hipErrorInvalidValue;
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        hipLaunchKernelGGL((dots_kernel<R, K>), dim3(nblk), dim3(KB), 0, s, V, ldv, w, n, part);
        return hipGetLastError();
    });
}

template <typename R>
hipError_t launch_update(const cpx<R>* V, long ldv, int k, const double* d, const CoefPack& scl, cpx<R>* w,
                         long n, double* part, int nblk, hipStream_t s)
{
    if (k < 1) return hipErrorInvalidValue;
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        hipLaunchKernelGGL((update_kernel<R, K>), dim3(nblk), dim3(KB), 0, s, V, ldv, d, scl, w, n, part);
        return hipGetLastError();
    });
}

template <typename R>
hipError_t launch_combine(const cpx<R>* V, long ldv, int k, const CoefPack& a, cpx<R>* x, long n, int nblk,
                          hipStream_t s)
{
    if (k <= 0) return hipSuccess;
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        hipLaunchKernelGGL((combine_kernel<R, K>), dim3(nblk), dim3(KB), 0, s, V, ldv, a, x, n);
        return hipGetLastError();
    });
}

hipError_t launch_reduce(const double* part, int nblk, int nq, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_kernel, dim3(nq), dim3(KB), 0, s, part, nblk, out);
    return hipGetLastError();
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
