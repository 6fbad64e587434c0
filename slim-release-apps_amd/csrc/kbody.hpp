// Device bodies of the streaming Krylov / BiCGStab vector kernels, shared by
// the single-vector kernels (krylov.hip, bicgstab.hip) and their work-list
// forms (batch.hip).  A body processes block `b` of a launch of `G` blocks
// over one vector set of length n.
//
// Reductions are deterministic: block b reduces a fixed index range in a
// fixed order into part[q*G + b], and a second kernel sums the partials of a
// column in block order.  Partial sums are FP64 for both storage precisions.
#pragma once

#include "helm3d_common.hpp"
#include "bicgstab.hpp"

namespace helm3d {
namespace kb {

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

// Contiguous range [lo, hi) of block b (even length).
__device__ inline void range_even(long n, int b, int G, long& lo, long& hi)
{
    const long per = ((n + G - 1) / G + 1) & ~1L;
    lo = long(b) * per;
    hi = min(n, lo + per);
}
__device__ inline void range_plain(long n, int b, int G, long& lo, long& hi)
{
    const long per = (n + G - 1) / G;
    lo = long(b) * per;
    hi = min(n, lo + per);
}

// Basis load: with NT the load carries the non-temporal hint, so a basis far
// larger than the caches streams through without evicting the vector being
// updated (measured: the K = 20 update pass at 240x240x140 goes from 4.9 to
// 5.3-5.4 TB/s).
template <bool NT, typename R>
__device__ inline cpx<R> ld_basis(const cpx<R>* p)
{
    if constexpr (NT) {
        typedef R r2 __attribute__((ext_vector_type(2)));
        const r2 v = __builtin_nontemporal_load(reinterpret_cast<const r2*>(p));
        return cpx<R>{v.x, v.y};
    } else {
        return *p;
    }
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

// d_c = sum conj(v_c) w, c < K: partial columns 2c (re), 2c+1 (im).
template <typename R, int K, bool NT = false>
__device__ __forceinline__ void dots_body(const cpx<R>* __restrict__ V, long ldv, const cpx<R>* __restrict__ w,
                                          long n, double* __restrict__ part, int b, int G)
{
    __shared__ double red[(KB / 64) * 2 * K];
    double acc[2 * K];
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) acc[q] = 0.0;
    long lo, hi;
    range_even(n, b, G, lo, hi);
    // two points per thread and iteration, t and t + KB: every load
    // instruction reads one contiguous run of KB complex values
    long t = lo + threadIdx.x;
    for (; t + KB < hi; t += 2 * KB) {
        const cpx<R> w0 = w[t], w1 = w[t + KB];
        cpx<R> v0[K], v1[K];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            v0[c] = ld_basis<NT>(&V[c * ldv + t]);
            v1[c] = ld_basis<NT>(&V[c * ldv + t + KB]);
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
        for (int c = 0; c < K; ++c) acc_dot(acc[2 * c], acc[2 * c + 1], ld_basis<NT>(&V[c * ldv + t]), w0);
    }
    block_sum<2 * K>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 2 * K; ++q) part[long(q) * G + b] = acc[q];
    }
}

// w <- w - sum_{c<K} a_c v_c; partial ||w||^2 in column 0.
template <typename R, int K, bool NT = false>
__device__ __forceinline__ void update_body(const cpx<R>* __restrict__ V, long ldv, const double (&ar)[K],
                                            const double (&ai)[K], cpx<R>* __restrict__ w, long n,
                                            double* __restrict__ part, int b, int G)
{
    __shared__ double red[KB / 64];
    double s2[1] = {0.0};
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
    range_even(n, b, G, lo, hi);
    long t = lo + threadIdx.x;
    for (; t + KB < hi; t += 2 * KB) {
        const cpx<R> w0 = w[t], w1 = w[t + KB];
        cpx<R> v0[K], v1[K];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            v0[c] = ld_basis<NT>(&V[c * ldv + t]);
            v1[c] = ld_basis<NT>(&V[c * ldv + t + KB]);
        }
        w[t] = upd(w0, v0);
        w[t + KB] = upd(w1, v1);
    }
    if (t < hi) {
        cpx<R> v0[K];
#pragma unroll
        for (int c = 0; c < K; ++c) v0[c] = ld_basis<NT>(&V[c * ldv + t]);
        w[t] = upd(w[t], v0);
    }
    block_sum<1>(s2, red);
    if (threadIdx.x == 0) part[b] = s2[0];
}

// x <- x + sum_{c<kk} a_c v_c (the solution update of FGMRES.m:115-119);
// columns kk..K-1 are neither read nor used (kk <= K, uniform).
template <typename R, int K, bool NT = false>
__device__ __forceinline__ void combine_body(const cpx<R>* __restrict__ V, long ldv, const double (&pr)[K],
                                             const double (&pi)[K], cpx<R>* __restrict__ x, long n, int b, int G,
                                             int kk = K)
{
    long lo, hi;
    range_even(n, b, G, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        cpx<R> v[K];
#pragma unroll
        for (int c = 0; c < K; ++c) v[c] = c < kk ? ld_basis<NT>(&V[c * ldv + t]) : cpx<R>{R(0), R(0)};
        const cpx<R> xt = x[t];
        double re = xt.re, im = xt.im;
#pragma unroll
        for (int c = 0; c < K; ++c) {
            if (c < kk) {
                re = fma(pr[c], double(v[c].re), re);
                re = fma(-pi[c], double(v[c].im), re);
                im = fma(pr[c], double(v[c].im), im);
                im = fma(pi[c], double(v[c].re), im);
            }
        }
        x[t] = cpx<R>{R(re), R(im)};
    }
}

// partial ||x||^2 in column 0
template <typename R>
__device__ __forceinline__ void norm2_body(const cpx<R>* __restrict__ x, long n, double* __restrict__ part, int b,
                                           int G)
{
    __shared__ double red[KB / 64];
    double s2[1] = {0.0};
    long lo, hi;
    range_even(n, b, G, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        const cpx<R> v = x[t];
        s2[0] += double(v.re) * double(v.re) + double(v.im) * double(v.im);
    }
    block_sum<1>(s2, red);
    if (threadIdx.x == 0) part[b] = s2[0];
}

// sum of column q of nblk partials, in block order (one block per column)
__device__ __forceinline__ double reduce_body(const double* __restrict__ part, int nblk, int q)
{
    __shared__ double red[KB];
    const double* __restrict__ p = part + long(q) * nblk;
    double s = 0.0;
    int b = threadIdx.x;
    // eight loads in flight per trip, added in the same order as one at a
    // time (long columns: the update pass leaves up to 16384 partials)
    for (; b + 7 * KB < nblk; b += 8 * KB) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[b + u * KB];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; b < nblk; b += KB) s += p[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = KB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    return red[0];
}

// ---- BiCGStab (BICGSTAB.m:37-74) scalar arithmetic on the device slots ------

struct C {
    double re, im;
};
__device__ inline C cdiv(C a, C b)
{
    const double d = b.re * b.re + b.im * b.im;
    if (d == 0.0) return C{0.0, 0.0};
    return C{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ inline C cmulc(C a, C b) { return C{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline C ld(const double* s, int off) { return C{s[off], s[off + 1]}; }

// Partial sums of conj(a_q) . b_q for NP pairs; optionally ||a_last||^2.
template <typename R, int NP, bool NORM>
__device__ __forceinline__ void dots_pairs_body(const PairArgs<R>& P, long n, double* __restrict__ part, int b,
                                                int G)
{
    constexpr int NV = 2 * NP + (NORM ? 1 : 0);
    __shared__ double red[(KB / 64) * NV];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    long lo, hi;
    range_plain(n, b, G, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const cpx<R> a = P.a[q][t], bb = P.b[q][t];
            acc[2 * q] += double(a.re) * double(bb.re) + double(a.im) * double(bb.im);
            acc[2 * q + 1] += double(a.re) * double(bb.im) - double(a.im) * double(bb.re);
            if (NORM && q == NP - 1) acc[NV - 1] += double(a.re) * double(a.re) + double(a.im) * double(a.im);
        }
    }
    block_sum<NV>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) part[long(q) * G + b] = acc[q];
    }
}

// w = (t, s) / (t, t); z += a p + w s; r = s - w t; partials of (rhat, r) and
// ||r||^2.  Block 0 records w.
template <typename R>
__device__ __forceinline__ void bicg_xr_body(const cpx<R>* __restrict__ p, const cpx<R>* __restrict__ s,
                                             const cpx<R>* __restrict__ tt, const cpx<R>* __restrict__ rhat,
                                             cpx<R>* __restrict__ z, cpx<R>* __restrict__ r, long n,
                                             double* __restrict__ sc, double* __restrict__ part, int b, int G)
{
    __shared__ double red[(KB / 64) * 3];
    const C a = ld(sc, BS_ALPHA);
    const double tn = sc[BS_TT];
    const C w = tn > 0.0 ? C{sc[BS_TS] / tn, sc[BS_TS + 1] / tn} : C{0.0, 0.0};
    if (b == 0 && threadIdx.x == 0) {
        sc[BS_OMEGA] = w.re;
        sc[BS_OMEGA + 1] = w.im;
    }
    double acc[3] = {0.0, 0.0, 0.0};
    long lo, hi;
    range_plain(n, b, G, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        const cpx<R> pv = p[t], sv = s[t], tv = tt[t], hv = rhat[t], zv = z[t];
        double zr = double(zv.re) + a.re * double(pv.re) - a.im * double(pv.im) + w.re * double(sv.re) -
                    w.im * double(sv.im);
        double zi = double(zv.im) + a.re * double(pv.im) + a.im * double(pv.re) + w.re * double(sv.im) +
                    w.im * double(sv.re);
        z[t] = cpx<R>{R(zr), R(zi)};
        const double rr = double(sv.re) - (w.re * double(tv.re) - w.im * double(tv.im));
        const double ri = double(sv.im) - (w.re * double(tv.im) + w.im * double(tv.re));
        const cpx<R> rn{R(rr), R(ri)};
        r[t] = rn;
        acc[0] += double(hv.re) * double(rn.re) + double(hv.im) * double(rn.im);
        acc[1] += double(hv.re) * double(rn.im) - double(hv.im) * double(rn.re);
        acc[2] += double(rn.re) * double(rn.re) + double(rn.im) * double(rn.im);
    }
    block_sum<3>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) part[long(q) * G + b] = acc[q];
    }
}

}  // namespace kb
}  // namespace helm3d
