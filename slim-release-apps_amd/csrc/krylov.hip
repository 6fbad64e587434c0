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
#include "kbody.hpp"
#include "krylov.hpp"

namespace helm3d {

namespace {

using kb::KB;

// d_c = sum conj(v_c) w, c < K.  NT: streaming (non-temporal) basis loads.
template <typename R, int K, bool NT>
__global__ __launch_bounds__(KB) void dots_kernel(const cpx<R>* __restrict__ V, long ldv,
                                                  const cpx<R>* __restrict__ w, long n,
                                                  double* __restrict__ part, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    kb::dots_body<R, K, NT>(V, ldv, w, n, part, blockIdx.x, gridDim.x);
}

// w <- w - sum_{c<K} a_c v_c with a_c = scl_c * d_c (d: reduced dots in device
// memory, so the host need not see them before this pass); partial ||w||^2.
// nrm != nullptr: scl_c = (1 / sqrt(nrm[c]))^2 from device memory, the square
// of the slot scale the host applies to the Hessenberg entries.
// dpart (short partial columns only): every block sums the dots partials
// dpart[q*dnblk + b] itself -- one wave per column, lanes strided over the
// blocks, a fixed order -- instead of reading d, and block 0 writes the sums
// to hrow; no separate reduction launch.
template <typename R, int K, bool NT>
__global__ __launch_bounds__(KB) void update_kernel(const cpx<R>* __restrict__ V, long ldv,
                                                    const double* __restrict__ d, CoefPack scl,
                                                    const double* __restrict__ nrm, cpx<R>* __restrict__ w,
                                                    long n, double* __restrict__ part,
                                                    const double* __restrict__ dpart, int dnblk,
                                                    double* __restrict__ hrow, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double dsum[2 * K];
    if (dpart) {
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        for (int q = wid; q < 2 * K; q += KB / 64) {
            double s = 0.0;
            for (int b = lane; b < dnblk; b += 64) s += dpart[long(q) * dnblk + b];
            s = kb::wave_sum(s);
            if (lane == 0) dsum[q] = s;
        }
        __syncthreads();
        if (blockIdx.x == 0 && threadIdx.x < 2 * K) hrow[threadIdx.x] = dsum[threadIdx.x];
    }
    double ar[K], ai[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const double inv = nrm ? 1.0 / sqrt(nrm[c]) : 0.0;
        const double sc = nrm ? inv * inv : scl.v[c];
        ar[c] = sc * (dpart ? dsum[2 * c] : d[2 * c]);
        ai[c] = sc * (dpart ? dsum[2 * c + 1] : d[2 * c + 1]);
    }
    kb::update_body<R, K, NT>(V, ldv, ar, ai, w, n, part, blockIdx.x, gridDim.x);
}

// x <- x + sum_{c<K} a_c v_c (the solution update of FGMRES.m:115-119).
template <typename R, int K, bool NT>
__global__ __launch_bounds__(KB) void combine_kernel(const cpx<R>* __restrict__ V, long ldv, CoefPack a,
                                                     cpx<R>* __restrict__ x, long n)
{
    double pr[K], pi[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        pr[c] = a.v[2 * c];
        pi[c] = a.v[2 * c + 1];
    }
    kb::combine_body<R, K, NT>(V, ldv, pr, pi, x, n, blockIdx.x, gridDim.x);
}

// x <- x + sum_{c<kept} a_c v_c with the coefficients in device memory (a:
// 2K doubles, written by gmres_lsq_kernel, which also writes *kept <= K);
// the slots of steps past the kept ones are not read, so values a step
// after a breakdown left there (0 * inf) do not reach x.
template <typename R, int K, bool NT>
__global__ __launch_bounds__(KB) void combine_dev_kernel(const cpx<R>* __restrict__ V, long ldv,
                                                         const double* __restrict__ a, const int* __restrict__ kept,
                                                         cpx<R>* __restrict__ x, long n, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int kk = min(*kept, K);
    double pr[K], pi[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        pr[c] = a[2 * c];
        pi[c] = a[2 * c + 1];
    }
    kb::combine_body<R, K, NT>(V, ldv, pr, pi, x, n, blockIdx.x, gridDim.x, kk);
}

// ---- result-free restarted GMRES driven from the device ----------------------
// The host enqueues every cycle a solve may need; the Hessenberg least
// squares of FGMRES.m:109-111 runs in gmres_lsq_kernel (one thread), and the
// stop decisions of the host loop (quasi-residual per step, true residual per
// restart) set st->stop; the kernels of later cycles read st->skip and return.

struct Cx {
    double re, im;
};
__device__ inline Cx cx_add(Cx a, Cx b) { return Cx{a.re + b.re, a.im + b.im}; }
__device__ inline Cx cx_sub(Cx a, Cx b) { return Cx{a.re - b.re, a.im - b.im}; }
__device__ inline Cx cx_mul(Cx a, Cx b) { return Cx{fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)}; }
__device__ inline Cx cx_conj(Cx a) { return Cx{a.re, -a.im}; }
__device__ inline double cx_norm(Cx a) { return fma(a.re, a.re, a.im * a.im); }
__device__ inline Cx cx_scale(Cx a, double s) { return Cx{a.re * s, a.im * s}; }
__device__ inline Cx cx_div(Cx a, Cx b)
{
    const double d = cx_norm(b);
    return Cx{fma(a.re, b.re, a.im * b.im) / d, fma(a.im, b.re, -(a.re * b.im)) / d};
}

// Start of a device-driven solve: ||r0|| from nrm[0]; a zero residual stops
// before the first step (x is already the answer).
__global__ void gmres_start_kernel(const double* __restrict__ nrm, GmresDevState* st)
{
    const double b = sqrt(nrm[0]);
    st->normr0 = b;
    st->beta = b;
    st->hlast = 1.0;
    st->stop = b == 0.0 ? 1 : 0;
    st->skip = st->stop;
}

// Start of a later cycle (restart): stop if an earlier step or restart
// test stopped the solve, else read ||r|| of the new residual from nrm[0]
// and stop on the true residual (FGMRES.m:120-124).
__global__ void gmres_gate_kernel(GmresDevState* st) { st->skip = st->stop; }
__global__ void gmres_restart_kernel(const double* __restrict__ nrm, double tol, GmresDevState* st)
{
    if (st->skip) return;
    const double rn = sqrt(nrm[0]);
    if (rn / st->normr0 < tol) {
        st->stop = 1;
        st->skip = 1;
        return;
    }
    st->beta = rn;
}

// The least squares of one cycle of m steps (FGMRES.m:109-111): rows[j]
// (stride ldr) hold the dots d_0..d_{j-1} of step j and ||w_j||^2; the
// Hessenberg column of step j is sig_q d_q (q < j) over ||w_j||, with the
// basis slot scales sig_q = 1/||v_q||.  The columns are reduced as they come
// by Givens rotations (the triangular factor T and the rotated right-hand
// side g), so the quasi-residual of step k is |g_k| / ||r0|| -- the residual
// of the k-step least squares -- after O(k) work.  Steps after the first one
// at or below tol are dropped (the host loop would not have run them); coef
// receives a_q = y_q sig_q for the kept steps, zeros after.  One thread.
// Cycles of at most MB steps; MB = kMaxBasis keeps T, g and the rotations in
// LDS, smaller MB in registers (every index fixed once the loops unroll).
template <int MB>
__device__ __forceinline__ void gmres_lsq_body(const double* __restrict__ rows, int ldr, int m, double tol,
                                               GmresDevState* st, Cx* T, Cx* cs, Cx* sn, Cx* g, Cx* y,
                                               double* sig)
{
    constexpr int ld = MB + 1;
    // unrolled (register arrays) for short cycles, rolled for the LDS form
    constexpr int kU = MB <= 5 ? MB + 1 : 1;
    const double normr0 = st->normr0;
    sig[0] = 1.0 / st->beta;
    g[0] = Cx{st->beta, 0.0};
    int kept = m;
    double hlast = st->hlast;
#pragma unroll kU
    for (int k = 1; k <= MB; ++k) {
        if (k > m) break;
        const double* h = rows + long(k) * ldr;
        Cx* col = T + (k - 1) * ld;
#pragma unroll kU
        for (int q = 0; q < k; ++q) col[q] = cx_scale(Cx{h[2 * q], h[2 * q + 1]}, sig[q]);
        const double hn = sqrt(h[2 * k]);
        sig[k] = 1.0 / hn;
#pragma unroll kU
        for (int q = 0; q + 1 < k; ++q) {
            const Cx u = col[q], v = col[q + 1];
            col[q] = cx_add(cx_mul(cx_conj(cs[q]), u), cx_mul(cx_conj(sn[q]), v));
            col[q + 1] = cx_sub(cx_mul(cs[q], v), cx_mul(sn[q], u));
        }
        // the rotation that zeroes ||w_k|| below the diagonal
        const Cx a = col[k - 1];
        const double den = sqrt(cx_norm(a) + hn * hn);
        Cx c{1.0, 0.0}, sv{0.0, 0.0};
        if (den != 0.0) {
            c = cx_scale(a, 1.0 / den);
            sv = Cx{hn / den, 0.0};
        }
        cs[k - 1] = c;
        sn[k - 1] = sv;
        col[k - 1] = Cx{den, 0.0};
        const Cx gk = g[k - 1];
        g[k - 1] = cx_mul(cx_conj(c), gk);
        g[k] = cx_scale(cx_mul(sv, gk), -1.0);
        hlast = sqrt(cx_norm(g[k])) / normr0;
        if (hlast <= tol) {
            kept = k;
            break;
        }
    }
#pragma unroll kU
    for (int r = MB - 1; r >= 0; --r) {
        if (r < kept) {
            Cx acc = g[r];
#pragma unroll kU
            for (int c = r + 1; c < MB; ++c)
                if (c < kept) acc = cx_sub(acc, cx_mul(T[r + c * ld], y[c]));
            y[r] = cx_div(acc, T[r + r * ld]);
        }
    }
#pragma unroll kU
    for (int q = 0; q < kMaxBasis; ++q) {
        const Cx a = (q < MB && q < kept) ? cx_scale(y[q < MB ? q : 0], sig[q < MB ? q : 0]) : Cx{0.0, 0.0};
        st->coef[2 * q] = a.re;
        st->coef[2 * q + 1] = a.im;
    }
    st->hlast = hlast;
    st->kept = kept;
    if (hlast <= tol) st->stop = 1;
}

template <int MB>
__global__ __launch_bounds__(64) void gmres_lsq_kernel(const double* __restrict__ rows, int ldr, int m, double tol,
                                                       GmresDevState* st)
{
    if (threadIdx.x != 0 || st->skip) return;
    if constexpr (MB == kMaxBasis) {
        __shared__ Cx T[(MB + 1) * MB];
        __shared__ Cx cs[MB], sn[MB], g[MB + 1], y[MB];
        __shared__ double sig[MB + 1];
        gmres_lsq_body<MB>(rows, ldr, m, tol, st, T, cs, sn, g, y, sig);
    } else {
        Cx T[(MB + 1) * MB], cs[MB], sn[MB], g[MB + 1], y[MB];
        double sig[MB + 1];
        gmres_lsq_body<MB>(rows, ldr, m, tol, st, T, cs, sn, g, y, sig);
    }
}

// out[q] = sum_b part[q*nblk + b] in block order; one block per column q
// (out2, when given, receives the same values).
__global__ __launch_bounds__(KB) void reduce_kernel(const double* __restrict__ part, int nblk,
                                                    double* __restrict__ out, double* __restrict__ out2,
                                                    const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const double s = kb::reduce_body(part, nblk, blockIdx.x);
    if (threadIdx.x == 0) {
        out[blockIdx.x] = s;
        if (out2) out2[blockIdx.x] = s;
    }
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
    kb::norm2_body<R>(x, n, part, blockIdx.x, gridDim.x);
}

// Start of a solve from x = 0 in one pass: w = b, x = 0 and the partial
// ||b||^2 of norm2_kernel (same blocks, same order).
template <typename R>
__global__ __launch_bounds__(KB) void start_kernel(const cpx<R>* __restrict__ b, cpx<R>* __restrict__ w,
                                                   cpx<R>* __restrict__ x, long n, double* __restrict__ part)
{
    __shared__ double red[KB / 64];
    double s2[1] = {0.0};
    long lo, hi;
    kb::range_even(n, blockIdx.x, gridDim.x, lo, hi);
    for (long t = lo + threadIdx.x; t < hi; t += KB) {
        const cpx<R> v = b[t];
        w[t] = v;
        x[t] = cpx<R>{R(0), R(0)};
        s2[0] += double(v.re) * double(v.re) + double(v.im) * double(v.im);
    }
    kb::block_sum<1>(s2, red);
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
    // About 8 points per thread; a vector too short to fill 256 blocks that
    // way (the coarse multigrid levels) gets up to 256 blocks of one or a few
    // points per thread, which shortens the per-thread loops of these
    // latency-bound launches.
    long b = (n + KB * 8 - 1) / (KB * 8);
    if (b < 256) b = std::min((n + KB - 1) / KB, 256L);
    return int(std::max(1L, std::min(b, long(kMaxPartials))));
}

int dots_blocks(long n)
{
    const long kb = krylov_blocks(n);
    if (kb <= kInBlockDots) return int(kb);
    const long b = std::max((n + KB * 32 - 1) / (KB * 32), std::min(kb, 512L));
    return int(std::min(b, long(kMaxPartials)));
}

int update_blocks(long n)
{
    const long kb = krylov_blocks(n);
    if (kb <= kInBlockDots) return int(kb);
    const long b = std::max((n + KB * 2 - 1) / (KB * 2), kb);
    return int(std::min(b, long(kMaxUpdateBlocks)));
}

// Streaming basis loads once the basis slots read by a pass exceed the
// last-level cache (256 MB Infinity Cache on MI355X); smaller bases (coarse
// multigrid levels, small grids) stay cacheable between the dots and update
// passes.
template <typename R>
bool stream_basis(int k, long n)
{
    return double(k) * double(n) * sizeof(cpx<R>) > 256.0 * (1 << 20);
}

template <typename R>
hipError_t launch_dots(const cpx<R>* V, long ldv, int k, const cpx<R>* w, long n, double* part, int nblk,
                       hipStream_t s, const int* skip)
{
    if (k < 1) return hipErrorInvalidValue;
    const bool nt = stream_basis<R>(k, n);
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        if (nt)
            hipLaunchKernelGGL((dots_kernel<R, K, true>), dim3(nblk), dim3(KB), 0, s, V, ldv, w, n, part, skip);
        else
            hipLaunchKernelGGL((dots_kernel<R, K, false>), dim3(nblk), dim3(KB), 0, s, V, ldv, w, n, part, skip);
        return hipGetLastError();
    });
}

template <typename R>
hipError_t launch_update(const cpx<R>* V, long ldv, int k, const double* d, const CoefPack& scl, cpx<R>* w,
                         long n, double* part, int nblk, hipStream_t s, const double* nrm, const double* dpart,
                         double* hrow, const int* skip)
{
    if (k < 1) return hipErrorInvalidValue;
    const bool nt = stream_basis<R>(k, n);
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        if (nt)
            hipLaunchKernelGGL((update_kernel<R, K, true>), dim3(nblk), dim3(KB), 0, s, V, ldv, d, scl, nrm, w, n,
                               part, dpart, nblk, hrow, skip);
        else
            hipLaunchKernelGGL((update_kernel<R, K, false>), dim3(nblk), dim3(KB), 0, s, V, ldv, d, scl, nrm, w, n,
                               part, dpart, nblk, hrow, skip);
        return hipGetLastError();
    });
}

template <typename R>
hipError_t launch_combine_dev(const cpx<R>* V, long ldv, int k, const double* a, const int* kept, cpx<R>* x,
                              long n, int nblk, hipStream_t s, const int* skip)
{
    if (k <= 0) return hipSuccess;
    const bool nt = stream_basis<R>(k, n);
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        if (nt)
            hipLaunchKernelGGL((combine_dev_kernel<R, K, true>), dim3(nblk), dim3(KB), 0, s, V, ldv, a, kept, x, n,
                               skip);
        else
            hipLaunchKernelGGL((combine_dev_kernel<R, K, false>), dim3(nblk), dim3(KB), 0, s, V, ldv, a, kept, x, n,
                               skip);
        return hipGetLastError();
    });
}

hipError_t launch_gmres_start(const double* nrm, GmresDevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(gmres_start_kernel, dim3(1), dim3(1), 0, s, nrm, st);
    return hipGetLastError();
}
hipError_t launch_gmres_gate(GmresDevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(gmres_gate_kernel, dim3(1), dim3(1), 0, s, st);
    return hipGetLastError();
}
hipError_t launch_gmres_restart(const double* nrm, double tol, GmresDevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(gmres_restart_kernel, dim3(1), dim3(1), 0, s, nrm, tol, st);
    return hipGetLastError();
}
hipError_t launch_gmres_lsq(const double* rows, int ldr, int m, double tol, GmresDevState* st, hipStream_t s)
{
    if (m < 1 || m >= kMaxBasis) return hipErrorInvalidValue;
    if (m <= 2)
        hipLaunchKernelGGL(gmres_lsq_kernel<2>, dim3(1), dim3(1), 0, s, rows, ldr, m, tol, st);
    else if (m <= 5)
        hipLaunchKernelGGL(gmres_lsq_kernel<5>, dim3(1), dim3(1), 0, s, rows, ldr, m, tol, st);
    else
        hipLaunchKernelGGL(gmres_lsq_kernel<kMaxBasis>, dim3(1), dim3(1), 0, s, rows, ldr, m, tol, st);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_combine(const cpx<R>* V, long ldv, int k, const CoefPack& a, cpx<R>* x, long n, int nblk,
                          hipStream_t s)
{
    if (k <= 0) return hipSuccess;
    const bool nt = stream_basis<R>(k, n);
    return dispatch_k<1>(k, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        if (nt) hipLaunchKernelGGL((combine_kernel<R, K, true>), dim3(nblk), dim3(KB), 0, s, V, ldv, a, x, n);
        else hipLaunchKernelGGL((combine_kernel<R, K, false>), dim3(nblk), dim3(KB), 0, s, V, ldv, a, x, n);
        return hipGetLastError();
    });
}

hipError_t launch_reduce(const double* part, int nblk, int nq, double* out, hipStream_t s, double* out2,
                         const int* skip)
{
    hipLaunchKernelGGL(reduce_kernel, dim3(nq), dim3(KB), 0, s, part, nblk, out, out2, skip);
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

template <typename R>
hipError_t launch_start(const cpx<R>* b, cpx<R>* w, cpx<R>* x, long n, double* part, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL((start_kernel<R>), dim3(nblk), dim3(KB), 0, s, b, w, x, n, part);
    return hipGetLastError();
}

#define HELM3D_KRYLOV_INST(R)                                                                               \
    template hipError_t launch_start<R>(const cpx<R>*, cpx<R>*, cpx<R>*, long, double*, int, hipStream_t);  \
    template hipError_t launch_dots<R>(const cpx<R>*, long, int, const cpx<R>*, long, double*, int,        \
                                       hipStream_t, const int*);                                            \
    template hipError_t launch_update<R>(const cpx<R>*, long, int, const double*, const CoefPack&, cpx<R>*,  \
                                         long, double*, int, hipStream_t, const double*, const double*,     \
                                         double*, const int*);                                              \
    template hipError_t launch_combine_dev<R>(const cpx<R>*, long, int, const double*, const int*, cpx<R>*, \
                                              long, int, hipStream_t, const int*);                          \
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
