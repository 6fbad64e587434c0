// FWI set-up and gradient support kernels.
//
//  * param2wavenum on the extended grid: the model on the physical grid is
//    extended into the PML by edge replication (opExtension.m:94-96 per axis,
//    composed in discrete_helmholtz.m:214-217 and applied at :240) and mapped
//    to wn = c v^-2 | c v | 1e-6 c v with c = (2 pi f)^2 and the derivatives
//    dwn, ddwn (param2wavenum.m:20-49).  One pass: each extended point reads
//    its clamped physical point (L2-resident along the PML slabs).
//  * the adjoint of the extension (opExtension.m:97-101 per axis): every
//    physical point sums the extended points that replicate it.  Interior
//    points have one preimage; face, edge and corner points sum their PML
//    line, sheet or block.
//  * a row-compressed operator with real weights for source injection and
//    receiver sampling (the sinc/Kaiser opInterp operators of DataGrid.m:22-44
//    reduce to a unit impulse on grid points, and are sparse in general).
//    Each output row is a gather, so the kernel needs no atomics; the host
//    builds the transposed rows for the injection direction.
#include <algorithm>

#include "fwi.hpp"

namespace helm3d {

namespace {

constexpr int KB = 256;

__device__ inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

__global__ __launch_bounds__(KB) void param2wavenum_kernel(ExtGrid G, const double* __restrict__ v, int unit,
                                                           double c, double* __restrict__ wn,
                                                           double* __restrict__ dwn, double* __restrict__ ddwn)
{
    const long N = long(G.nt[0]) * G.nt[1] * G.nt[2];
    const long plane = long(G.nt[0]) * G.nt[1];
    for (long p = long(blockIdx.x) * KB + threadIdx.x; p < N; p += long(gridDim.x) * KB) {
        const int k = int(p / plane);
        const long r = p - long(k) * plane;
        const int j = int(r / G.nt[0]);
        const int i = int(r - long(j) * G.nt[0]);
        const int pi = clampi(i - G.lo[0], 0, G.np[0] - 1);
        const int pj = clampi(j - G.lo[1], 0, G.np[1] - 1);
        const int pk = clampi(k - G.lo[2], 0, G.np[2] - 1);
        const double m = v[pi + long(pj) * G.np[0] + long(pk) * G.np[0] * G.np[1]];
        double a, d, dd;
        if (unit == 0) {  // m/s
            const double iv = 1.0 / m;
            const double iv2 = iv * iv;
            a = c * iv2;
            d = -2.0 * c * iv2 * iv;
            dd = 6.0 * c * iv2 * iv2;
        } else {  // s2/m2 (c' = c) or s2/km2 (c' = 1e-6 c): linear in m
            a = c * m;
            d = c;
            dd = 0.0;
        }
        wn[p] = a;
        if (dwn) dwn[p] = d;
        if (ddwn) ddwn[p] = dd;
    }
}

// Preimage of physical index q along one axis: [q+lo] for interior points,
// [0, lo] for q = 0 and [q+lo, nt-1] for q = np-1 (both for np = 1).
__device__ inline void preimage(int q, int np, int lo, int nt, int& a, int& b)
{
    a = q == 0 ? 0 : q + lo;
    b = q == np - 1 ? nt - 1 : q + lo;
}

__global__ __launch_bounds__(KB) void extension_adjoint_kernel(ExtGrid G, const double* __restrict__ gpml,
                                                               double* __restrict__ g, int accumulate)
{
    const long Np = long(G.np[0]) * G.np[1] * G.np[2];
    const long pplane = long(G.np[0]) * G.np[1];
    const long tplane = long(G.nt[0]) * G.nt[1];
    for (long p = long(blockIdx.x) * KB + threadIdx.x; p < Np; p += long(gridDim.x) * KB) {
        const int k = int(p / pplane);
        const long r = p - long(k) * pplane;
        const int j = int(r / G.np[0]);
        const int i = int(r - long(j) * G.np[0]);
        int ia, ib, ja, jb, ka, kb;
        preimage(i, G.np[0], G.lo[0], G.nt[0], ia, ib);
        preimage(j, G.np[1], G.lo[1], G.nt[1], ja, jb);
        preimage(k, G.np[2], G.lo[2], G.nt[2], ka, kb);
        double s = 0.0;
        for (int kk = ka; kk <= kb; ++kk)
            for (int jj = ja; jj <= jb; ++jj) {
                const double* row = gpml + long(kk) * tplane + long(jj) * G.nt[0];
                for (int ii = ia; ii <= ib; ++ii) s += row[ii];
            }
        g[p] = accumulate ? g[p] + s : s;
    }
}

__global__ __launch_bounds__(KB) void csr_apply_kernel(int nrows, const long long* __restrict__ rowptr,
                                                       const long long* __restrict__ col,
                                                       const double* __restrict__ w,
                                                       const long long* __restrict__ oidx, double alpha,
                                                       const cpx<double>* __restrict__ in,
                                                       cpx<double>* __restrict__ out, int accumulate)
{
    const int r = blockIdx.x * KB + threadIdx.x;
    if (r >= nrows) return;
    double re = 0.0, im = 0.0;
    for (long long e = rowptr[r]; e < rowptr[r + 1]; ++e) {
        const cpx<double> v = in[col[e]];
        re += w[e] * v.re;
        im += w[e] * v.im;
    }
    const long long o = oidx ? oidx[r] : r;
    cpx<double> y{alpha * re, alpha * im};
    if (accumulate) {
        y.re += out[o].re;
        y.im += out[o].im;
    }
    out[o] = y;
}

int grid_blocks(long n) { return int(std::min<long>(std::max<long>((n + KB - 1) / KB, 1), 8192)); }

}  // namespace

hipError_t launch_param2wavenum(const ExtGrid& G, const double* v, int unit, double c, double* wn, double* dwn,
                                double* ddwn, hipStream_t s)
{
    const long N = long(G.nt[0]) * G.nt[1] * G.nt[2];
    hipLaunchKernelGGL(param2wavenum_kernel, dim3(grid_blocks(N)), dim3(KB), 0, s, G, v, unit, c, wn, dwn, ddwn);
    return hipGetLastError();
}

hipError_t launch_extension_adjoint(const ExtGrid& G, const double* gpml, double* g, bool accumulate,
                                    hipStream_t s)
{
    const long Np = long(G.np[0]) * G.np[1] * G.np[2];
    hipLaunchKernelGGL(extension_adjoint_kernel, dim3(grid_blocks(Np)), dim3(KB), 0, s, G, gpml, g,
                       int(accumulate));
    return hipGetLastError();
}

hipError_t launch_csr_apply(int nrows, const long long* rowptr, const long long* col, const double* w,
                            const long long* oidx, double alpha, const cpx<double>* in, cpx<double>* out,
                            bool accumulate, hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(csr_apply_kernel, dim3((nrows + KB - 1) / KB), dim3(KB), 0, s, nrows, rowptr, col, w, oidx,
                       alpha, in, out, int(accumulate));
    return hipGetLastError();
}

}  // namespace helm3d
